// FP64 instantiations of the sweep kernel, plain and PD: the FP64 mode, the rescue of FP32 pairs
// and the per-region baseline (sweep.hpp).
#include "sweep.hpp"

namespace fgkl {

#define FGKL_F64_ROW(PD)                                                                    \
  {sweep_kernel<double, 1, PD, false>, sweep_kernel<double, 2, PD, false>,                  \
   sweep_kernel<double, 3, PD, false>, sweep_kernel<double, 4, PD, false>,                  \
   sweep_kernel<double, 5, PD, false>, sweep_kernel<double, 6, PD, false>,                  \
   sweep_kernel<double, 7, PD, false>, sweep_kernel<double, 8, PD, false>}

const KernelD kSweepF64[2][kMaxRowsPerLane] = {FGKL_F64_ROW(false), FGKL_F64_ROW(true)};

#undef FGKL_F64_ROW

const KernelD kLongF64[2] = {long_kernel<double, false>, long_kernel<double, true>};

}  // namespace fgkl
