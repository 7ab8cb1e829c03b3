// FP64 instantiations of the sweep kernel: plain and PD without suffix sharing (the FP64 mode
// without a suffix plan, the rescue of FP32 pairs, the per-region baseline) and plain with it
// (FP64 main passes of batches with a suffix plan) (sweep.hpp).
#include "sweep.hpp"

namespace fgkl {

#define FGKL_F64_ROW(PD, S)                                                                 \
  {sweep_kernel<double, 1, PD, false, S>, sweep_kernel<double, 2, PD, false, S>,            \
   sweep_kernel<double, 3, PD, false, S>, sweep_kernel<double, 4, PD, false, S>,            \
   sweep_kernel<double, 5, PD, false, S>, sweep_kernel<double, 6, PD, false, S>,            \
   sweep_kernel<double, 7, PD, false, S>, sweep_kernel<double, 8, PD, false, S>}

const KernelD kSweepF64[2][kMaxRowsPerLane] = {FGKL_F64_ROW(false, false),
                                               FGKL_F64_ROW(true, false)};
const KernelD kSweepF64Suffix[kMaxRowsPerLane] = FGKL_F64_ROW(false, true);

#undef FGKL_F64_ROW

const KernelD kLongF64[2] = {long_kernel<double, false>, long_kernel<double, true>};

}  // namespace fgkl
