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

// The uniform-gap builds of the plain sweep (see Rows): without and with suffix sharing.
#define FGKL_F64_UG_ROW(S)                                                                  \
  {sweep_kernel<double, 1, false, false, S, true>, sweep_kernel<double, 2, false, false, S, true>, \
   sweep_kernel<double, 3, false, false, S, true>, sweep_kernel<double, 4, false, false, S, true>, \
   sweep_kernel<double, 5, false, false, S, true>, sweep_kernel<double, 6, false, false, S, true>, \
   sweep_kernel<double, 7, false, false, S, true>, sweep_kernel<double, 8, false, false, S, true>}

const KernelD kSweepF64Uniform[2][kMaxRowsPerLane] = {FGKL_F64_UG_ROW(false), FGKL_F64_UG_ROW(true)};

// The uniform-gap build of the PD sweep.
const KernelD kSweepF64PdUniform[kMaxRowsPerLane] = {
    sweep_kernel<double, 1, true, false, false, true>, sweep_kernel<double, 2, true, false, false, true>,
    sweep_kernel<double, 3, true, false, false, true>, sweep_kernel<double, 4, true, false, false, true>,
    sweep_kernel<double, 5, true, false, false, true>, sweep_kernel<double, 6, true, false, false, true>,
    sweep_kernel<double, 7, true, false, false, true>, sweep_kernel<double, 8, true, false, false, true>};

#undef FGKL_F64_UG_ROW

#undef FGKL_F64_ROW

const KernelD kLongF64[2] = {long_kernel<double, false>, long_kernel<double, true>};

}  // namespace fgkl
