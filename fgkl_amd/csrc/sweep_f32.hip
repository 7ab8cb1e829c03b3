// FP32 instantiations of the sweep kernel: plain and PD, each in the natural build and in the
// occupancy-raised one the autotuner can pick (sweep.hpp).
#include "sweep.hpp"

namespace fgkl {

#define FGKL_F32_ROW(PD, B)                                                                 \
  {sweep_kernel<float, 1, PD, B, !PD>, sweep_kernel<float, 2, PD, B, !PD>,                  \
   sweep_kernel<float, 3, PD, B, !PD>, sweep_kernel<float, 4, PD, B, !PD>,                  \
   sweep_kernel<float, 5, PD, B, !PD>, sweep_kernel<float, 6, PD, B, !PD>,                  \
   sweep_kernel<float, 7, PD, B, !PD>, sweep_kernel<float, 8, PD, B, !PD>}

const KernelF kSweepF32[2][2][kMaxRowsPerLane] = {
    {FGKL_F32_ROW(false, false), FGKL_F32_ROW(false, true)},
    {FGKL_F32_ROW(true, false), FGKL_F32_ROW(true, true)}};

#undef FGKL_F32_ROW

const KernelF kLongF32[2] = {long_kernel<float, false>, long_kernel<float, true>};

}  // namespace fgkl
