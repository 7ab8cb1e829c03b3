// FP32 instantiations of the plain sweep kernel for items whose reads have uniform gap
// qualities (one value of each gap term per lane, see Rows), in the natural build and in the
// occupancy-raised one the autotuner can pick (sweep.hpp).
#include "sweep.hpp"

namespace fgkl {

#define FGKL_F32_UG_ROW(B)                                                                     \
  {sweep_kernel<float, 1, false, B, true, true>, sweep_kernel<float, 2, false, B, true, true>, \
   sweep_kernel<float, 3, false, B, true, true>, sweep_kernel<float, 4, false, B, true, true>, \
   sweep_kernel<float, 5, false, B, true, true>, sweep_kernel<float, 6, false, B, true, true>, \
   sweep_kernel<float, 7, false, B, true, true>, sweep_kernel<float, 8, false, B, true, true>}

const KernelF kSweepF32Uniform[2][kMaxRowsPerLane] = {FGKL_F32_UG_ROW(false),
                                                      FGKL_F32_UG_ROW(true)};

#undef FGKL_F32_UG_ROW

}  // namespace fgkl
