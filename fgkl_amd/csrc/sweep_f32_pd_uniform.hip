// FP32 instantiations of the PD sweep kernel for items whose reads have uniform gap qualities
// (one value of each gap term per lane, scaled gap values; see Rows), in the natural build and
// in the occupancy-raised one the autotuner can pick (sweep.hpp).
#include "sweep.hpp"

namespace fgkl {

#define FGKL_F32_PD_UG_ROW(B)                                                                  \
  {sweep_kernel<float, 1, true, B, false, true>, sweep_kernel<float, 2, true, B, false, true>, \
   sweep_kernel<float, 3, true, B, false, true>, sweep_kernel<float, 4, true, B, false, true>, \
   sweep_kernel<float, 5, true, B, false, true>, sweep_kernel<float, 6, true, B, false, true>, \
   sweep_kernel<float, 7, true, B, false, true>, sweep_kernel<float, 8, true, B, false, true>}

const KernelF kSweepF32PdUniform[2][kMaxRowsPerLane] = {FGKL_F32_PD_UG_ROW(false),
                                                        FGKL_F32_PD_UG_ROW(true)};

#undef FGKL_F32_PD_UG_ROW

}  // namespace fgkl
