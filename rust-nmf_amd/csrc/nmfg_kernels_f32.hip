// gfx950 kernels of the dense NMF update path, V stored as FP32.
#include "nmfg_kernels.h"

namespace nmfg {

const RankKernels *kernels_for_rank_f32(int r) { return lookup_rank<VP_F32>(r); }

}  // namespace nmfg
