// gfx950 kernels of the dense NMF update path, V stored as two bfloat16 planes (hi + lo).
#include "nmfg_kernels.h"

namespace nmfg {

const RankKernels *kernels_for_rank_b16x2(int r) { return lookup_rank<VP_B16X2>(r); }

}  // namespace nmfg
