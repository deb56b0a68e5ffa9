// gfx950 kernels of the dense NMF update path, V stored as FP64.
#include "nmfg_kernels.h"

namespace nmfg {

const RankKernels *kernels_for_rank_f64(int r) { return lookup_rank<VP_F64>(r); }

const RankKernels *kernels_for_rank(int r, int vp) {
    switch (vp) {
        case VP_F64: return kernels_for_rank_f64(r);
        case VP_F32: return kernels_for_rank_f32(r);
        case VP_B16X2: return kernels_for_rank_b16x2(r);
        default: return nullptr;
    }
}

int max_supported_rank() { return kMaxRank; }

hipError_t launch_consensus(const int32_t *labels, int runs, int p, double *C, hipStream_t st) {
    const unsigned nt = (unsigned)((p + CONS_T - 1) / CONS_T);
    hipLaunchKernelGGL(k_consensus, dim3(nt, nt), dim3(CONS_T * 8), 0, st, labels, runs, p, C);
    return hipGetLastError();
}

hipError_t launch_convert_v(int vp, const double *src, int64_t lds, int64_t rows, int64_t p,
                            void *dst, int64_t ldv, int64_t plane, hipStream_t st) {
    switch (vp) {
        case VP_F64: return convert_v<VP_F64>(src, lds, rows, p, dst, ldv, plane, st);
        case VP_F32: return convert_v<VP_F32>(src, lds, rows, p, dst, ldv, plane, st);
        case VP_B16X2: return convert_v<VP_B16X2>(src, lds, rows, p, dst, ldv, plane, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace nmfg
