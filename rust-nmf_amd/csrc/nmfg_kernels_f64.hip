// gfx950 kernels of the dense NMF update path, V stored as FP64.
#include <cfloat>
#include <climits>

#include "nmfg_kernels.h"

namespace nmfg {
namespace {

// ---------------------------------------------------------------------------
// Objective total and stop test on the device (see ObjFinArgs).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_objfin(ObjFinArgs a) {
    __shared__ double buf[256];
    if (*a.done) return;
    double s = 0.0;
    for (int b = threadIdx.x; b < a.nparts; b += blockDim.x) s += a.part[b];
    buf[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) buf[threadIdx.x] += buf[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double f = a.kl ? buf[0] : 0.5 * buf[0];
    // the iteration comes from a device counter, so that a captured chunk of
    // iterations replays unchanged
    const int64_t i = *a.it;
    *a.it = i + 1;
    a.devs[i] = f;
    bool stop = false;
    if (isnan(f)) {
        stop = true;
    } else if (a.stop == 1) {
        // StationaryStop (lib.rs:1378-1400): state = [-, max, min, count]
        double *st = a.state;
        if (i == 0) {
            st[1] = f;
            st[2] = f;
            st[3] = 1.0;
        } else if (!(st[3] == 0.0 && i % a.interval != 0)) {
            st[1] = st[3] == 0.0 ? f : fmax(st[1], f);
            st[2] = st[3] == 0.0 ? f : fmin(st[2], f);
            st[3] += 1.0;
            if (st[3] == (double)(a.niter + 1)) {
                const double crit = fabs(st[1] - st[2]) / (double)a.niter;
                if (crit <= a.th)
                    stop = true;
                else
                    st[3] = 0.0;
            }
        }
    } else {
        const double p = a.state[0];
        a.state[0] = f;
        if (i > 0 && i >= a.min_iter && fabs(p - f) <= a.tol * fabs(p)) stop = true;
    }
    if (stop) {
        *a.n_iter = i;
        *a.done = 1;
    }
}

// ---------------------------------------------------------------------------
// Consensus matrix of `runs` labelings (connectivity average): one thread per
// entry, the labels of the block's rows and columns staged in LDS.
// ---------------------------------------------------------------------------
constexpr int CONS_T = 32;   // tile edge
constexpr int CONS_RUNS = 64;  // labelings staged per pass

__global__ __launch_bounds__(CONS_T * 8) void k_consensus(const int32_t *labels, int runs, int p,
                                                          double *C) {
    __shared__ int32_t li[CONS_RUNS][CONS_T], lj[CONS_RUNS][CONS_T];
    const int i0 = blockIdx.y * CONS_T, j0 = blockIdx.x * CONS_T;
    const int tx = threadIdx.x % CONS_T, ty = threadIdx.x / CONS_T;  // ty in 0..7
    int cnt[4] = {0, 0, 0, 0};
    for (int r0 = 0; r0 < runs; r0 += CONS_RUNS) {
        const int nr = min(CONS_RUNS, runs - r0);
        __syncthreads();
        for (int e = threadIdx.x; e < nr * CONS_T; e += blockDim.x) {
            const int rr = e / CONS_T, c = e % CONS_T;
            li[rr][c] = i0 + c < p ? labels[(int64_t)(r0 + rr) * p + i0 + c] : -1;
            lj[rr][c] = j0 + c < p ? labels[(int64_t)(r0 + rr) * p + j0 + c] : -2;
        }
        __syncthreads();
        for (int rr = 0; rr < nr; ++rr) {
            const int b = lj[rr][tx];
#pragma unroll
            for (int q = 0; q < 4; ++q) cnt[q] += li[rr][ty + 8 * q] == b ? 1 : 0;
        }
    }
    const int j = j0 + tx;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = i0 + ty + 8 * q;
        if (i < p && j < p) C[(int64_t)i * p + j] = (double)cnt[q] / (double)runs;
    }
}

}  // namespace

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

hipError_t launch_objfin(const ObjFinArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_objfin, dim3(1), dim3(256), 0, st, a);
    return hipGetLastError();
}

namespace {

// One thread per panel entry (writing the panel) or per slice entry.
__global__ __launch_bounds__(256) void k_panel(PanelArgs a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a.to_panel) {
        const int64_t tot = a.colpanel ? (int64_t)a.R * a.ld : a.rows * a.R;
        if (e >= tot) return;
        const int64_t j = a.colpanel ? e % a.ld : e / a.R;
        const int kk = (int)(a.colpanel ? e / a.ld : e % a.R);
        a.P[e] = (kk < a.kc && j < a.rows) ? a.M[j * a.k + a.c0 + kk] : 0.0;
    } else {
        const int64_t tot = a.rows * a.kc;
        if (e >= tot) return;
        const int64_t j = e / a.kc;
        const int kk = (int)(e % a.kc);
        a.M[j * a.k + a.c0 + kk] = a.colpanel ? a.P[(int64_t)kk * a.ld + j] : a.P[j * a.R + kk];
    }
}

}  // namespace

namespace {
__global__ __launch_bounds__(256) void k_copy_unless_done(const double *src, double *dst,
                                                          int64_t count, const int *done) {
    if (done && *done) return;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < count; e += (int64_t)gridDim.x * 256)
        dst[e] = src[e];
}
}  // namespace

hipError_t launch_copy_unless_done(const double *src, double *dst, int64_t count, const int *done,
                                   hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const int64_t wgs = std::min<int64_t>(2048, (count + 255) / 256);
    hipLaunchKernelGGL(k_copy_unless_done, dim3((unsigned)wgs), dim3(256), 0, st, src, dst, count, done);
    return hipGetLastError();
}

namespace {
__global__ __launch_bounds__(256) void k_fill_f64(double *p, int64_t count, double v) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < count; e += (int64_t)gridDim.x * 256)
        p[e] = v;
}

// Reads `n16` 16-byte words (grid-stride, coalesced) so that the caches
// hold them instead of what was read before; the XOR of the words is
// stored only if it equals a constant, so the loads cannot be dropped.
__global__ __launch_bounds__(256) void k_evict(const uint4 *buf, int64_t n16, unsigned *out) {
    unsigned acc = 0;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n16; e += (int64_t)gridDim.x * 256) {
        const uint4 v = buf[e];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == 0x9e3779b9u) out[blockIdx.x * 256 + threadIdx.x] = acc;
}

__global__ __launch_bounds__(64) void k_stop_reset(double *state, int *done, int64_t *niter) {
    const int t = threadIdx.x;
    if (t < 4) state[t] = 0.0;
    if (t == 0) {
        *done = 0;
        niter[0] = -1;
        niter[1] = 0;
    }
}

// One thread per entry of the solve-layout W (n x R) or H (R x ldv).
__global__ __launch_bounds__(256) void k_factor_pack(FactorPack a) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t nw = a.n * a.R, nh = (int64_t)a.R * a.ldv;
    if (e < nw) {
        const int64_t i = e / a.R;
        const int k = (int)(e % a.R);
        if (a.to_solve)
            a.sW[e] = k < a.r ? a.W[i * a.r + k] : 0.0;
        else if (k < a.r)
            a.W[i * a.r + k] = a.sW[e];
    } else if (e < nw + nh) {
        const int64_t x = e - nw;
        const int k = (int)(x / a.ldv);
        const int64_t j = x % a.ldv;
        if (a.to_solve)
            a.sH[x] = (k < a.r && j < a.p) ? a.H[k * a.ldh + j] : 0.0;
        else if (k < a.r && j < a.p)
            a.H[k * a.ldh + j] = a.sH[x];
    }
}
}  // namespace

hipError_t launch_fill_f64(double *p, int64_t count, double v, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const int64_t wgs = std::min<int64_t>(2048, (count + 255) / 256);
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)wgs), dim3(256), 0, st, p, count, v);
    return hipGetLastError();
}

hipError_t launch_evict(const void *buf, int64_t n16, unsigned *out, hipStream_t st) {
    if (n16 <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_evict, dim3(EVICT_BLOCKS), dim3(256), 0, st,
                       static_cast<const uint4 *>(buf), n16, out);
    return hipGetLastError();
}

hipError_t launch_stop_reset(double *state, int *done, int64_t *niter, hipStream_t st) {
    hipLaunchKernelGGL(k_stop_reset, dim3(1), dim3(64), 0, st, state, done, niter);
    return hipGetLastError();
}

// Stop state of a step group's members in one buffer: out[2 i] = stop flag,
// out[2 i + 1] = stop iteration of member i.
__global__ __launch_bounds__(64) void k_gather_stop(StopPtrs p, int64_t *out) {
    const int i = threadIdx.x;
    if (i >= p.count) return;
    out[2 * i] = *p.done[i];
    out[2 * i + 1] = *p.niter[i];
}

hipError_t launch_gather_stop(const StopPtrs &p, int64_t *out, hipStream_t st) {
    if (p.count < 1 || p.count > GROUP_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gather_stop, dim3(1), dim3(64), 0, st, p, out);
    return hipGetLastError();
}

hipError_t launch_factor_pack(const FactorPack &a, hipStream_t st) {
    const int64_t tot = a.n * a.R + (int64_t)a.R * a.ldv;
    hipLaunchKernelGGL(k_factor_pack, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_panel(const PanelArgs &a, hipStream_t st) {
    const int64_t tot = a.to_panel ? (a.colpanel ? (int64_t)a.R * a.ld : a.rows * a.R) : a.rows * a.kc;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_panel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

namespace {

// ---------------------------------------------------------------------------
// Tall LU with partial pivoting into P L (see LuWork).  Step j: k_lu_select
// reduces the partial pivot candidates of column j, records the pivot row
// and swaps the positions; k_lu_step scales column j of the other
// candidate rows, updates their columns j+1.. and leaves the partial
// candidates of column j+1.  A candidate is a row not yet used as pivot;
// the better of two is the larger |x|, then the smaller position.
// ---------------------------------------------------------------------------
struct LuBest {
    double v;
    int pos, row;
};

__device__ __forceinline__ bool lu_better(const LuBest &a, const LuBest &b) {
    return a.v > b.v || (a.v == b.v && a.pos < b.pos);
}

// Work-group reduction of one candidate per thread into slot blockIdx.x.
__device__ void lu_reduce_store(LuBest b, const LuWork &ws) {
    __shared__ double sv[256];
    __shared__ int sp[256], sr[256];
    const int t = threadIdx.x;
    sv[t] = b.v;
    sp[t] = b.pos;
    sr[t] = b.row;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if (t < h) {
            const LuBest o{sv[t + h], sp[t + h], sr[t + h]};
            const LuBest c{sv[t], sp[t], sr[t]};
            if (lu_better(o, c)) {
                sv[t] = o.v;
                sp[t] = o.pos;
                sr[t] = o.row;
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        ws.part_val[blockIdx.x] = sv[0];
        ws.part_pos[blockIdx.x] = sp[0];
        ws.part_row[blockIdx.x] = sr[0];
    }
}

__global__ __launch_bounds__(256) void k_lu_init(const double *X, int64_t m, int w, LuWork ws) {
    LuBest b{-1.0, INT_MAX, -1};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256) {
        ws.pos[i] = (int)i;
        ws.row_at[i] = (int)i;
        ws.step_of[i] = -1;
        const LuBest c{fabs(X[i * w]), (int)i, (int)i};
        if (lu_better(c, b)) b = c;
    }
    lu_reduce_store(b, ws);
}

__global__ __launch_bounds__(256) void k_lu_select(const double *X, int w, int j, int nparts, LuWork ws) {
    __shared__ double sv[256];
    __shared__ int sp[256], sr[256];
    const int t = threadIdx.x;
    LuBest b{-1.0, INT_MAX, -1};
    for (int q = t; q < nparts; q += 256) {
        const LuBest c{ws.part_val[q], ws.part_pos[q], ws.part_row[q]};
        if (lu_better(c, b)) b = c;
    }
    sv[t] = b.v;
    sp[t] = b.pos;
    sr[t] = b.row;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if (t < h) {
            const LuBest o{sv[t + h], sp[t + h], sr[t + h]};
            const LuBest c{sv[t], sp[t], sr[t]};
            if (lu_better(o, c)) {
                sv[t] = o.v;
                sp[t] = o.pos;
                sr[t] = o.row;
            }
        }
        __syncthreads();
    }
    if (t != 0) return;
    // no candidate compares better (all NaN): the first position, as idamax
    const int p = sr[0] >= 0 ? sr[0] : ws.row_at[j];
    // swap position j (row q) with the pivot's position
    const int q = ws.row_at[j], pp = ws.pos[p];
    ws.row_at[pp] = q;
    ws.pos[q] = pp;
    ws.row_at[j] = p;
    ws.pos[p] = j;
    ws.step_of[p] = j;
    ws.prow[j] = p;
    const double piv = X[(int64_t)p * w + j];
    ws.sel[0] = piv;
    ws.sel[1] = 1.0 / piv;
}

__global__ __launch_bounds__(256) void k_lu_step(double *X, int64_t m, int w, int j, LuWork ws) {
    const int p = ws.prow[j];
    const double piv = ws.sel[0], rcp = ws.sel[1];
    const double *urow = X + (int64_t)p * w;
    LuBest b{-1.0, INT_MAX, -1};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256) {
        if (ws.step_of[i] >= 0) continue;
        double *x = X + i * w;
        double l = x[j];
        if (piv != 0.0) {
            l = fabs(piv) >= DBL_MIN ? l * rcp : l / piv;
            x[j] = l;
        }
        for (int c = j + 1; c < w; ++c) x[c] = fma(-l, urow[c], x[c]);
        if (j + 1 < w) {
            const LuBest cand{fabs(x[j + 1]), ws.pos[i], (int)i};
            if (lu_better(cand, b)) b = cand;
        }
    }
    if (j + 1 < w) lu_reduce_store(b, ws);
}

// Pivot row of step j: 1 at column j, 0 after it (its multipliers before j
// stay); one thread per (step, column).
__global__ __launch_bounds__(256) void k_lu_finish(double *X, int w, LuWork ws) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= w * w) return;
    const int j = e / w, c = e % w;
    if (c < j) return;
    X[(int64_t)ws.prow[j] * w + c] = c == j ? 1.0 : 0.0;
}

}  // namespace

size_t lu_work_bytes(int64_t m, int w) {
    return sizeof(int) * (size_t)(3 * m + 2 * LU_PARTS + w) + sizeof(double) * (LU_PARTS + 2) + 64;
}

LuWork lu_work_at(void *base, int64_t m, int w) {
    // doubles first (8-byte aligned base), then the ints
    double *d = static_cast<double *>(base);
    LuWork ws{};
    ws.part_val = d;
    ws.sel = d + LU_PARTS;
    int *ip = reinterpret_cast<int *>(d + LU_PARTS + 2);
    ws.pos = ip;
    ws.row_at = ip + m;
    ws.step_of = ip + 2 * m;
    ws.part_pos = ip + 3 * m;
    ws.part_row = ip + 3 * m + LU_PARTS;
    ws.prow = ip + 3 * m + 2 * LU_PARTS;
    (void)w;
    return ws;
}

hipError_t launch_lu_pl(double *X, int64_t m, int w, const LuWork &ws, hipStream_t st) {
    if (m < w || w < 1) return hipErrorInvalidValue;
    const int nparts = (int)std::min<int64_t>(LU_PARTS, (m + 255) / 256);
    hipLaunchKernelGGL(k_lu_init, dim3(nparts), dim3(256), 0, st, X, m, w, ws);
    for (int j = 0; j < w; ++j) {
        hipLaunchKernelGGL(k_lu_select, dim3(1), dim3(256), 0, st, X, w, j, nparts, ws);
        hipLaunchKernelGGL(k_lu_step, dim3(nparts), dim3(256), 0, st, X, m, w, j, ws);
    }
    hipLaunchKernelGGL(k_lu_finish, dim3((w * w + 255) / 256), dim3(256), 0, st, X, w, ws);
    return hipGetLastError();
}

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

// BF16 form: rows r0 .. r0 + rows - 1 of V (FP64 at src, row stride lds; r0
// a multiple of 8) into the column pass's copy Vc.  One thread per (8-row
// block, column): it reads the block's 8 values of its column (zeros past
// the chunk's rows) and writes them as one dwordx4 per plane.
__global__ __launch_bounds__(256) void k_convert_cb(const double *src, int64_t lds, int64_t rows,
                                                     int64_t r0, int64_t p, uint16_t *Vc, int64_t ldv,
                                                     int64_t cplane) {
    const int64_t nb = (rows + 7) / 8, tot = nb * p;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ib = e / p, j = e % p;
        uint16_t hi[8], lo[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int64_t i = 8 * ib + s;
            bf16_split(i < rows ? src[i * lds + j] : 0.0, hi[s], lo[s]);
        }
        const int64_t o = ((r0 / 8 + ib) * ldv + j) * 8;
        uint4 h, l;
        h.x = hi[0] | ((uint32_t)hi[1] << 16);
        h.y = hi[2] | ((uint32_t)hi[3] << 16);
        h.z = hi[4] | ((uint32_t)hi[5] << 16);
        h.w = hi[6] | ((uint32_t)hi[7] << 16);
        l.x = lo[0] | ((uint32_t)lo[1] << 16);
        l.y = lo[2] | ((uint32_t)lo[3] << 16);
        l.z = lo[4] | ((uint32_t)lo[5] << 16);
        l.w = lo[6] | ((uint32_t)lo[7] << 16);
        *reinterpret_cast<uint4 *>(Vc + o) = h;
        *reinterpret_cast<uint4 *>(Vc + cplane + o) = l;
    }
}

hipError_t launch_convert_vc(const double *src, int64_t lds, int64_t rows, int64_t r0, int64_t p,
                             uint16_t *Vc, int64_t ldv, int64_t cplane, hipStream_t st) {
    if (r0 % 8 != 0 || rows < 1) return hipErrorInvalidValue;
    const int64_t tot = (rows + 7) / 8 * p;
    const unsigned grid = (unsigned)std::min<int64_t>(16384, std::max<int64_t>(1, (tot + 255) / 256));
    hipLaunchKernelGGL(k_convert_cb, dim3(grid), dim3(256), 0, st, src, lds, rows, r0, p, Vc, ldv, cplane);
    return hipGetLastError();
}

hipError_t launch_rowsum_v(int vp, const void *V, int64_t vplane, int64_t ldv, int n, int p,
                           double *out, hipStream_t st) {
    switch (vp) {
        case VP_F64: return rowsum_v<VP_F64>(V, vplane, ldv, n, p, out, st);
        case VP_F32: return rowsum_v<VP_F32>(V, vplane, ldv, n, p, out, st);
        case VP_B16X2: return rowsum_v<VP_B16X2>(V, vplane, ldv, n, p, out, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace nmfg
