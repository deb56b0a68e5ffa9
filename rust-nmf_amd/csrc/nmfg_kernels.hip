// gfx950 kernels for the dense NMF update path (FP64 form).
//
// One iteration of every method is two passes over V, each a single kernel:
//
//   column pass  A = W^T V            (lee / hals / dnmf / ehals)
//                A = W^T (V ./ (W H)) (brunet, WH formed on the fly, never stored)
//   row pass     B = V H'^T           (lee / hals / dnmf / ehals)
//                B = (V ./ (W H')) H'^T (brunet)
//
// Each pass streams V once (coalesced dwordx4 rows, 2 doubles per lane) and
// keeps the rank-R factor panel of the streamed dimension in registers or
// SGPRs.  Work-groups own a column tile (column pass) or a row block (row
// pass) and one split of the contracted dimension; the last work-group to
// arrive at a tile sums the split partials in a fixed order and runs the
// method's epilogue (multiplicative update, Gauss-Seidel sweep, ...), then
// writes a small reduction record (Gram, sums, sums of squares).  The last
// tile to arrive sums the records into the statistics the next pass needs.
// All reductions run in a fixed order, so results are deterministic.
//
// Arithmetic contracts followed (rust/src/lib.rs):
//   lee     lib.rs:206-322, 1460-1470, 1588-1647
//   brunet  lib.rs:63-204, 1451-1458, 1528-1586
//   hals    lib.rs:581-688; ehals lib.rs:710-790; dnmf lib.rs:808-906, 1067-1119
#include "nmfg_internal.h"

#include <algorithm>

namespace nmfg {
namespace {

constexpr double BETA_EHALS = 0.5;  // lib.rs:740

__device__ __forceinline__ int wave_id() {
    return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

__device__ __forceinline__ double rust_max(double a, double b) {
    // f64::max: NaN-ignoring, as C fmax.
    return fmax(a, b);
}

// XCD-aware decode of a 1-D grid into (tile, split): the S splits of one
// tile land on the same XCD (work-groups are dealt to the 8 XCDs
// round-robin), so the tile's partials stay within one L2.
__device__ __forceinline__ bool decode_tile(int S, int ntiles, int &tile, int &split) {
    const int bid = blockIdx.x;
    const int xcd = bid & 7, slot = bid >> 3;
    tile = xcd + 8 * (slot / S);
    split = slot % S;
    return tile < ntiles;
}

// Arrival protocol: returns true in every thread of the last work-group to
// arrive at counter `cnt` (expected arrivals `total`).  The caller's global
// writes are released before the increment and acquired after it.
__device__ __forceinline__ bool arrive_last(int *cnt, int total, int *s_flag) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) *s_flag = (atomicAdd(cnt, 1) == total - 1);
    __syncthreads();
    const bool last = *s_flag != 0;
    if (last) __threadfence();
    return last;
}

// Reduce-scatter of a 64-entry per-lane vector across the wave: afterwards
// lane l holds the wave-wide sum of entry l.  63 shuffles instead of the
// 6*64 of a butterfly per entry.  Stage HALF keeps the upper or lower half
// of the entries according to lane bit HALF and adds the partner's copy.
template <int HALF>
__device__ __forceinline__ void rs_stage(const double (&in)[2 * HALF], double (&out)[HALF],
                                         int lane) {
    const bool up = (lane & HALF) != 0;
#pragma unroll
    for (int m = 0; m < HALF; ++m) {
        const double mine = up ? in[HALF + m] : in[m];
        const double other = up ? in[m] : in[HALF + m];
        out[m] = mine + __shfl_xor(other, HALF, 64);
    }
}

__device__ __forceinline__ double wave_reduce_scatter64(const double (&x)[64]) {
    const int lane = threadIdx.x & 63;
    double a32[32], a16[16], a8[8], a4[4], a2[2], a1[1];
    rs_stage<32>(x, a32, lane);
    rs_stage<16>(a32, a16, lane);
    rs_stage<8>(a16, a8, lane);
    rs_stage<4>(a8, a4, lane);
    rs_stage<2>(a4, a2, lane);
    rs_stage<1>(a2, a1, lane);
    return a1[0];
}

// ---------------------------------------------------------------------------
// Reduction record helpers: record = [Gram R*R | sum R | sumsq R] over the
// values vals[row][k] held in LDS (nv rows, row stride R).
// ---------------------------------------------------------------------------
template <int R>
__device__ void write_record(const double *vals, int nv, double *rec) {
    constexpr int NRED = nred(R);
    for (int o = threadIdx.x; o < NRED; o += blockDim.x) {
        double s = 0.0;
        if (o < R * R) {
            const int a = o / R, b = o % R;
            for (int t = 0; t < nv; ++t) s += vals[t * R + a] * vals[t * R + b];
        } else if (o < R * R + R) {
            const int a = o - R * R;
            for (int t = 0; t < nv; ++t) s += vals[t * R + a];
        } else {
            const int a = o - R * R - R;
            for (int t = 0; t < nv; ++t) s += vals[t * R + a] * vals[t * R + a];
        }
        rec[o] = s;
    }
}

template <int R>
__device__ void sum_records(const double *recs, int nrec, double *out) {
    constexpr int NRED = nred(R);
    for (int o = threadIdx.x; o < NRED; o += blockDim.x) {
        double s = 0.0;
        for (int t = 0; t < nrec; ++t) s += recs[(int64_t)t * NRED + o];
        out[o] = s;
    }
}

// Derive the W-side statistics from the raw record sums already stored at
// wstat[0 .. nred(R)) (Gu | cs | ss).
template <int R>
__device__ void derive_wstat(double *ws, const double *hs, int r, int kind, int rescale, int clamp,
                             bool fold_d = true) {
    __syncthreads();
    for (int k = threadIdx.x; k < R; k += blockDim.x) {
        double sc = 1.0, hdiv = 1.0, hmul = 1.0;
        if (k < r) {
            if (kind == K_LEE && rescale) {
                const double s = ws[ws_cs(R) + k];
                sc = s > 0.0 ? s : 1.0;  // col_rescale: divide only when s > 0
            } else if (kind == K_DNMF) {
                // extract_diagonal (lib.rs:808-824) then H *= d (lib.rs:880-884)
                const double nw = sqrt(ws[ws_ss(R) + k]);
                const double nh = sqrt(hs[hs_ss(R) + k]);
                sc = nw > 0.0 ? nw : 1.0;
                hdiv = nh > 0.0 ? nh : 1.0;
                hmul = fold_d ? nw * nh : 1.0;
            }
        }
        ws[ws_sc(R) + k] = sc;
        ws[ws_hdiv(R) + k] = hdiv;
        ws[ws_hmul(R) + k] = hmul;
    }
    if (threadIdx.x == 0) ws[ws_hclamp(R)] = clamp ? 1.0 : 0.0;
    __syncthreads();
    for (int o = threadIdx.x; o < R * R; o += blockDim.x) {
        const int a = o / R, b = o % R;
        ws[ws_G(R) + o] = ws[ws_Gu(R) + o] / (ws[ws_sc(R) + a] * ws[ws_sc(R) + b]);
    }
}

// Logical H entry from the stored one (see wstat layout).
__device__ __forceinline__ double logical_h(double h, const double *ws, int R, int k, int kind,
                                            double eps) {
    if (kind == K_DNMF) return (h / ws[ws_hdiv(R) + k]) * ws[ws_hmul(R) + k];
    if (kind == K_BRUNET && ws[ws_hclamp(R)] != 0.0) return h < eps ? eps : h;
    return h;
}

// ===========================================================================
// Column pass
// ===========================================================================
template <int R, bool KL>
__global__ __launch_bounds__(256) void k_col(PassArgs a) {
    constexpr int NW = 4;
    constexpr int KC = R < 8 ? R : 8;
    __shared__ double lds_red[NW / 2][KC][TJ];
    __shared__ double lds_vals[TJ * R];
    __shared__ int s_flag;

    int tile, split;
    if (!decode_tile(a.S, a.ntiles, tile, split)) return;
    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int64_t ldv = a.ldv;
    const int j0 = tile * TJ + lane * 2;
    const int rbeg = split * a.span;
    const int rend = min(a.n, rbeg + a.span);
    const int nrows = max(0, rend - rbeg);
    const int per = (nrows + NW - 1) / NW;
    const int wb = rbeg + wave * per;
    const int we = min(rend, wb + per);
    const double *__restrict__ V = a.V;
    const double *__restrict__ Wsrc = a.kind == K_EHALS ? a.Wx : a.W;

    double acc[R][2];
#pragma unroll
    for (int k = 0; k < R; ++k) acc[k][0] = acc[k][1] = 0.0;

    if constexpr (KL) {
        // brunet H pass: per column j, q_u = V[u,j] / (W H)[u,j] (0 when WH == 0),
        // A[k,j] = sum_u W[u,k] q_u   (divergence_update_h_inner, lib.rs:95-111)
        double h[R][2];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const double2 hv = *(const double2 *)(a.H + (int64_t)k * ldv + j0);
            h[k][0] = logical_h(hv.x, a.wstat, R, k, K_BRUNET, a.eps);
            h[k][1] = logical_h(hv.y, a.wstat, R, k, K_BRUNET, a.eps);
        }
        for (int i = wb; i < we; ++i) {
            const double2 v = *(const double2 *)(V + (int64_t)i * ldv + j0);
            const double *wr = Wsrc + (int64_t)i * R;
            double wh0 = 0.0, wh1 = 0.0;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                wh0 += wr[k] * h[k][0];
                wh1 += wr[k] * h[k][1];
            }
            const double q0 = wh0 != 0.0 ? v.x / wh0 : 0.0;
            const double q1 = wh1 != 0.0 ? v.y / wh1 : 0.0;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                acc[k][0] += wr[k] * q0;
                acc[k][1] += wr[k] * q1;
            }
        }
    } else {
        constexpr int U = 4;
        int i = wb;
        for (; i + U <= we; i += U) {
            double2 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = *(const double2 *)(V + (int64_t)(i + u) * ldv + j0);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const double *wr = Wsrc + (int64_t)(i + u) * R;
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    acc[k][0] += wr[k] * v[u].x;
                    acc[k][1] += wr[k] * v[u].y;
                }
            }
        }
        for (; i < we; ++i) {
            const double2 v = *(const double2 *)(V + (int64_t)i * ldv + j0);
            const double *wr = Wsrc + (int64_t)i * R;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                acc[k][0] += wr[k] * v.x;
                acc[k][1] += wr[k] * v.y;
            }
        }
    }

    // Cross-wave tree reduction (fixed order), KC ranks at a time.
#pragma unroll
    for (int k0 = 0; k0 < R; k0 += KC) {
#pragma unroll
        for (int half = NW / 2; half >= 1; half >>= 1) {
            if (wave >= half && wave < 2 * half) {
#pragma unroll
                for (int kk = 0; kk < KC; ++kk)
                    if (k0 + kk < R) {
                        lds_red[wave - half][kk][lane * 2] = acc[k0 + kk][0];
                        lds_red[wave - half][kk][lane * 2 + 1] = acc[k0 + kk][1];
                    }
            }
            __syncthreads();
            if (wave < half) {
#pragma unroll
                for (int kk = 0; kk < KC; ++kk)
                    if (k0 + kk < R) {
                        acc[k0 + kk][0] += lds_red[wave][kk][lane * 2];
                        acc[k0 + kk][1] += lds_red[wave][kk][lane * 2 + 1];
                    }
            }
            __syncthreads();
        }
    }
    if (wave == 0) {
        double *pp = a.part + (int64_t)split * R * ldv;
#pragma unroll
        for (int k = 0; k < R; ++k)
            *(double2 *)(pp + (int64_t)k * ldv + j0) = make_double2(acc[k][0], acc[k][1]);
    }
    if (!arrive_last(a.tile_cnt + tile, a.S, &s_flag)) return;

    // ---- tile epilogue: one thread per column ----------------------------
    const int t = threadIdx.x;
    if (t < TJ) {
        const int j = tile * TJ + t;
        const bool valid = j < a.p;
        const int r = a.r;
        double num[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            double s = 0.0;
            for (int sp = 0; sp < a.S; ++sp) s += a.part[((int64_t)sp * R + k) * ldv + j];
            num[k] = s;
        }
        double hn[R];
        const double *ws = a.wstat;
        const double eps = a.eps;
        if (a.kind == K_LEE) {
            // euclidean_update_h_inner (lib.rs:225-243) with the logical W = W / sc
            double h[R];
#pragma unroll
            for (int k = 0; k < R; ++k) h[k] = a.H[(int64_t)k * ldv + j];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (k < r) {
                    const double nm = num[k] / ws[ws_sc(R) + k];
                    double den = 0.0;
#pragma unroll
                    for (int l = R - 1; l >= 0; --l) den += ws[ws_G(R) + k * R + l] * h[l];
                    const double tt = h[k] * nm;
                    hn[k] = (tt > eps ? tt : eps) / (den + eps);
                } else {
                    hn[k] = 0.0;
                }
            }
        } else if (a.kind == K_PROD) {
#pragma unroll
            for (int k = 0; k < R; ++k) hn[k] = k < r ? num[k] : 0.0;
        } else if (a.kind == K_BRUNET) {
            // H'[k,j] = H[k,j] * A[k,j] / colsum(W)_k   (lib.rs:110)
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const double h = logical_h(a.H[(int64_t)k * ldv + j], ws, R, k, K_BRUNET, eps);
                hn[k] = k < r ? h * num[k] / ws[ws_cs(R) + k] : 0.0;
            }
        } else {
            // hals_update_h / _reg (lib.rs:581-607, 1067-1092): Gauss-Seidel
            // in k.  dnmf and ehals first form the logical start column.
            double h[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const double hs = a.H[(int64_t)k * ldv + j];
                if (a.kind == K_EHALS) {
                    const double hp = a.Hp[(int64_t)k * ldv + j];
                    const double e = hs + BETA_EHALS * (hs - hp);
                    h[k] = k < r ? (e > eps ? e : eps) : 0.0;
                } else {
                    h[k] = logical_h(hs, ws, R, k, a.kind, eps);
                }
                if (a.kind == K_EHALS && valid) a.Hp[(int64_t)k * ldv + j] = hs;
            }
            const bool reg = a.kind == K_DNMF;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (k < r) {
                    const double A = num[k] / ws[ws_sc(R) + k];
                    double s = reg ? A - a.lambda : A;
#pragma unroll
                    for (int l = 0; l < R; ++l)
                        if (l != k) s -= ws[ws_G(R) + k * R + l] * h[l];
                    const double b = reg ? rust_max(ws[ws_G(R) + k * R + k] + a.gamma, eps)
                                         : rust_max(ws[ws_G(R) + k * R + k], eps);
                    const double val = s / b;
                    h[k] = val > eps ? val : eps;
                }
            }
#pragma unroll
            for (int k = 0; k < R; ++k) hn[k] = k < r ? h[k] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (valid) a.H[(int64_t)k * ldv + j] = hn[k];
            lds_vals[t * R + k] = valid ? hn[k] : 0.0;
        }
    }
    __syncthreads();
    write_record<R>(lds_vals, TJ, a.red + (int64_t)tile * nred(R));
    if (threadIdx.x == 0) a.tile_cnt[tile] = 0;
    if (!arrive_last(a.glob_cnt, a.ntiles, &s_flag)) return;
    sum_records<R>(a.red, a.ntiles, a.hstat);
    if (threadIdx.x == 0) *a.glob_cnt = 0;
}

// ===========================================================================
// Row pass
// ===========================================================================
template <int R>
struct RowCfg {
    static constexpr int NW = R <= 16 ? 4 : 8;
    static constexpr int RW = RB / NW;        // rows per wave
    static constexpr int CPL = R <= 16 ? 2 : 1;  // columns per lane per step
    static_assert(RW * R <= 64, "reduce-scatter holds RW*R <= 64 entries");
};

template <int R, bool KL>
__global__ __launch_bounds__(RowCfg<R>::NW * 64) void k_row(PassArgs a) {
    constexpr int RW = RowCfg<R>::RW;
    constexpr int CPL = RowCfg<R>::CPL;
    constexpr int STEP = 64 * CPL;
    __shared__ double lds_vals[RB * R];
    __shared__ int s_flag;

    int blk, split;
    if (!decode_tile(a.S, a.nblocks, blk, split)) return;
    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int64_t ldv = a.ldv;
    const int cb = split * a.span;
    const int ce = min((int)ldv, cb + a.span);
    const int row0 = blk * RB + wave * RW;
    const double *__restrict__ V = a.V;

    int rows[RW];
#pragma unroll
    for (int ri = 0; ri < RW; ++ri) rows[ri] = min(row0 + ri, a.n - 1);

    double acc[RW][R];
#pragma unroll
    for (int ri = 0; ri < RW; ++ri)
#pragma unroll
        for (int k = 0; k < R; ++k) acc[ri][k] = 0.0;

    if constexpr (KL) {
        // brunet W pass: per row i, q_u = V[i,u] / (W H')[i,u],
        // B[i,k] = sum_u H'[k,u] q_u   (divergence_update_w_impl, lib.rs:187-200)
        for (int c = cb; c < ce; c += STEP) {
            const int j = c + lane * CPL;
            double hv[R][CPL];
#pragma unroll
            for (int k = 0; k < R; ++k)
#pragma unroll
                for (int cc = 0; cc < CPL; ++cc) hv[k][cc] = a.H[(int64_t)k * ldv + j + cc];
#pragma unroll
            for (int ri = 0; ri < RW; ++ri) {
                const double *wr = a.W + (int64_t)rows[ri] * R;
#pragma unroll
                for (int cc = 0; cc < CPL; ++cc) {
                    const double v = V[(int64_t)rows[ri] * ldv + j + cc];
                    double wh = 0.0;
#pragma unroll
                    for (int k = 0; k < R; ++k) wh += wr[k] * hv[k][cc];
                    const double q = wh != 0.0 ? v / wh : 0.0;
#pragma unroll
                    for (int k = 0; k < R; ++k) acc[ri][k] += hv[k][cc] * q;
                }
            }
        }
    } else {
        for (int c = cb; c < ce; c += STEP) {
            const int j = c + lane * CPL;
            double hv[R][CPL];
            double vv[RW][CPL];
            if constexpr (CPL == 2) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const double2 t = *(const double2 *)(a.H + (int64_t)k * ldv + j);
                    hv[k][0] = t.x;
                    hv[k][1] = t.y;
                }
#pragma unroll
                for (int ri = 0; ri < RW; ++ri) {
                    const double2 t = *(const double2 *)(V + (int64_t)rows[ri] * ldv + j);
                    vv[ri][0] = t.x;
                    vv[ri][1] = t.y;
                }
            } else {
#pragma unroll
                for (int k = 0; k < R; ++k) hv[k][0] = a.H[(int64_t)k * ldv + j];
#pragma unroll
                for (int ri = 0; ri < RW; ++ri) vv[ri][0] = V[(int64_t)rows[ri] * ldv + j];
            }
#pragma unroll
            for (int ri = 0; ri < RW; ++ri)
#pragma unroll
                for (int k = 0; k < R; ++k)
#pragma unroll
                    for (int cc = 0; cc < CPL; ++cc) acc[ri][k] += vv[ri][cc] * hv[k][cc];
        }
    }

    // Wave reduce-scatter: lane l ends with the sum of entry l = ri*R + k.
    double x[64];
#pragma unroll
    for (int m = 0; m < 64; ++m) x[m] = m < RW * R ? acc[m / R][m % R] : 0.0;
    const double mysum = wave_reduce_scatter64(x);
    const int nbr = a.nblocks * RB;
    if (lane < RW * R) {
        const int ri = lane / R, k = lane % R;
        a.part[((int64_t)split * nbr + row0 + ri) * R + k] = mysum;
    }
    if (!arrive_last(a.tile_cnt + blk, a.S, &s_flag)) return;

    // ---- row-block epilogue: one thread per row ---------------------------
    const int t = threadIdx.x;
    const int r = a.r;
    const double eps = a.eps;
    const double *ws = a.wstat;
    const double *hs = a.hstat;
    if (t < RB) {
        const int i = blk * RB + t;
        const bool valid = i < a.n;
        double wn[R];
        if (valid) {
            double num[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                double s = 0.0;
                for (int sp = 0; sp < a.S; ++sp) s += a.part[((int64_t)sp * nbr + i) * R + k];
                num[k] = s;
            }
            double w[R];
            if (a.kind == K_LEE) {
                // euclidean_update_w_impl (lib.rs:305-320) on the logical W
#pragma unroll
                for (int k = 0; k < R; ++k) w[k] = a.W[(int64_t)i * R + k] / ws[ws_sc(R) + k];
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    if (k < r) {
                        double den = 0.0;
#pragma unroll
                        for (int l = R - 1; l >= 0; --l) den += w[l] * hs[hs_K(R) + l * R + k];
                        const double tt = w[k] * num[k];
                        wn[k] = (tt < eps ? eps : tt) / (den + eps);
                    } else {
                        wn[k] = 0.0;
                    }
                }
            } else if (a.kind == K_PROD) {
#pragma unroll
                for (int k = 0; k < R; ++k) wn[k] = k < r ? num[k] : 0.0;
            } else if (a.kind == K_BRUNET) {
                // W'[i,k] = W[i,k] * B[i,k] / rowsum(H')_k (lib.rs:199), then the
                // every-10th-iteration clamp (lib.rs:1572-1575)
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const double wv = a.W[(int64_t)i * R + k];
                    double x2 = k < r ? wv * num[k] / hs[hs_rs(R) + k] : 0.0;
                    if (a.clamp && k < r && x2 < eps) x2 = eps;
                    wn[k] = x2;
                }
            } else {
                // hals_update_w / _reg (lib.rs:610-635, 1094-1119)
                const double *wsrc = a.kind == K_EHALS ? a.Wx : a.W;
#pragma unroll
                for (int k = 0; k < R; ++k) w[k] = wsrc[(int64_t)i * R + k] / ws[ws_sc(R) + k];
                const bool reg = a.kind == K_DNMF;
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    if (k < r) {
                        double s = reg ? num[k] - a.lambda : num[k];
#pragma unroll
                        for (int l = 0; l < R; ++l)
                            if (l != k) s -= hs[hs_K(R) + k * R + l] * w[l];
                        const double d = reg ? rust_max(hs[hs_K(R) + k * R + k] + a.gamma, eps)
                                             : rust_max(hs[hs_K(R) + k * R + k], eps);
                        const double val = s / d;
                        w[k] = val > eps ? val : eps;
                    }
                }
#pragma unroll
                for (int k = 0; k < R; ++k) wn[k] = k < r ? w[k] : 0.0;
            }
            if (a.kind == K_EHALS) {
                // next extrapolation point max(eps, W_t+1 + beta (W_t+1 - W_t))
                // (lib.rs:755-759); the record below is its Gram.
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const double wc = a.W[(int64_t)i * R + k];
                    a.W[(int64_t)i * R + k] = wn[k];
                    const double e = wn[k] + BETA_EHALS * (wn[k] - wc);
                    const double x2 = k < r ? (e > eps ? e : eps) : 0.0;
                    a.Wx[(int64_t)i * R + k] = x2;
                    wn[k] = x2;
                }
            } else {
#pragma unroll
                for (int k = 0; k < R; ++k) a.W[(int64_t)i * R + k] = wn[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < R; ++k) wn[k] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) lds_vals[t * R + k] = wn[k];
    }
    __syncthreads();
    write_record<R>(lds_vals, RB, a.red + (int64_t)blk * nred(R));
    if (threadIdx.x == 0) a.tile_cnt[blk] = 0;
    if (!arrive_last(a.glob_cnt, a.nblocks, &s_flag)) return;
    sum_records<R>(a.red, a.nblocks, a.wstat);
    derive_wstat<R>(a.wstat, a.hstat, r, a.kind, a.rescale, a.clamp);
    if (threadIdx.x == 0) *a.glob_cnt = 0;
}

// ===========================================================================
// Objective pass: lane owns 2 columns, W rows are wave-uniform.
// ===========================================================================
template <int R>
__global__ __launch_bounds__(256) void k_obj(ObjArgs a) {
    constexpr int NW = 4;
    __shared__ double lds[256];
    __shared__ int s_flag;
    int tile, split;
    if (!decode_tile(a.S, a.ntiles, tile, split)) return;
    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int64_t ldv = a.ldv;
    const int j0 = tile * TJ + lane * 2;
    const int rbeg = split * a.span;
    const int rend = min(a.n, rbeg + a.span);
    const int nrows = max(0, rend - rbeg);
    const int per = (nrows + NW - 1) / NW;
    const int wb = rbeg + wave * per;
    const int we = min(rend, wb + per);
    const double *ws = a.wstat;
    double h[R][2], isc[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        double h0 = a.H[(int64_t)k * ldv + j0], h1 = a.H[(int64_t)k * ldv + j0 + 1];
        if (a.hmode == 1) {
            h0 = (h0 / ws[ws_hdiv(R) + k]) * ws[ws_hmul(R) + k];
            h1 = (h1 / ws[ws_hdiv(R) + k]) * ws[ws_hmul(R) + k];
        } else if (a.hmode == 2 && ws[ws_hclamp(R)] != 0.0) {
            h0 = h0 < a.eps ? a.eps : h0;
            h1 = h1 < a.eps ? a.eps : h1;
        }
        const bool ok = k < a.r;
        h[k][0] = ok && j0 < a.p ? h0 : 0.0;
        h[k][1] = ok && j0 + 1 < a.p ? h1 : 0.0;
        isc[k] = 1.0 / ws[ws_sc(R) + k];
    }
    double s = 0.0;
    for (int i = wb; i < we; ++i) {
        const double2 v = *(const double2 *)(a.V + (int64_t)i * ldv + j0);
        const double *wr = a.W + (int64_t)i * R;
        double wh0 = 0.0, wh1 = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const double w = wr[k] * isc[k];
            wh0 += w * h[k][0];
            wh1 += w * h[k][1];
        }
        if (a.kl) {
            // kl_divergence term (lib.rs:1414-1426)
            if (j0 < a.p) s += v.x > 0.0 ? (wh0 > 0.0 ? v.x * log(v.x / wh0) - v.x + wh0 : 0.0) : wh0;
            if (j0 + 1 < a.p)
                s += v.y > 0.0 ? (wh1 > 0.0 ? v.y * log(v.y / wh1) - v.y + wh1 : 0.0) : wh1;
        } else {
            const double d0 = v.x - wh0, d1 = v.y - wh1;
            s += d0 * d0 + d1 * d1;
        }
    }
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st >= 1; st >>= 1) {
        if ((int)threadIdx.x < st) lds[threadIdx.x] += lds[threadIdx.x + st];
        __syncthreads();
    }
    const int nblk = a.ntiles * a.S;
    if (threadIdx.x == 0) a.part[tile * a.S + split] = lds[0];
    if (!arrive_last(a.glob_cnt, nblk, &s_flag)) return;
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int b = 0; b < nblk; ++b) tot += a.part[b];
        *a.out = a.kl ? tot : 0.5 * tot;
        *a.glob_cnt = 0;
    }
}

// ===========================================================================
// Statistics of raw factors: blocks [0, nblocks) reduce W row blocks,
// blocks [nblocks, nblocks + ntiles) reduce H column tiles.  The last block
// to arrive sums both sets of records.
// ===========================================================================
template <int R>
__global__ __launch_bounds__(256) void k_stats(StatArgs a) {
    __shared__ double vals[TJ * R];
    __shared__ int s_flag;
    const int b = blockIdx.x;
    const int t = threadIdx.x;
    if (b < a.nblocks) {
        if (t < RB) {
            const int i = b * RB + t;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                double w = i < a.n ? a.W[(int64_t)i * R + k] : 0.0;
                if (a.kind == K_EHALS && i < a.n) {
                    // first extrapolation point: W0 + beta*(W0 - W0) clamped (lib.rs:755-759)
                    const double e = w + BETA_EHALS * (w - w);
                    w = k < a.r ? (e > a.eps ? e : a.eps) : 0.0;
                    a.Wx[(int64_t)i * R + k] = w;
                }
                vals[t * R + k] = w;
            }
        }
        __syncthreads();
        write_record<R>(vals, RB, a.wred + (int64_t)b * nred(R));
    } else {
        const int tile = b - a.nblocks;
        if (t < TJ) {
            const int j = tile * TJ + t;
#pragma unroll
            for (int k = 0; k < R; ++k) vals[t * R + k] = j < a.p ? a.H[(int64_t)k * a.ldv + j] : 0.0;
        }
        __syncthreads();
        write_record<R>(vals, TJ, a.hred + (int64_t)tile * nred(R));
    }
    if (!arrive_last(a.glob_cnt, a.nblocks + a.ntiles, &s_flag)) return;
    sum_records<R>(a.wred, a.nblocks, a.wstat);
    sum_records<R>(a.hred, a.ntiles, a.hstat);
    // Solve start: the stored W is the logical W (no pending rescale).  A dnmf
    // solve stopped before its first iteration normalises W columns and H
    // rows without folding d into H (lib.rs:849-853).
    if (a.dnmf_norm)
        derive_wstat<R>(a.wstat, a.hstat, a.r, K_DNMF, 0, 0, false);
    else
        derive_wstat<R>(a.wstat, a.hstat, a.r, K_HALS, 0, 0);
    if (threadIdx.x == 0) *a.glob_cnt = 0;
}

// Apply the pending logical transforms to the stored factors.
template <int R>
__global__ __launch_bounds__(256) void k_finalize(Finalize f) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double *ws = f.wstat;
    if (idx < (int64_t)f.n * R) {
        const int k = (int)(idx % R);
        f.W[idx] = f.W[idx] / ws[ws_sc(R) + k];
    }
    if (idx < (int64_t)R * f.ldv) {
        const int k = (int)(idx / f.ldv);
        const int64_t j = idx % f.ldv;
        if (j < f.p && k < f.r) f.H[idx] = logical_h(f.H[idx], ws, R, k, f.kind, f.eps);
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
inline int round8(int x) { return (x + 7) / 8 * 8; }

template <int R>
hipError_t launch_col(const PassArgs &a, hipStream_t st) {
    const dim3 grid(round8(a.ntiles) * a.S);
    if (a.kind == K_BRUNET)
        hipLaunchKernelGGL((k_col<R, true>), grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_col<R, false>), grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int R>
hipError_t launch_row(const PassArgs &a, hipStream_t st) {
    const dim3 grid(round8(a.nblocks) * a.S);
    const dim3 block(RowCfg<R>::NW * 64);
    if (a.kind == K_BRUNET)
        hipLaunchKernelGGL((k_row<R, true>), grid, block, 0, st, a);
    else
        hipLaunchKernelGGL((k_row<R, false>), grid, block, 0, st, a);
    return hipGetLastError();
}

template <int R>
hipError_t launch_obj(const ObjArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_obj<R>), dim3(round8(a.ntiles) * a.S), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int R>
hipError_t launch_stats(const StatArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_stats<R>), dim3(a.nblocks + a.ntiles), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int R>
hipError_t launch_fin(const Finalize &f, hipStream_t st) {
    const int64_t tot = std::max<int64_t>((int64_t)f.n * R, (int64_t)R * f.ldv);
    hipLaunchKernelGGL((k_finalize<R>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, f);
    return hipGetLastError();
}

template <int R>
constexpr RankKernels table_entry() {
    return RankKernels{R, launch_col<R>, launch_row<R>, launch_obj<R>, launch_stats<R>, launch_fin<R>};
}

// Supported template ranks; a rank r runs on the smallest R >= r with the
// factors zero-padded to R (padded entries stay exactly zero).
const RankKernels kTable[] = {
    table_entry<1>(),  table_entry<2>(),  table_entry<3>(),  table_entry<4>(),
    table_entry<5>(),  table_entry<6>(),  table_entry<8>(),  table_entry<10>(),
    table_entry<12>(), table_entry<16>(), table_entry<20>(), table_entry<24>(),
    table_entry<32>(),
};

}  // namespace

const RankKernels *kernels_for_rank(int r) {
    for (const auto &e : kTable)
        if (e.R >= r) return &e;
    return nullptr;
}

int max_supported_rank() { return kTable[sizeof(kTable) / sizeof(kTable[0]) - 1].R; }

}  // namespace nmfg
