// gfx950 kernels for the dense NMF update path, templated on the storage of V.
//
// One iteration of every method is three kernels:
//
//   k_col   column pass, one work-group per column tile over ALL rows of V:
//             A = W^T V          (lee / hals / dnmf / ehals: FP64 MFMA)
//             A = W^T (V ./ WH)  (brunet: VALU, WH formed per element)
//           plus, inside the work-group, the Gram W^T W, column sums and
//           sums of squares of W (read anyway), then the H epilogue for the
//           tile and a per-tile record [Gram of H' | row sums | sumsq].
//   k_red   sums the per-tile records in a fixed order.
//   k_row   row pass, one work-group per block of rows over ALL columns:
//             B = V H'^T or (V ./ WH') H'^T, then the W epilogue.
//
// No work-group waits on another inside a kernel (no atomics, no fences):
// every reduction is inside one work-group or across a kernel boundary, in a
// fixed order, so results are deterministic.  On MI355X an agent-scope fence
// writes back and invalidates the XCD's L2 and costs far more than a kernel
// boundary inside a HIP graph.
//
// Arithmetic contracts followed (rust/src/lib.rs):
//   lee     lib.rs:206-322, 1460-1470, 1588-1647
//   brunet  lib.rs:63-204, 1451-1458, 1528-1586
//   hals    lib.rs:581-688; ehals lib.rs:710-790; dnmf lib.rs:808-906, 1067-1119
//   offset  lib.rs:324-450, 1649-1759;  nsNMF lib.rs:454-466, 1761-1833
#pragma once
#include "nmfg_internal.h"

#include <algorithm>
#include <type_traits>

namespace nmfg {
namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
constexpr double BETA_EHALS = 0.5;  // lib.rs:740
constexpr int RF = 32;  // rows per work-group of the split row passes' second kernels

// Sum of the nsplit partials p[sp * stride], sp = 0 .. nsplit - 1, in split
// order, with 8 independent loads in flight per step.
__device__ __forceinline__ double split_sum(const double *p, int nsplit, int64_t stride) {
    double s = 0.0;
    int sp = 0;
    for (; sp + 8 <= nsplit; sp += 8) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = p[(int64_t)(sp + u) * stride];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += x[u];
    }
    for (; sp < nsplit; ++sp) s += p[(int64_t)sp * stride];
    return s;
}

__device__ __forceinline__ int wave_id() {
    return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// Per-solve views of the arguments of a batched launch (blockIdx.y = solve).
__device__ __forceinline__ PassArgs for_solve(PassArgs a) {
    const int64_t b = blockIdx.y;
    a.W += b * a.bt.w;
    if (a.kind == K_EHALS || a.kind == K_NSNMF) {
        a.Wx += b * a.bt.w;
        a.Hp += b * a.bt.h;
    }
    if (a.off) a.off += b * a.bt.off;
    a.H += b * a.bt.h;
    a.red += b * a.bt.red;
    a.hst += b * a.bt.hst;
    a.wsx += b * a.bt.wsx;
    if (a.gram) a.gram += b * a.bt.gram;
    if (a.rpart) a.rpart += b * a.bt.rpart;
    if (a.kpart) a.kpart += b * a.bt.kpart;
    if (a.Wf) a.Wf += b * a.bt.w;
    if (a.Hf) a.Hf += b * a.bt.h;
    if (a.num) a.num += b * a.bt.num;
    if (a.Wb) a.Wb += b * a.bt.wb;
    if (a.Hb) a.Hb += b * a.bt.hb;
    if (a.bt.ranks) a.r = a.bt.ranks[b];
    return a;
}

__device__ __forceinline__ WGramArgs for_solve(WGramArgs a) {
    const int64_t b = blockIdx.y;
    a.W += b * a.bt.w;
    if (a.Wf) a.Wf += b * a.bt.w;
    if (a.Wb) a.Wb += b * a.bt.wb;
    if (a.off) a.off += b * a.bt.off;
    a.part += b * a.bt.gpart;
    a.out += b * a.bt.gram;
    return a;
}

__device__ __forceinline__ RedArgs for_solve(RedArgs a) {
    const int64_t b = blockIdx.y;
    a.red += b * a.bt.red;
    a.hst += b * a.bt.hst;
    return a;
}

__device__ __forceinline__ WStatArgs for_solve(WStatArgs a) {
    const int64_t b = blockIdx.y;
    a.W += b * a.bt.w;
    a.part += b * a.bt.wpart;
    a.out += b * a.bt.wsc;
    if (a.bt.ranks) a.r = a.bt.ranks[b];
    return a;
}

__device__ __forceinline__ PrepArgs for_solve(PrepArgs a) {
    const int64_t b = blockIdx.y;
    a.W += b * a.bt.w;
    if (a.Wf) a.Wf += b * a.bt.w;
    if (a.kind == K_EHALS) a.Wx += b * a.bt.w;
    a.wsx += b * a.bt.wsx;
    if (a.bt.ranks) a.r = a.bt.ranks[b];
    return a;
}

__device__ __forceinline__ NsPrepArgs for_solve(NsPrepArgs a) {
    const int64_t b = blockIdx.y;
    a.W += b * a.bt.w;
    a.Wx += b * a.bt.w;
    a.wsc += b * a.bt.wsc;
    if (a.bt.ranks) a.r = a.bt.ranks[b];
    return a;
}

__device__ __forceinline__ Finalize for_solve(Finalize f) {
    const int64_t b = blockIdx.y;
    f.W += b * f.bt.w;
    f.H += b * f.bt.h;
    f.wsc += b * f.bt.wsc;
    f.hst += b * f.bt.hst;
    if (f.bt.ranks) f.r = f.bt.ranks[b];
    return f;
}

// float -> bfloat16, round to nearest even (inf and NaN kept).
__device__ __forceinline__ uint16_t bf16_rne(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan kept
    const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

// x as two bfloat16 values: f = float(x), hi = bf16(f), lo = bf16(f - hi)
// (hi + lo carries about 16 significant bits of f).
__device__ __forceinline__ void bf16_split(double x, uint16_t &hi, uint16_t &lo) {
    const float f = (float)x;
    hi = bf16_rne(f);
    lo = bf16_rne(f - __uint_as_float((uint32_t)hi << 16));
}

// f64::max: NaN-ignoring, as C fmax.
__device__ __forceinline__ double rust_max(double a, double b) { return fmax(a, b); }

// Reads of V elements in each storage form, widened to FP64 (exactly: every
// float and every bf16 pair sum is an FP64 value).  two() reads elements i
// and i + 1 (i even).
template <int VP>
struct VElem;

template <>
struct VElem<VP_F64> {
    __device__ static double one(const void *V, int64_t, int64_t i) {
        return static_cast<const double *>(V)[i];
    }
    __device__ static double2 two(const void *V, int64_t, int64_t i) {
        return *reinterpret_cast<const double2 *>(static_cast<const double *>(V) + i);
    }
};

template <>
struct VElem<VP_F32> {
    __device__ static double one(const void *V, int64_t, int64_t i) {
        return static_cast<const float *>(V)[i];
    }
    __device__ static double2 two(const void *V, int64_t, int64_t i) {
        const float2 f = *reinterpret_cast<const float2 *>(static_cast<const float *>(V) + i);
        return make_double2(f.x, f.y);
    }
};

__device__ __forceinline__ double bf16_to_double(uint32_t b) {
    return (double)__uint_as_float(b << 16);
}

template <>
struct VElem<VP_B16X2> {
    __device__ static double one(const void *V, int64_t plane, int64_t i) {
        const uint16_t *h = static_cast<const uint16_t *>(V);
        return bf16_to_double(h[i]) + bf16_to_double(h[plane + i]);
    }
    __device__ static double2 two(const void *V, int64_t plane, int64_t i) {
        const uint16_t *h = static_cast<const uint16_t *>(V);
        const uint32_t hi = *reinterpret_cast<const uint32_t *>(h + i);
        const uint32_t lo = *reinterpret_cast<const uint32_t *>(h + plane + i);
        return make_double2(bf16_to_double(hi & 0xffffu) + bf16_to_double(lo & 0xffffu),
                            bf16_to_double(hi >> 16) + bf16_to_double(lo >> 16));
    }
};

// D += A * B on a 16x16x4 FP64 tile.  Lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; it holds D[4*m + l/16][l%16] in element m (layout measured
// on gfx950: the 4 rows a lane holds are interleaved, stride 4).
__device__ __forceinline__ dbl4 mfma(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Reduce-scatter of a 64-entry per-lane vector across the wave: afterwards
// lane l holds the wave-wide sum of entry l.  Stage HALF keeps the upper or
// lower half of the entries according to lane bit HALF and adds the
// partner's copy (63 shuffles in all).
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

// Record (layout in nmfg_internal.h) of NV vectors of R entries held in LDS
// (vals[t * R + k]); written by all threads of the work-group.  Each output
// is summed as 4 interleaved partial sums (t mod 4, in order of t), added
// at the end in a fixed order: independent chains keep the LDS loads in
// flight.
template <int R, int NV>
__device__ void write_record(const double *vals, double *rec) {
    static_assert(NV % 4 == 0, "record over a multiple of 4 vectors");
    constexpr int NRED = nred(R);
    for (int o = threadIdx.x; o < NRED; o += blockDim.x) {
        int ia, ib;  // entries multiplied (ib < 0: plain sum of entry ia)
        if (o < ntri(R)) {
            int x = 0, rem = o;
            while (rem >= R - x) {
                rem -= R - x;
                ++x;
            }
            ia = x;
            ib = x + rem;
        } else if (o < ntri(R) + R) {
            ia = o - ntri(R);
            ib = -1;
        } else {
            ia = ib = o - ntri(R) - R;
        }
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
        if (ib < 0) {
#pragma unroll 4
            for (int t = 0; t < NV; t += 4) {
                p0 += vals[t * R + ia];
                p1 += vals[(t + 1) * R + ia];
                p2 += vals[(t + 2) * R + ia];
                p3 += vals[(t + 3) * R + ia];
            }
        } else {
#pragma unroll 4
            for (int t = 0; t < NV; t += 4) {
                p0 += vals[t * R + ia] * vals[t * R + ib];
                p1 += vals[(t + 1) * R + ia] * vals[(t + 1) * R + ib];
                p2 += vals[(t + 2) * R + ia] * vals[(t + 2) * R + ib];
                p3 += vals[(t + 3) * R + ia] * vals[(t + 3) * R + ib];
            }
        }
        rec[o] = (p0 + p1) + (p2 + p3);
    }
}

// Gauss-Seidel sweep over k of one column (H) or row (W):
// hals_update_{h,w} (lib.rs:581-635) and _reg (lib.rs:1067-1119).
// G is the R x R Gram (row stride gs), A the right-hand side.
template <int R>
__device__ __forceinline__ void gs_sweep(double (&x)[R], const double (&A)[R], const double *G,
                                         int gs, int r, bool reg, double lambda, double gamma,
                                         double eps) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (k < r) {
            double s = reg ? A[k] - lambda : A[k];
#pragma unroll
            for (int l = 0; l < R; ++l)
                if (l != k) s -= G[k * gs + l] * x[l];
            const double b = reg ? rust_max(G[k * gs + k] + gamma, eps) : rust_max(G[k * gs + k], eps);
            const double val = s / b;
            x[k] = val > eps ? val : eps;
        }
    }
}

// H epilogue of one column tile of TJ columns, after the column pass has
// combined its waves' sums: numL (R16 x (TJ + 1), row stride TJ + 1) holds
// W^T V of the stored W, GL (R16 x (R16 + 1)) its Gram, csL / ssL its column
// sums / sums of squares, hL (R x TJ) the tile's current H (lee / offset).
// Derives the stored -> logical divisor of W (and dnmf's H transform), runs
// the H update of the tile's columns (one thread per (k, column) for lee /
// offset / products, one per column for the Gauss-Seidel sweeps), writes H'
// (and its float copy Hf when given) and the tile's record.  Called by
// every thread of the work-group.
template <int R, bool GS, int TJ, typename NT = double>
__device__ void col_tile_epilogue(const PassArgs &a, int tile, NT *numL, double *GL,
                                  const double *csL, const double *ssL, double *scL, double *hdivL,
                                  double *hmulL, double *hnL, const double *hL, float *Hf = nullptr,
                                  uint16_t *Hb = nullptr) {
    constexpr int KB = (R + 15) / 16;
    constexpr int R16 = 16 * KB;
    constexpr int NS = TJ + 1;   // numL row stride
    constexpr int GS1 = R16 + 1;  // GL row stride
    const int r = a.r;
    const int64_t ldv = a.ldv;
    const int j0 = tile * TJ;
    // Divisor from the stored to the logical W, and the dnmf H transform.
    const int t = threadIdx.x;
    if (t < R16) {
        double sc = 1.0, hdiv = 1.0, hmul = 1.0;
        if (t < r && !a.first) {
            if (a.kind == K_LEE && a.rescale) {
                sc = csL[t] > 0.0 ? csL[t] : 1.0;  // col_rescale (lib.rs:1460-1470)
            } else if (a.kind == K_DNMF) {
                // extract_diagonal then H *= d (lib.rs:808-824, 880-884)
                const double nw = sqrt(ssL[t]);
                const double nh = sqrt(a.hst[hs_ss(R) + t]);
                sc = nw > 0.0 ? nw : 1.0;
                hdiv = nh > 0.0 ? nh : 1.0;
                hmul = nw * nh;
            }
        }
        scL[t] = sc;
        hdivL[t] = hdiv;
        hmulL[t] = hmul;
        if (tile == 0 && t < R) a.wsx[t] = sc;
    }
    __syncthreads();
    for (int o = t; o < R16 * R16; o += blockDim.x) {
        const int ra = o / R16, cb = o % R16;
        GL[ra * GS1 + cb] = GL[ra * GS1 + cb] / (scL[ra] * scL[cb]);
    }
    __syncthreads();

    if (a.kind == K_LEE || a.kind == K_PROD || a.kind == K_OFFSET) {
        // one thread per (k, column): euclidean_update_h_inner (lib.rs:225-243);
        // offset adds den_addon = (W^T off)_k, which k_wgram leaves in the
        // third section of the W record (offset_update_h_inner, lib.rs:351-352)
        const bool offs = a.kind == K_OFFSET;
        for (int e = t; e < R * TJ; e += blockDim.x) {
            const int k = e / TJ, c = e % TJ;
            const int j = j0 + c;
            const bool valid = j < a.p;
            double hn = 0.0;
            if (k < r) {
                if (a.kind == K_PROD) {
                    hn = (double)numL[k * NS + c];
                } else {
                    const double eps = a.eps;
                    double den = 0.0;
#pragma unroll
                    for (int l = R - 1; l >= 0; --l) den += GL[k * GS1 + l] * hL[l * TJ + c];
                    if (offs) den += ssL[k];
                    const double tt = hL[k * TJ + c] * ((double)numL[k * NS + c] / scL[k]);
                    hn = (tt > eps ? tt : eps) / (den + eps);
                }
            }
            if (valid) {
                a.H[(int64_t)k * ldv + j] = hn;
                if (Hf) Hf[(int64_t)k * ldv + j] = (float)hn;
                if (Hb) bf16_split(hn, Hb[(int64_t)k * ldv + j], Hb[(int64_t)(R + k) * ldv + j]);
            }
            hnL[c * R + k] = valid ? hn : 0.0;
        }
    } else if (GS && t < TJ) {
        // one thread per column: the Gauss-Seidel sweeps are sequential in k
        // (compiled only into the HALS-family instances, GS)
        const int j = j0 + t;
        const bool valid = j < a.p;
        const double eps = a.eps;
        double A[R], h[R], hn[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            A[k] = (double)numL[k * NS + t] / scL[k];
            h[k] = a.H[(int64_t)k * ldv + j];
        }
        if (a.kind == K_EHALS) {
            // extrapolated start max(eps, H + beta (H - H_prev)) (lib.rs:760-764)
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const double hp = a.Hp[(int64_t)k * ldv + j];
                const double e = h[k] + BETA_EHALS * (h[k] - hp);
                if (valid) a.Hp[(int64_t)k * ldv + j] = h[k];
                h[k] = k < r ? (e > eps ? e : eps) : 0.0;
            }
        } else if (a.kind == K_DNMF) {
#pragma unroll
            for (int k = 0; k < R; ++k) h[k] = (h[k] / hdivL[k]) * hmulL[k];
        }
        gs_sweep<R>(h, A, GL, GS1, r, a.kind == K_DNMF, a.lambda, a.gamma, eps);
#pragma unroll
        for (int k = 0; k < R; ++k) hn[k] = k < r ? h[k] : 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (valid) {
                a.H[(int64_t)k * ldv + j] = hn[k];
                if (Hf) Hf[(int64_t)k * ldv + j] = (float)hn[k];
            }
            hnL[t * R + k] = valid ? hn[k] : 0.0;
        }
    }
    if (a.kind == K_PROD) return;
    __syncthreads();
    write_record<R, TJ>(hnL, a.red + (int64_t)tile * nred(R));
}

// ===========================================================================
// Column pass, MFMA path (lee, hals, dnmf, ehals, products).
//
// Work-group = TJC = 32 columns x all n rows; NWC waves split the rows.  Per
// 4-row step lane l loads V[i + l/16][j0 + 2(l%16) .. +1] (one dwordx4) and
// W[i + l/16][16 kb + l%16]; two MFMAs per k-block accumulate W^T V for the
// even and odd columns, one more per block pair accumulates W^T W.
// ===========================================================================
template <int R, int VP, bool XG, bool GS, int NW = NWC, int U = 8>
__global__ __launch_bounds__(NW * 64) void k_col_mfma(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    constexpr int KB = (R + 15) / 16;
    constexpr int R16 = 16 * KB;
    // Gram blocks formed here (none when the Gram comes from k_wgram: XG)
    constexpr int NG = XG ? 0 : KB * (KB + 1) / 2;
    constexpr int SLOT = 8 * KB + (XG ? 0 : 4 * NG + 2 * KB);  // doubles per lane to combine
    // U = 4-row steps per unrolled block
    __shared__ double slot[NW / 2][SLOT][64];
    __shared__ double numL[R16][TJC + 1];
    __shared__ double GL[R16][R16 + 1];
    __shared__ double csL[R16], ssL[R16], scL[R16], hdivL[R16], hmulL[R16];
    __shared__ double hnL[TJC * R];
    __shared__ double hL[R][TJC];  // the tile's current H (lee / product epilogue)

    const int tile = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int q = lane >> 4, c = lane & 15;
    const int64_t ldv = a.ldv;
    const int j0 = tile * TJC;
    const void *__restrict__ V = a.V;
    const int64_t vplane = a.vplane;
    const double *__restrict__ Ws = a.kind == K_EHALS ? a.Wx : a.W;
    const bool with_gram = !XG && a.kind != K_PROD;

    const int per = (((a.n + NW - 1) / NW) + 3) & ~3;
    const int rb = min(a.n, wave * per);
    const int re = min(a.n, rb + per);
    if (a.kind == K_LEE || a.kind == K_OFFSET)
        for (int e = threadIdx.x; e < R * TJC; e += blockDim.x)
            hL[e / TJC][e % TJC] = a.H[(int64_t)(e / TJC) * ldv + j0 + e % TJC];

    dbl4 d0[KB], d1[KB], g[NG > 0 ? NG : 1];
    double cs[KB], ss[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        d0[kb] = dbl4{0, 0, 0, 0};
        d1[kb] = dbl4{0, 0, 0, 0};
        cs[kb] = ss[kb] = 0.0;
    }
#pragma unroll
    for (int t = 0; t < NG; ++t) g[t] = dbl4{0, 0, 0, 0};

    // Blocks of U 4-row steps; rows past re read row rb and are zeroed
    // (addresses stay in bounds).
    for (int i = rb; i < re; i += 4 * U) {
        double2 v[U];
        double w[U][KB];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int row = i + 4 * u + q;
            const bool ok = row < re;
            const int rowc = ok ? row : rb;
            v[u] = VElem<VP>::two(V, vplane, (int64_t)rowc * ldv + j0 + 2 * c);
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) {
                const int k = 16 * kb + c;
                const double wv = Ws[(int64_t)rowc * R + (k < R ? k : R - 1)];
                w[u][kb] = (ok && k < R) ? wv : 0.0;
            }
            if (!ok) v[u] = make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) {
                d0[kb] = mfma(w[u][kb], v[u].x, d0[kb]);
                d1[kb] = mfma(w[u][kb], v[u].y, d1[kb]);
            }
            if constexpr (!XG) {
                if (with_gram) {
                    int gi = 0;
#pragma unroll
                    for (int ka = 0; ka < KB; ++ka)
#pragma unroll
                        for (int kb = ka; kb < KB; ++kb, ++gi)
                            g[gi] = mfma(w[u][ka], w[u][kb], g[gi]);
#pragma unroll
                    for (int kb = 0; kb < KB; ++kb) {
                        cs[kb] += w[u][kb];
                        ss[kb] += w[u][kb] * w[u][kb];
                    }
                }
            }
        }
    }

    // Combine the waves' partials in a fixed tree order.
    double vals[SLOT];
    {
        int s = 0;
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                vals[s++] = d0[kb][m];
                vals[s++] = d1[kb][m];
            }
        if constexpr (!XG) {
#pragma unroll
            for (int t = 0; t < NG; ++t)
#pragma unroll
                for (int m = 0; m < 4; ++m) vals[s++] = g[t][m];
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) {
                vals[s++] = cs[kb];
                vals[s++] = ss[kb];
            }
        }
    }
#pragma unroll
    for (int half = NW / 2; half >= 1; half >>= 1) {
        if (wave >= half && wave < 2 * half)
#pragma unroll
            for (int s = 0; s < SLOT; ++s) slot[wave - half][s][lane] = vals[s];
        __syncthreads();
        if (wave < half)
#pragma unroll
            for (int s = 0; s < SLOT; ++s) vals[s] += slot[wave][s][lane];
        __syncthreads();
    }
    if (wave == 0) {
        int s = 0;
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int k = 16 * kb + 4 * m + q;
                numL[k][2 * c] = vals[s++];
                numL[k][2 * c + 1] = vals[s++];
            }
        if constexpr (!XG) {
            int gi = 0;
#pragma unroll
            for (int ka = 0; ka < KB; ++ka)
#pragma unroll
                for (int kb = ka; kb < KB; ++kb, ++gi)
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const double x = vals[s++];
                        const int ra = 16 * ka + 4 * m + q, cb = 16 * kb + c;
                        GL[ra][cb] = x;
                        if (ka != kb) GL[cb][ra] = x;
                    }
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) {
                double c1 = vals[s++];
                double s1 = vals[s++];
                c1 += __shfl_xor(c1, 16, 64);
                s1 += __shfl_xor(s1, 16, 64);
                c1 += __shfl_xor(c1, 32, 64);
                s1 += __shfl_xor(s1, 32, 64);
                if (q == 0) {
                    csL[16 * kb + c] = c1;
                    ssL[16 * kb + c] = s1;
                }
            }
        }
    }
    if constexpr (XG) {
        // Gram, column sums and sums of squares of the streamed W (k_wgram)
        for (int o = threadIdx.x; o < R16 * R16; o += blockDim.x) {
            const int ra = o / R16, cb = o % R16;
            GL[ra][cb] = (ra < R && cb < R) ? a.gram[tri_index(R, ra, cb)] : 0.0;
        }
        if ((int)threadIdx.x < R16) {
            const int k = threadIdx.x;
            csL[k] = k < R ? a.gram[ntri(R) + k] : 0.0;
            ssL[k] = k < R ? a.gram[ntri(R) + R + k] : 0.0;
        }
    }
    __syncthreads();

    col_tile_epilogue<R, GS, TJC>(a, tile, &numL[0][0], &GL[0][0], csL, ssL, scL, hdivL, hmulL, hnL,
                                  &hL[0][0]);
}

// ===========================================================================
// FP32 form (lee): the two products W^T V and V H'^T on
// v_mfma_f32_16x16x4_f32 with FP32 accumulation over float V and the float
// copies of the streamed factors (Wf from k_wgram, Hf from the column
// pass); Grams, epilogues, the objective and the stop test stay FP64.
// Lane layouts measured on gfx950 with an in-bounds probe: lane l supplies
// A[l%16][l/16] and B[l/16][l%16] and holds D[4(l/16) + m][l%16] in element
// m.
// ===========================================================================
typedef float flt4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ flt4 mfma32(float a, float b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Column pass, streaming kernel: work-group = TJS = 128 columns x all n
// rows, NW waves split the rows.  Per 4-row step lane l loads
// V[i + l/16][j0 + 64 h + 4(l%16) .. +3] for h = 0, 1 (two dwordx4) and
// Wf[i + l/16][16 kb + l%16]; MFMA (h, e) takes element e of load h, so it
// covers the columns j0 + 64 h + 4c + e, c = 0..15.  The waves' sums are
// combined in a fixed tree and written as FP32 numerators num[k][j] (R x
// ldv): for each (k-block, m, h) a lane stores one dwordx4,
// num[16 kb + 4 (l/16) + m][j0 + 64 h + 4 (l%16) ..].  The H update runs in
// k_colfin_f32, so this kernel keeps few registers and little LDS (high
// occupancy while streaming).
template <int R, int NW = 4, int U = 4>
__global__ __launch_bounds__(NW * 64) void k_colnum_f32(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    constexpr int KB = (R + 15) / 16;
    constexpr int SLOT = 32 * KB;  // floats per lane: KB k-blocks x 8 MFMAs x 4
    __shared__ float slot[NW / 2][SLOT][64];
    const int tile = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int q = lane >> 4, c = lane & 15;
    const int64_t ldv = a.ldv;
    const int j0 = tile * TJS;
    const float *__restrict__ V = static_cast<const float *>(a.V);
    const float *__restrict__ Wf = a.Wf;

    const int per = (((a.n + NW - 1) / NW) + 3) & ~3;
    const int rb = min(a.n, wave * per);
    const int re = min(a.n, rb + per);
    flt4 d[KB][8];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int e = 0; e < 8; ++e) d[kb][e] = flt4{0, 0, 0, 0};

    // Blocks of U 4-row steps; rows past re read row rb and are zeroed
    // (addresses stay in bounds); W entries past rank R are not loaded.
    for (int i = rb; i < re; i += 4 * U) {
        flt4 v[U][2];
        float w[U][KB];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int row = i + 4 * u + q;
            const bool ok = row < re;
            const int rowc = ok ? row : rb;
            const float *vr = V + (int64_t)rowc * ldv + j0 + 4 * c;
            v[u][0] = *reinterpret_cast<const flt4 *>(vr);
            v[u][1] = *reinterpret_cast<const flt4 *>(vr + 64);
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) {
                const int k = 16 * kb + c;
                float wv = 0.0f;
                if (ok && k < R) wv = Wf[(int64_t)rowc * R + k];
                w[u][kb] = wv;
            }
            if (!ok) v[u][0] = v[u][1] = flt4{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int kb = 0; kb < KB; ++kb)
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    d[kb][e] = mfma32(w[u][kb], v[u][e >> 2][e & 3], d[kb][e]);
    }

    // Combine the waves' partials in a fixed tree order (FP32 sums).
    float vals[SLOT];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int e = 0; e < 8; ++e)
#pragma unroll
            for (int m = 0; m < 4; ++m) vals[((kb * 4 + m) * 2 + (e >> 2)) * 4 + (e & 3)] = d[kb][e][m];
#pragma unroll
    for (int half = NW / 2; half >= 1; half >>= 1) {
        if (wave >= half && wave < 2 * half)
#pragma unroll
            for (int s = 0; s < SLOT; ++s) slot[wave - half][s][lane] = vals[s];
        __syncthreads();
        if (wave < half)
#pragma unroll
            for (int s = 0; s < SLOT; ++s) vals[s] += slot[wave][s][lane];
        __syncthreads();
    }
    if (wave == 0)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int k = 16 * kb + 4 * q + m;
                if (k < R)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int o = ((kb * 4 + m) * 2 + h) * 4;
                        *reinterpret_cast<flt4 *>(a.num + (int64_t)k * ldv + j0 + 64 * h + 4 * c) =
                            flt4{vals[o], vals[o + 1], vals[o + 2], vals[o + 3]};
                    }
            }
}

// Column pass, H update: one work-group per TJF-column tile reads the
// tile's FP32 numerators (k_colnum_f32), the Gram, column sums and sums of
// squares of the streamed W (k_wgram, FP64) and runs the FP64 epilogue
// (col_tile_epilogue), which also writes H's float copy for the row pass.
template <int R>
__global__ __launch_bounds__(256) void k_colfin_f32(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    constexpr int KB = (R + 15) / 16;
    constexpr int R16 = 16 * KB;
    __shared__ double numL[R16][TJF + 1];
    __shared__ double GL[R16][R16 + 1];
    __shared__ double csL[R16], ssL[R16], scL[R16], hdivL[R16], hmulL[R16];
    __shared__ double hnL[TJF * R];
    __shared__ double hL[R][TJF];
    const int tile = blockIdx.x;
    const int64_t ldv = a.ldv;
    const int j0 = tile * TJF;
    // Every global load of the tile is issued before the first LDS store
    // (fixed trip counts at 256 threads), so their latencies overlap instead
    // of adding up loop trip by loop trip.
    constexpr int NT = 256, LN = (R16 * TJF + NT - 1) / NT, LH = (R * TJF + NT - 1) / NT,
                  LG = (R16 * R16 + NT - 1) / NT;
    const int t = threadIdx.x;
    double nv[LN], hv[LH], gv[LG];
#pragma unroll
    for (int i = 0; i < LN; ++i) {
        const int e = t + i * NT, k = e / TJF, c = e % TJF;
        nv[i] = (e < R16 * TJF && k < R) ? (double)a.num[(int64_t)k * ldv + j0 + c] : 0.0;
    }
    // BF16 form with row splits: the other splits' numerators, in split order
    for (int sp = 1; sp < a.kl_split; ++sp)
#pragma unroll
        for (int i = 0; i < LN; ++i) {
            const int e = t + i * NT, k = e / TJF, c = e % TJF;
            if (e < R16 * TJF && k < R) nv[i] += (double)a.num[(int64_t)(sp * R + k) * ldv + j0 + c];
        }
#pragma unroll
    for (int i = 0; i < LH; ++i) {
        const int e = t + i * NT;
        hv[i] = e < R * TJF ? a.H[(int64_t)(e / TJF) * ldv + j0 + e % TJF] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < LG; ++i) {
        const int o = t + i * NT, ra = o / R16, cb = o % R16;
        gv[i] = (o < R16 * R16 && ra < R && cb < R) ? a.gram[tri_index(R, ra, cb)] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < LN; ++i) {
        const int e = t + i * NT;
        if (e < R16 * TJF) numL[e / TJF][e % TJF] = nv[i];
    }
#pragma unroll
    for (int i = 0; i < LH; ++i) {
        const int e = t + i * NT;
        if (e < R * TJF) hL[e / TJF][e % TJF] = hv[i];
    }
#pragma unroll
    for (int i = 0; i < LG; ++i) {
        const int o = t + i * NT;
        if (o < R16 * R16) GL[o / R16][o % R16] = gv[i];
    }
    if ((int)threadIdx.x < R16) {
        const int k = threadIdx.x;
        csL[k] = k < R ? a.gram[ntri(R) + k] : 0.0;
        ssL[k] = k < R ? a.gram[ntri(R) + R + k] : 0.0;
    }
    __syncthreads();
    col_tile_epilogue<R, false, TJF>(a, tile, &numL[0][0], &GL[0][0], csL, ssL, scL, hdivL, hmulL,
                                     hnL, &hL[0][0], a.Hf, a.Hb);
}

// Row pass: partial numerators V H'^T (rpart; k_rowfin adds the column
// splits and runs the FP64 W epilogue).  Work-group = NW waves x RT 16-row
// tiles each (wave w owns rows row0 + 16 (w RT + t) + l%16) over one column
// split, walked in chunks of CHF = 64 columns.  The chunk's R rows of Hf are
// staged in LDS once per work-group and read by every wave (double
// buffered: the next chunk's global loads are issued before the current
// chunk's MFMAs and stored after them).  Per 16-column step lane l loads,
// for each of its tiles, V[row][c0 + 4(l/16) .. +3] (one dwordx4) and reads
// H'[16 kb + l%16][c0 + 4(l/16) .. +3] from LDS; MFMA e takes element e of
// both: D[k][row] += H'[k][c] V[row][c], c = c0 + 4(l/16) + e (even and odd
// e in separate accumulators).  Each wave owns its rows, so its sums go to
// rpart directly.  Grid: nrb x nsplit work-groups.
// Work-groups go to the 8 XCDs round-robin by flattened id.  With gridDim.x
// a multiple of 8 the flattened id mod 8 is blockIdx.x mod 8, and this
// renumbering gives consecutive logical indices to one XCD: the row passes'
// work-groups of one column split (consecutive indices) then share that
// XCD's L2, so H' of the split comes from HBM about once instead of once per
// row block.  Otherwise the identity.
__device__ __forceinline__ int xcd_local_index() {
    const int b = blockIdx.x, g = gridDim.x;
    return (g & 7) ? b : (b & 7) * (g >> 3) + (b >> 3);
}

constexpr int CHF = 64;
constexpr int HLD4 = CHF / 4 + 1;  // LDS row stride of a staged chunk, in float4 (padded)
template <int R, int RT, int NW = 4>
__global__ __launch_bounds__(NW * 64) void k_row_f32(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    constexpr int KB = (R + 15) / 16;
    constexpr int NF4 = R * (CHF / 4);                   // float4s of a staged chunk
    constexpr int PER = (NF4 + NW * 64 - 1) / (NW * 64);  // staged float4s per thread
    __shared__ flt4 hs[2][R * HLD4];
    const int xl = xcd_local_index();
    const int rb = xl % a.nrb, split = xl / a.nrb;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = wave_id();
    const int q = lane >> 4, c = lane & 15;
    const int64_t ldv = a.ldv;
    const int row0 = (rb * NW + wave) * RT * 16;
    int64_t vrow[RT];
    bool rok[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int row = row0 + 16 * t + c;
        rok[t] = row < a.n;
        vrow[t] = (int64_t)(rok[t] ? row : a.n - 1) * ldv;
    }
    const int nch = (int)(ldv / CHF);
    const int per = (nch + a.nsplit - 1) / a.nsplit;
    const int cs = min(nch, split * per), ce = min(nch, cs + per);
    const float *__restrict__ V = static_cast<const float *>(a.V);
    const float *__restrict__ H = a.Hf;
    flt4 hreg[PER];
    auto gload = [&](int ch) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int f = tid + u * NW * 64;
            if (f < NF4)
                hreg[u] = *reinterpret_cast<const flt4 *>(H + (int64_t)(f >> 4) * ldv + ch * CHF +
                                                          4 * (f & 15));
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int f = tid + u * NW * 64;
            if (f < NF4) hs[buf][(f >> 4) * HLD4 + (f & 15)] = hreg[u];
        }
    };
    flt4 d0[RT][KB], d1[RT][KB];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) d0[t][kb] = d1[t][kb] = flt4{0, 0, 0, 0};
    if (cs < ce) {
        gload(cs);
        sstore(0);
    }
    __syncthreads();
    for (int ch = cs; ch < ce; ++ch) {
        const int buf = (ch - cs) & 1;
        if (ch + 1 < ce) gload(ch + 1);
        flt4 v[4][RT];
#pragma unroll
        for (int st = 0; st < 4; ++st)
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                v[st][t] = *reinterpret_cast<const flt4 *>(V + vrow[t] + ch * CHF + 16 * st + 4 * q);
                if (!rok[t]) v[st][t] = flt4{0, 0, 0, 0};
            }
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            flt4 h[KB];
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) {
                const int k = 16 * kb + c;
                h[kb] = k < R ? hs[buf][k * HLD4 + 4 * st + q] : flt4{0, 0, 0, 0};
            }
#pragma unroll
            for (int e = 0; e < 4; e += 2)
#pragma unroll
                for (int t = 0; t < RT; ++t)
#pragma unroll
                    for (int kb = 0; kb < KB; ++kb) {
                        d0[t][kb] = mfma32(h[kb][e], v[st][t][e], d0[t][kb]);
                        d1[t][kb] = mfma32(h[kb][e + 1], v[st][t][e + 1], d1[t][kb]);
                    }
        }
        if (ch + 1 < ce) sstore(buf ^ 1);
        __syncthreads();
    }
    // lane holds D[k = 16 kb + 4 q + m][row = 16 t + c] of its wave's rows
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int k = 16 * kb + 4 * q + m, ri = row0 + 16 * t + c;
                if (k < R && ri < a.n)
                    a.rpart[((int64_t)split * a.n + ri) * R + k] = (double)(d0[t][kb][m] + d1[t][kb][m]);
            }
}

// ===========================================================================
// BF16 hi + lo form (lee): W^T V and V H'^T on v_mfma_f32_16x16x32_bf16
// with FP32 accumulation.  Every operand x is carried as two bfloat16
// values, hi = bf16(float(x)) and lo = bf16(float(x) - hi) (about 16
// significant bits), and a product a b is formed as hi_a hi_b + hi_a lo_b +
// lo_a hi_b (lo_a lo_b is below the representation error).  The MFMA
// contracts 32 indices, 8 per lane: lane l supplies A[l%16][slots of group
// l/16] and B[slots of group l/16][l%16], the same 8 slots in both, and
// holds D[4(l/16) + m][l%16] in element m.  A lane's 8 values therefore lie
// along the contracted index: the column pass (a sum over rows) streams a
// copy of V in blocks of 8 rows with each column's 8 values contiguous
// (PassArgs::Vc) and a transposed copy of W (Wb); the row pass (a sum over
// columns) streams V's row-major planes and H' (Hb).  Grams, epilogues,
// the objective and the stop test stay FP64.
// ===========================================================================
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
// 8 bfloat16 values as one dwordx4 (a vector type, so arrays of it stay in
// registers)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ flt4 mfma_bf16(u32x4 a, u32x4 b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Column pass, numerators: work-group = NW waves over one tile of TJF = 16 NB
// columns and one row split (kl_split splits of the nc / 32 row steps);
// the waves divide the split's steps.  Per 32-row step lane l (c = l%16,
// g = l/16) loads, for each 16-column block b, rows 32 s + 8 g .. + 7 of
// column j0 + 16 b + c from both planes of Vc (two dwordx4), and per
// k-block the same rows of W's column 16 kb + c from both planes of Wb.  The
// waves' sums are combined in a fixed tree; the split's numerators go to
// num[split][k][j] (R x ldv floats per split), summed by k_colfin_f32.
// FUSE (one split): the work-group runs the H epilogue itself instead
// (k_colfin_f32's, from the same numerators widened to FP64), so that it
// overlaps with the other work-groups' streaming; the tile's current H is
// staged in LDS before the stream, and the combine's slots share their LDS
// with the epilogue's arrays.
template <int R, bool FUSE>
struct ColB16Lds {
    static constexpr int KB = (R + 15) / 16, R16 = 16 * KB;
    static constexpr int SLOTB = 2 * KB * 4 * 4 * 64 * 4;  // NW / 2 x SLOT x 64 floats (NW = 4, NB = 4)
    static constexpr int EPIB = FUSE ? 4 * R16 * (TJF + 2) + 8 * (R16 * (R16 + 1) + TJF * R + 5 * R16) : 0;
    static constexpr int UNION = SLOTB > EPIB ? SLOTB : EPIB;
    static constexpr int HLB = FUSE ? 8 * R * TJF : 8;
};

template <int R, int NB = 4, int NW = 4, int U = 2, bool FUSE = false>
__global__ __launch_bounds__(NW * 64, FUSE ? 3 : 1) void k_colnum_b16(PassArgs args) {
    static_assert(16 * NB == TJF, "a tile of the numerator pass is a tile of k_colfin_f32");
    static_assert(NB == 4 && NW == 4, "ColB16Lds sizes the combine slots for NB = NW = 4");
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    constexpr int KB = (R + 15) / 16;
    constexpr int R16 = 16 * KB;
    constexpr int SLOT = KB * NB * 4;  // floats per lane
    using L = ColB16Lds<R, FUSE>;
    __shared__ __align__(16) char lds_u[L::UNION];
    __shared__ __align__(16) double hL[L::HLB / 8];
    float(*slot)[SLOT][64] = reinterpret_cast<float(*)[SLOT][64]>(lds_u);
    const int tile = blockIdx.x % a.ntiles, split = blockIdx.x / a.ntiles;
    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int g = lane >> 4, c = lane & 15;
    const int64_t ldv = a.ldv;
    const int j0 = tile * TJF;
    const int steps = a.nc / 32;
    const int sper = (steps + a.kl_split - 1) / a.kl_split;
    const int s0 = min(steps, split * sper), s1 = min(steps, s0 + sper);
    const int wper = (s1 - s0 + NW - 1) / NW;
    const int ws0 = min(s1, s0 + wave * wper), ws1 = min(s1, ws0 + wper);
    const uint16_t *__restrict__ Vc = a.Vc;
    const uint16_t *__restrict__ Wb = a.Wb;
    const int64_t wplane = (int64_t)R * a.nc;
    if constexpr (FUSE)  // the tile's current H, for the epilogue
        for (int e = threadIdx.x; e < R * TJF; e += NW * 64)
            hL[e] = a.H[(int64_t)(e / TJF) * ldv + j0 + e % TJF];
    flt4 d[KB][NB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int b = 0; b < NB; ++b) d[kb][b] = flt4{0, 0, 0, 0};
    const u32x4 z = {0u, 0u, 0u, 0u};
    // U steps per trip; a step past ws1 re-reads step ws0 with W taken as 0.
    // Every load is issued unconditionally (in-bounds addresses) and zeroed
    // afterwards: a load under a lane condition merges with its alternative
    // at once, which waits for it and for every load issued before it.
    for (int s = ws0; s < ws1; s += U) {
        u32x4 vh[U][NB], vl[U][NB], wh[U][KB], wl[U][KB];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool ok = s + u < ws1;
            const int sc = ok ? s + u : ws0;
            const int64_t vb = ((int64_t)(4 * sc + g) * ldv + j0 + c) * 8;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                vh[u][b] = *reinterpret_cast<const u32x4 *>(Vc + vb + 128 * b);
                vl[u][b] = *reinterpret_cast<const u32x4 *>(Vc + a.cplane + vb + 128 * b);
            }
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) {
                const int k = 16 * kb + c;
                const int64_t wo = (int64_t)min(k, R - 1) * a.nc + 32 * sc + 8 * g;
                wh[u][kb] = *reinterpret_cast<const u32x4 *>(Wb + wo);
                wl[u][kb] = *reinterpret_cast<const u32x4 *>(Wb + wplane + wo);
            }
        }
        __builtin_amdgcn_sched_barrier(0);  // all of the trip's loads in flight first
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int kb = 0; kb < KB; ++kb)
                if (s + u >= ws1 || 16 * kb + c >= R) wh[u][kb] = wl[u][kb] = z;
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int kb = 0; kb < KB; ++kb)
#pragma unroll
                for (int b = 0; b < NB; ++b) d[kb][b] = mfma_bf16(wh[u][kb], vh[u][b], d[kb][b]);
#pragma unroll
            for (int kb = 0; kb < KB; ++kb)
#pragma unroll
                for (int b = 0; b < NB; ++b) d[kb][b] = mfma_bf16(wh[u][kb], vl[u][b], d[kb][b]);
#pragma unroll
            for (int kb = 0; kb < KB; ++kb)
#pragma unroll
                for (int b = 0; b < NB; ++b) d[kb][b] = mfma_bf16(wl[u][kb], vh[u][b], d[kb][b]);
        }
    }

    // Combine the waves' partials in a fixed tree order (FP32 sums).
    float vals[SLOT];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int m = 0; m < 4; ++m) vals[(kb * NB + b) * 4 + m] = d[kb][b][m];
#pragma unroll
    for (int half = NW / 2; half >= 1; half >>= 1) {
        if (wave >= half && wave < 2 * half)
#pragma unroll
            for (int q = 0; q < SLOT; ++q) slot[wave - half][q][lane] = vals[q];
        __syncthreads();
        if (wave < half)
#pragma unroll
            for (int q = 0; q < SLOT; ++q) vals[q] += slot[wave][q][lane];
        __syncthreads();
    }
    // lane holds D[k = 16 kb + 4 g + m][column j0 + 16 b + c]
    if constexpr (!FUSE) {
        if (wave == 0) {
            float *num = a.num + (int64_t)split * R * ldv;
#pragma unroll
            for (int kb = 0; kb < KB; ++kb)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int k = 16 * kb + 4 * g + m;
                    if (k < R)
#pragma unroll
                        for (int b = 0; b < NB; ++b)
                            num[(int64_t)k * ldv + j0 + 16 * b + c] = vals[(kb * NB + b) * 4 + m];
                }
        }
    } else {
        // the combine's slots are free now: the epilogue's arrays take their LDS
        // numerators as the FP32 sums (widened in the epilogue), row stride TJF + 1
        float *numL = reinterpret_cast<float *>(lds_u);             // [R16][TJF + 1]
        double *GL = reinterpret_cast<double *>(lds_u + 4 * R16 * (TJF + 2));  // [R16][R16 + 1]
        double *hnL = GL + R16 * (R16 + 1);                         // [TJF * R]
        double *csL = hnL + TJF * R, *ssL = csL + R16, *scL = ssL + R16, *hdivL = scL + R16,
               *hmulL = hdivL + R16;
        if (wave == 0)
#pragma unroll
            for (int kb = 0; kb < KB; ++kb)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int k = 16 * kb + 4 * g + m;
#pragma unroll
                    for (int b = 0; b < NB; ++b)
                        numL[k * (TJF + 1) + 16 * b + c] = vals[(kb * NB + b) * 4 + m];
                }
        for (int o = threadIdx.x; o < R16 * R16; o += NW * 64) {
            const int ra = o / R16, cb = o % R16;
            GL[ra * (R16 + 1) + cb] = (ra < R && cb < R) ? a.gram[tri_index(R, ra, cb)] : 0.0;
        }
        if ((int)threadIdx.x < R16) {
            const int k = threadIdx.x;
            csL[k] = k < R ? a.gram[ntri(R) + k] : 0.0;
            ssL[k] = k < R ? a.gram[ntri(R) + R + k] : 0.0;
        }
        __syncthreads();
        col_tile_epilogue<R, false, TJF, float>(a, tile, numL, GL, csL, ssL, scL, hdivL, hmulL, hnL,
                                                hL, nullptr, a.Hb);
    }
}

// Row pass: partial numerators V H'^T (rpart; k_rowfin adds the column
// splits and runs the FP64 W epilogue).  Work-group = NW waves x RT 16-row
// tiles each (wave w owns rows row0 + 16 (w RT + t) + l%16) over one column
// split, walked in chunks of CHF = 64 columns whose R rows of both planes
// of Hb are staged in LDS once per work-group (double buffered, as
// k_row_f32).  Per 32-column step s of a chunk lane l (c = l%16, g = l/16)
// loads, for each of its tiles, columns c0 + 32 s + 8 g .. + 7 of its row
// from both planes of V (two dwordx4) and reads the same columns of H' row
// 16 kb + c from LDS: D[k][row] += H'[k][col] V[row][col].
constexpr int HLB = CHF / 8 + 1;  // LDS row stride of a staged chunk, in 8-value groups (padded)
template <int R, int RT, int NW = 4>
__global__ __launch_bounds__(NW * 64) void k_row_b16(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    constexpr int KB = (R + 15) / 16;
    constexpr int NU = 2 * R * (CHF / 8);                // 8-value groups of a staged chunk
    constexpr int PER = (NU + NW * 64 - 1) / (NW * 64);  // staged groups per thread
    __shared__ u32x4 hs[2][2 * R * HLB];
    const int xl = xcd_local_index();
    const int rb = xl % a.nrb, split = xl / a.nrb;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = wave_id();
    const int g = lane >> 4, c = lane & 15;
    const int64_t ldv = a.ldv;
    const int row0 = (rb * NW + wave) * RT * 16;
    int64_t vrow[RT];
    bool rok[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int row = row0 + 16 * t + c;
        rok[t] = row < a.n;
        vrow[t] = (int64_t)(rok[t] ? row : a.n - 1) * ldv;
    }
    const int nch = (int)(ldv / CHF);
    const int per = (nch + a.nsplit - 1) / a.nsplit;
    const int cs = min(nch, split * per), ce = min(nch, cs + per);
    const uint16_t *__restrict__ V = static_cast<const uint16_t *>(a.V);
    const uint16_t *__restrict__ H = a.Hb;
    const u32x4 z = {0u, 0u, 0u, 0u};
    u32x4 hreg[PER];
    // staged group f: plane f / (8 R), H' row (f / 8) % R, columns 8 (f % 8) ..
    // (loads unconditional, from group min(f, NU - 1); only f < NU is stored)
    auto gload = [&](int ch) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int f = min(tid + u * NW * 64, NU - 1);
            hreg[u] = *reinterpret_cast<const u32x4 *>(H + (int64_t)(f >> 3) * ldv + ch * CHF +
                                                       8 * (f & 7));
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int f = tid + u * NW * 64;
            if (f < NU) hs[buf][(f >> 3) * HLB + (f & 7)] = hreg[u];
        }
    };
    flt4 d[RT][KB];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) d[t][kb] = flt4{0, 0, 0, 0};
    if (cs < ce) {
        gload(cs);
        sstore(0);
    }
    __syncthreads();
    for (int ch = cs; ch < ce; ++ch) {
        const int buf = (ch - cs) & 1;
        if (ch + 1 < ce) gload(ch + 1);
        u32x4 vh[2][RT], vl[2][RT];
#pragma unroll
        for (int st = 0; st < 2; ++st)
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const int64_t o = vrow[t] + ch * CHF + 32 * st + 8 * g;
                vh[st][t] = *reinterpret_cast<const u32x4 *>(V + o);
                vl[st][t] = *reinterpret_cast<const u32x4 *>(V + a.vplane + o);
            }
        // every V load of the chunk is in flight before the first MFMA (the
        // scheduler would otherwise trade loads in flight for occupancy)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int st = 0; st < 2; ++st)
#pragma unroll
            for (int t = 0; t < RT; ++t)
                if (!rok[t]) vh[st][t] = vl[st][t] = z;
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            u32x4 hh[KB], hl[KB];
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) {
                const int k = min(16 * kb + c, R - 1);
                hh[kb] = hs[buf][k * HLB + 4 * st + g];
                hl[kb] = hs[buf][(R + k) * HLB + 4 * st + g];
                if (16 * kb + c >= R) hh[kb] = hl[kb] = z;
            }
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) d[t][kb] = mfma_bf16(hh[kb], vh[st][t], d[t][kb]);
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) d[t][kb] = mfma_bf16(hh[kb], vl[st][t], d[t][kb]);
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) d[t][kb] = mfma_bf16(hl[kb], vh[st][t], d[t][kb]);
        }
        if (ch + 1 < ce) sstore(buf ^ 1);
        __syncthreads();
    }
    // lane holds D[k = 16 kb + 4 g + m][row = 16 t + c] of its wave's rows
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int k = 16 * kb + 4 * g + m, ri = row0 + 16 * t + c;
                if (k < R && ri < a.n) a.rpart[((int64_t)split * a.n + ri) * R + k] = (double)d[t][kb][m];
            }
}

// ===========================================================================
// Column pass, KL (brunet, nsNMF).  Work-group = TJK = 64 columns x all rows,
// lane owns one column (its H column in VGPRs), W rows are wave-uniform.
// divergence_update_h_inner (lib.rs:74-115):
//   q_u = V[u,j] / (W H)[u,j] (0 when WH == 0);  H'[k,j] = H[k,j] (W^T q)_k / colsum(W)_k
// nsNMF streams Wx = W S in place of W (lib.rs:1805-1812) and also writes
// S H' to Hp, whose row sums the tile records carry for the W update
// (lib.rs:1813-1814).
// ===========================================================================
// colsum(W) over rows r0 .. n - 1 (lib.rs:1451-1458) by the NT threads of a
// work-group: thread t sums column t % R over rows r0 + t / R, + G, ...; the
// G partials are then added in order into csL.  Ends with a barrier.
template <int R, int NT>
__device__ void kl_colsum(const double *W, int n, double *csP, double *csL, int r0 = 0) {
    constexpr int G = NT / R;
    const int t = threadIdx.x;
    double s = 0.0;
    if (t < G * R) {
        // 8 independent loads in flight per step, added in row order
        const int k = t % R;
        int i = r0 + t / R;
        for (; i + 7 * G < n; i += 8 * G) {
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = W[(int64_t)(i + u * G) * R + k];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += x[u];
        }
        for (; i < n; i += G) s += W[(int64_t)i * R + k];
    }
    csP[t] = s;
    __syncthreads();
    if (t < R) {
        double c = 0.0;
        for (int g = 0; g < G; ++g) c += csP[g * R + t];
        csL[t] = c;
    }
    __syncthreads();
}

// The column's current H with brunet's pending every-10th-iteration clamp
// (lib.rs:1572-1575).
template <int R>
__device__ __forceinline__ void kl_load_h(const PassArgs &a, int j, double (&h)[R]) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const double hv = a.H[(int64_t)k * a.ldv + j];
        h[k] = (a.clamp_h && k < a.r) ? (hv < a.eps ? a.eps : hv) : hv;
    }
}

// H' = H (W^T q) / colsum(W) for the tile's columns (wave 0, lane = column),
// nsNMF's S H', and the tile record; called by every thread.
template <int R>
__device__ void kl_col_finish(const PassArgs &a, int tile, const double (&h)[R],
                              const double (&acc)[R], const double *csL, double *hnL) {
    const int lane = threadIdx.x & 63;
    const int j = tile * TJK + lane;
    const int r = a.r;
    const int64_t ldv = a.ldv;
    if (wave_id() == 0) {
        const bool valid = j < a.p;
        double hn[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            hn[k] = k < r ? h[k] * acc[k] / csL[k] : 0.0;
            if (valid) {
                a.H[(int64_t)k * ldv + j] = hn[k];
                if (a.Hf) a.Hf[(int64_t)k * ldv + j] = (float)hn[k];  // FP32 form's row pass
            }
        }
        if (a.kind == K_NSNMF) {
            // S H' column (S = (1 - theta) I + (theta / r) J, lib.rs:457-466),
            // forward over l
            const double so = a.theta / (double)r, sd = (1.0 - a.theta) + so;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                double s = 0.0;
#pragma unroll
                for (int l = 0; l < R; ++l)
                    if (l < r) s += (l == k ? sd : so) * hn[l];
                const double sh = k < r ? s : 0.0;
                if (valid) a.Hp[(int64_t)k * ldv + j] = sh;
                hnL[lane * R + k] = valid ? sh : 0.0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < R; ++k) hnL[lane * R + k] = valid ? hn[k] : 0.0;
        }
    }
    __syncthreads();
    write_record<R, TJK>(hnL, a.red + (int64_t)tile * nred(R));
}

// Work-group = TJK columns x one row range (all rows unless the pass is
// split: blockIdx.x = tile + ntiles * split, and each split writes its
// partial W^T q to kpart for k_colfin_kl).  OBJ (brunet with an
// objective-based stop): the pass also sums the KL terms of the factors it
// reads (those of the previous iteration, pending clamp included) from the
// W H and V / W H it forms anyway, one partial per work-group to opart.
// T = float (FP32 form, brunet): W H, the ratio and W^T q in FP32 over float
// V and W's float copy; the H update (kl_col_finish) stays FP64.
// V rows in flight per lane in the KL column pass, by template rank
// (2000 x 7750 brunet column pass against one row: rank 5 35.9 -> 29.7 us,
// rank 10 38.8 -> 37.0 us with 4 rows; rank 16 49.4 us with 4, 45.1 with 2;
// 1500 x 3000 rank 10 27.1 -> 23.0 us).
template <int R>
constexpr int kl_rows_in_flight() { return R <= 10 ? 4 : 2; }
template <int R, int VP, int NWK = (R <= 6 ? 16 : 8), bool OBJ = false, typename T = double>
__global__ __launch_bounds__(NWK * 64) void k_col_kl(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    constexpr int KC = R < 8 ? R : 8;  // sums combined per LDS round
    __shared__ double slot[NWK / 2][KC][64];
    __shared__ double csP[NWK * 64];
    __shared__ double csL[R];
    __shared__ double hnL[TJK * R];
    const int tile = blockIdx.x % a.ntiles, split = blockIdx.x / a.ntiles;
    const bool part = a.kl_split > 1;
    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int64_t ldv = a.ldv;
    const int j = tile * TJK + lane;
    const void *__restrict__ V = a.V;
    const int64_t vplane = a.vplane;
    const double *__restrict__ W = a.kind == K_NSNMF ? a.Wx : a.W;

    const int rps = (a.n + a.kl_split - 1) / a.kl_split;
    const int r0 = min(a.n, split * rps), r1 = min(a.n, r0 + rps);
    if (!part) kl_colsum<R, NWK * 64>(W, a.n, csP, csL);
    double h[R];
    kl_load_h<R>(a, j, h);
    const int per = (r1 - r0 + NWK - 1) / NWK;
    const int rb = min(r1, r0 + wave * per);
    const int re = min(r1, rb + per);
    double acc[R];
    double fo = 0.0;  // OBJ: this lane's KL terms
    if constexpr (std::is_same<T, float>::value) {
        // FP32 form: float V, W's float copy, FP32 sums (FP64 sums of the
        // same terms, tried, cost 25 % of the rank sweep's time for a 4x
        // smaller worst-case objective difference)
        const float *__restrict__ Vf = static_cast<const float *>(V);
        float hf[R], accf[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            hf[k] = (float)h[k];
            accf[k] = 0.0f;
        }
        auto rowf = [&](int i, float v) {
            const float *wr = a.Wf + (int64_t)i * R;
            float wh = 0.0f;
#pragma unroll
            for (int k = 0; k < R; ++k) wh += wr[k] * hf[k];
            const float qv = wh != 0.0f ? v / wh : 0.0f;
#pragma unroll
            for (int k = 0; k < R; ++k) accf[k] += wr[k] * qv;
        };
        constexpr int KLU = kl_rows_in_flight<R>();
        int i = rb;
        for (; i + KLU <= re; i += KLU) {
            float vu[KLU];
#pragma unroll
            for (int u = 0; u < KLU; ++u) vu[u] = Vf[(int64_t)(i + u) * ldv + j];
#pragma unroll
            for (int u = 0; u < KLU; ++u) rowf(i + u, vu[u]);
        }
        for (; i < re; ++i) rowf(i, Vf[(int64_t)i * ldv + j]);
#pragma unroll
        for (int k = 0; k < R; ++k) acc[k] = (double)accf[k];
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) acc[k] = 0.0;
        auto row = [&](int i, double v) {
            const double *wr = W + (int64_t)i * R;
            double wh = 0.0;
#pragma unroll
            for (int k = 0; k < R; ++k) wh += wr[k] * h[k];
            const double qv = wh != 0.0 ? v / wh : 0.0;
            if constexpr (OBJ) {
                // the objective's term (lib.rs:1406-1445): qv is v / wh when wh > 0
                const double t = v > 0.0 ? (wh > 0.0 ? v * log(qv) - v + wh : 0.0) : wh;
                if (j < a.p) fo += t;
            }
#pragma unroll
            for (int k = 0; k < R; ++k) acc[k] += wr[k] * qv;
        };
        // KLU rows' V loads issued together, then the rows in order (the
        // sums keep the row order of the one-row loop)
        constexpr int KLU = kl_rows_in_flight<R>();
        int i = rb;
        for (; i + KLU <= re; i += KLU) {
            double vu[KLU];
#pragma unroll
            for (int u = 0; u < KLU; ++u) vu[u] = VElem<VP>::one(V, vplane, (int64_t)(i + u) * ldv + j);
#pragma unroll
            for (int u = 0; u < KLU; ++u) row(i + u, vu[u]);
        }
        for (; i < re; ++i) row(i, VElem<VP>::one(V, vplane, (int64_t)i * ldv + j));
    }
    if constexpr (OBJ) {
        // fixed tree over the work-group (csP is free here)
        csP[threadIdx.x] = fo;
        __syncthreads();
        for (int st = NWK * 32; st >= 1; st >>= 1) {
            if ((int)threadIdx.x < st) csP[threadIdx.x] += csP[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) a.opart[blockIdx.x] = csP[0];
        __syncthreads();
    }
    // Waves combine in a fixed tree, KC sums per LDS round.
#pragma unroll
    for (int k0 = 0; k0 < R; k0 += KC) {
#pragma unroll
        for (int half = NWK / 2; half >= 1; half >>= 1) {
            if (wave >= half && wave < 2 * half)
#pragma unroll
                for (int kk = 0; kk < KC; ++kk)
                    if (k0 + kk < R) slot[wave - half][kk][lane] = acc[k0 + kk];
            __syncthreads();
            if (wave < half)
#pragma unroll
                for (int kk = 0; kk < KC; ++kk)
                    if (k0 + kk < R) acc[k0 + kk] += slot[wave][kk][lane];
            __syncthreads();
        }
    }
    if (part) {
        if (wave == 0)
#pragma unroll
            for (int k = 0; k < R; ++k)
                a.kpart[(((int64_t)split * a.ntiles + tile) * R + k) * 64 + lane] = acc[k];
        if (tile == 0) {
            // the split's rows' column sums, for k_colfin_kl (after the
            // partials; here, past the streaming loop, they add no registers)
            kl_colsum<R, NWK * 64>(W, r1, csP, csL, r0);
            if ((int)threadIdx.x < R)
                a.kpart[(int64_t)a.kl_split * a.ntiles * R * 64 + split * R + threadIdx.x] =
                    csL[threadIdx.x];
        }
        return;
    }
    kl_col_finish<R>(a, tile, h, acc, csL, hnL);
}

// Split KL column pass, second kernel: one work-group per tile sums the
// splits' partials (W^T q, and colsum(W) from each split's tile-0
// work-group) in split order and runs the H epilogue.
template <int R>
__global__ __launch_bounds__(256) void k_colfin_kl(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    __shared__ double csL[R];
    __shared__ double hnL[TJK * R];
    const int tile = blockIdx.x;
    const int lane = threadIdx.x & 63;
    double h[R], acc[R];
#pragma unroll
    for (int k = 0; k < R; ++k) acc[k] = 0.0;
    if (wave_id() == 0) {
        kl_load_h<R>(a, tile * TJK + lane, h);
        for (int sp = 0; sp < a.kl_split; ++sp)
#pragma unroll
            for (int k = 0; k < R; ++k)
                acc[k] += a.kpart[(((int64_t)sp * a.ntiles + tile) * R + k) * 64 + lane];
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) h[k] = 0.0;
        // colsum(W), the splits' partial column sums in split order, by
        // wave 1 while wave 0 loads its partials
        if (wave_id() == 1 && lane < R) {
            const double *cs = a.kpart + (int64_t)a.kl_split * a.ntiles * R * 64;
            double c = 0.0;
            for (int sp = 0; sp < a.kl_split; ++sp) c += cs[sp * R + lane];
            csL[lane] = c;
        }
    }
    __syncthreads();
    kl_col_finish<R>(a, tile, h, acc, csL, hnL);
}

// ===========================================================================
// lsNMF (weighted Frobenius MU, lib.rs:946-1043).  Per entry (i, j) with
// weight s = S[i,j]: est = (W H)[i,j], vw = V[i,j] s, ew = est s;
//   column pass  num_H[k,j] += W[i,k] vw,  den_H[k,j] += W[i,k] ew
//                H' = max(H num, eps) / (den + eps)
//   row pass     num_W[i,k] += vw H'[k,j], den_W[i,k] += ew H'[k,j]
//                W' = max(W num, eps) / (den + eps)
// The weight is stored as FP64 (WF_F64) or, when every entry is 0 or 1, as
// one byte per entry (WF_U8).  Same tiling as the KL passes.
// ===========================================================================
template <int WF>
__device__ __forceinline__ double weight_at(const void *S, int64_t i) {
    if constexpr (WF == WF_U8)
        return static_cast<const uint8_t *>(S)[i] ? 1.0 : 0.0;
    else
        return static_cast<const double *>(S)[i];
}

// H' epilogue of one column (wave 0, lane = column): lib.rs:999-1006.
template <int R>
__device__ __forceinline__ void ls_h_finish(const PassArgs &a, int j, const double (&h)[R],
                                            const double (&num)[R], const double (&den)[R]) {
    if (j >= a.p) return;
    const double eps = a.eps;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        double x = 0.0;
        if (k < a.r) {
            const double t = h[k] * num[k];
            x = (t > eps ? t : eps) / (den[k] + eps);
        }
        a.H[(int64_t)k * a.ldv + j] = x;
    }
}

template <int R, int VP, int WF, int NWK = (R <= 6 ? 16 : 8)>
__global__ __launch_bounds__(NWK * 64) void k_col_ls(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    constexpr int KC = R < 8 ? R : 8;
    __shared__ double slot[NWK / 2][2 * KC][64];
    const int tile = blockIdx.x % a.ntiles, split = blockIdx.x / a.ntiles;
    const bool part = a.kl_split > 1;
    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int64_t ldv = a.ldv;
    const int j = tile * TJK + lane;
    const void *__restrict__ V = a.V;
    const void *__restrict__ S = a.wgt;
    const int64_t vplane = a.vplane;
    const double *__restrict__ W = a.W;
    double h[R];
#pragma unroll
    for (int k = 0; k < R; ++k) h[k] = a.H[(int64_t)k * ldv + j];
    const int rps = (a.n + a.kl_split - 1) / a.kl_split;
    const int r0 = min(a.n, split * rps), r1 = min(a.n, r0 + rps);
    const int per = (r1 - r0 + NWK - 1) / NWK;
    const int rb = min(r1, r0 + wave * per);
    const int re = min(r1, rb + per);
    double num[R], den[R];
#pragma unroll
    for (int k = 0; k < R; ++k) num[k] = den[k] = 0.0;
    for (int i = rb; i < re; ++i) {
        const int64_t e = (int64_t)i * ldv + j;
        const double s = weight_at<WF>(S, e);
        const double v = VElem<VP>::one(V, vplane, e);
        const double *wr = W + (int64_t)i * R;
        double est = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) est += wr[k] * h[k];
        const double vw = v * s, ew = est * s;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            num[k] += wr[k] * vw;
            den[k] += wr[k] * ew;
        }
    }
    // waves combine in a fixed tree, KC pairs per LDS round
#pragma unroll
    for (int k0 = 0; k0 < R; k0 += KC) {
#pragma unroll
        for (int half = NWK / 2; half >= 1; half >>= 1) {
            if (wave >= half && wave < 2 * half)
#pragma unroll
                for (int kk = 0; kk < KC; ++kk)
                    if (k0 + kk < R) {
                        slot[wave - half][2 * kk][lane] = num[k0 + kk];
                        slot[wave - half][2 * kk + 1][lane] = den[k0 + kk];
                    }
            __syncthreads();
            if (wave < half)
#pragma unroll
                for (int kk = 0; kk < KC; ++kk)
                    if (k0 + kk < R) {
                        num[k0 + kk] += slot[wave][2 * kk][lane];
                        den[k0 + kk] += slot[wave][2 * kk + 1][lane];
                    }
            __syncthreads();
        }
    }
    if (wave != 0) return;
    if (part) {
        double *kp = a.kpart + (((int64_t)split * a.ntiles + tile) * 2 * R) * 64 + lane;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            kp[(int64_t)k * 64] = num[k];
            kp[(int64_t)(R + k) * 64] = den[k];
        }
        return;
    }
    ls_h_finish<R>(a, j, h, num, den);
}

// Split lsNMF column pass, second kernel: one wave per tile sums the
// splits' partials in split order and runs the H epilogue.
template <int R>
__global__ __launch_bounds__(64) void k_colfin_ls(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    const int tile = blockIdx.x, lane = threadIdx.x;
    const int j = tile * TJK + lane;
    double h[R], num[R], den[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        h[k] = a.H[(int64_t)k * a.ldv + j];
        num[k] = den[k] = 0.0;
    }
    for (int sp = 0; sp < a.kl_split; ++sp) {
        const double *kp = a.kpart + (((int64_t)sp * a.ntiles + tile) * 2 * R) * 64 + lane;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            num[k] += kp[(int64_t)k * 64];
            den[k] += kp[(int64_t)(R + k) * 64];
        }
    }
    ls_h_finish<R>(a, j, h, num, den);
}

// W' epilogue of rows row0 .. + nrows from numerators / denominators in
// LDS (nd[ri * 2R + k], nd[ri * 2R + R + k]): lib.rs:1017-1024.
template <int R>
__device__ void ls_w_finish(const PassArgs &a, int row0, int nrows, const double *nd) {
    const double eps = a.eps;
    for (int e = threadIdx.x; e < nrows * R; e += blockDim.x) {
        const int ri = e / R, k = e % R, i = row0 + ri;
        if (i >= a.n) continue;
        double x = 0.0;
        if (k < a.r) {
            const double t = a.W[(int64_t)i * R + k] * nd[ri * 2 * R + k];
            x = (t > eps ? t : eps) / (nd[ri * 2 * R + R + k] + eps);
        }
        a.W[(int64_t)i * R + k] = x;
    }
}

// Work-group = ls_rows(R) rows x a column range (all columns unless split:
// blockIdx.x = block + nrowblk * split, partials to rpart, k_rowfin_ls
// finishes); NWR waves split the columns, one column per lane and step.
template <int R, int VP, int WF>
__global__ __launch_bounds__(NWR * 64) void k_row_ls(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    constexpr int RW = ls_rows(R);
    constexpr int NE = 2 * RW * R;
    constexpr int NE64 = (NE + 63) / 64;
    __shared__ double wred[NWR][NE64 * 64];
    __shared__ double nd[NE64 * 64];
    const bool part = a.nsplit > 1;
    const int blk = blockIdx.x % a.nrowblk, split = blockIdx.x / a.nrowblk;
    const int row0 = blk * RW;
    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int64_t ldv = a.ldv;
    const int nall = (int)(ldv / 64);
    const int sps = (nall + a.nsplit - 1) / a.nsplit;
    const int s0 = min(nall, split * sps), s1 = min(nall, s0 + sps);
    const int per = (s1 - s0 + NWR - 1) / NWR;
    const int cb = min(s1, s0 + wave * per) * 64, ce = min(s1, s0 + wave * per + per) * 64;
    const void *__restrict__ V = a.V;
    const void *__restrict__ S = a.wgt;
    const int64_t vplane = a.vplane;
    const double *__restrict__ H = a.H;
    int rows[RW];
#pragma unroll
    for (int ri = 0; ri < RW; ++ri) rows[ri] = min(row0 + ri, a.n - 1);
    double num[RW][R], den[RW][R];
#pragma unroll
    for (int ri = 0; ri < RW; ++ri)
#pragma unroll
        for (int k = 0; k < R; ++k) num[ri][k] = den[ri][k] = 0.0;
    for (int c0 = cb; c0 < ce; c0 += 64) {
        const int j = c0 + lane;
        double hv[R], vv[RW], sv[RW];
#pragma unroll
        for (int k = 0; k < R; ++k) hv[k] = H[(int64_t)k * ldv + j];
#pragma unroll
        for (int ri = 0; ri < RW; ++ri) {
            const int64_t e = (int64_t)rows[ri] * ldv + j;
            vv[ri] = VElem<VP>::one(V, vplane, e);
            sv[ri] = weight_at<WF>(S, e);
        }
#pragma unroll
        for (int ri = 0; ri < RW; ++ri) {
            const double *wr = a.W + (int64_t)rows[ri] * R;
            double est = 0.0;
#pragma unroll
            for (int k = 0; k < R; ++k) est += wr[k] * hv[k];
            const double vw = vv[ri] * sv[ri], ew = est * sv[ri];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                num[ri][k] += vw * hv[k];
                den[ri][k] += ew * hv[k];
            }
        }
    }
    // lane -> wave (reduce-scatter) -> work-group (waves in order)
#pragma unroll
    for (int rr = 0; rr < NE64; ++rr) {
        double x[64];
#pragma unroll
        for (int m = 0; m < 64; ++m) {
            const int e = rr * 64 + m;
            const int ri = e / (2 * R), q = e % (2 * R);
            x[m] = e < NE ? (q < R ? num[ri][q] : den[ri][q - R]) : 0.0;
        }
        wred[wave][rr * 64 + lane] = wave_reduce_scatter64(x);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NE; e += blockDim.x) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NWR; ++w) s += wred[w][e];
        nd[e] = s;
        if (part && row0 + e / (2 * R) < a.n)
            a.rpart[((int64_t)split * a.n + row0 + e / (2 * R)) * 2 * R + e % (2 * R)] = s;
    }
    if (part) return;
    __syncthreads();
    ls_w_finish<R>(a, row0, RW, nd);
}

// Split lsNMF row pass, second kernel: RF rows per work-group.
template <int R>
__global__ __launch_bounds__(256) void k_rowfin_ls(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    __shared__ double nd[RF * 2 * R];
    const int row0 = blockIdx.x * RF;
    for (int e = threadIdx.x; e < RF * 2 * R; e += blockDim.x) {
        const int i = row0 + e / (2 * R);
        nd[e] = i < a.n ? split_sum(a.rpart + (int64_t)i * 2 * R + e % (2 * R), a.nsplit,
                                    (int64_t)a.n * 2 * R)
                        : 0.0;
    }
    __syncthreads();
    ls_w_finish<R>(a, row0, RF, nd);
}

// ===========================================================================
// Tile records -> H statistics, fixed order.  One wave per output: lane l
// sums tiles l, l+64, l+128, ... then the 64 lane sums combine by a fixed
// shuffle tree.
// ===========================================================================
template <int R>
__global__ __launch_bounds__(256) void k_red(RedArgs args) {
    if (args.done && *args.done) return;
    const RedArgs a = for_solve(args);
    constexpr int NRED = nred(R);
    const int o = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (o >= NRED) return;
    // lane's tiles lane, lane + 64, ... in order, 8 loads in flight per step
    double s = 0.0;
    int t = lane;
    for (; t + 7 * 64 < a.ntiles; t += 8 * 64) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = a.red[(int64_t)(t + 64 * u) * NRED + o];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += x[u];
    }
    for (; t < a.ntiles; t += 64) s += a.red[(int64_t)t * NRED + o];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) a.hst[o] = s;
}

// The same sums for many tiles: one work-group of 4 waves per output; wave w
// sums tiles w, w + 4, ... (lane l: tiles 4 l + w, 4 l + w + 256, ...), then
// the 4 wave totals are added in order.
template <int R>
__global__ __launch_bounds__(256) void k_red_wide(RedArgs args) {
    if (args.done && *args.done) return;
    const RedArgs a = for_solve(args);
    constexpr int NRED = nred(R);
    __shared__ double ws[4];
    const int o = blockIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double s = 0.0;
    int t = 4 * lane + w;
    for (; t + 7 * 256 < a.ntiles; t += 8 * 256) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = a.red[(int64_t)(t + 256 * u) * NRED + o];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += x[u];
    }
    for (; t < a.ntiles; t += 256) s += a.red[(int64_t)t * NRED + o];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) ws[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) a.hst[o] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

// ===========================================================================
// Row pass (VALU; brunet at every rank, the others up to rank 10).
// Work-group = RW rows x all columns; NW waves split the columns (one wave
// when there are enough row blocks to fill the GPU).  Lane owns CPL adjacent columns per step; the R values of H' per
// column are reused across the RW rows.
// ===========================================================================
template <int R>
struct RowCfg {
    static constexpr int RW = rows_per_rowblk(R);
    static constexpr int CPL = R <= 16 ? 2 : 1;
};

// W epilogue of rows row0 .. row0 + nrows - 1 from their numerators
// numL[ri * R + k] (V H'^T or (V ./ WH') H'^T), with HHt = H'H'^T (R x R) and,
// for lee and offset, wL = the rows' logical W, all in LDS; called by every
// thread of the work-group.  Offset overwrites numL with W'.
template <int R, bool KL_ONLY = false, bool GS = true>
__device__ void w_epilogue(const PassArgs &a, int row0, int nrows, double *numL,
                           const double *HHt, const double *wL) {
    const int t = threadIdx.x;
    const int r = a.r;
    const double eps = a.eps;
    if constexpr (KL_ONLY) {
        // brunet: W'[i,k] = W[i,k] B[i,k] / rowsum(H')_k (lib.rs:199), then the
        // every-10th-iteration clamp (lib.rs:1572-1575); one thread per entry
        for (int e = t; e < nrows * R; e += blockDim.x) {
            const int i = row0 + e / R, k = e % R;
            if (i >= a.n) continue;
            double x = k < r ? a.W[(int64_t)i * R + k] * numL[e] / a.hst[hs_rs(R) + k] : 0.0;
            if (a.clamp_w && k < r && x < eps) x = eps;
            a.W[(int64_t)i * R + k] = x;
            if (a.Wf) a.Wf[(int64_t)i * R + k] = (float)x;  // FP32 form's next column pass
        }
        return;
    }
    if (a.kind == K_LEE || a.kind == K_PROD || a.kind == K_OFFSET) {
        // one thread per (row, k): euclidean_update_w_impl (lib.rs:305-320)
        // on the logical W; offset adds off[i] rowsum(H')_k
        // (offset_update_w_impl, lib.rs:437-438)
        const bool offs = a.kind == K_OFFSET;
        for (int e = t; e < nrows * R; e += blockDim.x) {
            const int i = row0 + e / R, k = e % R;
            if (i >= a.n) continue;
            double wn = 0.0;
            if (k < r) {
                if (a.kind == K_PROD) {
                    wn = numL[e];
                } else {
                    const double *w = &wL[(e / R) * R];
                    double den = 0.0;
#pragma unroll
                    for (int l = R - 1; l >= 0; --l) den += w[l] * HHt[l * R + k];
                    if (offs) den += a.off[i] * a.hst[hs_rs(R) + k];
                    const double tt = w[k] * numL[e];
                    wn = (tt < eps ? eps : tt) / (den + eps);
                }
            }
            a.W[(int64_t)i * R + k] = wn;
            if (offs) numL[e] = wn;
        }
        if (offs) {
            // offset update (lib.rs:1729-1741): rowsum(W' H' + off)_i =
            // sum_k W'[i,k] rowsum(H')_k + p off[i]
            __syncthreads();
            for (int ri = t; ri < nrows; ri += blockDim.x) {
                const int i = row0 + ri;
                if (i >= a.n) continue;
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < R; ++k)
                    if (k < r) s += numL[ri * R + k] * a.hst[hs_rs(R) + k];
                const double o = a.off[i];
                const double rw = s + (double)a.p * o;
                const double rv = a.rsv[i];
                const double num = rv > eps ? rv : eps;
                a.off[i] = o * num / (rw + eps);
            }
        }
        return;
    }
    if constexpr (!GS) return;  // instances without the Gauss-Seidel registers
    // one thread per row: the Gauss-Seidel sweeps are sequential in k
    for (int ri = t; ri < nrows; ri += blockDim.x) {
        const int i = row0 + ri;
        if (i >= a.n) continue;
        double num[R], wn[R];
#pragma unroll
        for (int k = 0; k < R; ++k) num[k] = numL[ri * R + k];
        if (a.kind == K_BRUNET) {
            // W'[i,k] = W[i,k] B[i,k] / rowsum(H')_k (lib.rs:199), then the
            // every-10th-iteration clamp (lib.rs:1572-1575)
#pragma unroll
            for (int k = 0; k < R; ++k) {
                double x = k < r ? a.W[(int64_t)i * R + k] * num[k] / a.hst[hs_rs(R) + k] : 0.0;
                if (a.clamp_w && k < r && x < eps) x = eps;
                wn[k] = x;
            }
        } else {
            double w[R];
            const double *ws = a.kind == K_EHALS ? a.Wx : a.W;
#pragma unroll
            for (int k = 0; k < R; ++k) w[k] = ws[(int64_t)i * R + k] / a.wsx[k];
            gs_sweep<R>(w, num, HHt, R, r, a.kind == K_DNMF, a.lambda, a.gamma, eps);
#pragma unroll
            for (int k = 0; k < R; ++k) wn[k] = k < r ? w[k] : 0.0;
        }
        if (a.kind == K_EHALS) {
            // next extrapolation point max(eps, W_t+1 + beta (W_t+1 - W_t)) (lib.rs:755-759)
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const double wc = a.W[(int64_t)i * R + k];
                const double e = wn[k] + BETA_EHALS * (wn[k] - wc);
                a.Wx[(int64_t)i * R + k] = k < r ? (e > eps ? e : eps) : 0.0;
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) a.W[(int64_t)i * R + k] = wn[k];
    }
}

// H'H'^T (from hst) and, for lee, the logical W of rows row0 .. + nrows into LDS.
template <int R>
__device__ void w_epilogue_stage(const PassArgs &a, int row0, int nrows, double *HHt, double *wL) {
    for (int o = threadIdx.x; o < R * R; o += blockDim.x)
        HHt[o] = a.hst[hs_K(R) + tri_index(R, o / R, o % R)];
    if (a.kind == K_LEE || a.kind == K_OFFSET)
        for (int e = threadIdx.x; e < nrows * R; e += blockDim.x) {
            const int i = row0 + e / R;
            wL[e] = i < a.n ? a.W[(int64_t)i * R + e % R] / a.wsx[e % R] : 0.0;
        }
}

// T = float (FP32 form, brunet): the KL branch forms W H', the ratio and its
// products with H' in FP32 over float V and the float copies of W and H'.
template <int R, bool KL, int VP, int NW = NWR, bool GS = true, int RW = RowCfg<R>::RW,
          int CPL = RowCfg<R>::CPL, typename T = double>
__global__ __launch_bounds__(NW * 64) void k_row(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    constexpr int NE = RW * R;
    constexpr int NE64 = (NE + 63) / 64;
    constexpr int STEP = 64 * CPL;
    __shared__ double wred[NW][NE64 * 64];
    __shared__ double numL[NE64 * 64];
    __shared__ double HHt[R * R];  // H H^T, full (hst holds the upper triangle)
    __shared__ double wL[RW * R];  // lee: the block's logical W rows
    // KL row pass with column splits (nsplit > 1): blockIdx.x = block +
    // nrowblk * split; each split writes partial numerators to rpart and
    // k_rowfin finishes.
    const bool part = KL && a.nsplit > 1;
    const int blk = part ? blockIdx.x % a.nrowblk : blockIdx.x;
    const int split = part ? blockIdx.x / a.nrowblk : 0;
    const int row0 = blk * RW;
    if (!part) w_epilogue_stage<R>(a, row0, RW, HHt, wL);

    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int64_t ldv = a.ldv;
    const int nall = (int)(ldv / STEP);
    const int sps = part ? (nall + a.nsplit - 1) / a.nsplit : nall;
    const int s0 = min(nall, split * sps), s1 = min(nall, s0 + sps);
    const int per = (s1 - s0 + NW - 1) / NW;
    const int cb = min(s1, s0 + wave * per) * STEP;
    const int ce = min(s1, s0 + wave * per + per) * STEP;
    const void *__restrict__ V = a.V;
    const int64_t vplane = a.vplane;
    const double *__restrict__ H = a.H;

    int rows[RW];
#pragma unroll
    for (int ri = 0; ri < RW; ++ri) rows[ri] = min(row0 + ri, a.n - 1);
    T acc[RW][R];
#pragma unroll
    for (int ri = 0; ri < RW; ++ri)
#pragma unroll
        for (int k = 0; k < R; ++k) acc[ri][k] = T(0);

    if constexpr (KL && std::is_same<T, float>::value) {
        // FP32 form: float V, H' and W (their float copies)
        const float *__restrict__ Vf = static_cast<const float *>(V);
        for (int cc0 = cb; cc0 < ce; cc0 += STEP) {
            const int j = cc0 + lane * CPL;
            float hv[R][CPL], vv[RW][CPL];
#pragma unroll
            for (int k = 0; k < R; ++k)
#pragma unroll
                for (int cc = 0; cc < CPL; ++cc) hv[k][cc] = a.Hf[(int64_t)k * ldv + j + cc];
#pragma unroll
            for (int ri = 0; ri < RW; ++ri)
#pragma unroll
                for (int cc = 0; cc < CPL; ++cc) vv[ri][cc] = Vf[(int64_t)rows[ri] * ldv + j + cc];
#pragma unroll
            for (int ri = 0; ri < RW; ++ri) {
                const float *wr = a.Wf + (int64_t)rows[ri] * R;
#pragma unroll
                for (int cc = 0; cc < CPL; ++cc) {
                    float wh = 0.0f;
#pragma unroll
                    for (int k = 0; k < R; ++k) wh += wr[k] * hv[k][cc];
                    const float qv = wh != 0.0f ? vv[ri][cc] / wh : 0.0f;
#pragma unroll
                    for (int k = 0; k < R; ++k) acc[ri][k] += hv[k][cc] * qv;
                }
            }
        }
    } else if constexpr (KL) {
        // divergence_update_w_impl (lib.rs:187-200): q = V[i,u] / (W H')[i,u],
        // B[i,k] = sum_u H'[k,u] q_u
        for (int cc0 = cb; cc0 < ce; cc0 += STEP) {
            const int j = cc0 + lane * CPL;
            double hv[R][CPL], vv[RW][CPL];
#pragma unroll
            for (int k = 0; k < R; ++k)
#pragma unroll
                for (int cc = 0; cc < CPL; ++cc) hv[k][cc] = H[(int64_t)k * ldv + j + cc];
#pragma unroll
            for (int ri = 0; ri < RW; ++ri)
#pragma unroll
                for (int cc = 0; cc < CPL; ++cc)
                    vv[ri][cc] = VElem<VP>::one(V, vplane, (int64_t)rows[ri] * ldv + j + cc);
#pragma unroll
            for (int ri = 0; ri < RW; ++ri) {
                const double *wr = a.W + (int64_t)rows[ri] * R;
#pragma unroll
                for (int cc = 0; cc < CPL; ++cc) {
                    const double v = vv[ri][cc];
                    double wh = 0.0;
#pragma unroll
                    for (int k = 0; k < R; ++k) wh += wr[k] * hv[k][cc];
                    const double qv = wh != 0.0 ? v / wh : 0.0;
#pragma unroll
                    for (int k = 0; k < R; ++k) acc[ri][k] += hv[k][cc] * qv;
                }
            }
        }
    } else {
        for (int cc0 = cb; cc0 < ce; cc0 += STEP) {
            const int j = cc0 + lane * CPL;
            double hv[R][CPL], vv[RW][CPL];
            if constexpr (CPL == 2) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const double2 x = *(const double2 *)(H + (int64_t)k * ldv + j);
                    hv[k][0] = x.x;
                    hv[k][1] = x.y;
                }
#pragma unroll
                for (int ri = 0; ri < RW; ++ri) {
                    const double2 x = VElem<VP>::two(V, vplane, (int64_t)rows[ri] * ldv + j);
                    vv[ri][0] = x.x;
                    vv[ri][1] = x.y;
                }
            } else {
#pragma unroll
                for (int k = 0; k < R; ++k) hv[k][0] = H[(int64_t)k * ldv + j];
#pragma unroll
                for (int ri = 0; ri < RW; ++ri)
                    vv[ri][0] = VElem<VP>::one(V, vplane, (int64_t)rows[ri] * ldv + j);
            }
#pragma unroll
            for (int ri = 0; ri < RW; ++ri)
#pragma unroll
                for (int k = 0; k < R; ++k)
#pragma unroll
                    for (int cc = 0; cc < CPL; ++cc) acc[ri][k] += vv[ri][cc] * hv[k][cc];
        }
    }

    // Lane -> wave (reduce-scatter, 64 entries per round) -> work-group
    // (waves summed in order).
#pragma unroll
    for (int rr = 0; rr < NE64; ++rr) {
        double x[64];
#pragma unroll
        for (int m = 0; m < 64; ++m) {
            const int e = rr * 64 + m;
            x[m] = e < NE ? (double)acc[e / R][e % R] : 0.0;
        }
        wred[wave][rr * 64 + lane] = wave_reduce_scatter64(x);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NE; e += blockDim.x) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += wred[w][e];
        numL[e] = s;
        if (part && row0 + e / R < a.n)
            a.rpart[((int64_t)split * a.n + row0 + e / R) * R + e % R] = s;
    }
    if (part) return;
    __syncthreads();
    w_epilogue<R, KL, GS>(a, row0, RW, numL, HHt, wL);
}

// ===========================================================================
// Row pass with MFMA (ranks above 10, not KL): partial numerators V H'^T of
// a block of 16 RT rows over one column split.  Work-group = NW waves taking
// 8-column steps in turn over the split's columns; every wave covers all RT
// 16-row tiles of the block.  Per step lane l loads, for each row tile t,
// V[row0 + 16 t + l%16][c0 + 2(l/16) .. +1] (one dwordx4), and once for all
// tiles the same two columns of H'[16 kb + l%16]; two MFMAs per (tile,
// k-block) accumulate D[k][row] += H'[k][c] V[row][c] (even and odd columns
// in separate accumulators).  Each H' value loaded thus feeds RT row tiles:
// H' is re-read once per 16 RT rows (at rank 20 H' has twice V's bytes per
// 16-row tile, so RT = 1 would move more H' than V).  Each wave issues the
// loads of U steps before their MFMAs (the bytes in flight per CU set the
// streaming rate at one to two waves per SIMD).
// Grid: nrb x nsplit work-groups (logical index xcd_local_index() = rb +
// nrb * split).
// ===========================================================================
// Column steps in flight per wave of the 4-tile instance, by storage form.
constexpr int row4_steps(int vp) { return vp == VP_B16X2 ? 2 : 4; }

template <int R, int VP, int RT, int NW, int U = 2>
__global__ __launch_bounds__(NW * 64) void k_row_mfma(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    constexpr int KB = (R + 15) / 16;
    __shared__ double red[NW][RT * KB * 4][64];
    const int xl = xcd_local_index();
    const int rb = xl % a.nrb, split = xl / a.nrb;
    const int lane = threadIdx.x & 63;
    const int wave = wave_id();
    const int q = lane >> 4, c = lane & 15;
    const int64_t ldv = a.ldv;
    const int row0 = rb * RB_MFMA * RT;
    int64_t vrow[RT];
    bool rok[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        const int row = row0 + 16 * t + c;
        rok[t] = row < a.n;
        vrow[t] = (int64_t)(rok[t] ? row : a.n - 1) * ldv;
    }
    const int nsteps = (int)(ldv / 8);
    const int per = (nsteps + a.nsplit - 1) / a.nsplit;
    const int s0 = min(nsteps, split * per), s1 = min(nsteps, s0 + per);
    const void *__restrict__ V = a.V;
    const int64_t vplane = a.vplane;
    const double *__restrict__ H = a.H;
    int64_t hrow[KB];
    bool kok[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        const int k = 16 * kb + c;
        kok[kb] = k < R;
        hrow[kb] = (int64_t)(k < R ? k : R - 1) * ldv;
    }
    dbl4 d0[RT][KB], d1[RT][KB];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) d0[t][kb] = d1[t][kb] = dbl4{0, 0, 0, 0};

    for (int st = s0 + wave; st < s1; st += NW * U) {
        double2 v[U][RT], h[U][KB];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int su = st + u * NW;
            const bool ok = su < s1;
            const int col = (ok ? su : s0) * 8 + 2 * q;
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                v[u][t] = VElem<VP>::two(V, vplane, vrow[t] + col);
                if (!ok || !rok[t]) v[u][t] = make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int kb = 0; kb < KB; ++kb) {
                h[u][kb] = *reinterpret_cast<const double2 *>(H + hrow[kb] + col);
                if (!kok[kb]) h[u][kb] = make_double2(0.0, 0.0);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) {
                    d0[t][kb] = mfma(h[u][kb].x, v[u][t].x, d0[t][kb]);
                    d1[t][kb] = mfma(h[u][kb].y, v[u][t].y, d1[t][kb]);
                }
    }
    // waves summed in order; lane holds D[k = 16 kb + 4 m + q][row = 16 t + c]
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int m = 0; m < 4; ++m)
                red[wave][(t * KB + kb) * 4 + m][lane] = d0[t][kb][m] + d1[t][kb][m];
    __syncthreads();
    for (int e = threadIdx.x; e < RT * KB * 4 * 64; e += blockDim.x) {
        const int slot = e / 64, l = e % 64;
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += red[w][slot][l];
        const int t = slot / (KB * 4), kb = (slot / 4) % KB, m = slot % 4;
        const int k = 16 * kb + 4 * m + (l >> 4), ri = row0 + 16 * t + (l & 15);
        if (k < R && ri < a.n) a.rpart[((int64_t)split * a.n + ri) * R + k] = s;
    }
}

// Sum of the column-split partial numerators (fixed order) and the W
// epilogue, RF rows per work-group (MFMA row pass, or the split KL row pass).

template <int R, bool KL = false, bool GS = true>
__global__ __launch_bounds__(256) void k_rowfin(PassArgs args) {
    if (args.done && *args.done) return;
    const PassArgs a = for_solve(args);
    __shared__ double numL[RF * R];
    __shared__ double HHt[R * R];
    __shared__ double wL[RF * R];
    const int row0 = blockIdx.x * RF;
    w_epilogue_stage<R>(a, row0, RF, HHt, wL);
    for (int e = threadIdx.x; e < RF * R; e += blockDim.x) {
        const int i = row0 + e / R, k = e % R;
        numL[e] = i < a.n ? split_sum(a.rpart + (int64_t)i * R + k, a.nsplit, (int64_t)a.n * R) : 0.0;
    }
    __syncthreads();
    w_epilogue<R, KL, GS>(a, row0, RF, numL, HHt, wL);
}

// ===========================================================================
// Column sums / sums of squares of the stored W.  Phase 0: one partial per
// WSTAT_ROWS rows; phase 1 (one work-group): sum the partials in order and
// derive the divisor to the logical W.
// ===========================================================================
template <int R>
__global__ __launch_bounds__(256) void k_wstats(WStatArgs args, int phase) {
    if (args.done && *args.done) return;
    const WStatArgs a = for_solve(args);
    __shared__ double buf[256][2];
    const int t = threadIdx.x;
    if (phase == 0) {
        const int b = blockIdx.x;
        const int i0 = b * WSTAT_ROWS;
        const int i1 = min(a.n, i0 + WSTAT_ROWS);
        // thread t: column t % R, rows i0 + t / R, i0 + t / R + 256 / R, ...
        const int k = t % R, lanes = (256 / R) * R, stride = 256 / R;
        double s = 0.0, s2 = 0.0;
        if (t < lanes) {
            // 8 independent loads in flight per step, added in row order
            int i = i0 + t / R;
            for (; i + 7 * stride < i1; i += 8 * stride) {
                double x[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) x[u] = a.W[(int64_t)(i + u * stride) * R + k];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    s += x[u];
                    s2 += x[u] * x[u];
                }
            }
            for (; i < i1; i += stride) {
                const double w = a.W[(int64_t)i * R + k];
                s += w;
                s2 += w * w;
            }
        }
        buf[t][0] = s;
        buf[t][1] = s2;
        __syncthreads();
        if (t < R) {
            double c1 = 0.0, c2 = 0.0;
            for (int u = t; u < lanes; u += R) {
                c1 += buf[u][0];
                c2 += buf[u][1];
            }
            a.part[(int64_t)b * 2 * R + t] = c1;
            a.part[(int64_t)b * 2 * R + R + t] = c2;
        }
    } else {
        if (t < R) {
            double c1 = 0.0, c2 = 0.0;
            for (int b = 0; b < a.nparts; ++b) {
                c1 += a.part[(int64_t)b * 2 * R + t];
                c2 += a.part[(int64_t)b * 2 * R + R + t];
            }
            double sc = 1.0;
            if (t < a.r && (!a.first || a.dnmf_norm)) {
                if ((a.kind == K_LEE && a.rescale) || a.kind == K_NSNMF) sc = c1 > 0.0 ? c1 : 1.0;
                if (a.kind == K_DNMF) sc = sqrt(c2) > 0.0 ? sqrt(c2) : 1.0;
            }
            a.out[t] = c1;
            a.out[R + t] = c2;
            a.out[2 * R + t] = sc;
        }
    }
}

// ===========================================================================
// Objective over the logical factors: per work-group partial sums of
// ||V - WH||^2 or of the KL terms (lib.rs:1406-1445).  Work-group = TJK
// columns x one row split (osplit splits keep several work-groups per CU:
// the KL terms are bound by the FP64 division and log, so balance matters),
// NWO waves over the split's rows; k_objfin sums the partials.  One instance
// per objective form (MODE): the per-row branches of a general loop keep
// the compiler from scheduling it (2x slower at 2000 x 7750).
// ===========================================================================
// Column j's logical H (the pending transforms of hmode), divided by the
// logical W's divisors sc (W H = sum_k W[i,k] (h_k / sc_k); exact for the
// forms whose divisors are 1).
template <int R>
__device__ __forceinline__ void obj_column_h(const ObjArgs &a, const double *sc, const double *ss,
                                             int j, bool valid, double (&h)[R]) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
        double hv = valid ? a.H[(int64_t)k * a.ldv + j] : 0.0;
        if (a.hmode == 1) {
            const double nw = sqrt(ss[k]);
            const double nh = sqrt(a.hst[hs_ss(R) + k]);
            hv = (hv / (nh > 0.0 ? nh : 1.0)) * (nw * nh);
        } else if (a.hmode == 2 && a.clamp_h) {
            hv = hv < a.eps ? a.eps : hv;
        }
        h[k] = (k < a.r && valid) ? hv : 0.0;
    }
    if (a.hmode == 3) {
        // nsNMF estimate W S H (lib.rs:1784-1788) as W (S H)
        const double so = a.theta / (double)a.r, sd = (1.0 - a.theta) + so;
        double sh[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            double s = 0.0;
#pragma unroll
            for (int l = 0; l < R; ++l) s += (l == k ? sd : so) * h[l];
            sh[k] = k < a.r ? s : 0.0;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) h[k] = sh[k];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) h[k] *= a.unit_sc ? 1.0 : 1.0 / sc[k];
}

// One element's objective term from v = V[i, j] and wh = (W H)[i, j].
template <int MODE>
__device__ __forceinline__ double obj_term(const ObjArgs &a, double v, double wh, int64_t e) {
    if (MODE == OBJ_KL) return v > 0.0 ? (wh > 0.0 ? v * log(v / wh) - v + wh : 0.0) : wh;
    if (MODE == OBJ_WGT) {
        // lsNMF: ((V - W H) S)^2 (lib.rs:962-973)
        const double sw = a.wf == WF_U8 ? weight_at<WF_U8>(a.wgt, e) : weight_at<WF_F64>(a.wgt, e);
        const double d = (v - wh) * sw;
        return d * d;
    }
    return (v - wh) * (v - wh);
}

// Work-group = TJO columns (2 per lane, loaded together) x one row split;
// NWO waves over the split's rows.
template <int R, int VP, int MODE>
__global__ __launch_bounds__(NWO * 64) void k_obj(ObjArgs a) {
    if (a.done && *a.done) return;
    __shared__ double red[NWO * 64];
    __shared__ double wdiv[2 * R];  // [sum of squares | divisor]
    // blockIdx.x = tile + ntiles * split: the split's rows of a TJO-column tile
    const int tile = blockIdx.x % a.ntiles, split = blockIdx.x / a.ntiles;
    const int wave = wave_id();
    const int64_t ldv = a.ldv;
    const int j = tile * TJO + 2 * (threadIdx.x & 63);
    const bool v0 = j < a.p, v1 = j + 1 < a.p;
    // divisors of the logical W: from k_wstats's output, or reduced here from
    // its phase-0 partials with phase 1's arithmetic and order
    if (a.wpart) {
        if ((int)threadIdx.x < R) {
            const int t = threadIdx.x;
            double c1 = 0.0, c2 = 0.0;
            for (int b = 0; b < a.wnparts; ++b) {
                c1 += a.wpart[(int64_t)b * 2 * R + t];
                c2 += a.wpart[(int64_t)b * 2 * R + R + t];
            }
            double d = 1.0;
            if (t < a.r && !a.wfirst) {
                if ((a.wkind == K_LEE && a.wrescale) || a.wkind == K_NSNMF) d = c1 > 0.0 ? c1 : 1.0;
                if (a.wkind == K_DNMF) d = sqrt(c2) > 0.0 ? sqrt(c2) : 1.0;
            }
            wdiv[t] = c2;
            wdiv[R + t] = d;
        }
        __syncthreads();
    }
    const double *sc = a.wpart ? wdiv + R : a.wsc + 2 * R;
    const double *ss = a.wpart ? wdiv : a.wsc + R;
    double h0[R], h1[R];
    obj_column_h<R>(a, sc, ss, j, v0, h0);
    obj_column_h<R>(a, sc, ss, j + 1, v1, h1);
    const int rps = (a.n + a.osplit - 1) / a.osplit;
    const int r0 = min(a.n, split * rps), r1 = min(a.n, r0 + rps);
    const int per = (r1 - r0 + NWO - 1) / NWO;
    const int rb = min(r1, r0 + wave * per);
    const int re = min(r1, rb + per);
    double s = 0.0;
    for (int i = rb; i < re; ++i) {
        const int64_t e = (int64_t)i * ldv + j;
        const double2 v = VElem<VP>::two(a.V, a.vplane, e);
        const double *wr = a.W + (int64_t)i * R;
        double wh0 = 0.0, wh1 = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            wh0 += wr[k] * h0[k];
            wh1 += wr[k] * h1[k];
        }
        if (MODE == OBJ_FROB_OFF) {
            // offset estimate W H + off (lib.rs:1744-1750)
            wh0 += a.off[i];
            wh1 += a.off[i];
        }
        const double t0 = obj_term<MODE>(a, v.x, wh0, e);
        const double t1 = obj_term<MODE>(a, v.y, wh1, e + 1);
        if (v0) s += t0;
        if (v1) s += t1;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = NWO * 32; st >= 1; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.part[blockIdx.x] = red[0];
}

// The same partial sums with W H formed on the FP64 matrix cores.  Same
// work-groups (TJO columns x one row split, NWO waves).  The logical H of
// the TJO columns (R4 = ceil(R / 4) blocks of 4 ranks, zero beyond r) is
// staged in LDS once; each wave then takes 16-row blocks of the split in
// turn: lane l loads W[i0 + l%16][4 kb + l/16] (the A operand, R4 values,
// kept for the block) and for each 16-column tile t chains R4
// v_mfma_f64_16x16x4 with B = H[4 kb + l/16][16 t + l%16] from LDS; the lane
// then holds (W H)[i0 + 4 m + l/16][16 t + l%16], m = 0..3, and adds those
// elements' terms.  Every element's W H is the same FP64 sum of r products
// as in k_obj, in another order.
template <int R, int VP, int MODE>
__global__ __launch_bounds__(NWO * 64) void k_obj_mfma(ObjArgs a) {
    if (a.done && *a.done) return;
    constexpr int R4 = (R + 3) / 4;
    constexpr int NT = TJO / 16;  // column tiles per work-group
    // tiles per round: fewer live registers for the KL terms' log and above
    // rank 24
    constexpr int TPR = (MODE == OBJ_KL || R > 24) ? 4 : 8;
    __shared__ double hL[R4 * 4][TJO];
    __shared__ double red[NWO * 64];
    __shared__ double wdiv[2 * R];
    const int tile = blockIdx.x % a.ntiles, split = blockIdx.x / a.ntiles;
    const int wave = wave_id();
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    const int64_t ldv = a.ldv;
    const int j0 = tile * TJO;
    if (a.wpart) {
        if ((int)threadIdx.x < R) {
            const int t = threadIdx.x;
            double c1 = 0.0, c2 = 0.0;
            for (int b = 0; b < a.wnparts; ++b) {
                c1 += a.wpart[(int64_t)b * 2 * R + t];
                c2 += a.wpart[(int64_t)b * 2 * R + R + t];
            }
            double d = 1.0;
            if (t < a.r && !a.wfirst) {
                if ((a.wkind == K_LEE && a.wrescale) || a.wkind == K_NSNMF) d = c1 > 0.0 ? c1 : 1.0;
                if (a.wkind == K_DNMF) d = sqrt(c2) > 0.0 ? sqrt(c2) : 1.0;
            }
            wdiv[t] = c2;
            wdiv[R + t] = d;
        }
        __syncthreads();
    }
    const double *sc = a.wpart ? wdiv + R : a.wsc + 2 * R;
    const double *ss = a.wpart ? wdiv : a.wsc + R;
    for (int jj = threadIdx.x; jj < TJO; jj += NWO * 64) {
        double h[R];
        const int j = j0 + jj;
        obj_column_h<R>(a, sc, ss, j, j < a.p, h);
#pragma unroll
        for (int k = 0; k < R4 * 4; ++k) hL[k][jj] = k < R ? h[k] : 0.0;
    }
    __syncthreads();
    const int rps = (a.n + a.osplit - 1) / a.osplit;
    const int r0 = min(a.n, split * rps), r1 = min(a.n, r0 + rps);
    double s = 0.0;
    // A operand of row block i0: row i0 + c of W (rows past n read row
    // n - 1; their terms are dropped below); loaded one block ahead
    auto load_a = [&](int i0, double (&av)[R4]) {
        const double *wr = a.W + (int64_t)min(i0 + c, a.n - 1) * R;
#pragma unroll
        for (int kb = 0; kb < R4; ++kb) av[kb] = 4 * kb + g < R ? wr[4 * kb + g] : 0.0;
    };
    double av[R4];
    load_a(r0 + 16 * wave, av);
    for (int i0 = r0 + 16 * wave; i0 < r1; i0 += 16 * NWO) {
        double avn[R4];
        load_a(i0 + 16 * NWO, avn);
        int rowc[4];
        bool rowok[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int i = i0 + 4 * m + g;
            rowok[m] = i < r1;
            rowc[m] = min(i, a.n - 1);
        }
        // per round of TPR tiles: the V loads first, then TPR independent
        // MFMA chains interleaved (rank block outer), then the terms
#pragma unroll
        for (int t0 = 0; t0 < NT; t0 += TPR) {
            double v[TPR][4];
#pragma unroll
            for (int t = 0; t < TPR; ++t)
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    v[t][m] = VElem<VP>::one(a.V, a.vplane,
                                             (int64_t)rowc[m] * ldv + j0 + 16 * (t0 + t) + c);
            dbl4 acc[TPR];
#pragma unroll
            for (int t = 0; t < TPR; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kb = 0; kb < R4; ++kb)
#pragma unroll
                for (int t = 0; t < TPR; ++t)
                    acc[t] = mfma(av[kb], hL[4 * kb + g][16 * (t0 + t) + c], acc[t]);
#pragma unroll
            for (int t = 0; t < TPR; ++t) {
                const int j = j0 + 16 * (t0 + t) + c;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    double wh = acc[t][m];
                    if (MODE == OBJ_FROB_OFF) wh += a.off[rowc[m]];
                    const double term = obj_term<MODE>(a, v[t][m], wh, (int64_t)rowc[m] * ldv + j);
                    if (rowok[m] && j < a.p) s += term;
                }
            }
        }
#pragma unroll
        for (int kb = 0; kb < R4; ++kb) av[kb] = avn[kb];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = NWO * 32; st >= 1; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.part[blockIdx.x] = red[0];
}

// Apply the pending logical transforms to the stored factors at the end of
// a solve (lee rescale, dnmf normalisation, brunet's pending H clamp).
template <int R>
__global__ __launch_bounds__(256) void k_finalize(Finalize args) {
    const Finalize f = for_solve(args);
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double *sc = f.wsc + 2 * R;
    if (idx < (int64_t)f.n * R) {
        const int k = (int)(idx % R);
        f.W[idx] = f.W[idx] / sc[k];
    }
    if (idx < (int64_t)R * f.ldv) {
        const int k = (int)(idx / f.ldv);
        const int64_t j = idx % f.ldv;
        if (j < f.p && k < f.r) {
            if (f.kind == K_DNMF) {
                const double nw = sqrt(f.wsc[R + k]);
                const double nh = sqrt(f.hst[hs_ss(R) + k]);
                const double hd = f.H[idx] / (nh > 0.0 ? nh : 1.0);
                f.H[idx] = f.dnmf_fold ? hd * (nw * nh) : hd;
            } else if (f.kind == K_BRUNET && f.clamp_h) {
                f.H[idx] = f.H[idx] < f.eps ? f.eps : f.H[idx];
            }
        }
    }
}

template <int R>
__global__ __launch_bounds__(256) void k_prep(PrepArgs args) {
    const PrepArgs a = for_solve(args);
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < R) a.wsx[idx] = 1.0;
    if (a.Wf && idx < (int64_t)a.n * R) a.Wf[idx] = (float)a.W[idx];  // FP32 form
    if (a.kind == K_EHALS && idx < (int64_t)a.n * R) {
        const int k = (int)(idx % R);
        const double w = a.W[idx];
        const double e = w + BETA_EHALS * (w - w);
        a.Wx[idx] = k < a.r ? (e > a.eps ? e : a.eps) : 0.0;
    }
}

template <int R>
__global__ __launch_bounds__(256) void k_hrec(HRecArgs a) {
    __shared__ double vals[TJC * R];
    const int tile = blockIdx.x;
    const int t = threadIdx.x;
    if (t < TJC) {
        const int j = tile * TJC + t;
#pragma unroll
        for (int k = 0; k < R; ++k) vals[t * R + k] = j < a.p ? a.H[(int64_t)k * a.ldv + j] : 0.0;
    }
    __syncthreads();
    write_record<R, TJC>(vals, a.red + (int64_t)tile * nred(R));
}

// Gram (upper triangle), column sums and sums of squares of the n x R matrix
// W as one record.  Phase 0: one partial record per WGRAM_ROWS rows, the
// rows staged in LDS and each output summed in row order; phase 1: one wave
// per output sums the partials (lane l takes partials l, l + 64, ...) and
// combines the lanes by a fixed shuffle tree.  With an offset vector the
// third section is W^T off (den_addon, lib.rs:1715-1720).
template <int R>
__global__ __launch_bounds__(256) void k_wgram(WGramArgs args, int phase) {
    if (args.done && *args.done) return;
    const WGramArgs a = for_solve(args);
    constexpr int NRED = nred(R);
    if (phase == 0) {
        __shared__ double wL[WGRAM_ROWS * R];
        __shared__ double oL[WGRAM_ROWS];
        const int i0 = blockIdx.x * WGRAM_ROWS;
        const int rows = min(a.n - i0, WGRAM_ROWS);
        for (int e = threadIdx.x; e < rows * R; e += blockDim.x) {
            const double x = a.W[(int64_t)i0 * R + e];
            wL[e] = x;
            if (a.Wf) a.Wf[(int64_t)i0 * R + e] = (float)x;  // FP32 form: the column pass's W
        }
        if (a.off)
            for (int e = threadIdx.x; e < rows; e += blockDim.x) oL[e] = a.off[i0 + e];
        __syncthreads();
        if (a.Wb)  // BF16 form: the column pass's W, transposed, as hi + lo planes
            for (int e = threadIdx.x; e < rows * R; e += blockDim.x) {
                const int k = e / rows, i = e % rows;
                uint16_t hi, lo;
                bf16_split(wL[i * R + k], hi, lo);
                a.Wb[(int64_t)k * a.nc + i0 + i] = hi;
                a.Wb[(int64_t)(R + k) * a.nc + i0 + i] = lo;
            }
        for (int o = threadIdx.x; o < NRED; o += blockDim.x) {
            double s = 0.0;
            if (o < ntri(R)) {
                int ka = 0, rem = o;
                while (rem >= R - ka) {
                    rem -= R - ka;
                    ++ka;
                }
                const int kb = ka + rem;
                for (int i = 0; i < rows; ++i) s += wL[i * R + ka] * wL[i * R + kb];
            } else if (o < ntri(R) + R) {
                const int k = o - ntri(R);
                for (int i = 0; i < rows; ++i) s += wL[i * R + k];
            } else {
                const int k = o - ntri(R) - R;
                if (a.off)
                    for (int i = 0; i < rows; ++i) s += wL[i * R + k] * oL[i];
                else
                    for (int i = 0; i < rows; ++i) s += wL[i * R + k] * wL[i * R + k];
            }
            a.part[(int64_t)blockIdx.x * NRED + o] = s;
        }
    } else {
        const int o = blockIdx.x * 4 + (threadIdx.x >> 6);
        const int lane = threadIdx.x & 63;
        if (o >= NRED) return;
        double s = 0.0;
        for (int b = lane; b < a.nparts; b += 64) s += a.part[(int64_t)b * NRED + o];
#pragma unroll
        for (int sh = 32; sh >= 1; sh >>= 1) s += __shfl_xor(s, sh, 64);
        if (lane == 0) a.out[o] = s;
    }
}

// nsNMF, start of an iteration: col_rescale of the stored W (lib.rs:1816,
// divisors from k_wstats; 1 on the first iteration) written back, and
// Wx = W S (lib.rs:1805), forward over l.  One thread per row.
template <int R>
__global__ __launch_bounds__(256) void k_nsprep(NsPrepArgs args) {
    if (args.done && *args.done) return;
    const NsPrepArgs a = for_solve(args);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double *sc = a.wsc + 2 * R;
    const int r = a.r;
    const double so = a.theta / (double)r, sd = (1.0 - a.theta) + so;
    double w[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        w[k] = k < r ? a.W[(int64_t)i * R + k] / sc[k] : 0.0;
        a.W[(int64_t)i * R + k] = w[k];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < R; ++l)
            if (l < r) s += w[l] * (l == k ? sd : so);
        a.Wx[(int64_t)i * R + k] = k < r ? s : 0.0;
    }
}

// Row sums of V: one wave per row, lane l sums columns l, l + 64, ... in
// order, then a fixed shuffle tree.
template <int VP>
__global__ __launch_bounds__(256) void k_rowsum_v(const void *V, int64_t vplane, int64_t ldv, int n,
                                                   int p, double *out) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    double s = 0.0;
    for (int j = lane; j < p; j += 64) s += VElem<VP>::one(V, vplane, (int64_t)i * ldv + j);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) out[i] = s;
}

// ---------------------------------------------------------------------------
// V storage conversion (FP64 source, row stride lds) -> form VP.
// ---------------------------------------------------------------------------
template <int VP>
__global__ __launch_bounds__(256) void k_convert(const double *src, int64_t lds, int64_t rows,
                                                  int64_t p, void *dst, int64_t ldv, int64_t plane) {
    const int64_t tot = rows * p;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / p, j = e % p;
        const double x = src[i * lds + j];
        const int64_t o = i * ldv + j;
        if constexpr (VP == VP_F64) {
            static_cast<double *>(dst)[o] = x;
        } else if constexpr (VP == VP_F32) {
            static_cast<float *>(dst)[o] = (float)x;
        } else {
            const float f = (float)x;
            const uint16_t hi = bf16_rne(f);
            const float rest = f - __uint_as_float((uint32_t)hi << 16);
            static_cast<uint16_t *>(dst)[o] = hi;
            static_cast<uint16_t *>(dst)[plane + o] = bf16_rne(rest);
        }
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int R, int VP>
hipError_t launch_col(const PassArgs &a, hipStream_t st) {
    if constexpr (VP == VP_B16X2) {
        if (a.b16) {
            // lee: numerators per row split to a.num, the Gram from k_wgram
            if (a.kind != K_LEE || !a.num || !a.xgram || !a.gram || !a.Vc || !a.Wb || !a.Hb ||
                a.kl_split < 1)
                return hipErrorInvalidValue;
            if (a.kl_split == 1) {  // the epilogue inside the numerator pass
                hipLaunchKernelGGL((k_colnum_b16<R, 4, 4, 2, true>), dim3(a.ntiles, a.bt.count),
                                   dim3(4 * 64), 0, st, a);
                return hipGetLastError();
            }
            hipLaunchKernelGGL((k_colnum_b16<R>), dim3(a.ntiles * a.kl_split, a.bt.count),
                               dim3(4 * 64), 0, st, a);
            hipLaunchKernelGGL((k_colfin_f32<R>), dim3(a.ntiles, a.bt.count), dim3(256), 0, st, a);
            return hipGetLastError();
        }
    }
    if constexpr (VP == VP_F32) {
        if (a.f32 && !kl_kind(a.kind)) {
            // lee: numerators to a.num, the Gram from k_wgram
            if (a.kind != K_LEE || !a.num || !a.xgram || !a.gram || !a.Wf || !a.Hf)
                return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_colnum_f32<R>), dim3((a.ntiles + 1) / 2, a.bt.count), dim3(4 * 64),
                               0, st, a);
            hipLaunchKernelGGL((k_colfin_f32<R>), dim3(a.ntiles, a.bt.count), dim3(256), 0, st, a);
            return hipGetLastError();
        }
    }
    if constexpr (VP == VP_F32) {
        if (a.f32 && kl_kind(a.kind)) {
            if (!a.Wf || !a.Hf || a.opart) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_col_kl<R, VP, (R <= 6 ? 16 : 8), false, float>),
                               dim3(a.ntiles * a.kl_split, a.bt.count), dim3((R <= 6 ? 16 : 8) * 64),
                               0, st, a);
            return hipGetLastError();
        }
    }
    if (kl_kind(a.kind) && a.opart)
        hipLaunchKernelGGL((k_col_kl<R, VP, (R <= 6 ? 16 : 8), true>), dim3(a.ntiles * a.kl_split, a.bt.count),
                           dim3((R <= 6 ? 16 : 8) * 64), 0, st, a);
    else if (kl_kind(a.kind))
        hipLaunchKernelGGL((k_col_kl<R, VP>), dim3(a.ntiles * a.kl_split, a.bt.count),
                           dim3((R <= 6 ? 16 : 8) * 64), 0, st, a);
    else if (a.kind == K_HALS || a.kind == K_EHALS || a.kind == K_DNMF) {
        if (a.xgram)
            hipLaunchKernelGGL((k_col_mfma<R, VP, true, true>), dim3(a.ntiles, a.bt.count),
                               dim3(NWC * 64), 0, st, a);
        else
            hipLaunchKernelGGL((k_col_mfma<R, VP, false, true>), dim3(a.ntiles, a.bt.count),
                               dim3(NWC * 64), 0, st, a);
    } else if (a.xgram) {
        hipLaunchKernelGGL((k_col_mfma<R, VP, true, false>), dim3(a.ntiles, a.bt.count),
                           dim3(NWC * 64), 0, st, a);
    } else {
        hipLaunchKernelGGL((k_col_mfma<R, VP, false, false>), dim3(a.ntiles, a.bt.count),
                           dim3(NWC * 64), 0, st, a);
    }
    return hipGetLastError();
}

template <int R>
hipError_t launch_red(const RedArgs &a, hipStream_t st) {
    if (a.ntiles >= RED_WIDE_TILES)
        hipLaunchKernelGGL((k_red_wide<R>), dim3(nred(R), a.bt.count), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_red<R>), dim3((nred(R) + 3) / 4, a.bt.count), dim3(256), 0, st, a);
    return hipGetLastError();
}

// Row pass: the VALU kernel (one block of RowCfg<R>::RW rows over all
// columns), or for ranks >= ROW_MFMA_MIN_R (not KL) the MFMA partials over
// nrb x nsplit blocks; the caller then launches rowfin.
template <int R, int VP>
hipError_t launch_row(const PassArgs &a, hipStream_t st) {
    const bool one = a.row_waves == 1;
    if (kl_kind(a.kind)) {
        constexpr int KRW = kl_rows_per_rowblk(R);
        constexpr int CPL = RowCfg<R>::CPL;
        const unsigned g = (unsigned)(a.nrowblk * std::max(1, a.nsplit));
        if constexpr (VP == VP_F32) {
            if (a.f32) {
                if (!a.Wf || !a.Hf) return hipErrorInvalidValue;
                if (one)
                    hipLaunchKernelGGL((k_row<R, true, VP, 1, true, KRW, CPL, float>),
                                       dim3(g, a.bt.count), dim3(64), 0, st, a);
                else
                    hipLaunchKernelGGL((k_row<R, true, VP, NWR, true, KRW, CPL, float>),
                                       dim3(g, a.bt.count), dim3(NWR * 64), 0, st, a);
                return hipGetLastError();
            }
        }
        if (one)
            hipLaunchKernelGGL((k_row<R, true, VP, 1, true, KRW>), dim3(g, a.bt.count), dim3(64), 0,
                               st, a);
        else
            hipLaunchKernelGGL((k_row<R, true, VP, NWR, true, KRW>), dim3(g, a.bt.count),
                               dim3(NWR * 64), 0, st, a);
    } else if (a.rowmfma && a.b16) {
        if constexpr (VP == VP_B16X2) {
            if (!a.Hb || !a.rpart) return hipErrorInvalidValue;
            const dim3 g(a.nrb * a.nsplit, a.bt.count);
            if (a.row_tiles == 4)
                hipLaunchKernelGGL((k_row_b16<R, 4>), g, dim3(4 * 64), 0, st, a);
            else if (a.row_tiles == 2)
                hipLaunchKernelGGL((k_row_b16<R, 2>), g, dim3(4 * 64), 0, st, a);
            else
                hipLaunchKernelGGL((k_row_b16<R, 1>), g, dim3(4 * 64), 0, st, a);
        } else {
            return hipErrorInvalidValue;
        }
    } else if (a.rowmfma && a.f32) {
        if constexpr (VP == VP_F32) {
            if (!a.Hf || !a.rpart) return hipErrorInvalidValue;
            const dim3 g(a.nrb * a.nsplit, a.bt.count);
            if (a.row_tiles == 4)
                hipLaunchKernelGGL((k_row_f32<R, 4>), g, dim3(4 * 64), 0, st, a);
            else if (a.row_tiles == 2)
                hipLaunchKernelGGL((k_row_f32<R, 2>), g, dim3(4 * 64), 0, st, a);
            else
                hipLaunchKernelGGL((k_row_f32<R, 1>), g, dim3(4 * 64), 0, st, a);
        } else {
            return hipErrorInvalidValue;
        }
    } else if (a.rowmfma) {
        // 4 tiles (large p): 4 column steps per wave in flight for FP64 and
        // FP32 storage (2000 x 500000 rank 20 row pass: FP64 2.13 -> 1.78 ms,
        // FP32 2.26 -> 2.17 ms against 2 steps); BF16 hi+lo stays at 2 (2.39
        // against 2.49 ms), as does the 2-tile instance (slower at p = 7750)
        if (a.row_tiles == 4)
            hipLaunchKernelGGL((k_row_mfma<R, VP, 4, 4, row4_steps(VP)>),
                               dim3(a.nrb * a.nsplit, a.bt.count), dim3(4 * 64), 0, st, a);
        else if (a.row_tiles == 2)
            hipLaunchKernelGGL((k_row_mfma<R, VP, 2, 4>), dim3(a.nrb * a.nsplit, a.bt.count),
                               dim3(4 * 64), 0, st, a);
        else
            hipLaunchKernelGGL((k_row_mfma<R, VP, 1, 8, 4>), dim3(a.nrb * a.nsplit, a.bt.count),
                               dim3(8 * 64), 0, st, a);
    } else if (a.kind == K_HALS || a.kind == K_EHALS || a.kind == K_DNMF) {
        if (one)
            hipLaunchKernelGGL((k_row<R, false, VP, 1>), dim3(a.nrowblk, a.bt.count), dim3(64), 0, st, a);
        else
            hipLaunchKernelGGL((k_row<R, false, VP>), dim3(a.nrowblk, a.bt.count), dim3(NWR * 64), 0,
                               st, a);
    } else if (one) {
        hipLaunchKernelGGL((k_row<R, false, VP, 1, false>), dim3(a.nrowblk, a.bt.count), dim3(64), 0,
                           st, a);
    } else {
        hipLaunchKernelGGL((k_row<R, false, VP, NWR, false>), dim3(a.nrowblk, a.bt.count),
                           dim3(NWR * 64), 0, st, a);
    }
    return hipGetLastError();
}

template <int R>
hipError_t launch_colfin(const PassArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_colfin_kl<R>), dim3(a.ntiles, a.bt.count), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int R>
hipError_t launch_rowfin(const PassArgs &a, hipStream_t st) {
    if (kl_kind(a.kind))
        hipLaunchKernelGGL((k_rowfin<R, true>), dim3((a.n + RF - 1) / RF, a.bt.count), dim3(256), 0,
                           st, a);
    else if (a.kind == K_HALS || a.kind == K_EHALS || a.kind == K_DNMF)
        hipLaunchKernelGGL((k_rowfin<R>), dim3((a.n + RF - 1) / RF, a.bt.count), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_rowfin<R, false, false>), dim3((a.n + RF - 1) / RF, a.bt.count),
                           dim3(256), 0, st, a);
    return hipGetLastError();
}

// Work-groups of the KL column pass (which = 0) or of the 4-wave KL row pass
// (which = 1) that fit on one CU at once.
template <int R, int VP>
int kl_occupancy(int which, int f32) {
    int nb = 0;
    if (which == 2) {  // the BF16 form's numerator pass (bf16 hi + lo storage only)
        if constexpr (VP == VP_B16X2) {
            hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_colnum_b16<R>, 4 * 64, 0);
            return e == hipSuccess && nb > 0 ? nb : 1;
        }
        return 1;
    }
    if constexpr (VP == VP_F32) {
        if (f32) {
            hipError_t e = which == 0
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                      &nb, k_col_kl<R, VP, (R <= 6 ? 16 : 8), false, float>, (R <= 6 ? 16 : 8) * 64, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                      &nb, k_row<R, true, VP, NWR, true, kl_rows_per_rowblk(R), RowCfg<R>::CPL, float>,
                      NWR * 64, 0);
            return e == hipSuccess && nb > 0 ? nb : 1;
        }
    }
    hipError_t e = which == 0
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_col_kl<R, VP>, (R <= 6 ? 16 : 8) * 64, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &nb, k_row<R, true, VP, NWR, true, kl_rows_per_rowblk(R)>, NWR * 64, 0);
    return e == hipSuccess && nb > 0 ? nb : 1;
}

// Work-groups of the MFMA row pass with `row_tiles` row tiles that fit on
// one CU at once (f32: the FP32 form's kernel, or the BF16 form's for
// bf16 hi + lo storage).
template <int R, int VP>
int rowmfma_occupancy(int row_tiles, int f32) {
    int nb = 0;
    if constexpr (VP == VP_B16X2) {
        if (f32) {
            hipError_t e = row_tiles == 4
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_row_b16<R, 4>, 4 * 64, 0)
                : row_tiles == 2
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_row_b16<R, 2>, 4 * 64, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_row_b16<R, 1>, 4 * 64, 0);
            return e == hipSuccess && nb > 0 ? nb : 1;
        }
    }
    if constexpr (VP == VP_F32) {
        if (f32) {
            hipError_t e = row_tiles == 4
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_row_f32<R, 4>, 4 * 64, 0)
                : row_tiles == 2
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_row_f32<R, 2>, 4 * 64, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_row_f32<R, 1>, 4 * 64, 0);
            return e == hipSuccess && nb > 0 ? nb : 1;
        }
    }
    hipError_t e = row_tiles == 4   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                          &nb, k_row_mfma<R, VP, 4, 4, row4_steps(VP)>, 4 * 64, 0)
                   : row_tiles == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                          &nb, k_row_mfma<R, VP, 2, 4>, 4 * 64, 0)
                                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                          &nb, k_row_mfma<R, VP, 1, 8, 4>, 8 * 64, 0);
    return e == hipSuccess && nb > 0 ? nb : 1;
}

template <int R>
hipError_t launch_wgram(const WGramArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_wgram<R>), dim3(a.nparts, a.bt.count), dim3(256), 0, st, a, 0);
    hipLaunchKernelGGL((k_wgram<R>), dim3((nred(R) + 3) / 4, a.bt.count), dim3(256), 0, st, a, 1);
    return hipGetLastError();
}

template <int R>
hipError_t launch_wstats(const WStatArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_wstats<R>), dim3(a.nparts, a.bt.count), dim3(256), 0, st, a, 0);
    if (!a.parts_only) hipLaunchKernelGGL((k_wstats<R>), dim3(1, a.bt.count), dim3(256), 0, st, a, 1);
    return hipGetLastError();
}

template <int R, int VP>
hipError_t launch_obj(const ObjArgs &a, hipStream_t st) {
    const dim3 grid(a.ntiles * a.osplit), block(NWO * 64);
    if (a.mfma) {
        if (a.kl) {
            if (a.off || a.wgt) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_obj_mfma<R, VP, OBJ_KL>), grid, block, 0, st, a);
        } else if (a.wgt) {
            if (a.off) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_obj_mfma<R, VP, OBJ_WGT>), grid, block, 0, st, a);
        } else if (a.off) {
            hipLaunchKernelGGL((k_obj_mfma<R, VP, OBJ_FROB_OFF>), grid, block, 0, st, a);
        } else {
            hipLaunchKernelGGL((k_obj_mfma<R, VP, OBJ_FROB>), grid, block, 0, st, a);
        }
        return hipGetLastError();
    }
    if (a.kl) {
        if (a.off || a.wgt) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_obj<R, VP, OBJ_KL>), grid, block, 0, st, a);
    } else if (a.wgt) {
        if (a.off) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_obj<R, VP, OBJ_WGT>), grid, block, 0, st, a);
    } else if (a.off) {
        hipLaunchKernelGGL((k_obj<R, VP, OBJ_FROB_OFF>), grid, block, 0, st, a);
    } else {
        hipLaunchKernelGGL((k_obj<R, VP, OBJ_FROB>), grid, block, 0, st, a);
    }
    return hipGetLastError();
}

template <int R>
hipError_t launch_fin(const Finalize &f, hipStream_t st) {
    const int64_t tot = std::max<int64_t>((int64_t)f.n * R, (int64_t)R * f.ldv);
    hipLaunchKernelGGL((k_finalize<R>), dim3((unsigned)((tot + 255) / 256), f.bt.count), dim3(256), 0, st, f);
    return hipGetLastError();
}

template <int R>
hipError_t launch_prep(const PrepArgs &a, hipStream_t st) {
    const int64_t tot = std::max<int64_t>((int64_t)a.n * R, R);
    hipLaunchKernelGGL((k_prep<R>), dim3((unsigned)((tot + 255) / 256), a.bt.count), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int R>
hipError_t launch_hrec(const HRecArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_hrec<R>), dim3(a.ntiles), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int R>
hipError_t launch_nsprep(const NsPrepArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_nsprep<R>), dim3((a.n + 255) / 256, a.bt.count), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int VP>
hipError_t rowsum_v(const void *V, int64_t vplane, int64_t ldv, int n, int p, double *out,
                    hipStream_t st) {
    hipLaunchKernelGGL((k_rowsum_v<VP>), dim3((n + 3) / 4), dim3(256), 0, st, V, vplane, ldv, n, p, out);
    return hipGetLastError();
}

template <int R, int VP>
hipError_t launch_ls(const PassArgs &a, hipStream_t st, int which) {
    const unsigned nb = (unsigned)a.bt.count;
    switch (which) {
        case 0:  // column pass
            if (a.wf == WF_U8)
                hipLaunchKernelGGL((k_col_ls<R, VP, WF_U8>), dim3(a.ntiles * a.kl_split, nb),
                                   dim3((R <= 6 ? 16 : 8) * 64), 0, st, a);
            else
                hipLaunchKernelGGL((k_col_ls<R, VP, WF_F64>), dim3(a.ntiles * a.kl_split, nb),
                                   dim3((R <= 6 ? 16 : 8) * 64), 0, st, a);
            if (a.kl_split > 1)
                hipLaunchKernelGGL((k_colfin_ls<R>), dim3(a.ntiles, nb), dim3(64), 0, st, a);
            break;
        default: {  // row pass
            const unsigned g = (unsigned)(a.nrowblk * std::max(1, a.nsplit));
            if (a.wf == WF_U8)
                hipLaunchKernelGGL((k_row_ls<R, VP, WF_U8>), dim3(g, nb), dim3(NWR * 64), 0, st, a);
            else
                hipLaunchKernelGGL((k_row_ls<R, VP, WF_F64>), dim3(g, nb), dim3(NWR * 64), 0, st, a);
            if (a.nsplit > 1)
                hipLaunchKernelGGL((k_rowfin_ls<R>), dim3((a.n + RF - 1) / RF, nb), dim3(256), 0, st, a);
        }
    }
    return hipGetLastError();
}

template <int R, int VP>
int ls_occupancy(int which) {
    int nb = 0;
    hipError_t e = which == 0
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_col_ls<R, VP, WF_U8>, (R <= 6 ? 16 : 8) * 64, 0)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_row_ls<R, VP, WF_U8>, NWR * 64, 0);
    return e == hipSuccess && nb > 0 ? nb : 1;
}

template <int R, int VP>
constexpr RankKernels table_entry() {
    return RankKernels{R,
                       launch_col<R, VP>,
                       launch_red<R>,
                       launch_row<R, VP>,
                       launch_wstats<R>,
                       launch_obj<R, VP>,
                       launch_fin<R>,
                       launch_prep<R>,
                       launch_hrec<R>,
                       launch_wgram<R>,
                       launch_rowfin<R>,
                       launch_nsprep<R>,
                       launch_colfin<R>,
                       kl_occupancy<R, VP>,
                       launch_ls<R, VP>,
                       ls_occupancy<R, VP>,
                       rowmfma_occupancy<R, VP>};
}

// Supported template ranks; a rank r runs on the smallest R >= r with the
// factors zero-padded to R (padded entries stay exactly zero).
template <int VP>
const RankKernels *lookup_rank(int r) {
    static const RankKernels table[] = {
        table_entry<1, VP>(),  table_entry<2, VP>(),  table_entry<3, VP>(),
        table_entry<4, VP>(),  table_entry<5, VP>(),  table_entry<6, VP>(),
        table_entry<8, VP>(),  table_entry<10, VP>(), table_entry<12, VP>(),
        table_entry<16, VP>(), table_entry<20, VP>(), table_entry<24, VP>(),
        table_entry<32, VP>(),
    };
    for (const auto &e : table)
        if (e.R >= r) return &e;
    return nullptr;
}

constexpr int kMaxRank = 32;

template <int VP>
hipError_t convert_v(const double *src, int64_t lds, int64_t rows, int64_t p, void *dst,
                     int64_t ldv, int64_t plane, hipStream_t st) {
    const int64_t tot = rows * p;
    const unsigned grid = (unsigned)std::min<int64_t>(16384, std::max<int64_t>(1, (tot + 255) / 256));
    hipLaunchKernelGGL((k_convert<VP>), dim3(grid), dim3(256), 0, st, src, lds, rows, p, dst, ldv,
                       plane);
    return hipGetLastError();
}

}  // namespace
}  // namespace nmfg
