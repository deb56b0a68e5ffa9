// Internal interface between the host runtime (nmfg_runtime.cpp) and the
// gfx950 kernels (nmfg_kernels.hip).  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nmfg {

// Update rule run by a pass.  The column pass forms W^T V (or, for KL,
// W^T (V ./ WH)) per column of V; the row pass forms V H^T (or
// (V ./ WH) H^T) per row of V.  The epilogue run by the last work-group
// of each tile differs per rule.
enum Kind : int { K_LEE = 0, K_BRUNET = 1, K_HALS = 2, K_DNMF = 3, K_EHALS = 4, K_PROD = 5 };
// K_PROD: plain products for the randomized range finder -- the column pass
// writes A = W^T V into H, the row pass writes B = V H^T into W.

// Column tile width of both passes: 64 lanes x 2 doubles (one dwordx4
// load per lane per row of V).
constexpr int TJ = 128;
// Rows per work-group in the row pass (4 waves x 4 rows).
constexpr int RB = 16;

// Reduction record written per tile / row block and summed globally:
//   [ Gram (R*R) | sum (R) | sumsq (R) ]
__host__ __device__ constexpr int nred(int R) { return R * R + 2 * R; }

// W-side statistics consumed by the next column pass (wstat):
//   Gu[R*R]  raw Gram of the stored W
//   cs[R]    column sums of the stored W
//   ss[R]    column sums of squares of the stored W
//   sc[R]    divisor taking stored W to the logical W (lee rescale, dnmf norm)
//   G[R*R]   Gram of the logical W
//   hdiv[R], hmul[R]  logical H = (stored H / hdiv) * hmul  (dnmf)
//   hclamp   brunet: 1 when the logical H is max(stored H, eps)
__host__ __device__ constexpr int ws_Gu(int) { return 0; }
__host__ __device__ constexpr int ws_cs(int R) { return R * R; }
__host__ __device__ constexpr int ws_ss(int R) { return R * R + R; }
__host__ __device__ constexpr int ws_sc(int R) { return R * R + 2 * R; }
__host__ __device__ constexpr int ws_G(int R) { return R * R + 3 * R; }
__host__ __device__ constexpr int ws_hdiv(int R) { return 2 * R * R + 3 * R; }
__host__ __device__ constexpr int ws_hmul(int R) { return 2 * R * R + 4 * R; }
__host__ __device__ constexpr int ws_hclamp(int R) { return 2 * R * R + 5 * R; }
__host__ __device__ constexpr int wstat_len(int R) { return 2 * R * R + 5 * R + 1; }

// H-side statistics consumed by the row pass (hstat): [K | rowsum | sumsq].
__host__ __device__ constexpr int hs_K(int) { return 0; }
__host__ __device__ constexpr int hs_rs(int R) { return R * R; }
__host__ __device__ constexpr int hs_ss(int R) { return R * R + R; }
__host__ __device__ constexpr int hstat_len(int R) { return R * R + 2 * R; }

// Everything a pass needs; passed by value as the kernel argument.
struct PassArgs {
    const double *V;   // n x ldv, zero-padded columns p..ldv-1
    int64_t ldv;       // == ldh, multiple of TJ
    int n, p, r;       // r <= R (template rank); factors zero-padded to R
    int ntiles;        // ceil(p / TJ)
    int nblocks;       // ceil(n / RB)
    int S;             // row splits (column pass) or column splits (row pass)
    int span;          // rows per split (column pass) / columns per split (row pass)
    double *W;         // stored W, n x R        (ehals: current iterate)
    double *Wx;        // ehals: extrapolated iterate (read by the column pass)
    double *H;         // stored H, R x ldv      (ehals: current iterate)
    double *Hp;        // ehals: previous iterate
    double *part;      // split partials
    double *red;       // per tile / per row block reduction records
    double *wstat;     // W-side statistics (see above)
    double *hstat;     // H-side statistics
    int *tile_cnt;     // arrival counters, one per tile / row block
    int *glob_cnt;     // one global arrival counter for this pass
    double eps;
    double lambda, gamma;  // dnmf L1 / L2 terms for this factor
    int rescale;           // lee column rescale
    int clamp;             // brunet: this iteration clamps (iter % 10 == 0)
    int kind;
};

// Objective pass: 0.5 ||V - WH||^2 or KL(V || WH) over the logical factors.
struct ObjArgs {
    const double *V;
    int64_t ldv;
    int n, p, r;
    int ntiles;
    int S, span;
    const double *W;
    const double *H;
    const double *wstat;
    int hmode;          // 0 identity, 1 dnmf transform, 2 brunet clamp
    double eps;
    double *part;       // per block partial sums (ntiles*S)
    int *glob_cnt;
    double *out;        // scalar result
    int kl;
};

// Statistics of raw factors (solve start / single-step calls).
struct StatArgs {
    const double *W;   // n x R
    double *Wx;        // ehals: receives the first extrapolation point
    const double *H;   // R x ldv
    int64_t ldv;
    int n, p, r;
    int ntiles, nblocks;
    double *wred, *hred;     // scratch reduction records
    double *wstat, *hstat;
    int *glob_cnt;           // [2]
    int kind;
    double eps;
    int dnmf_norm;  // dnmf stopped at iteration 0: normalise W and H, no d fold
};

struct Finalize {
    double *W;
    double *H;
    int64_t ldv;
    int n, p, r;
    const double *wstat;
    int kind;
    double eps;
};

// Launchers (one instantiation per supported template rank R).
using col_pass_fn = hipError_t (*)(const PassArgs &, hipStream_t);
using row_pass_fn = hipError_t (*)(const PassArgs &, hipStream_t);
using obj_fn = hipError_t (*)(const ObjArgs &, hipStream_t);
using stat_fn = hipError_t (*)(const StatArgs &, hipStream_t);
using fin_fn = hipError_t (*)(const Finalize &, hipStream_t);

struct RankKernels {
    int R;
    col_pass_fn col;
    row_pass_fn row;
    obj_fn obj;
    stat_fn stats;
    fin_fn fin;
};

// Returns the kernel table for the smallest supported R >= r, or nullptr.
const RankKernels *kernels_for_rank(int r);
int max_supported_rank();

}  // namespace nmfg
