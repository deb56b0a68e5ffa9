// Internal interface between the host runtime (nmfg_runtime.cpp) and the
// gfx950 kernels (nmfg_kernels.hip).  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nmfg {

// Update rule run by a pass.  The column pass forms W^T V (or, for KL,
// W^T (V ./ WH)) per column of V; the row pass forms V H^T (or
// (V ./ WH) H^T) per row of V.  K_PROD writes the plain products (used by
// the randomized range finder of nndsvd_init).  K_OFFSET is lee with a
// per-row offset (lee passes plus the offset terms); K_NSNMF is brunet on
// W S and S H (KL passes on the smoothed factors).
enum Kind : int {
    K_LEE = 0, K_BRUNET = 1, K_HALS = 2, K_DNMF = 3, K_EHALS = 4, K_PROD = 5, K_OFFSET = 6,
    K_NSNMF = 7, K_LSNMF = 8
};
inline bool kl_kind(int kind) { return kind == K_BRUNET || kind == K_NSNMF; }

// Storage of V on the device (values of enum nmfg_precision).  Factors and
// all arithmetic are FP64 in every form; the forms differ in V's bytes:
//   VP_F64    double
//   VP_F32    float (V rounded to nearest)
//   VP_B16X2  two bfloat16 planes hi, lo with f = float(V), hi = bf16(f),
//             lo = bf16(f - hi) (both round to nearest even); V ~ hi + lo.
//             The lo plane starts vplane elements after the hi plane.
constexpr int VP_F64 = 0, VP_F32 = 1, VP_B16X2 = 2;
// Storage of the lsNMF weight matrix (n x ldv, zero padded): FP64, or one
// byte per entry when every weight is 0 or 1.
constexpr int WF_F64 = 0, WF_U8 = 1;
inline int64_t v_bytes_per_elem(int vp) { return vp == VP_F64 ? 8 : 4; }

// V and H are stored with leading dimension ldv = ceil(p / LDV_ALIGN) * LDV_ALIGN,
// zero padded.
constexpr int LDV_ALIGN = 128;
// Column pass (MFMA path): one work-group per TJC columns, NWC waves
// splitting the rows.
constexpr int TJC = 32;
constexpr int NWC = 8;
// Column pass for KL (VALU path): one work-group per TJK columns, waves
// splitting the rows (16 at ranks up to 6, else 8).  The objective kernel
// uses the same tiles with NWK waves.
constexpr int TJK = 64;
// Column pass of the FP32 form: one work-group per TJF columns.
constexpr int TJF = 64;
constexpr int TJS = 128;  // FP32 numerator pass (k_colnum_f32): columns per work-group
constexpr int NWK = 16;
// Objective pass: waves per work-group and the objective forms.
constexpr int NWO = 4;
constexpr int TJO = 128;  // objective pass: columns per work-group (2 per lane)
enum { OBJ_FROB = 0, OBJ_FROB_OFF = 1, OBJ_KL = 2, OBJ_WGT = 3 };
// Row pass: one work-group per RWR(R) rows, NWR waves splitting the columns.
constexpr int NWR = 4;
__host__ __device__ constexpr int rows_per_rowblk(int R) {
    return R <= 10 ? 8 : (80 / R < 1 ? 1 : 80 / R);
}
// The KL row pass takes 4-row blocks up to rank 16 (fewer accumulator
// registers, more resident waves; measured at 2000 x 7750: rank 10 37.9 us vs
// 41.4 with 8 rows, rank 16 45.5 vs 64.0 with 5, rank 12 38.2 vs 51.7 with 6,
// rank 5 29.1 vs 31.1 with 8; 2-row blocks at rank 24 were slower, 97.7 vs
// 87.2 us), while lee and HALS are faster with 8 rows.
__host__ __device__ constexpr int kl_rows_per_rowblk(int R) {
    return R <= 16 ? 4 : rows_per_rowblk(R);
}
// Rows per block of the lsNMF row pass (2 R accumulators per row).  Each
// block reads all of H once, so more rows mean less L2 traffic per V row,
// as far as the registers allow: at 2000 x 7750 with a 5 % held-out mask,
// 3 rows at rank 10 (163 VGPRs, 3 waves per SIMD) took 100 iterations in
// 8.4 ms against 9.0 with 2; at rank 12 (191 VGPRs, 2 waves) 17.2 against
// 14.0, so ranks 11..16 keep 2.
__host__ __device__ constexpr int ls_rows(int R) {
    return R <= 4 ? 8 : (R <= 8 ? 4 : (R <= 10 ? 3 : (R <= 16 ? 2 : 1)));
}
// Rows per work-group of the W statistics kernel.
constexpr int WSTAT_ROWS = 1024;
// Rows per partial record of the W Gram kernel (k_wgram).
constexpr int WGRAM_ROWS = 64;

// Per-tile record written by the column pass and summed by k_red into hst,
// over the tile's new H columns h_j:
//   [ Gram sum_j h_j h_j^T, upper triangle row by row (R(R+1)/2) |
//     row sums (R) | row sums of squares (R) ]
__host__ __device__ constexpr int ntri(int R) { return R * (R + 1) / 2; }
__host__ __device__ constexpr int nred(int R) { return ntri(R) + 2 * R; }
__host__ __device__ constexpr int tri_index(int R, int a, int b) {
    return a <= b ? a * R - a * (a - 1) / 2 + (b - a) : b * R - b * (b - 1) / 2 + (a - b);
}
__host__ __device__ constexpr int hs_K(int) { return 0; }
__host__ __device__ constexpr int hs_rs(int R) { return ntri(R); }
__host__ __device__ constexpr int hs_ss(int R) { return ntri(R) + R; }

// Batched solves on one V (restarts of a rank sweep): solve b of `count`
// uses the per-solve buffers at b * (per-solve size, in doubles) from the
// base pointers; launches run one grid slice per solve (blockIdx.y = b).
struct Batch {
    int count;
    int64_t w, h, red, hst, wsx, wpart, wsc;
    int64_t gpart, gram;  // W Gram partials / sums (external-Gram column pass)
    int64_t rpart;        // row-pass partial numerators (column-split row pass)
    int64_t off;          // offset vector (n)
    int64_t kpart;        // split KL column pass partials
    int64_t num;          // FP32 / BF16 forms: column-pass numerators (floats)
    int64_t wb, hb;       // BF16 form: bf16 hi + lo copies of W and H' (uint16 elements)
    // Per-member ranks of a ragged batch (a rank sweep's members padded to
    // one template rank), or null when every member has rank r.
    const int *ranks;
};

struct PassArgs {
    const void *V;    // n x ldv, zero-padded, storage by the kernel's VP
    int64_t vplane;   // VP_B16X2: elements from the hi plane to the lo plane
    int64_t ldv;
    int n, p, r;      // r <= R (template rank); W, H zero-padded to R
    int ntiles;       // column-pass work-groups
    int nrowblk;      // row-pass work-groups
    double *W;        // stored W, n x R     (ehals: current iterate)
    double *Wx;       // ehals: extrapolated iterate, nsNMF: W S; streamed by the column pass
    double *H;        // stored H, R x ldv   (ehals: current iterate)
    double *Hp;       // ehals: previous H iterate; nsNMF: S H' (written by the column pass)
    double *off;      // offset: the offset vector (n), updated by the row pass
    const double *rsv;  // offset: row sums of V (n)
    const void *wgt;  // lsNMF: weight matrix (n x ldv, form wf)
    int wf;
    double *red;      // column-pass tile records (ntiles x nred(R))
    double *wsx;      // [R] divisor stored W -> logical W, written by column WG 0
    double *hst;      // H statistics summed by k_red (nred(R))
    const double *gram;  // xgram: record of the streamed W (k_wgram), read by the column pass
    int xgram;           // column pass takes the Gram / sums of W from `gram`
    double *rpart;       // column-split row pass: nsplit x n x R partial numerators
    int nsplit;          // column splits of the row pass (MFMA row pass, split KL row pass)
    int rowmfma;         // row pass on the MFMA kernel (+ k_rowfin)
    int nrb;             // row blocks of the MFMA row pass
    int row_tiles;       // MFMA row pass: 16-row tiles per block (1 or 4)
    double *opart;       // brunet objective-based stop: KL-term partials of the column pass
    const int *done;     // device stop flag: every kernel of the loop returns at once when set
    int row_waves;       // VALU row pass: waves per work-group (1 or NWR)
    int kl_split;        // KL column pass: row splits (> 1: partials to kpart, then k_colfin_kl)
    double *kpart;       // kl_split x ntiles x R x TJK partial W^T q
    int f32;             // FP32 form: products on FP32 MFMA over float V and the float copies
    float *Wf;           // FP32 form: float copy of the streamed W (n x R), from k_wgram
                         // (lee) or the KL row pass (brunet)
    float *Hf;           // FP32 form: float copy of H' (R x ldv), written by the column pass
    float *num;          // FP32 form: W^T V numerators of the column pass (R x ldv);
                         // BF16 form: one R x ldv block per row split (kl_split)
    int b16;             // BF16 form: products on bf16 MFMA over hi + lo operands
    const uint16_t *Vc;  // BF16 form: V's copy for the column pass, blocks of 8 rows with
                         // each column's 8 values contiguous: plane x ((i / 8) ldv + j) 8 + i % 8
    int64_t cplane;      // BF16 form: elements from Vc's hi plane to its lo plane (nc ldv)
    int nc;              // BF16 form: rows of Vc and of Wb's planes (n rounded up to 32)
    const uint16_t *Wb;  // BF16 form: the streamed W transposed, hi plane R x nc, then lo
    uint16_t *Hb;        // BF16 form: H' hi plane R x ldv, then lo (written by the column pass)
    double eps, lambda, gamma;
    double theta;     // nsNMF smoothing
    int rescale;      // lee column rescale
    int first;        // first iteration: stored W and H are the logical ones
    int clamp_w;      // brunet: clamp W' in this row pass (iteration % 10 == 0)
    int clamp_h;      // brunet: the stored H is pending a clamp (previous iteration)
    int kind;
    Batch bt;
};

struct RedArgs {
    const double *red;
    int ntiles;
    double *hst;
    Batch bt;
    const int *done;
};

// Column sums and sums of squares of the stored W (partials per
// WSTAT_ROWS rows, then summed in order by one work-group).
struct WStatArgs {
    const double *W;
    int n, r;
    double *part;     // nparts x 2R
    int nparts;
    double *out;      // [cs R | ss R | sc R]
    int kind, rescale;
    int first;        // stored W is already logical: sc = 1
    int dnmf_norm;    // dnmf stopped before its first iteration: sc = column norm
    int parts_only;   // phase 0 only (the consumer reduces the partials itself)
    Batch bt;
    const int *done;
};

// Start of a solve: wsx = 1; ehals: Wx = max(eps, W) (first extrapolation
// point W0 + beta (W0 - W0), lib.rs:755-759).
struct PrepArgs {
    const double *W;
    double *Wx;
    double *wsx;
    float *Wf;        // FP32 form: float copy of W0 (else null)
    int n, r;
    int kind;
    double eps;
    Batch bt;
};

// Gram, column sums and sums of squares of an n x R matrix (the W the
// column pass streams) as one record (layout above): phase 0 writes one
// partial record per WGRAM_ROWS rows, phase 1 sums them in a fixed order.
struct WGramArgs {
    const double *W;
    const double *off;  // offset: the third section holds W^T off instead of the sums of squares
    int n;
    double *part;     // nparts x nred(R), nparts = ceil(n / WGRAM_ROWS)
    int nparts;
    double *out;      // nred(R)
    int kind;
    float *Wf;        // FP32 form: float copy of W written alongside (else null)
    uint16_t *Wb;     // BF16 form: W transposed as bf16 hi (R x nc) and lo planes (else null)
    int nc;           // BF16 form: row stride of Wb's planes
    Batch bt;
    const int *done;
};

// Records of the stored H (for a W update that does not follow a column pass).
struct HRecArgs {
    const double *H;
    int64_t ldv;
    int p;
    int ntiles;       // ceil(p / TJC)
    double *red;
};

// Objective: per work-group partial sums, summed in order on the host.
struct ObjArgs {
    const void *V;
    int64_t vplane;
    int64_t ldv;
    int n, p, r;
    int ntiles;
    const double *W;
    const double *H;
    const double *wsc;  // [cs | ss | sc] of the stored W (unused when unit_sc or wpart)
    int unit_sc;        // the logical W is the stored W (divisors 1)
    const double *wpart;  // else, when set: k_wstats phase-0 partials, reduced here
    int wnparts, wkind, wrescale, wfirst;
    const double *hst;  // dnmf transform uses hst.ss
    int hmode;          // 0 identity, 1 dnmf transform, 2 brunet pending clamp, 3 nsNMF S H
    int clamp_h;
    const double *off;  // offset: added to every entry of row i of W H
    double theta;
    const void *wgt;    // lsNMF: objective 0.5 sum ((V - W H) S)^2
    int wf;
    double eps;
    double *part;       // ntiles x osplit partial sums
    int osplit;         // row splits per tile
    int kl;
    const int *done;
    int mfma;           // W H on the FP64 matrix cores (k_obj_mfma)
};

// Device-side objective total and stop test (objective-based stops run
// without a host round trip per iteration).  One work-group: sums the k_obj
// partials in a fixed order, records f in devs[it], applies the stop rule
// (reference stationary window, lib.rs:1378-1400, or the relative change)
// and sets *done / *n_iter when it passes (or f is NaN).
struct ObjFinArgs {
    const double *part;
    int nparts;
    int kl;
    int64_t *it;                // device counter: this evaluation's iteration (advanced here)
    int stop;                   // NMFG_STOP_STATIONARY or NMFG_STOP_REL_CHANGE
    double tol;                 // relative change
    int64_t min_iter;
    int64_t interval, niter;    // stationary window
    double th;
    double *devs;               // [max_iter + 1]
    double *state;              // [f_prev, max, min, count]
    int *done;
    int64_t *n_iter;
};
hipError_t launch_objfin(const ObjFinArgs &a, hipStream_t st);

struct Finalize {
    double *W;
    double *H;
    int64_t ldv;
    int n, p, r;
    const double *wsc;  // [cs | ss | sc] of the stored W
    const double *hst;
    int kind, clamp_h;
    int dnmf_fold;      // dnmf: 1 = H <- (H / nh) * (nw * nh); 0 = H <- H / nh
    double eps;
    Batch bt;
};

using pass_fn = hipError_t (*)(const PassArgs &, hipStream_t);
using red_fn = hipError_t (*)(const RedArgs &, hipStream_t);
using wstat_fn = hipError_t (*)(const WStatArgs &, hipStream_t);
using obj_fn = hipError_t (*)(const ObjArgs &, hipStream_t);
using fin_fn = hipError_t (*)(const Finalize &, hipStream_t);
using prep_fn = hipError_t (*)(const PrepArgs &, hipStream_t);
using hrec_fn = hipError_t (*)(const HRecArgs &, hipStream_t);
using wgram_fn = hipError_t (*)(const WGramArgs &, hipStream_t);

// nsNMF start of an iteration: the stored W becomes the logical W (divided
// by wsc's divisors, col_rescale) and Wx = W S.
struct NsPrepArgs {
    double *W;
    double *Wx;
    const double *wsc;
    int n, r;
    double theta;
    Batch bt;
    const int *done;
};
using nsprep_fn = hipError_t (*)(const NsPrepArgs &, hipStream_t);

struct RankKernels {
    int R;
    pass_fn col;      // column pass (MFMA or KL variant by kind)
    red_fn red;
    pass_fn row;
    wstat_fn wstats;
    obj_fn obj;
    fin_fn fin;
    prep_fn prep;
    hrec_fn hrec;
    wgram_fn wgram;
    pass_fn rowfin;   // sums the column-split partials and runs the W epilogue
    nsprep_fn nsprep;
    pass_fn colfin;   // split KL column pass: sums the partials and runs the H epilogue
    int (*kl_occupancy)(int which, int f32);  // work-groups per CU: 0 KL column pass, 1 KL row
                                              // pass, 2 the BF16 form's numerator pass
    hipError_t (*ls)(const PassArgs &, hipStream_t, int which);  // lsNMF: 0 column, 1 row pass
    int (*ls_occupancy)(int which);
    int (*rowmfma_occupancy)(int row_tiles, int f32);  // work-groups per CU of the MFMA row pass
};

// Record reductions over at least this many tiles use 4 waves per output.
constexpr int RED_WIDE_TILES = 4096;
// Column passes with more tiles than this take the Gram of W from k_wgram
// (one per iteration) instead of forming it in every tile.
constexpr int XGRAM_TILES = 1024;
// Ranks above this use the MFMA row pass (16-row blocks, column splits).
constexpr int ROW_MFMA_MIN_R = 11;
// VALU row pass: one-wave work-groups from this many row blocks (x batch).
constexpr int ROW_ONE_WAVE_BLOCKS = 1024;
// Rows of one MFMA row block.
constexpr int RB_MFMA = 16;

// Kernel table for the smallest supported R >= r and V storage vp, or nullptr.
const RankKernels *kernels_for_rank(int r, int vp);
const RankKernels *kernels_for_rank_f64(int r);
const RankKernels *kernels_for_rank_f32(int r);
const RankKernels *kernels_for_rank_b16x2(int r);
int max_supported_rank();

// Moves the kc columns c0 .. c0 + kc - 1 of M (rows x k, row-major) to a
// factor panel (to_panel = 1) or back (to_panel = 0).  Panel layouts: column
// panel P[kk * ld + j] (R x ld, H's), row panel P[i * R + kk] (rows x R,
// W's).  Writing a panel fills every entry outside the slice with 0.
struct PanelArgs {
    double *M;
    int64_t rows;
    int k, c0, kc;
    double *P;
    int R;
    int64_t ld;  // column panel: row stride (>= rows)
    int colpanel, to_panel;
};
hipError_t launch_panel(const PanelArgs &a, hipStream_t st);

// dst[e] = src[e], e < count, unless *done (a copy that a stopped solve's
// loop skips, for capture in its graphs).
hipError_t launch_copy_unless_done(const double *src, double *dst, int64_t count, const int *done,
                                   hipStream_t st);

// Bytes of the statically allocated device scratch of the no-allocation
// entry point (one copy per device, held until the process ends), and the
// device address of the current device's copy.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
hipError_t noalloc_scratch_address(void **out);

// p[e] = v for e < count.
hipError_t launch_fill_f64(double *p, int64_t count, double v, hipStream_t st);
// Cache eviction read of n16 16-byte words at buf; out holds EVICT_BLOCKS *
// 256 unsigned (written only in the unlikely case the kernel's XOR matches).
constexpr int EVICT_BLOCKS = 2048;
hipError_t launch_evict(const void *buf, int64_t n16, unsigned *out, hipStream_t st);
// Start state of an objective-based stop: state [f_prev, max, min, count] =
// 0, *done = 0, niter = [-1 (stop iteration), 0 (evaluation counter)].
hipError_t launch_stop_reset(double *state, int *done, int64_t *niter, hipStream_t st);
// Members of a step group (nmfg_group) and the pointers to their stop flags
// and stop iterations, gathered by one kernel per step.
constexpr int GROUP_MAX = 64;
struct StopPtrs {
    const int *done[GROUP_MAX];
    const int64_t *niter[GROUP_MAX];
    int count;
};
hipError_t launch_gather_stop(const StopPtrs &p, int64_t *out, hipStream_t st);

// Caller factors in device memory, W (n x r, row-major) and H (r x p, row
// stride ldh), to a solve's layout sW (n x R) and sH (R x ldv), zero padded
// (to_solve = 1), or the first r columns / rows back (to_solve = 0).
struct FactorPack {
    double *W, *H;
    int64_t n, p, ldh;
    int r;
    double *sW, *sH;
    int R;
    int64_t ldv;
    int to_solve;
};
hipError_t launch_factor_pack(const FactorPack &a, hipStream_t st);

// LU with partial pivoting of a tall matrix X (m x w, row-major, m >= w) in
// place into P L, the first output of scipy.linalg.lu(X, permute_l=True)
// (LAPACK dgetrf's pivots: largest |x| of the column, the first in the
// row order after the earlier swaps on ties; L scaled by the pivot's
// reciprocal unless |pivot| < DBL_MIN).  Rows are never moved: pos / row_at
// track the swapped order, step_of marks the pivot rows.  Work buffers:
// 3 m ints (pos, row_at, step_of), LU_PARTS partial triples, w pivot rows,
// and the selected pivot.
constexpr int LU_PARTS = 1024;
struct LuWork {
    int *pos, *row_at, *step_of;  // m each
    double *part_val;             // LU_PARTS
    int *part_pos, *part_row;     // LU_PARTS each
    int *prow;                    // w: pivot row of each step
    double *sel;                  // [pivot value, its reciprocal]
};
size_t lu_work_bytes(int64_t m, int w);
LuWork lu_work_at(void *base, int64_t m, int w);
hipError_t launch_lu_pl(double *X, int64_t m, int w, const LuWork &ws, hipStream_t st);

// Consensus (connectivity average) of `runs` clusterings of p items:
// C[i][j] = (number of runs with label_i == label_j) / runs, p x p.
hipError_t launch_consensus(const int32_t *labels, int runs, int p, double *C, hipStream_t st);

// Converts rows x p FP64 values (row stride lds; host or device memory) into
// V's storage at dst (row stride ldv, lo plane at +plane elements).  src must
// be device-accessible.
hipError_t launch_convert_v(int vp, const double *src, int64_t lds, int64_t rows, int64_t p,
                            void *dst, int64_t ldv, int64_t plane, hipStream_t st);
// The same rows (rows x p FP64 at src, V's rows r0 .. r0 + rows - 1, r0 a
// multiple of 8) into the BF16 form's column-pass copy Vc (layout of
// PassArgs::Vc; rows past the chunk's end within its last 8-row block are
// written as zeros).
hipError_t launch_convert_vc(const double *src, int64_t lds, int64_t rows, int64_t r0, int64_t p,
                             uint16_t *Vc, int64_t ldv, int64_t cplane, hipStream_t st);

// Column-pass tile width for a kind at template rank R: TJK columns per
// work-group in the KL and lsNMF passes (k_col_kl, k_col_ls, their
// finishing kernels), TJC in the MFMA column pass.  The grid is
// ceil(p / width) tiles, so the width must be the one the kernel indexes
// with: a tile's last column, ntiles x TJK - 1, then stays below ldv (a
// multiple of 128).
inline int col_tile_width(int kind, int /*R*/, bool f32 = false) {
    return kl_kind(kind) || kind == K_LSNMF ? TJK : (f32 ? TJF : TJC);
}

// BF16 form: rows of the column pass's copy of V (and of Wb), a multiple of
// the pass's 32-row step; the rows past n hold zeros.
inline int64_t b16_rows(int64_t n) { return (n + 31) / 32 * 32; }

// Row sums of V (storage form vp, n x ldv), forward partial sums per lane
// then a fixed tree: out[i] for i < n.
hipError_t launch_rowsum_v(int vp, const void *V, int64_t vplane, int64_t ldv, int n, int p,
                           double *out, hipStream_t st);

}  // namespace nmfg
