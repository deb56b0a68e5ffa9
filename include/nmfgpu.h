/*
 * nmfgpu -- C ABI of the MI355X (gfx950) dense NMF update library.
 *
 * This is the drop-in boundary for omicverse/rust-NMF's native module
 * `nmf_rs._rust` (rust/src/lib.rs:1835-2187).  Each entry point below names
 * the reference function it replaces.  Plain pointers and sizes only; all
 * matrices are row-major float64 (the reference's `_f64_c` contract,
 * python/nmf_rs/__init__.py:235-240): V is n x p, W is n x r, H is r x p.
 *
 * Return codes: NMFG_OK (0) or a negative category.  nmfg_last_error()
 * returns a message for the calling thread's last failure.  A binding maps
 * NMFG_EINVAL to ValueError and everything else to RuntimeError, as the
 * reference does (lib.rs:1499-1502, 1520-1523, 1967-1971, 1981-1986).
 *
 * All computation runs on the GPU.  There is no CPU fallback: without a
 * usable gfx950 device every compute entry point returns NMFG_EDEVICE.
 */
#ifndef NMFGPU_H
#define NMFGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NMFG_OK 0
#define NMFG_EINVAL (-1)
#define NMFG_ENOMEM (-2)
#define NMFG_EDEVICE (-3)
#define NMFG_EUNSUPPORTED (-4)
#define NMFG_EBUDGET (-5)

/* Method codes; names and aliases follow Algo::from_str (lib.rs:1472-1505). */
enum nmfg_algo {
    NMFG_BRUNET = 0,
    NMFG_LEE = 1,
    NMFG_OFFSET = 2,  /* lee + per-row offset (lib.rs:1649-1759) */
    NMFG_NSNMF = 3,   /* brunet on W S, S H (lib.rs:1761-1833) */
    NMFG_HALS = 4,
    NMFG_EHALS = 5,
    NMFG_DNMF = 6,
    NMFG_LSNMF = 7,   /* weighted Frobenius MU (lib.rs:946-1043); needs a weight */
    NMFG_SNMF_R = 8,  /* Kim-Park sparse NMF (lib.rs:1198-1334): nmfg_run only */
    NMFG_SNMF_L = 9   /* the same on V^T with W and H swapped */
};

/* Stop::from_str (lib.rs:1507-1526).  NMFG_STOP_REL_CHANGE has no reference
 * counterpart (served solves): after iteration t >= min_iter, stop when
 * |f(t-1) - f(t)| <= rel_tol * |f(t-1)| for the objective f the deviances
 * hold (lib.rs:1406-1445); W and H are those of iteration t. */
enum nmfg_stop { NMFG_STOP_MAX_ITER = 0, NMFG_STOP_STATIONARY = 1, NMFG_STOP_REL_CHANGE = 2 };

/* Precision forms of a matrix (V's storage on the device and the arithmetic
 * of the two big products).  No reference counterpart (the reference is
 * FP64 only).
 *   NMFG_FP64          8 B per element; FP64 arithmetic throughout
 *   NMFG_FP32_STORE    4 B, V rounded to float; FP64 arithmetic on the
 *                      stored values (exactly the FP64 algorithm on
 *                      float(V))
 *   NMFG_BF16X2_STORE  4 B as two bfloat16 planes hi + lo of float(V)
 *                      (about 16 significant bits); FP64 arithmetic
 *   NMFG_FP32          float V; lee forms W^T V and V H'^T on FP32 MFMA
 *                      (v_mfma_f32_16x16x4_f32) and brunet its KL passes
 *                      in FP32, all with FP32 accumulation over float
 *                      copies of the streamed factors; W and H, the Grams,
 *                      the update epilogues, the objective and the stop
 *                      test stay FP64.  The HALS family (whose updates
 *                      subtract nearly equal products, which FP32
 *                      accumulation cannot resolve), the other methods and
 *                      the products of nndsvd's range finder run FP64
 *                      arithmetic on the float V
 *   NMFG_BF16X2        V as the bfloat16 hi + lo planes of NMFG_BF16X2_STORE
 *                      plus a second copy of them laid out for the column
 *                      pass (8 B per element in HBM, 4 B streamed per
 *                      element per pass); lee forms W^T V and V H'^T on
 *                      bf16 MFMA (v_mfma_f32_16x16x32_bf16) as
 *                      hi hi + hi lo + lo hi with FP32 accumulation, over
 *                      hi + lo copies of the streamed factors; W and H, the
 *                      Grams, the epilogues, the objective and the stop
 *                      test stay FP64.  The other methods run FP64
 *                      arithmetic on the hi + lo planes
 * The arithmetic form belongs to the matrix: every solve on it uses it. */
enum nmfg_precision {
    NMFG_FP64 = 0,
    NMFG_FP32_STORE = 1,
    NMFG_BF16X2_STORE = 2,
    NMFG_FP32 = 3,
    NMFG_BF16X2 = 4
};

/* Keyword arguments of py_nmf_run (lib.rs:1912-1933). */
typedef struct nmfg_params {
    int32_t algo;           /* enum nmfg_algo */
    int64_t max_iter;
    double eps;             /* 1e-9; brunet always uses DBL_EPSILON (lib.rs:1539) */
    int32_t rescale;        /* 1: lee column rescale (lib.rs:1635) */
    double theta;           /* nsNMF smoothing (0.5) */
    double sparsity_h, sparsity_w;     /* dnmf L1 terms */
    double smoothness_h, smoothness_w; /* dnmf L2 terms */
    int32_t stop;           /* enum nmfg_stop */
    double stationary_th;   /* 2.220446049250313e-16 */
    int64_t check_interval; /* 50 */
    int64_t check_niter;    /* 10 */
    int32_t num_threads;    /* accepted for API parity; the GPU ignores it */
    int32_t precision;      /* enum nmfg_precision: V's form for nmfg_run */
    double rel_tol;         /* NMFG_STOP_REL_CHANGE tolerance (1e-6) */
    int64_t min_iter;       /* NMFG_STOP_REL_CHANGE: iterations before a stop (1) */
} nmfg_params;

/* Fills the reference defaults of py_nmf_run (lib.rs:1912-1933). */
void nmfg_params_init(nmfg_params *prm);

typedef struct nmfg_ctx nmfg_ctx;

/* One context per device; entry points taking a context are thread-safe.
 * Device memory of the context's matrices and solves is kept by the context
 * when they are destroyed and handed to its next allocations (all of them
 * ordered on the context's stream), so destroying a matrix or solve never
 * waits for other contexts' work the way hipFree does.  Up to NMFG_CACHE_MB
 * (environment, read at create; default 8192, 0 frees at once) idle MiB per
 * context; nmfg_ctx_destroy frees them.  Destroy a context's matrices and
 * solves before the context. */
int nmfg_ctx_create(int32_t device, nmfg_ctx **out);
int nmfg_ctx_destroy(nmfg_ctx *ctx);
/* The hipStream_t every call on this context is ordered on. */
void *nmfg_ctx_stream(nmfg_ctx *ctx);
/* Waits for the context's stream and frees the idle device memory it keeps
 * (no reference counterpart: the reference holds no device memory). */
int nmfg_ctx_release_cached(nmfg_ctx *ctx);

/* Device-memory budget (no reference counterpart; for serving).  Every
 * block a context attached to the budget holds -- matrices, solves, and the
 * released blocks it keeps for reuse -- is charged to it.  An allocation
 * that would exceed it first frees idle blocks (the context's own, then
 * those of other attached contexts that are not busy) and otherwise fails
 * with NMFG_EBUDGET.  nmfg_footprint gives the bytes a matrix of n x p in
 * prm->precision plus one solve of rank r (prm->algo) allocate, plus, when
 * nndsvd_width > 0, nndsvd_init's device range finder of that width
 * (rank + oversamples) -- what a server admits requests against. */
typedef struct nmfg_budget nmfg_budget;
int nmfg_budget_create(int64_t bytes, nmfg_budget **out);
/* NMFG_EINVAL while contexts are still attached. */
int nmfg_budget_destroy(nmfg_budget *b);
/* Attaches (b != NULL) or detaches the context; the blocks it already holds
 * are charged to / released from the budget. */
int nmfg_ctx_set_budget(nmfg_ctx *ctx, nmfg_budget *b);
int nmfg_budget_query(nmfg_budget *b, int64_t *cap, int64_t *used, int64_t *peak);
/* hipMemGetInfo of the device: the figure a budget is set from. */
int nmfg_device_memory(int32_t device, int64_t *free_bytes, int64_t *total_bytes);
int nmfg_footprint(const nmfg_params *prm, int64_t n, int64_t p, int32_t r, int32_t nndsvd_width,
                   int64_t *bytes);

/* Replaces py_nmf_run (lib.rs:1934-2169) for lee, brunet, offset, nsNMF,
 * hals, ehals, dnmf, snmf/r and snmf/l (lsNMF: nmfg_run_lsnmf; NMFG_EINVAL
 * here, as the reference's ValueError without a weight, lib.rs:1967-1971).
 * snmf records a deviance every iteration whatever the stop and runs
 * max_iter iterations (lib.rs:1254-1321); its eta / beta are smoothness_w /
 * sparsity_h (snmf/r) or smoothness_h / sparsity_w (snmf/l), eta < 0
 * meaning max(V), and eps is FCNNLS's tolerance.  W and H carry W0/H0 in and the result out.  When
 * stop is NMFG_STOP_STATIONARY, up to dev_cap objective values are written
 * to `deviances` and their count to *n_dev (the reference returns
 * n_iter + 1 of them).  V is read with leading dimension ldv (>= p).
 * offset starts from rowMeans(V) here; nmfg_run_offset takes and returns
 * the offset vector. */
int nmfg_run(nmfg_ctx *ctx, const nmfg_params *prm, const double *V, int64_t n, int64_t p,
             int64_t ldv, double *W, double *H, int32_t r, int64_t *n_iter, double *deviances,
             int64_t dev_cap, int64_t *n_dev);
/* nmfg_run with py_nmf_run's `offset` argument and "offset" result
 * (lib.rs:1944, 1963, 2023-2036, 2160-2164): for NMFG_OFFSET, `offset` (n
 * doubles) receives the final offset; when offset_in is nonzero it also
 * holds the initial offset on entry, else the run starts from rowMeans(V)
 * (lib.rs:1669-1683).  offset may be NULL (not returned); other methods
 * leave it untouched. */
int nmfg_run_offset(nmfg_ctx *ctx, const nmfg_params *prm, const double *V, int64_t n, int64_t p,
                    int64_t ldv, double *W, double *H, int32_t r, double *offset,
                    int32_t offset_in, int64_t *n_iter, double *deviances, int64_t dev_cap,
                    int64_t *n_dev);

/* py_nmf_run with `weight` (lib.rs:1945, 1964, 2063-2079) for NMFG_LSNMF:
 * weight is n x p (leading dimension ldw >= p).  A weight whose entries are
 * all 0 or 1 is held as one byte per entry on the device. */
int nmfg_run_lsnmf(nmfg_ctx *ctx, const nmfg_params *prm, const double *V, int64_t n, int64_t p,
                   int64_t ldv, const double *weight, int64_t ldw, double *W, double *H, int32_t r,
                   int64_t *n_iter, double *deviances, int64_t dev_cap, int64_t *n_dev);

/* nmfg_run without device allocation.  No counterpart in the reference: for
 * callers that cannot call hipMalloc (code running while another stream of
 * the process is under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  V, W and H are device memory, written before
 * the call: V is n x p FP64 with row stride ldv, a multiple of 128 elements
 * (>= p), its padding columns p .. ldv - 1 finite (zeros, say); W (n x r,
 * row-major) and H (r x p, row stride ldh >= p) carry W0 / H0 in and the
 * result out.  Every working buffer of the solve is carved from a
 * statically allocated device array of NMFG_NOALLOC_SCRATCH_BYTES per
 * device, held from library load to process end; the call makes no
 * hipMalloc / hipFree and captures no graph.  A solve that needs more
 * scratch returns NMFG_ENOMEM (2000 x 7750 at rank 10 needs about 5 MB).
 * Methods lee, brunet, hals, ehals, dnmf, nsNMF; precision NMFG_FP64; every
 * stop.  Same kernels and result as nmfg_run.  Calls on one device share
 * its scratch array, which keeps the last call's contents (each call
 * initialises what it reads): concurrent calls on one device would race on
 * it, so the library serialises them with a per-device lock and returns
 * only after the call's device work has finished. */
#define NMFG_NOALLOC_SCRATCH_BYTES ((int64_t)64 << 20)
int nmfg_run_device_noalloc(nmfg_ctx *ctx, const nmfg_params *prm, const double *V, int64_t n,
                            int64_t p, int64_t ldv, double *W, double *H, int64_t ldh, int32_t r,
                            int64_t *n_iter, double *deviances, int64_t dev_cap, int64_t *n_dev);

/* Replace py_euclidean_update_h / _w (lib.rs:1879-1907, impls 247-322). */
int nmfg_euclidean_update_h(nmfg_ctx *ctx, const double *V, int64_t n, int64_t p,
                            const double *W, const double *H, int32_t r, double eps, double *H_out);
int nmfg_euclidean_update_w(nmfg_ctx *ctx, const double *V, int64_t n, int64_t p,
                            const double *W, const double *H, int32_t r, double eps, double *W_out);
/* Replace py_divergence_update_h / _w (lib.rs:1849-1877, impls 117-204). */
int nmfg_divergence_update_h(nmfg_ctx *ctx, const double *V, int64_t n, int64_t p,
                             const double *W, const double *H, int32_t r, double *H_out);
int nmfg_divergence_update_w(nmfg_ctx *ctx, const double *V, int64_t n, int64_t p,
                             const double *W, const double *H, int32_t r, double *W_out);

/* Replaces py_smoothing_matrix (lib.rs:2171-2175, smoothing_matrix 457-466). */
int nmfg_smoothing_matrix(int32_t r, double theta, double *S);

/* Replaces set_num_threads (lib.rs:1841-1847): returns 1 on the first call
 * and 0 afterwards, like rayon's build_global.  The GPU path ignores it. */
int nmfg_set_num_threads(int64_t n);

/* FCNNLS on the normal equations (replaces fcnnls::fcnnls, fcnnls.rs:270-501,
 * for X^T X and X^T Y already formed): ctc = X^T X (l x l, row-major),
 * cta = (X^T Y)^T (p x l: row j for right-hand side j), K out p x l.
 * Host FP64 (the snmf drivers' k x k solves); l in 1..32. */
int nmfg_fcnnls(const double *ctc, const double *cta, int32_t l, int64_t p, double eps, double *K);

/* Thread-local message for the last failed call on this thread. */
const char *nmfg_last_error(void);

/* Highest supported factorisation rank. */
int32_t nmfg_max_rank(void);

/* ------------------------------------------------------------------------
 * Device-resident matrices and solves.  No reference counterpart: the
 * reference copies V per call (lib.rs:1960-1964).  These keep V, W and H on
 * the GPU across calls, for repeated solves, benchmarking and serving.
 * ---------------------------------------------------------------------- */
typedef struct nmfg_matrix nmfg_matrix;
typedef struct nmfg_solve nmfg_solve;

int nmfg_matrix_create(nmfg_ctx *ctx, int64_t n, int64_t p, int32_t precision, nmfg_matrix **out);
/* Copies FP64 V (leading dimension ldv; host memory, or device memory such
 * as a torch tensor's) into the device layout and storage form, and waits
 * for the copy.  The copy runs on the context's stream, which is created
 * non-blocking and so is not ordered after other streams: device-memory V
 * must be completely written before the call (synchronize the stream that
 * produced it, or wait on an event of it). */
int nmfg_matrix_upload(nmfg_matrix *m, const double *V, int64_t ldv);
int nmfg_matrix_destroy(nmfg_matrix *m);
/* Dense products with the resident V for nndsvd_init's randomized SVD
 * (python/nmf_rs/__init__.py:172-174 -> sklearn randomized_svd, whose
 * A @ Q / A.T @ Q products this replaces): trans == 0 gives C (n x k) =
 * V B with B p x k; trans == 1 gives C (p x k) = V^T B with B n x k.
 * Host buffers, row-major. */
int nmfg_matrix_mul(nmfg_matrix *m, int32_t trans, const double *B, int32_t k, double *C);
/* The range finder of that randomized SVD with its LU power-iteration
 * normaliser (sklearn.utils.extmath.randomized_range_finder, n_iter > 2),
 * all on the device: A = V^T when flip, else V (a x b); Q0 (b x width,
 * host) the Gaussian sketch; n_iter times Q = P L of lu(A Q), then Q = P L
 * of lu(A^T Q); Y (a x width, host) = A Q for the final QR.  Both
 * dimensions of V must be >= width. */
int nmfg_matrix_range_finder(nmfg_matrix *m, int32_t flip, const double *Q0, int32_t width,
                             int32_t n_iter, double *Y);
/* P L of the LU with partial pivoting of X (m x w, m >= w, row-major host
 * buffers), the first output of scipy.linalg.lu(X, permute_l=True) that the
 * range finder above uses, computed on the device. */
int nmfg_lu_permute_l(nmfg_ctx *ctx, const double *X, int64_t m, int32_t w, double *PL);

int nmfg_solve_create(nmfg_matrix *m, int32_t r, const nmfg_params *prm, nmfg_solve **out);
/* `batch` independent solves of rank r on the same V, advanced together:
 * every kernel launch covers all of them (one grid slice per solve).  For
 * the restarts of a rank sweep (examples/build_scaling_notebook.py:157-208
 * loops over runs calling nmf_rs.nmf once per run).  Fixed iteration count
 * only (stop = NMFG_STOP_MAX_ITER).  Solve b equals a single solve from the
 * same W0/H0 to roundoff: the row and column splits of the passes (and so
 * the order of their partial sums) are chosen for the whole batch's grid;
 * where they coincide with the single solve's, the bits are the same. */
int nmfg_solve_create_batched(nmfg_matrix *m, int32_t r, const nmfg_params *prm, int32_t batch,
                              nmfg_solve **out);
int32_t nmfg_solve_batch(const nmfg_solve *s);
/* Host W0 (n x r) and H0 (r x p) -- for a batched solve, `batch`
 * consecutive such pairs; stays on the device until replaced. */
int nmfg_solve_set_factors(nmfg_solve *s, const double *W0, const double *H0);
/* Offset solves: initial offset (`batch` consecutive n-vectors, host),
 * default rowMeans(V); and the current offset after a solve (waits).
 * NMFG_EINVAL for other methods. */
int nmfg_solve_set_offset(nmfg_solve *s, const double *offset);
/* lsNMF solves: the weight matrix (host, n x p with leading dimension ldw),
 * shared by every solve of a batch; required before the first launch. */
int nmfg_solve_set_weight(nmfg_solve *s, const double *weight, int64_t ldw);
int nmfg_solve_get_offset(nmfg_solve *s, double *offset);
/* Enqueues one full solve from the stored W0/H0 on the context stream and
 * returns without waiting (max_iter stop: one replay of a captured HIP
 * graph).  An objective-based stop (stationary, relative change) enqueues
 * 10 iterations at a time (after the first 10, one replay of a captured
 * graph) and waits for the device's stop flag between them, so the call
 * returns once the solve has stopped. */
int nmfg_solve_launch(nmfg_solve *s);
/* Waits for the solve and copies W (n x r) and H (r x p) to the host
 * (batched: `batch` consecutive pairs). */
int nmfg_solve_fetch(nmfg_solve *s, double *W, double *H, int64_t *n_iter, double *deviances,
                     int64_t dev_cap, int64_t *n_dev);
/* Pass timing: after one real iteration from W0/H0, captures max_iter
 * back-to-back column-pass launches in a HIP graph and times one replay with
 * a HIP event pair, then the same for the row pass; returns the milliseconds
 * and the launch counts.  Leaves the working factors in an unspecified state
 * (the next nmfg_solve_launch starts again from W0/H0). */
int nmfg_solve_profile(nmfg_solve *s, double *col_ms, int64_t *col_launches, double *row_ms,
                       int64_t *row_launches);
/* Pass timing with cold caches (no reference counterpart; measurement):
 * as nmfg_solve_profile, with a read of evict_bytes (1 MiB .. 64 GiB; more
 * than the L2 and MALL together, e.g. 1 GiB) of a scratch buffer before
 * every pass, so that each pass reads V from HBM.  The returned times are
 * the graph's time less that of the reads alone. */
int nmfg_solve_profile_evicted(nmfg_solve *s, int64_t evict_bytes, double *col_ms,
                               int64_t *col_launches, double *row_ms, int64_t *row_launches);
int nmfg_solve_destroy(nmfg_solve *s);

/* Rank sweeps.  No reference counterpart: the reference's notebook loops over
 * ranks x restarts calling nmf_rs.nmf once per solve
 * (examples/build_scaling_notebook.py:157-208).  `count` solves of ranks
 * ranks[i] on the same V, fixed iteration count (prm->max_iter; lee,
 * brunet, hals, ehals, dnmf; the matrix's precision form), advanced by ONE
 * replay of one captured HIP graph: the members are grouped by rank into
 * batched solves padded to a template rank (each member keeps its own rank;
 * its padding columns stay zero), every kernel launch of a group covers all
 * its members, and the groups sit on parallel branches of the graph.  With
 * tune != 0 the plan (grouping, branch count, split counts) is the fastest
 * of a set of candidates timed on the device over 10 iterations, cached per
 * (shape, form, method, ranks) for the process; tune == 0 takes the cached
 * plan or a default.  A member equals the single solve from the same start
 * to roundoff (FP64: within 1e-12 of the oracle). */
typedef struct nmfg_sweep nmfg_sweep;
int nmfg_sweep_create(nmfg_matrix *m, const int32_t *ranks, int32_t count, const nmfg_params *prm,
                      int32_t tune, nmfg_sweep **out);
/* W0[i] (n x ranks[i]) and H0[i] (ranks[i] x p), host, row-major. */
int nmfg_sweep_set_factors(nmfg_sweep *sw, const double *const *W0, const double *const *H0);
/* Enqueues the whole sweep (one graph replay) on the context stream. */
int nmfg_sweep_launch(nmfg_sweep *sw);
/* Waits and copies member i's W and H to W[i], H[i] (NULL entries skipped). */
int nmfg_sweep_fetch(nmfg_sweep *sw, double *const *W, double *const *H);
/* The chosen plan as JSON text (cap bytes including the terminator) and,
 * when autotuned in this call's create, the tuning run's milliseconds. */
int nmfg_sweep_plan(const nmfg_sweep *sw, char *buf, int64_t cap, double *tuned_ms);
int nmfg_sweep_destroy(nmfg_sweep *sw);

/* Step groups (continuous batching of concurrent solves).  No reference
 * counterpart: the reference runs one solve per call.  Solves of any
 * method, rank, matrix, precision form and context on the group's device
 * (factors set) are advanced together on `branches` (1..8) streams: a
 * member joins the branch with the least work and stays on it.  Each
 * nmfg_group_step starts a round on every branch that has members and no
 * round in flight -- one or more chunks of 10 iterations of each of the
 * branch's members (a member whose chunk costs a fraction of the branch's
 * costliest member's runs up to 4 chunks back to back; NMFG_GROUP_CHUNKS,
 * read at create, sets that limit), then one kernel gathers the branch's
 * stop flags and one copy reads them back -- and returns once some round
 * has ended (the branches advance independently, so a short round does not
 * wait for a long one).  Members that stopped (objective-based stop) or
 * reached max_iter get their finishing work and leave; solves added
 * meanwhile join their branch's next round.
 * A member runs exactly the kernels of its solve run alone, in the same
 * order: nmfg_solve_fetch then returns the same bits, n_iter and deviances
 * as nmfg_solve_launch would.  At most 64 members. */
typedef struct nmfg_group nmfg_group;
int nmfg_group_create(nmfg_ctx *ctx, int32_t branches, nmfg_group **out);
/* Adds a solve; it starts at the next step.  Callable from any thread.  The
 * group's work waits for what is queued on the solve's context stream at
 * this call (its upload and factors); from here until a step returns it the
 * caller must not use the solve or free its matrix. */
int nmfg_group_add(nmfg_group *g, nmfg_solve *s);
/* One step (see above): returns after at least one branch's round has
 * ended.  The members that finished are removed and stored
 * in finished[0 .. *n_finished - 1] (cap must be at least the member count,
 * nmfg_group_size); their results are ready for nmfg_solve_fetch.  With no
 * member, returns at once with *n_finished = 0. */
int nmfg_group_step(nmfg_group *g, nmfg_solve **finished, int32_t cap, int32_t *n_finished);
/* Members plus solves waiting to join. */
int nmfg_group_size(nmfg_group *g, int32_t *members);
/* Steps taken, member-steps (the sum over steps of the members advanced)
 * and the most members in one step. */
int nmfg_group_stats(nmfg_group *g, int64_t *steps, int64_t *member_steps, int32_t *max_members);
/* Destroys the group (members still in it are left to their owners, not run further). */
int nmfg_group_destroy(nmfg_group *g);

/* Consensus matrix of `runs` clusterings of p items (labels: runs x p,
 * row-major): C[i][j] = #{runs with label_i == label_j} / runs, p x p host
 * buffer.  Replaces the connectivity_matrix sum and division of
 * examples/build_scaling_notebook.py:121-124, 195-200. */
int nmfg_consensus(nmfg_ctx *ctx, const int32_t *labels, int32_t runs, int64_t p, double *C);

#ifdef __cplusplus
}
#endif
#endif
