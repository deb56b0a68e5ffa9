/*
 * CPU oracle for the dense NMF update path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the arithmetic of omicverse/rust-NMF's CPU path
 * (rust/src/lib.rs).  It is the parity checker for the HIP library and the
 * timed CPU baseline of the benchmark.  Nothing in the product library links
 * or calls it.
 *
 * Every sum starts at 0.0 and adds one term at a time in the order the
 * reference loop uses; products are formed as written; no FMA contraction
 * (built with -ffp-contract=off); every division is an IEEE division.
 * For lee and brunet this makes the oracle bit-identical to the reference
 * arithmetic.  For the HALS family the reference's W^T V and V H^T go
 * through the third-party `matrixmultiply` 0.3.10 crate, whose summation
 * order is not pinned by anything the reference ships; the oracle fixes a
 * forward order there (documented per function below).  The same holds for
 * the dense products of offset (rowSums(W H) in the offset update) and of
 * nsNMF (W S, S H and (W S) H); their other sums follow the reference.
 *
 * Parallelism mirrors the reference's rayon axes: H updates split columns,
 * W updates split rows.  Results do not depend on the thread count because
 * every output element is computed serially by one thread.
 */
#ifndef NMF_ORACLE_H
#define NMF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Method codes (same numbering as the product C ABI). */
enum { ORC_BRUNET = 0, ORC_LEE = 1, ORC_OFFSET = 2, ORC_NSNMF = 3, ORC_HALS = 4, ORC_EHALS = 5,
       ORC_DNMF = 6, ORC_LSNMF = 7, ORC_SNMF_R = 8, ORC_SNMF_L = 9 };
enum { ORC_STOP_MAXITER = 0, ORC_STOP_STATIONARY = 1 };

typedef struct orc_params {
    int32_t algo;
    int64_t max_iter;
    double eps;
    int32_t rescale;
    double sparsity_h, sparsity_w, smoothness_h, smoothness_w;
    int32_t stop;
    double stationary_th;
    int64_t check_interval, check_niter;
    int32_t num_threads;
    double theta;             /* nsNMF smoothing (lib.rs:1921, default 0.5) */
} orc_params;

/* Full driver (lib.rs:1528-1647 lee/brunet, 637-688 hals, 710-790 ehals,
 * 839-906 dnmf).  W (n x r) and H (r x p), row-major, are overwritten with
 * the result.  Deviances are written only for stop == stationary.
 * Returns 0, or -1 on an unsupported algo / bad argument. */
int orc_run(const orc_params *prm, const double *V, int64_t n, int64_t p,
            double *W, double *H, int32_t r, int64_t *n_iter,
            double *dev, int64_t dev_cap, int64_t *n_dev);

/* Full driver with the offset vector (lib.rs:1649-1759 for ORC_OFFSET; any
 * other algo ignores `off`).  off (n) is in/out: with off_given == 0 it is
 * first set to rowMeans(V) (lib.rs:1669-1683), else taken as the initial
 * offset.  orc_run runs ORC_OFFSET from rowMeans(V) and drops the offset. */
int orc_run_offset(const orc_params *prm, const double *V, int64_t n, int64_t p,
                   double *W, double *H, int32_t r, double *off, int32_t off_given,
                   int64_t *n_iter, double *dev, int64_t dev_cap, int64_t *n_dev);

/* lsNMF driver (lib.rs:908-1043): weighted Frobenius MU with the per-entry
 * weight matrix S (n x p, row-major).  W^T (V.S), W^T (WH.S), (V.S) H^T and
 * (WH.S) H^T go through matrixmultiply in the reference; oracle order:
 * forward over the inner index, W H forward over k. */
int orc_run_lsnmf(const orc_params *prm, const double *V, const double *S, int64_t n, int64_t p,
                  double *W, double *H, int32_t r, int64_t *n_iter, double *dev, int64_t dev_cap,
                  int64_t *n_dev);

/* FCNNLS (rust/src/fcnnls.rs:270-501): min ||Y - X K|| s.t. K >= 0 with
 * X (nobs x l), Y (nobs x p), row-major; K (l x p) out, Pset (l x p, 0/1)
 * out if not NULL.  X^T X, X^T Y and the KKT product C^T C K go through
 * ndarray dot in the reference; oracle order: forward over the inner index.
 * Returns 0, or -1 for l > 64. */
int orc_fcnnls(const double *X, const double *Y, int64_t nobs, int32_t l, int64_t p, double eps,
               double *K, uint8_t *pset);

/* Kim-Park sparse NMF driver (lib.rs:1198-1334) for ORC_SNMF_R / ORC_SNMF_L
 * through orc_run (eta = smoothness_w, beta = sparsity_h for snmf/R;
 * smoothness_h, sparsity_w for snmf/L; eps is FCNNLS's tolerance).  A
 * deviance 0.5 ||A - W H||^2 (row-major sum) is recorded every iteration
 * whatever the stop; there is no early stop (lib.rs:1313-1321). */

/* Single steps (lib.rs:117-144, 146-204, 247-276, 278-322). */
void orc_euclidean_update_h(const double *V, int64_t n, int64_t p,
                            const double *W, const double *H, int32_t r,
                            double eps, double *Hout, int32_t nthreads);
void orc_euclidean_update_w(const double *V, int64_t n, int64_t p,
                            const double *W, const double *H, int32_t r,
                            double eps, double *Wout, int32_t nthreads);
void orc_divergence_update_h(const double *V, int64_t n, int64_t p,
                             const double *W, const double *H, int32_t r,
                             double *Hout, int32_t nthreads);
void orc_divergence_update_w(const double *V, int64_t n, int64_t p,
                             const double *W, const double *H, int32_t r,
                             double *Wout, int32_t nthreads);

/* Objectives over a dense W H (lib.rs:1403-1445). */
double orc_euclidean_distance(const double *V, int64_t n, int64_t p,
                              const double *W, const double *H, int32_t r);
double orc_kl_divergence(const double *V, int64_t n, int64_t p,
                         const double *W, const double *H, int32_t r);

/* nsNMF smoothing matrix (lib.rs:457-466). */
void orc_smoothing_matrix(int32_t r, double theta, double *S);

#ifdef __cplusplus
}
#endif
#endif
