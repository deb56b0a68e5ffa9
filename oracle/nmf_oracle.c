/*
 * CPU oracle for the dense NMF update path -- TEST INFRASTRUCTURE ONLY.
 * Used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg; never by the product library.  See nmf_oracle.h for the contract.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math -fPIC -shared
 *        -pthread (oracle/Makefile).
 */
#include "nmf_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Static-chunk parallel-for over [0, N): thread t gets [t*N/T, (t+1)*N/T).  */
/* ------------------------------------------------------------------------ */
typedef void (*range_fn)(void *ctx, int64_t lo, int64_t hi);

typedef struct {
    range_fn fn;
    void *ctx;
    int64_t lo, hi;
} par_job;

static void *par_tramp(void *a) {
    par_job *j = (par_job *)a;
    j->fn(j->ctx, j->lo, j->hi);
    return NULL;
}

static void par_for(int64_t N, int nthreads, range_fn fn, void *ctx) {
    if (nthreads <= 1 || N < 2) {
        fn(ctx, 0, N);
        return;
    }
    if (nthreads > N) nthreads = (int)N;
    if (nthreads > 256) nthreads = 256;
    pthread_t tid[256];
    par_job job[256];
    for (int t = 0; t < nthreads; ++t) {
        job[t].fn = fn;
        job[t].ctx = ctx;
        job[t].lo = N * t / nthreads;
        job[t].hi = N * (t + 1) / nthreads;
    }
    for (int t = 1; t < nthreads; ++t) pthread_create(&tid[t], NULL, par_tramp, &job[t]);
    par_tramp(&job[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(tid[t], NULL);
}

static int default_threads(int req) { return req > 0 ? req : 1; }

/* Row-major transposed copy, rows x cols -> cols x rows (lib.rs:25-38). */
static double *transpose_new(const double *a, int64_t rows, int64_t cols) {
    double *t = (double *)malloc(sizeof(double) * (size_t)(rows * cols > 0 ? rows * cols : 1));
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t j = 0; j < cols; ++j) t[j * rows + i] = a[i * cols + j];
    return t;
}

static double *dup(const double *a, int64_t len) {
    double *d = (double *)malloc(sizeof(double) * (size_t)(len > 0 ? len : 1));
    if (len > 0) memcpy(d, a, sizeof(double) * (size_t)len);
    return d;
}

/* ======================================================================== */
/* lee (Frobenius multiplicative update)                                    */
/* ======================================================================== */

/* G = W^T W with reverse-u summation (lib.rs:1620-1631, also 263-273). */
static void lee_wtw(const double *Wt, int64_t n, int r, double *G) {
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < r; ++j) {
            double s = 0.0;
            for (int64_t u = n - 1; u >= 0; --u) s += Wt[i * n + u] * Wt[j * n + u];
            G[i * r + j] = s;
        }
}

typedef struct {
    const double *Vt, *Wt, *H, *G;
    double *Hp;
    int64_t n, p;
    int r;
    double eps;
    const double *addon; /* offset: W^T off (r), added to den; NULL for lee */
} lee_h_ctx;

/* euclidean_update_h_inner (lib.rs:210-245): per column j, for each i,
 *   num = sum_{u=n-1..0} W[u,i] V[u,j];  den = sum_{l=r-1..0} G[i,l] H[l,j];
 *   H'[i,j] = max(H[i,j]*num, eps) / (den + eps)   (t > eps ? t : eps). */
static void lee_h_range(void *c, int64_t lo, int64_t hi) {
    lee_h_ctx *x = (lee_h_ctx *)c;
    const int64_t n = x->n, p = x->p;
    const int r = x->r;
    for (int64_t j = lo; j < hi; ++j) {
        const double *vt = x->Vt + j * n;
        for (int i = 0; i < r; ++i) {
            const double *wt = x->Wt + (int64_t)i * n;
            double num = 0.0;
            for (int64_t u = n - 1; u >= 0; --u) num += wt[u] * vt[u];
            double den = 0.0;
            for (int l = r - 1; l >= 0; --l) den += x->G[i * r + l] * x->H[(int64_t)l * p + j];
            if (x->addon) den += x->addon[i]; /* offset_update_h_inner (lib.rs:351-352) */
            double t = x->H[(int64_t)i * p + j] * num;
            x->Hp[(int64_t)i * p + j] = (t > x->eps ? t : x->eps) / (den + x->eps);
        }
    }
}

typedef struct {
    const double *V, *W, *H, *K;
    double *Wp;
    int64_t n, p;
    int r;
    double eps;
    const double *off, *rsh; /* offset: off (n) and rowSums(H) (r); NULL for lee */
} lee_w_ctx;

/* euclidean_update_w_impl row body (lib.rs:305-320): per row i, for each j,
 *   num = sum_{u=p-1..0} V[i,u] H[j,u];  den = sum_{l=r-1..0} W[i,l] K[l,j];
 *   W'[i,j] = (t < eps ? eps : t) / (den + eps),  t = W[i,j]*num. */
static void lee_w_range(void *c, int64_t lo, int64_t hi) {
    lee_w_ctx *x = (lee_w_ctx *)c;
    const int64_t p = x->p;
    const int r = x->r;
    for (int64_t i = lo; i < hi; ++i) {
        const double *v = x->V + i * p;
        for (int j = 0; j < r; ++j) {
            const double *h = x->H + (int64_t)j * p;
            double num = 0.0;
            for (int64_t u = p - 1; u >= 0; --u) num += v[u] * h[u];
            double den = 0.0;
            for (int l = r - 1; l >= 0; --l) den += x->W[i * r + l] * x->K[l * r + j];
            if (x->off) den += x->off[i] * x->rsh[j]; /* offset_update_w_impl (lib.rs:437-438) */
            double t = x->W[i * r + j] * num;
            x->Wp[i * r + j] = (t < x->eps ? x->eps : t) / (den + x->eps);
        }
    }
}

/* K = H H^T, K[i][j] = sum_{u=p-1..0} H[j,u] H[i,u] (lib.rs:292-302). */
static void lee_hht(const double *H, int64_t p, int r, double *K) {
    for (int j = 0; j < r; ++j)
        for (int i = 0; i < r; ++i) {
            double s = 0.0;
            for (int64_t u = p - 1; u >= 0; --u) s += H[(int64_t)j * p + u] * H[(int64_t)i * p + u];
            K[i * r + j] = s;
        }
}

static void lee_update_w(const double *V, int64_t n, int64_t p, const double *W, const double *H,
                         int r, double eps, double *Wp, int nt) {
    double *K = (double *)malloc(sizeof(double) * (size_t)(r * r));
    lee_hht(H, p, r, K);
    lee_w_ctx c = {V, W, H, K, Wp, n, p, r, eps, NULL, NULL};
    par_for(n, nt, lee_w_range, &c);
    free(K);
}

/* col_rescale (lib.rs:1460-1470): forward column sum, divide when > 0. */
static void col_rescale(double *W, int64_t n, int r) {
    for (int j = 0; j < r; ++j) {
        double s = 0.0;
        for (int64_t i = 0; i < n; ++i) s += W[i * r + j];
        if (s > 0.0)
            for (int64_t i = 0; i < n; ++i) W[i * r + j] = W[i * r + j] / s;
    }
}

/* ======================================================================== */
/* brunet (KL multiplicative update)                                        */
/* ======================================================================== */
typedef struct {
    const double *Vt, *W, *Wt, *H, *sum_w;
    double *Hp;
    int64_t n, p;
    int r;
} div_h_ctx;

/* divergence_update_h_inner (lib.rs:74-115). */
static void div_h_range(void *c, int64_t lo, int64_t hi) {
    div_h_ctx *x = (div_h_ctx *)c;
    const int64_t n = x->n, p = x->p;
    const int r = x->r;
    double *q = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double hcol[64];
    double *hc = r <= 64 ? hcol : (double *)malloc(sizeof(double) * (size_t)r);
    for (int64_t j = lo; j < hi; ++j) {
        for (int k = 0; k < r; ++k) hc[k] = x->H[(int64_t)k * p + j];
        const double *vt = x->Vt + j * n;
        for (int64_t u = 0; u < n; ++u) {
            double wh = 0.0;
            for (int k = 0; k < r; ++k) wh += x->W[u * r + k] * hc[k];
            q[u] = wh != 0.0 ? vt[u] / wh : 0.0;
        }
        for (int i = 0; i < r; ++i) {
            const double *wt = x->Wt + (int64_t)i * n;
            double tmp = 0.0;
            for (int64_t u = 0; u < n; ++u) tmp += wt[u] * q[u];
            x->Hp[(int64_t)i * p + j] = hc[i] * tmp / x->sum_w[i];
        }
    }
    if (hc != hcol) free(hc);
    free(q);
}

typedef struct {
    const double *V, *W, *H, *Ht, *sum_h;
    double *Wp;
    int64_t n, p;
    int r;
} div_w_ctx;

/* divergence_update_w_impl row body (lib.rs:184-202). */
static void div_w_range(void *c, int64_t lo, int64_t hi) {
    div_w_ctx *x = (div_w_ctx *)c;
    const int64_t p = x->p;
    const int r = x->r;
    double *q = (double *)malloc(sizeof(double) * (size_t)(p > 0 ? p : 1));
    for (int64_t i = lo; i < hi; ++i) {
        const double *v = x->V + i * p;
        const double *w = x->W + i * r;
        for (int64_t u = 0; u < p; ++u) {
            double wh = 0.0;
            for (int k = 0; k < r; ++k) wh += w[k] * x->Ht[u * r + k];
            q[u] = wh != 0.0 ? v[u] / wh : 0.0;
        }
        for (int j = 0; j < r; ++j) {
            const double *h = x->H + (int64_t)j * p;
            double tmp = 0.0;
            for (int64_t u = 0; u < p; ++u) tmp += h[u] * q[u];
            x->Wp[i * r + j] = w[j] * tmp / x->sum_h[j];
        }
    }
    free(q);
}

static void colsum_fwd(const double *W, int64_t n, int r, double *s) {
    for (int k = 0; k < r; ++k) {
        double a = 0.0;
        for (int64_t u = 0; u < n; ++u) a += W[u * r + k];
        s[k] = a;
    }
}

static void div_update_h_with(const double *Vt, const double *W, const double *H, int64_t n,
                              int64_t p, int r, double *Hp, int nt) {
    double *sum_w = (double *)malloc(sizeof(double) * (size_t)r);
    colsum_fwd(W, n, r, sum_w);
    double *Wt = transpose_new(W, n, r);
    div_h_ctx c = {Vt, W, Wt, H, sum_w, Hp, n, p, r};
    par_for(p, nt, div_h_range, &c);
    free(Wt);
    free(sum_w);
}

/* lib.rs:146-204: rowsum H (forward), H^T copy, per-row update. */
static void div_update_w(const double *V, int64_t n, int64_t p, const double *W, const double *H,
                         int r, double *Wp, int nt) {
    double *sum_h = (double *)malloc(sizeof(double) * (size_t)r);
    for (int k = 0; k < r; ++k) {
        double s = 0.0;
        for (int64_t u = 0; u < p; ++u) s += H[(int64_t)k * p + u];
        sum_h[k] = s;
    }
    double *Ht = transpose_new(H, r, p);
    div_w_ctx c = {V, W, H, Ht, sum_h, Wp, n, p, r};
    par_for(n, nt, div_w_range, &c);
    free(Ht);
    free(sum_h);
}

/* pmax_inplace (lib.rs:1451-1458). */
static void pmax(double *x, int64_t len, double eps) {
    for (int64_t i = 0; i < len; ++i)
        if (x[i] < eps) x[i] = eps;
}

/* ======================================================================== */
/* HALS family                                                              */
/* ======================================================================== */
typedef struct {
    const double *Wt, *Vt;
    double *A;
    int64_t n, p;
    int r;
} wtv_ctx;

/* A = W^T V (lib.rs:490-517).  The reference reaches matrixmultiply's dgemm
 * here; its order is not pinned.  Oracle order: A[k,u] = sum_{i=0..n-1}
 * W[i,k] V[i,u], forward. */
static void wtv_range(void *c, int64_t lo, int64_t hi) {
    wtv_ctx *x = (wtv_ctx *)c;
    const int64_t n = x->n, p = x->p;
    for (int64_t u = lo; u < hi; ++u) {
        const double *vt = x->Vt + u * n;
        for (int k = 0; k < x->r; ++k) {
            const double *wt = x->Wt + (int64_t)k * n;
            double s = 0.0;
            for (int64_t i = 0; i < n; ++i) s += wt[i] * vt[i];
            x->A[(int64_t)k * p + u] = s;
        }
    }
}

typedef struct {
    const double *V, *H;
    double *B;
    int64_t n, p;
    int r;
} vht_ctx;

/* B = V H^T (lib.rs:521-540), matrixmultiply in the reference.  Oracle
 * order: B[i,k] = sum_{u=0..p-1} V[i,u] H[k,u], forward. */
static void vht_range(void *c, int64_t lo, int64_t hi) {
    vht_ctx *x = (vht_ctx *)c;
    const int64_t p = x->p;
    for (int64_t i = lo; i < hi; ++i) {
        const double *v = x->V + i * p;
        for (int k = 0; k < x->r; ++k) {
            const double *h = x->H + (int64_t)k * p;
            double s = 0.0;
            for (int64_t u = 0; u < p; ++u) s += v[u] * h[u];
            x->B[i * x->r + k] = s;
        }
    }
}

/* compute_wtw_from_wt (lib.rs:542-560): forward sums. */
static void hals_wtw(const double *Wt, int64_t n, int r, double *G) {
    for (int k = 0; k < r; ++k)
        for (int j = 0; j < r; ++j) {
            double s = 0.0;
            for (int64_t i = 0; i < n; ++i) s += Wt[(int64_t)k * n + i] * Wt[(int64_t)j * n + i];
            G[k * r + j] = s;
        }
}

/* compute_hht (lib.rs:562-579): forward sums. */
static void hals_hht(const double *H, int64_t p, int r, double *K) {
    for (int k = 0; k < r; ++k)
        for (int j = 0; j < r; ++j) {
            double s = 0.0;
            for (int64_t u = 0; u < p; ++u) s += H[(int64_t)k * p + u] * H[(int64_t)j * p + u];
            K[k * r + j] = s;
        }
}

typedef struct {
    double *H;
    const double *A, *G;
    int64_t p;
    int r;
    double lambda, gamma, eps;
    int reg;
} hals_h_ctx;

/* hals_update_h (lib.rs:581-607) and hals_update_h_reg (lib.rs:1067-1092):
 * Gauss-Seidel over k within each column.  Rust f64::max == C fmax. */
static void hals_h_range(void *c, int64_t lo, int64_t hi) {
    hals_h_ctx *x = (hals_h_ctx *)c;
    const int64_t p = x->p;
    const int r = x->r;
    for (int64_t u = lo; u < hi; ++u) {
        for (int k = 0; k < r; ++k) {
            double s = x->reg ? x->A[(int64_t)k * p + u] - x->lambda : x->A[(int64_t)k * p + u];
            for (int j = 0; j < r; ++j)
                if (j != k) s -= x->G[k * r + j] * x->H[(int64_t)j * p + u];
            double b = x->reg ? fmax(x->G[k * r + k] + x->gamma, x->eps) : fmax(x->G[k * r + k], x->eps);
            double val = s / b;
            x->H[(int64_t)k * p + u] = val > x->eps ? val : x->eps;
        }
    }
}

typedef struct {
    double *W;
    const double *B, *K;
    int r;
    double lambda, gamma, eps;
    int reg;
} hals_w_ctx;

/* hals_update_w (lib.rs:610-635) and hals_update_w_reg (lib.rs:1094-1119). */
static void hals_w_range(void *c, int64_t lo, int64_t hi) {
    hals_w_ctx *x = (hals_w_ctx *)c;
    const int r = x->r;
    for (int64_t i = lo; i < hi; ++i) {
        double *w = x->W + i * r;
        for (int k = 0; k < r; ++k) {
            double s = x->reg ? x->B[i * r + k] - x->lambda : x->B[i * r + k];
            for (int j = 0; j < r; ++j)
                if (j != k) s -= x->K[k * r + j] * w[j];
            double d = x->reg ? fmax(x->K[k * r + k] + x->gamma, x->eps) : fmax(x->K[k * r + k], x->eps);
            double val = s / d;
            w[k] = val > x->eps ? val : x->eps;
        }
    }
}

/* One HALS iteration: H sweep then W sweep (lib.rs:666-677, 866-872). */
static void hals_iter(const double *V, const double *Vt, int64_t n, int64_t p, double *W, double *H,
                      int r, double eps, int reg, double lh, double lw, double gh, double gw, int nt,
                      double *A, double *B, double *G, double *K) {
    double *Wt = transpose_new(W, n, r);
    wtv_ctx wc = {Wt, Vt, A, n, p, r};
    par_for(p, nt, wtv_range, &wc);
    hals_wtw(Wt, n, r, G);
    free(Wt);
    hals_h_ctx hc = {H, A, G, p, r, lh, gh, eps, reg};
    par_for(p, nt, hals_h_range, &hc);
    vht_ctx vc = {V, H, B, n, p, r};
    par_for(n, nt, vht_range, &vc);
    hals_hht(H, p, r, K);
    hals_w_ctx wc2 = {W, B, K, r, lw, gw, eps, reg};
    par_for(n, nt, hals_w_range, &wc2);
}

/* extract_diagonal (lib.rs:808-824). */
static void extract_diagonal(double *W, double *H, int64_t n, int64_t p, int r, double *d) {
    for (int k = 0; k < r; ++k) {
        double sw = 0.0, sh = 0.0;
        for (int64_t i = 0; i < n; ++i) sw += W[i * r + k] * W[i * r + k];
        for (int64_t j = 0; j < p; ++j) sh += H[(int64_t)k * p + j] * H[(int64_t)k * p + j];
        double nw = sqrt(sw), nh = sqrt(sh);
        double dk = nw * nh;
        if (nw > 0.0)
            for (int64_t i = 0; i < n; ++i) W[i * r + k] /= nw;
        if (nh > 0.0)
            for (int64_t j = 0; j < p; ++j) H[(int64_t)k * p + j] /= nh;
        d[k] = dk;
    }
}

/* ======================================================================== */
/* Objectives and stationary stop                                           */
/* ======================================================================== */

/* Dense W H.  The reference uses matrixmultiply (wh_dense, lib.rs:40-61);
 * oracle order: WH[i,j] = sum_{k=0..r-1} W[i,k] H[k,j], forward. */
static double wh_at(const double *W, const double *H, int64_t p, int r, int64_t i, int64_t j) {
    double s = 0.0;
    for (int k = 0; k < r; ++k) s += W[i * r + k] * H[(int64_t)k * p + j];
    return s;
}

/* euclidean_distance (lib.rs:1432-1445): column-major, 0.5 * sum d^2. */
double orc_euclidean_distance(const double *V, int64_t n, int64_t p, const double *W,
                              const double *H, int32_t r) {
    double s = 0.0;
    for (int64_t j = 0; j < p; ++j)
        for (int64_t i = 0; i < n; ++i) {
            double d = V[i * p + j] - wh_at(W, H, p, r, i, j);
            s += d * d;
        }
    return 0.5 * s;
}

/* kl_divergence (lib.rs:1406-1429): v>0 & wh>0 adds v ln(v/wh) - v + wh;
 * v>0 & wh<=0 adds nothing; v<=0 adds wh. */
double orc_kl_divergence(const double *V, int64_t n, int64_t p, const double *W, const double *H,
                         int32_t r) {
    double s = 0.0;
    for (int64_t j = 0; j < p; ++j)
        for (int64_t i = 0; i < n; ++i) {
            double vv = V[i * p + j];
            double ww = wh_at(W, H, p, r, i, j);
            if (vv > 0.0) {
                if (ww > 0.0) s += vv * log(vv / ww) - vv + ww;
            } else {
                s += ww;
            }
        }
    return s;
}

/* StationaryStop (lib.rs:1336-1401). */
typedef struct {
    int64_t interval, niter;
    double th, last_max, last_min;
    int64_t n_collected;
} stat_stop;

static void ss_reset(stat_stop *s) {
    s->last_max = -INFINITY;
    s->last_min = INFINITY;
    s->n_collected = 0;
}

static void ss_record(stat_stop *s, double v) {
    if (v > s->last_max) s->last_max = v;
    if (v < s->last_min) s->last_min = v;
    s->n_collected += 1;
}

static int ss_should_stop(stat_stop *s, int64_t i, double d) {
    if (isnan(d)) return 1;
    if (i == 0) {
        ss_reset(s);
        ss_record(s, d);
        return 0;
    }
    if (s->n_collected == 0 && i % s->interval != 0) return 0;
    ss_record(s, d);
    if (s->n_collected == s->niter + 1) {
        double crit = fabs(s->last_max - s->last_min) / (double)s->niter;
        if (crit <= s->th) return 1;
        ss_reset(s);
    }
    return 0;
}

/* ======================================================================== */
/* offset: lee with a per-row offset (Badea 2008)                           */
/* ======================================================================== */
typedef struct {
    const double *V, *W, *H;
    double *off;
    int64_t p;
    int r;
    double eps;
} off_upd_ctx;

/* Offset update (lib.rs:1729-1741): per row, forward sums over j of V[i,j]
 * and of (W H)[i,j] + off[i];  off[i] = off[i] * max(rowsum V, eps) /
 * (rowsum(WH + off) + eps).  W H goes through wh_dense (matrixmultiply) in
 * the reference; oracle order: wh_at. */
static void off_upd_range(void *c, int64_t lo, int64_t hi) {
    off_upd_ctx *x = (off_upd_ctx *)c;
    for (int64_t i = lo; i < hi; ++i) {
        double rv = 0.0, rw = 0.0;
        for (int64_t j = 0; j < x->p; ++j) {
            rv += x->V[i * x->p + j];
            rw += wh_at(x->W, x->H, x->p, x->r, i, j) + x->off[i];
        }
        const double num = rv > x->eps ? rv : x->eps;
        x->off[i] = x->off[i] * num / (rw + x->eps);
    }
}

/* One offset iteration (lib.rs:1711-1741): H via offset_update_h_inner with
 * den_addon = W^T off (reverse u, lib.rs:1715-1720), W via
 * offset_update_w_impl (lib.rs:393-450), then the offset update. */
static void offset_iter(const double *V, const double *Vt, int64_t n, int64_t p, double *W,
                        double *H, int r, double eps, double *off, int nt, double *G, double *K,
                        double *Hn, double *Wn, double *addon, double *rsh) {
    double *Wt = transpose_new(W, n, r);
    lee_wtw(Wt, n, r, G);
    for (int i = 0; i < r; ++i) {
        double s = 0.0;
        for (int64_t u = n - 1; u >= 0; --u) s += Wt[(int64_t)i * n + u] * off[u];
        addon[i] = s;
    }
    lee_h_ctx hc = {Vt, Wt, H, G, Hn, n, p, r, eps, addon};
    par_for(p, nt, lee_h_range, &hc);
    free(Wt);
    memcpy(H, Hn, sizeof(double) * (size_t)(r * p));
    lee_hht(H, p, r, K);
    for (int i = 0; i < r; ++i) { /* rowSumsH, reverse u (lib.rs:417-425) */
        double s = 0.0;
        for (int64_t u = p - 1; u >= 0; --u) s += H[(int64_t)i * p + u];
        rsh[i] = s;
    }
    lee_w_ctx wc = {V, W, H, K, Wn, n, p, r, eps, off, rsh};
    par_for(n, nt, lee_w_range, &wc);
    memcpy(W, Wn, sizeof(double) * (size_t)(n * r));
    off_upd_ctx oc = {V, W, H, off, p, r, eps};
    par_for(n, nt, off_upd_range, &oc);
}

/* euclidean_distance (lib.rs:1432-1445) of V against W H + off[i] per row
 * (lib.rs:1688-1694, 1744-1750). */
static double euclid_offset(const double *V, int64_t n, int64_t p, const double *W, const double *H,
                            int r, const double *off) {
    double s = 0.0;
    for (int64_t j = 0; j < p; ++j)
        for (int64_t i = 0; i < n; ++i) {
            double d = V[i * p + j] - (wh_at(W, H, p, r, i, j) + off[i]);
            s += d * d;
        }
    return 0.5 * s;
}

/* ======================================================================== */
/* nsNMF: brunet on W S and S H (Pascual-Montano 2006)                      */
/* ======================================================================== */

/* W S and S H are ndarray `dot` (matrixmultiply) in the reference
 * (lib.rs:1784-1788, 1805, 1815); oracle order: forward over the inner
 * index. */
static void dense_ws(const double *W, const double *S, int64_t n, int r, double *WS) {
    for (int64_t u = 0; u < n; ++u)
        for (int k = 0; k < r; ++k) {
            double s = 0.0;
            for (int l = 0; l < r; ++l) s += W[u * r + l] * S[l * r + k];
            WS[u * r + k] = s;
        }
}

static void dense_sh(const double *S, const double *H, int64_t p, int r, double *SH) {
    for (int k = 0; k < r; ++k)
        for (int64_t j = 0; j < p; ++j) {
            double s = 0.0;
            for (int l = 0; l < r; ++l) s += S[k * r + l] * H[(int64_t)l * p + j];
            SH[(int64_t)k * p + j] = s;
        }
}

/* One nsNMF iteration (lib.rs:1799-1820): H by divergence_update_h_inner
 * with W S as the basis and its forward column sums, W by
 * divergence_update_w_impl with S H, then col_rescale(W). */
static void nsnmf_iter(const double *V, const double *Vt, int64_t n, int64_t p, double *W,
                       double *H, int r, const double *S, int nt, double *WS, double *SH,
                       double *Hn, double *Wn) {
    dense_ws(W, S, n, r, WS);
    div_update_h_with(Vt, WS, H, n, p, r, Hn, nt);
    memcpy(H, Hn, sizeof(double) * (size_t)(r * p));
    dense_sh(S, H, p, r, SH);
    div_update_w(V, n, p, W, SH, r, Wn, nt);
    memcpy(W, Wn, sizeof(double) * (size_t)(n * r));
    col_rescale(W, n, r);
}

/* kl_divergence(V, (W S) H) (lib.rs:1784-1788, 1791-1792). */
static double kl_nsnmf(const double *V, int64_t n, int64_t p, const double *W, const double *H,
                       int r, const double *S, double *WS) {
    dense_ws(W, S, n, r, WS);
    return orc_kl_divergence(V, n, p, WS, H, r);
}

/* ======================================================================== */
/* lsNMF: weighted Frobenius MU (Wang, Kossenkov, Ochs 2006)                */
/* ======================================================================== */
typedef struct {
    const double *V, *S, *W, *H;
    double *H2;
    int64_t n, p;
    int r;
    double eps;
} ls_h_ctx;

/* H update (lib.rs:985-1006) for columns [lo, hi): est = W H, num = W^T
 * (V.S), den = W^T (est.S); H' = max(H num, eps) / (den + eps). */
static void ls_h_range(void *c, int64_t lo, int64_t hi) {
    ls_h_ctx *x = (ls_h_ctx *)c;
    const int64_t n = x->n, p = x->p;
    const int r = x->r;
    double numd[128];
    for (int64_t u = lo; u < hi; ++u) {
        double *num = numd, *den = numd + 64;
        for (int k = 0; k < r; ++k) num[k] = den[k] = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            const double s = x->S[i * p + u];
            const double vw = x->V[i * p + u] * s;
            const double ew = wh_at(x->W, x->H, p, r, i, u) * s;
            for (int k = 0; k < r; ++k) {
                num[k] += x->W[i * r + k] * vw;
                den[k] += x->W[i * r + k] * ew;
            }
        }
        for (int k = 0; k < r; ++k) {
            const double t = x->H[(int64_t)k * p + u] * num[k];
            const double nm = t > x->eps ? t : x->eps;
            x->H2[(int64_t)k * p + u] = nm / (den[k] + x->eps);
        }
    }
}

typedef struct {
    const double *V, *S, *W, *H;
    double *W2;
    int64_t n, p;
    int r;
    double eps;
} ls_w_ctx;

/* W update (lib.rs:1008-1024) for rows [lo, hi): est = W H', num = (V.S)
 * H'^T, den = (est.S) H'^T; W' = max(W num, eps) / (den + eps). */
static void ls_w_range(void *c, int64_t lo, int64_t hi) {
    ls_w_ctx *x = (ls_w_ctx *)c;
    const int64_t p = x->p;
    const int r = x->r;
    double numd[128];
    for (int64_t i = lo; i < hi; ++i) {
        double *num = numd, *den = numd + 64;
        for (int k = 0; k < r; ++k) num[k] = den[k] = 0.0;
        for (int64_t u = 0; u < p; ++u) {
            const double s = x->S[i * p + u];
            const double vw = x->V[i * p + u] * s;
            const double ew = wh_at(x->W, x->H, p, r, i, u) * s;
            for (int k = 0; k < r; ++k) {
                num[k] += vw * x->H[(int64_t)k * p + u];
                den[k] += ew * x->H[(int64_t)k * p + u];
            }
        }
        for (int k = 0; k < r; ++k) {
            const double t = x->W[i * r + k] * num[k];
            const double nm = t > x->eps ? t : x->eps;
            x->W2[i * r + k] = nm / (den[k] + x->eps);
        }
    }
}

/* 0.5 sum ((V - W H) S)^2, row-major order (lib.rs:962-973). */
static double ls_objective(const double *V, const double *S, int64_t n, int64_t p, const double *W,
                           const double *H, int r) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < p; ++j) {
            const double d = (V[i * p + j] - wh_at(W, H, p, r, i, j)) * S[i * p + j];
            s += d * d;
        }
    return 0.5 * s;
}

int orc_run_lsnmf(const orc_params *prm, const double *V, const double *S, int64_t n, int64_t p,
                  double *W, double *H, int32_t r, int64_t *n_iter, double *dev, int64_t dev_cap,
                  int64_t *n_dev) {
    if (!prm || !S || r <= 0 || r > 64 || n <= 0 || p <= 0) return -1;
    const int nt = default_threads(prm->num_threads);
    const double eps = prm->eps;
    stat_stop ss;
    ss.interval = prm->check_interval > 0 ? prm->check_interval : 1;
    ss.niter = prm->check_niter;
    ss.th = prm->stationary_th;
    ss_reset(&ss);
    int64_t nd = 0, iter = 0;
    const int stationary = prm->stop == ORC_STOP_STATIONARY;
    double *H2 = (double *)malloc(sizeof(double) * (size_t)(r * p));
    double *W2 = (double *)malloc(sizeof(double) * (size_t)(n * r));
    int stop0 = 0;
    if (stationary) {
        double d = ls_objective(V, S, n, p, W, H, r);
        if (dev && nd < dev_cap) dev[nd] = d;
        nd++;
        stop0 = ss_should_stop(&ss, 0, d);
    }
    while (!stop0 && iter < prm->max_iter) {
        iter += 1;
        ls_h_ctx hc = {V, S, W, H, H2, n, p, r, eps};
        par_for(p, nt, ls_h_range, &hc);
        memcpy(H, H2, sizeof(double) * (size_t)(r * p));
        ls_w_ctx wc = {V, S, W, H, W2, n, p, r, eps};
        par_for(n, nt, ls_w_range, &wc);
        memcpy(W, W2, sizeof(double) * (size_t)(n * r));
        if (stationary) {
            double d = ls_objective(V, S, n, p, W, H, r);
            if (dev && nd < dev_cap) dev[nd] = d;
            nd++;
            if (ss_should_stop(&ss, iter, d)) break;
        }
    }
    free(H2);
    free(W2);
    if (n_iter) *n_iter = iter;
    if (n_dev) *n_dev = nd;
    return 0;
}

/* ======================================================================== */
/* FCNNLS (fcnnls.rs) and Kim-Park sparse NMF                               */
/* ======================================================================== */

/* cholesky_inplace (fcnnls.rs:40-64): lower factor of the r x r matrix a
 * (row stride r) in place; 0 when a diagonal is <= 0. */
static int chol_inplace(double *a, int r) {
    for (int j = 0; j < r; ++j) {
        double s = a[j * r + j];
        for (int k = 0; k < j; ++k) s -= a[j * r + k] * a[j * r + k];
        if (s <= 0.0) return 0;
        const double l = sqrt(s);
        a[j * r + j] = l;
        for (int i = j + 1; i < r; ++i) {
            double t = a[i * r + j];
            for (int k = 0; k < j; ++k) t -= a[i * r + k] * a[j * r + k];
            a[i * r + j] = t / l;
        }
    }
    return 1;
}

/* cholesky_solve (fcnnls.rs:67-86). */
static void chol_solve(const double *L, int r, double *b) {
    for (int i = 0; i < r; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[i * r + k] * b[k];
        b[i] = s / L[i * r + i];
    }
    for (int i = r - 1; i >= 0; --i) {
        double s = b[i];
        for (int k = i + 1; k < r; ++k) s -= L[k * r + i] * b[k];
        b[i] = s / L[i * r + i];
    }
}

/* Solve A x = b for one right-hand side b (in place): Cholesky, or the
 * regularised fallback of pseudoinv_solve (fcnnls.rs:95-131: add tol, 10 tol,
 * 100 tol, ... to the diagonal until the factor exists; zeros past 1e6). */
static void spd_solve(const double *A, int r, double *b, int pseudo, double *work) {
    memcpy(work, A, sizeof(double) * (size_t)(r * r));
    if (!pseudo && chol_inplace(work, r)) {
        chol_solve(work, r, b);
        return;
    }
    double *areg = work + r * r;
    memcpy(areg, A, sizeof(double) * (size_t)(r * r));
    double er = 1e-12;
    for (;;) {
        memcpy(work, areg, sizeof(double) * (size_t)(r * r));
        if (chol_inplace(work, r)) break;
        for (int i = 0; i < r; ++i) areg[i * r + i] += er;
        er *= 10.0;
        if (er > 1e6) {
            for (int i = 0; i < r; ++i) b[i] = 0.0;
            return;
        }
    }
    chol_solve(work, r, b);
}

/* cssls (fcnnls.rs:143-266) for the columns listed in cols: K[:, j] =
 * restricted solve on the passive rows of column j (all rows when pset is
 * NULL), zero elsewhere.  Grouping columns by pattern in the reference only
 * shares the factorisation; each column's arithmetic is the same. */
static void cssls_cols(const double *ctc, const double *cta, int l, int64_t p, const uint8_t *pset,
                       const int64_t *cols, int64_t ncols, double *K) {
    double sub[64 * 64], rhs[64], work[2 * 64 * 64];
    int act[64];
    for (int64_t c = 0; c < ncols; ++c) {
        const int64_t j = cols[c];
        int na = 0;
        for (int i = 0; i < l; ++i)
            if (!pset || pset[(int64_t)i * p + j]) act[na++] = i;
        for (int i = 0; i < l; ++i) K[(int64_t)i * p + j] = 0.0;
        if (na == 0) continue;
        for (int a = 0; a < na; ++a) {
            for (int b = 0; b < na; ++b) sub[a * na + b] = ctc[act[a] * l + act[b]];
            rhs[a] = cta[(int64_t)act[a] * p + j];
        }
        spd_solve(sub, na, rhs, 0, work);
        for (int a = 0; a < na; ++a) K[(int64_t)act[a] * p + j] = rhs[a];
    }
}

int orc_fcnnls(const double *X, const double *Y, int64_t nobs, int32_t l, int64_t p, double eps,
               double *K, uint8_t *pset_out) {
    if (l < 1 || l > 64) return -1;
    double *ctc = (double *)malloc(sizeof(double) * (size_t)(l * l));
    double *cta = (double *)malloc(sizeof(double) * (size_t)(l * p));
    double *D = (double *)malloc(sizeof(double) * (size_t)(l * p));
    double *Wk = (double *)malloc(sizeof(double) * (size_t)l);
    uint8_t *pset = (uint8_t *)malloc((size_t)(l * p));
    int64_t *all = (int64_t *)calloc((size_t)p, sizeof(int64_t));
    int64_t *fset = (int64_t *)malloc(sizeof(int64_t) * (size_t)p);
    int64_t *hset = (int64_t *)malloc(sizeof(int64_t) * (size_t)p);
    int64_t *nf = (int64_t *)malloc(sizeof(int64_t) * (size_t)p);
    double *amin = (double *)malloc(sizeof(double) * (size_t)p);
    int *midx = (int *)malloc(sizeof(int) * (size_t)p);
    /* normal equations, forward over the observations */
    for (int a = 0; a < l; ++a)
        for (int b = 0; b < l; ++b) {
            double s = 0.0;
            for (int64_t o = 0; o < nobs; ++o) s += X[o * l + a] * X[o * l + b];
            ctc[a * l + b] = s;
        }
    for (int a = 0; a < l; ++a)
        for (int64_t j = 0; j < p; ++j) {
            double s = 0.0;
            for (int64_t o = 0; o < nobs; ++o) s += X[o * l + a] * Y[o * p + j];
            cta[(int64_t)a * p + j] = s;
        }
    for (int64_t j = 0; j < p; ++j) all[j] = j;
    /* unconstrained start, Pset = K > 0, K = 0 elsewhere (fcnnls.rs:290-302) */
    cssls_cols(ctc, cta, l, p, NULL, all, p, K);
    for (int64_t e = 0; e < (int64_t)l * p; ++e) {
        pset[e] = K[e] > 0.0;
        if (!pset[e]) K[e] = 0.0;
    }
    memcpy(D, K, sizeof(double) * (size_t)(l * p));
    int64_t nfs = 0;
    for (int64_t j = 0; j < p; ++j) {
        int cnt = 0;
        for (int i = 0; i < l; ++i) cnt += pset[(int64_t)i * p + j];
        if (cnt != l) fset[nfs++] = j;
    }
    const int max_iter = 3 * l;
    int iter = 0;
    while (nfs > 0) {
        cssls_cols(ctc, cta, l, p, pset, fset, nfs, K);
        int64_t nh = 0;
        for (int64_t c = 0; c < nfs; ++c) {
            const int64_t j = fset[c];
            int neg = 0;
            for (int i = 0; i < l; ++i) neg |= K[(int64_t)i * p + j] < eps;
            if (neg) hset[nh++] = j;
        }
        /* inner loop (fcnnls.rs:322-392) */
        while (nh > 0 && iter < max_iter) {
            iter += 1;
            int any = 0;
            for (int64_t c = 0; c < nh && !any; ++c)
                for (int i = 0; i < l; ++i) {
                    const int64_t e = (int64_t)i * p + hset[c];
                    if (pset[e] && K[e] < eps) {
                        any = 1;
                        break;
                    }
                }
            if (!any) break;
            for (int64_t c = 0; c < nh; ++c) {
                const int64_t j = hset[c];
                double best = INFINITY;
                int bi = 0;
                for (int i = 0; i < l; ++i) {
                    const int64_t e = (int64_t)i * p + j;
                    double al = INFINITY;
                    if (pset[e] && K[e] < eps) {
                        const double den = D[e] - K[e];
                        if (den != 0.0) al = D[e] / den;
                    }
                    if (al < best) {
                        best = al;
                        bi = i;
                    }
                }
                amin[c] = best;
                midx[c] = bi;
            }
            for (int64_t c = 0; c < nh; ++c) {
                const int64_t j = hset[c];
                for (int i = 0; i < l; ++i) {
                    const int64_t e = (int64_t)i * p + j;
                    D[e] -= amin[c] * (D[e] - K[e]);
                }
                D[(int64_t)midx[c] * p + j] = 0.0;
                pset[(int64_t)midx[c] * p + j] = 0;
            }
            cssls_cols(ctc, cta, l, p, pset, hset, nh, K);
            /* Hset over all columns (fcnnls.rs:387-391) */
            nh = 0;
            for (int64_t j = 0; j < p; ++j) {
                int neg = 0;
                for (int i = 0; i < l; ++i) neg |= K[(int64_t)i * p + j] < eps;
                if (neg) hset[nh++] = j;
            }
        }
        /* KKT check on Fset: W = CtA - CtC K (fcnnls.rs:400-427) */
        int64_t nn = 0;
        for (int64_t c = 0; c < nfs; ++c) {
            const int64_t j = fset[c];
            for (int a = 0; a < l; ++a) {
                double s = 0.0;
                for (int b = 0; b < l; ++b) s += ctc[a * l + b] * K[(int64_t)b * p + j];
                Wk[a] = cta[(int64_t)a * p + j] - s;
            }
            double mv = -INFINITY;
            int mr = 0;
            for (int i = 0; i < l; ++i)
                if (!pset[(int64_t)i * p + j] && Wk[i] > mv) {
                    mv = Wk[i];
                    mr = i;
                }
            if (mv > eps) {
                pset[(int64_t)mr * p + j] = 1;
                for (int i = 0; i < l; ++i) D[(int64_t)i * p + j] = K[(int64_t)i * p + j];
                nf[nn++] = j;
            }
        }
        memcpy(fset, nf, sizeof(int64_t) * (size_t)nn);
        nfs = nn;
        if (iter >= max_iter) break;
    }
    if (pset_out) memcpy(pset_out, pset, (size_t)(l * p));
    free(ctc);
    free(cta);
    free(D);
    free(Wk);
    free(pset);
    free(all);
    free(fset);
    free(hset);
    free(nf);
    free(amin);
    free(midx);
    return 0;
}

/* Kim-Park driver (lib.rs:1198-1334) on A (m x n), internal W (m x k),
 * H (k x n), in place; deviances appended each iteration. */
static void snmf_kim_park(const double *A, int64_t m, int64_t n, double *W, double *H, int k,
                          int64_t max_iter, double eta, double beta, double eps, double *dev,
                          int64_t dev_cap, int64_t *nd, int64_t *n_iter) {
    double eta_eff = eta;
    if (eta < 0.0) { /* max(A) folded from 0.0 with f64::max */
        eta_eff = 0.0;
        for (int64_t e = 0; e < m * n; ++e) eta_eff = fmax(eta_eff, A[e]);
    }
    const double beta_eff = fmax(beta, 1e-30);
    const double sb = sqrt(beta_eff);
    for (int j = 0; j < k; ++j) { /* col_l2_normalize (lib.rs:1188-1197) */
        double s = 0.0;
        for (int64_t i = 0; i < m; ++i) s += W[i * k + j] * W[i * k + j];
        const double nrm = sqrt(s);
        if (nrm > 0.0)
            for (int64_t i = 0; i < m; ++i) W[i * k + j] = W[i * k + j] / nrm;
    }
    double *Xh = (double *)malloc(sizeof(double) * (size_t)((m + 1) * k));
    double *Yh = (double *)calloc((size_t)((m + 1) * n), sizeof(double));
    double *Xw = (double *)calloc((size_t)((n + k) * k), sizeof(double));
    double *Yw = (double *)calloc((size_t)((n + k) * m), sizeof(double));
    double *Wt = (double *)malloc(sizeof(double) * (size_t)(k * m));
    memcpy(Yh, A, sizeof(double) * (size_t)(m * n));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < m; ++j) Yw[i * m + j] = A[j * n + i];
    int64_t it = 0;
    while (it < max_iter) {
        it += 1;
        memcpy(Xh, W, sizeof(double) * (size_t)(m * k));
        for (int j = 0; j < k; ++j) Xh[m * k + j] = sb;
        orc_fcnnls(Xh, Yh, m + 1, k, n, eps, H, NULL);
        for (int64_t i = 0; i < n; ++i)
            for (int j = 0; j < k; ++j) Xw[i * k + j] = H[(int64_t)j * n + i];
        for (int j = 0; j < k; ++j) Xw[(n + j) * k + j] = eta_eff;
        orc_fcnnls(Xw, Yw, n + k, k, m, eps, Wt, NULL);
        for (int64_t i = 0; i < m; ++i)
            for (int j = 0; j < k; ++j) W[i * k + j] = Wt[(int64_t)j * m + i];
        double s = 0.0;
        for (int64_t i = 0; i < m; ++i)
            for (int64_t j = 0; j < n; ++j) {
                double wh = 0.0;
                for (int q = 0; q < k; ++q) wh += W[i * k + q] * H[(int64_t)q * n + j];
                const double d = A[i * n + j] - wh;
                s += d * d;
            }
        if (dev && *nd < dev_cap) dev[*nd] = 0.5 * s;
        *nd += 1;
    }
    *n_iter = it;
    free(Xh);
    free(Yh);
    free(Xw);
    free(Yw);
    free(Wt);
}

static int run_snmf(const orc_params *prm, const double *V, int64_t n, int64_t p, double *W,
                    double *H, int r, int64_t *n_iter, double *dev, int64_t dev_cap,
                    int64_t *n_dev) {
    const int left = prm->algo == ORC_SNMF_L;
    int64_t nd = 0, it = 0;
    if (!left) {
        snmf_kim_park(V, n, p, W, H, r, prm->max_iter, prm->smoothness_w, prm->sparsity_h,
                      prm->eps, dev, dev_cap, &nd, &it);
    } else {
        /* A = V^T (p x n); internal W = H^T (p x r), internal H = W^T (r x n) */
        double *At = transpose_new(V, n, p);
        double *Wi = transpose_new(H, r, p);
        double *Hi = transpose_new(W, n, r);
        snmf_kim_park(At, p, n, Wi, Hi, r, prm->max_iter, prm->smoothness_h, prm->sparsity_w,
                      prm->eps, dev, dev_cap, &nd, &it);
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < r; ++k) W[i * r + k] = Hi[(int64_t)k * n + i];
        for (int k = 0; k < r; ++k)
            for (int64_t j = 0; j < p; ++j) H[(int64_t)k * p + j] = Wi[j * r + k];
        free(At);
        free(Wi);
        free(Hi);
    }
    if (n_iter) *n_iter = it;
    if (n_dev) *n_dev = nd;
    return 0;
}

/* ======================================================================== */
/* Drivers                                                                  */
/* ======================================================================== */
int orc_run(const orc_params *prm, const double *V, int64_t n, int64_t p, double *W, double *H,
            int32_t r, int64_t *n_iter, double *dev, int64_t dev_cap, int64_t *n_dev) {
    return orc_run_offset(prm, V, n, p, W, H, r, NULL, 0, n_iter, dev, dev_cap, n_dev);
}

int orc_run_offset(const orc_params *prm, const double *V, int64_t n, int64_t p, double *W,
                   double *H, int32_t r, double *off_io, int32_t off_given, int64_t *n_iter,
                   double *dev, int64_t dev_cap, int64_t *n_dev) {
    if (!prm || r <= 0 || n <= 0 || p <= 0) return -1;
    const int algo = prm->algo;
    if (algo == ORC_SNMF_R || algo == ORC_SNMF_L) {
        if (r > 64) return -1;
        return run_snmf(prm, V, n, p, W, H, r, n_iter, dev, dev_cap, n_dev);
    }
    if (algo != ORC_BRUNET && algo != ORC_LEE && algo != ORC_HALS && algo != ORC_EHALS &&
        algo != ORC_DNMF && algo != ORC_OFFSET && algo != ORC_NSNMF)
        return -1;
    const int nt = default_threads(prm->num_threads);
    const int stationary = prm->stop == ORC_STOP_STATIONARY;
    const int kl = algo == ORC_BRUNET || algo == ORC_NSNMF;
    /* brunet ignores the caller's eps (lib.rs:1539). */
    const double eps = kl ? DBL_EPSILON : prm->eps;
    stat_stop ss;
    ss.interval = prm->check_interval > 0 ? prm->check_interval : 1;
    ss.niter = prm->check_niter;
    ss.th = prm->stationary_th;
    ss_reset(&ss);
    int64_t nd = 0;
#define PUSH_DEV(val)                          \
    do {                                       \
        if (dev && nd < dev_cap) dev[nd] = (val); \
        nd++;                                  \
    } while (0)

    double *Vt = transpose_new(V, n, p); /* once per run, as the reference */
    double *Hn = (double *)malloc(sizeof(double) * (size_t)(r * p));
    double *Wn = (double *)malloc(sizeof(double) * (size_t)(n * r));
    double *G = (double *)malloc(sizeof(double) * (size_t)(r * r));
    double *K = (double *)malloc(sizeof(double) * (size_t)(r * r));
    double *A = NULL, *B = NULL, *Wprev = NULL, *Hprev = NULL, *Wsave = NULL, *Hsave = NULL;
    double *d = (double *)malloc(sizeof(double) * (size_t)r);
    double *off = NULL, *addon = NULL, *rsh = NULL, *S = NULL, *WS = NULL, *SH = NULL;
    if (algo == ORC_OFFSET) {
        off = off_io ? off_io : (double *)malloc(sizeof(double) * (size_t)n);
        if (!off_io || !off_given)
            for (int64_t i = 0; i < n; ++i) { /* rowMeans(V) (lib.rs:1669-1683) */
                double s = 0.0;
                for (int64_t j = 0; j < p; ++j) s += V[i * p + j];
                off[i] = s / (double)p;
            }
        addon = (double *)malloc(sizeof(double) * (size_t)r);
        rsh = (double *)malloc(sizeof(double) * (size_t)r);
    }
    if (algo == ORC_NSNMF) {
        S = (double *)malloc(sizeof(double) * (size_t)(r * r));
        orc_smoothing_matrix(r, prm->theta, S);
        WS = (double *)malloc(sizeof(double) * (size_t)(n * r));
        SH = (double *)malloc(sizeof(double) * (size_t)(r * p));
    }
    if (algo == ORC_HALS || algo == ORC_EHALS || algo == ORC_DNMF) {
        A = (double *)malloc(sizeof(double) * (size_t)(r * p));
        B = (double *)malloc(sizeof(double) * (size_t)(n * r));
    }
    if (algo == ORC_EHALS) {
        Wprev = dup(W, n * r);
        Hprev = dup(H, (int64_t)r * p);
        Wsave = (double *)malloc(sizeof(double) * (size_t)(n * r));
        Hsave = (double *)malloc(sizeof(double) * (size_t)(r * p));
    }

    int64_t iter = 0;
    int stopped_at_zero = 0;
    if (stationary) {
        double dv = algo == ORC_NSNMF   ? kl_nsnmf(V, n, p, W, H, r, S, WS)
                    : algo == ORC_OFFSET ? euclid_offset(V, n, p, W, H, r, off)
                    : kl                 ? orc_kl_divergence(V, n, p, W, H, r)
                                         : orc_euclidean_distance(V, n, p, W, H, r);
        PUSH_DEV(dv);
        if (ss_should_stop(&ss, 0, dv)) stopped_at_zero = 1;
    }
    if (stopped_at_zero) {
        if (algo == ORC_DNMF) extract_diagonal(W, H, n, p, r, d); /* lib.rs:849-853 */
    } else {
        while (iter < prm->max_iter) {
            iter += 1;
            if (algo == ORC_LEE) {
                double *Wt = transpose_new(W, n, r);
                lee_wtw(Wt, n, r, G);
                lee_h_ctx hc = {Vt, Wt, H, G, Hn, n, p, r, eps, NULL};
                par_for(p, nt, lee_h_range, &hc);
                free(Wt);
                memcpy(H, Hn, sizeof(double) * (size_t)(r * p));
                lee_update_w(V, n, p, W, H, r, eps, Wn, nt);
                memcpy(W, Wn, sizeof(double) * (size_t)(n * r));
                if (prm->rescale) col_rescale(W, n, r);
            } else if (algo == ORC_BRUNET) {
                div_update_h_with(Vt, W, H, n, p, r, Hn, nt);
                memcpy(H, Hn, sizeof(double) * (size_t)(r * p));
                div_update_w(V, n, p, W, H, r, Wn, nt);
                memcpy(W, Wn, sizeof(double) * (size_t)(n * r));
                if (iter % 10 == 0) {
                    pmax(H, (int64_t)r * p, eps);
                    pmax(W, n * r, eps);
                }
            } else if (algo == ORC_OFFSET) {
                offset_iter(V, Vt, n, p, W, H, r, eps, off, nt, G, K, Hn, Wn, addon, rsh);
            } else if (algo == ORC_NSNMF) {
                nsnmf_iter(V, Vt, n, p, W, H, r, S, nt, WS, SH, Hn, Wn);
            } else if (algo == ORC_HALS) {
                hals_iter(V, Vt, n, p, W, H, r, eps, 0, 0, 0, 0, 0, nt, A, B, G, K);
            } else if (algo == ORC_EHALS) {
                /* lib.rs:751-778, beta = 0.5 */
                const double beta = 0.5;
                memcpy(Wsave, W, sizeof(double) * (size_t)(n * r));
                memcpy(Hsave, H, sizeof(double) * (size_t)(r * p));
                for (int64_t i = 0; i < n * r; ++i) {
                    double e = W[i] + beta * (W[i] - Wprev[i]);
                    W[i] = e > eps ? e : eps;
                }
                for (int64_t i = 0; i < (int64_t)r * p; ++i) {
                    double e = H[i] + beta * (H[i] - Hprev[i]);
                    H[i] = e > eps ? e : eps;
                }
                hals_iter(V, Vt, n, p, W, H, r, eps, 0, 0, 0, 0, 0, nt, A, B, G, K);
                memcpy(Wprev, Wsave, sizeof(double) * (size_t)(n * r));
                memcpy(Hprev, Hsave, sizeof(double) * (size_t)(r * p));
            } else { /* ORC_DNMF, lib.rs:862-884 */
                hals_iter(V, Vt, n, p, W, H, r, eps, 1, prm->sparsity_h, prm->sparsity_w,
                          prm->smoothness_h, prm->smoothness_w, nt, A, B, G, K);
                extract_diagonal(W, H, n, p, r, d);
                for (int k = 0; k < r; ++k)
                    for (int64_t j = 0; j < p; ++j) H[(int64_t)k * p + j] *= d[k];
            }
            if (stationary) {
                double dv = algo == ORC_NSNMF   ? kl_nsnmf(V, n, p, W, H, r, S, WS)
                            : algo == ORC_OFFSET ? euclid_offset(V, n, p, W, H, r, off)
                            : kl                 ? orc_kl_divergence(V, n, p, W, H, r)
                                                 : orc_euclidean_distance(V, n, p, W, H, r);
                PUSH_DEV(dv);
                if (ss_should_stop(&ss, iter, dv)) break;
            }
        }
    }
#undef PUSH_DEV
    if (n_iter) *n_iter = iter;
    if (n_dev) *n_dev = nd;
    free(Vt);
    free(Hn);
    free(Wn);
    free(G);
    free(K);
    free(A);
    free(B);
    free(Wprev);
    free(Hprev);
    free(Wsave);
    free(Hsave);
    free(d);
    if (off != off_io) free(off);
    free(addon);
    free(rsh);
    free(S);
    free(WS);
    free(SH);
    return 0;
}

/* ======================================================================== */
/* Single steps                                                             */
/* ======================================================================== */
void orc_euclidean_update_h(const double *V, int64_t n, int64_t p, const double *W, const double *H,
                            int32_t r, double eps, double *Hout, int32_t nthreads) {
    /* euclidean_update_h_impl (lib.rs:247-276) */
    double *Vt = transpose_new(V, n, p);
    double *Wt = transpose_new(W, n, r);
    double *G = (double *)malloc(sizeof(double) * (size_t)(r * r));
    lee_wtw(Wt, n, r, G);
    lee_h_ctx hc = {Vt, Wt, H, G, Hout, n, p, r, eps, NULL};
    par_for(p, default_threads(nthreads), lee_h_range, &hc);
    free(G);
    free(Wt);
    free(Vt);
}

void orc_euclidean_update_w(const double *V, int64_t n, int64_t p, const double *W, const double *H,
                            int32_t r, double eps, double *Wout, int32_t nthreads) {
    lee_update_w(V, n, p, W, H, r, eps, Wout, default_threads(nthreads));
}

void orc_divergence_update_h(const double *V, int64_t n, int64_t p, const double *W,
                             const double *H, int32_t r, double *Hout, int32_t nthreads) {
    /* divergence_update_h_impl (lib.rs:117-144) */
    double *Vt = transpose_new(V, n, p);
    div_update_h_with(Vt, W, H, n, p, r, Hout, default_threads(nthreads));
    free(Vt);
}

void orc_divergence_update_w(const double *V, int64_t n, int64_t p, const double *W,
                             const double *H, int32_t r, double *Wout, int32_t nthreads) {
    div_update_w(V, n, p, W, H, r, Wout, default_threads(nthreads));
}

void orc_smoothing_matrix(int32_t r, double theta, double *S) {
    double off = theta / (double)r;
    double diag = (1.0 - theta) + off;
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < r; ++j) S[i * r + j] = i == j ? diag : off;
}
