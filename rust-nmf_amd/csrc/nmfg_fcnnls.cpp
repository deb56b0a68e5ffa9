// Host FP64 FCNNLS (fast combinatorial non-negative least squares) on the
// normal equations, for the Kim-Park sparse NMF drivers.  The k x k solves
// stay in FP64 on the host; the products with V that feed them (W^T V,
// V H^T) run on the GPU.
//
// Semantics follow rust/src/fcnnls.rs:270-501 (itself R NMF::.fcnnls):
//   start from the unconstrained solve, Pset = K > 0;
//   outer loop over the columns whose passive set is incomplete (Fset):
//     restricted solves; inner rounds (a budget of 3 l rounds shared by all
//     columns and outer passes) drop the variable with the smallest step
//     alpha = D / (D - K) from every column of Hset -- the columns with any
//     entry below eps, which with eps > 0 includes every column with an
//     inactive variable, those without a passive entry below eps dropping
//     row 0 (alpha = inf, default argmin) -- and Hset is then recomputed over
//     all columns;
//   KKT: add the inactive variable with the largest C^T a - C^T C k > eps.
// Restricted solves use the Cholesky factor of the passive sub-Gram, or the
// reference's regularised fallback (diagonal + 1e-12, 1e-11, ...; zeros
// past 1e6).  Factors are cached per passive pattern, which is what the
// reference's grouping of columns by pattern shares.
//
// Built with -ffp-contract=off: every product and sum rounds as written.
#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "nmfg_fcnnls.h"

namespace nmfg {
namespace {

constexpr int kMaxVars = 32;

// Lower Cholesky factor in place (row stride na); false on a diagonal <= 0.
bool cholesky(double *a, int na) {
    for (int j = 0; j < na; ++j) {
        double s = a[j * na + j];
        for (int q = 0; q < j; ++q) s -= a[j * na + q] * a[j * na + q];
        if (s <= 0.0) return false;
        const double d = std::sqrt(s);
        a[j * na + j] = d;
        for (int i = j + 1; i < na; ++i) {
            double t = a[i * na + j];
            for (int q = 0; q < j; ++q) t -= a[i * na + q] * a[j * na + q];
            a[i * na + j] = t / d;
        }
    }
    return true;
}

// Restricted normal-equation solver for one passive pattern.
struct Pattern {
    int na = 0;
    int act[kMaxVars];
    bool zero = false;          // regularisation gave up: solution 0
    std::vector<double> L;      // na x na lower factor

    void build(const double *ctc, int l, uint32_t mask) {
        na = 0;
        for (int i = 0; i < l; ++i)
            if (mask >> i & 1u) act[na++] = i;
        if (na == 0) return;
        std::vector<double> g((size_t)na * na);
        for (int a = 0; a < na; ++a)
            for (int b = 0; b < na; ++b) g[(size_t)a * na + b] = ctc[act[a] * l + act[b]];
        L = g;
        if (cholesky(L.data(), na)) return;
        std::vector<double> reg = g;
        double er = 1e-12;
        for (;;) {
            L = reg;
            if (cholesky(L.data(), na)) return;
            for (int i = 0; i < na; ++i) reg[(size_t)i * na + i] += er;
            er *= 10.0;
            if (er > 1e6) {
                zero = true;
                return;
            }
        }
    }

    // k (l entries) = restricted solution for right-hand side rhs (l entries).
    void solve(const double *rhs, int l, double *k) const {
        for (int i = 0; i < l; ++i) k[i] = 0.0;
        if (na == 0 || zero) return;
        double x[kMaxVars];
        for (int a = 0; a < na; ++a) x[a] = rhs[act[a]];
        const double *F = L.data();
        for (int i = 0; i < na; ++i) {
            double s = x[i];
            for (int q = 0; q < i; ++q) s -= F[i * na + q] * x[q];
            x[i] = s / F[i * na + i];
        }
        for (int i = na - 1; i >= 0; --i) {
            double s = x[i];
            for (int q = i + 1; q < na; ++q) s -= F[q * na + i] * x[q];
            x[i] = s / F[i * na + i];
        }
        for (int a = 0; a < na; ++a) k[act[a]] = x[a];
    }
};

class Solver {
   public:
    Solver(const double *ctc, int l) : ctc_(ctc), l_(l) {}
    void solve(uint32_t mask, const double *rhs, double *k) {
        auto it = cache_.find(mask);
        if (it == cache_.end()) {
            Pattern pt;
            pt.build(ctc_, l_, mask);
            it = cache_.emplace(mask, std::move(pt)).first;
        }
        it->second.solve(rhs, l_, k);
    }

   private:
    const double *ctc_;
    int l_;
    std::unordered_map<uint32_t, Pattern> cache_;
};

bool any_below(const double *k, int l, double eps) {
    for (int i = 0; i < l; ++i)
        if (k[i] < eps) return true;
    return false;
}

// Fork-join pool: run(n, fn) calls fn(thread, lo, hi) on contiguous slices of
// [0, n), one per worker, and returns when all are done.  Workers persist
// for the process; calls are serialised by the caller's lock.
class Pool {
   public:
    static Pool &get() {
        static Pool *p = new Pool();  // never destroyed: workers block for the process
        return *p;
    }
    int size() const { return (int)threads_.size() + 1; }
    template <class F>
    void run(int64_t n, F &&fn) {
        const int nt = size();
        if (nt == 1 || n < 256) {
            fn(0, (int64_t)0, n);
            return;
        }
        std::function<void(int, int64_t, int64_t)> job = fn;
        {
            std::unique_lock<std::mutex> lk(mu_);
            job_ = &job;
            n_ = n;
            pending_ = nt - 1;
            ++gen_;
        }
        cv_.notify_all();
        fn(0, 0, n / nt);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ == 0; });
        job_ = nullptr;
    }

   private:
    Pool() {
        unsigned hw = std::thread::hardware_concurrency();
        int nt = (int)std::min<unsigned>(hw ? hw : 1, 16u);
        if (const char *e = std::getenv("NMFG_HOST_THREADS")) nt = std::max(1, std::atoi(e));
        for (int t = 1; t < nt; ++t) threads_.emplace_back([this, t, nt] { loop(t, nt); });
        for (auto &th : threads_) th.detach();
    }
    void loop(int t, int nt) {
        uint64_t seen = 0;
        for (;;) {
            std::function<void(int, int64_t, int64_t)> *job;
            int64_t n;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                job = job_;
                n = n_;
            }
            (*job)(t, n * t / nt, n * (t + 1) / nt);
            std::unique_lock<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    std::function<void(int, int64_t, int64_t)> *job_ = nullptr;
    int64_t n_ = 0;
    int pending_ = 0;
    uint64_t gen_ = 0;
};

std::mutex g_pool_mu;  // one FCNNLS on the pool at a time

}  // namespace

int fcnnls_normal(const double *ctc, const double *cta, int l, int64_t p, double eps, double *K) {
    if (l < 1 || l > kMaxVars) return -1;
    const uint32_t full = l == 32 ? 0xffffffffu : ((1u << l) - 1u);
    std::lock_guard<std::mutex> plk(g_pool_mu);
    Pool &pool = Pool::get();
    // one factor cache per worker (same factors, computed where first needed)
    std::vector<Solver> sol;
    for (int t = 0; t < pool.size(); ++t) sol.emplace_back(ctc, l);
    std::vector<uint32_t> P((size_t)p);
    std::vector<double> D((size_t)p * l);
    std::vector<uint8_t> flag((size_t)p);
    // unconstrained start; Pset = K > 0 and K = 0 elsewhere
    pool.run(p, [&](int t, int64_t lo, int64_t hi) {
        for (int64_t j = lo; j < hi; ++j) {
            double *k = K + j * l;
            sol[t].solve(full, cta + j * l, k);
            uint32_t m = 0;
            for (int i = 0; i < l; ++i) {
                if (k[i] > 0.0)
                    m |= 1u << i;
                else
                    k[i] = 0.0;
            }
            P[(size_t)j] = m;
        }
    });
    std::memcpy(D.data(), K, sizeof(double) * (size_t)p * l);
    std::vector<int64_t> fset, hset, next;
    for (int64_t j = 0; j < p; ++j)
        if (P[(size_t)j] != full) fset.push_back(j);
    const int budget = 3 * l;
    int rounds = 0;
    while (!fset.empty()) {
        // restricted solves of Fset; Hset = its columns with an entry < eps
        pool.run((int64_t)fset.size(), [&](int t, int64_t lo, int64_t hi) {
            for (int64_t c = lo; c < hi; ++c) {
                const int64_t j = fset[(size_t)c];
                sol[t].solve(P[(size_t)j], cta + j * l, K + j * l);
                flag[(size_t)c] = any_below(K + j * l, l, eps);
            }
        });
        hset.clear();
        for (size_t c = 0; c < fset.size(); ++c)
            if (flag[c]) hset.push_back(fset[c]);
        while (!hset.empty() && rounds < budget) {
            ++rounds;
            bool drop = false;
            for (int64_t j : hset) {
                const double *k = K + j * l;
                for (int i = 0; i < l && !drop; ++i) drop = (P[(size_t)j] >> i & 1u) && k[i] < eps;
                if (drop) break;
            }
            if (!drop) break;
            // per column: the step to the first passive entry reaching 0 (or
            // the default row 0), D and Pset update, restricted re-solve
            pool.run((int64_t)hset.size(), [&](int t, int64_t lo, int64_t hi) {
                for (int64_t c = lo; c < hi; ++c) {
                    const int64_t j = hset[(size_t)c];
                    double *k = K + j * l;
                    double *d = D.data() + j * l;
                    double best = std::numeric_limits<double>::infinity();
                    int bi = 0;
                    for (int i = 0; i < l; ++i) {
                        double a = std::numeric_limits<double>::infinity();
                        if ((P[(size_t)j] >> i & 1u) && k[i] < eps) {
                            const double den = d[i] - k[i];
                            if (den != 0.0) a = d[i] / den;
                        }
                        if (a < best) {
                            best = a;
                            bi = i;
                        }
                    }
                    for (int i = 0; i < l; ++i) d[i] -= best * (d[i] - k[i]);
                    d[bi] = 0.0;
                    P[(size_t)j] &= ~(1u << bi);
                    sol[t].solve(P[(size_t)j], cta + j * l, k);
                }
            });
            // Hset over all columns
            pool.run(p, [&](int, int64_t lo, int64_t hi) {
                for (int64_t j = lo; j < hi; ++j) flag[(size_t)j] = any_below(K + j * l, l, eps);
            });
            hset.clear();
            for (int64_t j = 0; j < p; ++j)
                if (flag[(size_t)j]) hset.push_back(j);
        }
        // KKT on Fset: g = C^T a - C^T C k; add the largest inactive g > eps
        pool.run((int64_t)fset.size(), [&](int, int64_t lo, int64_t hi) {
            for (int64_t c = lo; c < hi; ++c) {
                const int64_t j = fset[(size_t)c];
                const double *k = K + j * l;
                const double *rhs = cta + j * l;
                double mv = -std::numeric_limits<double>::infinity();
                int mr = 0;
                for (int i = 0; i < l; ++i) {
                    if (P[(size_t)j] >> i & 1u) continue;
                    double s = 0.0;
                    for (int b = 0; b < l; ++b) s += ctc[i * l + b] * k[b];
                    const double g = rhs[i] - s;
                    if (g > mv) {
                        mv = g;
                        mr = i;
                    }
                }
                flag[(size_t)c] = mv > eps;
                if (mv > eps) {
                    P[(size_t)j] |= 1u << mr;
                    std::memcpy(D.data() + j * l, k, sizeof(double) * l);
                }
            }
        });
        next.clear();
        for (size_t c = 0; c < fset.size(); ++c)
            if (flag[c]) next.push_back(fset[c]);
        fset.swap(next);
        if (rounds >= budget) break;
    }
    return 0;
}

}  // namespace nmfg
