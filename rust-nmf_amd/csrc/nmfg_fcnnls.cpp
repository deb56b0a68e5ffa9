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
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
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

}  // namespace

int fcnnls_normal(const double *ctc, const double *cta, int l, int64_t p, double eps, double *K) {
    if (l < 1 || l > kMaxVars) return -1;
    const uint32_t full = l == 32 ? 0xffffffffu : ((1u << l) - 1u);
    Solver sol(ctc, l);
    std::vector<uint32_t> P((size_t)p);
    std::vector<double> D((size_t)p * l);
    // unconstrained start; Pset = K > 0 and K = 0 elsewhere
    for (int64_t j = 0; j < p; ++j) {
        double *k = K + j * l;
        sol.solve(full, cta + j * l, k);
        uint32_t m = 0;
        for (int i = 0; i < l; ++i) {
            if (k[i] > 0.0)
                m |= 1u << i;
            else
                k[i] = 0.0;
        }
        P[(size_t)j] = m;
    }
    std::memcpy(D.data(), K, sizeof(double) * (size_t)p * l);
    std::vector<int64_t> fset, hset, next;
    for (int64_t j = 0; j < p; ++j)
        if (P[(size_t)j] != full) fset.push_back(j);
    const int budget = 3 * l;
    int rounds = 0;
    std::vector<double> amin;
    std::vector<int> arg;
    while (!fset.empty()) {
        hset.clear();
        for (int64_t j : fset) {
            sol.solve(P[(size_t)j], cta + j * l, K + j * l);
            if (any_below(K + j * l, l, eps)) hset.push_back(j);
        }
        while (!hset.empty() && rounds < budget) {
            ++rounds;
            bool drop = false;
            for (int64_t j : hset) {
                const double *k = K + j * l;
                for (int i = 0; i < l && !drop; ++i) drop = (P[(size_t)j] >> i & 1u) && k[i] < eps;
                if (drop) break;
            }
            if (!drop) break;
            amin.assign(hset.size(), 0.0);
            arg.assign(hset.size(), 0);
            for (size_t c = 0; c < hset.size(); ++c) {
                const int64_t j = hset[c];
                const double *k = K + j * l;
                const double *d = D.data() + j * l;
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
                amin[c] = best;
                arg[c] = bi;
            }
            for (size_t c = 0; c < hset.size(); ++c) {
                const int64_t j = hset[c];
                const double *k = K + j * l;
                double *d = D.data() + j * l;
                for (int i = 0; i < l; ++i) d[i] -= amin[c] * (d[i] - k[i]);
                d[arg[c]] = 0.0;
                P[(size_t)j] &= ~(1u << arg[c]);
            }
            for (int64_t j : hset) sol.solve(P[(size_t)j], cta + j * l, K + j * l);
            hset.clear();
            for (int64_t j = 0; j < p; ++j)
                if (any_below(K + j * l, l, eps)) hset.push_back(j);
        }
        // KKT on Fset: g = C^T a - C^T C k; add the largest inactive g > eps
        next.clear();
        for (int64_t j : fset) {
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
            if (mv > eps) {
                P[(size_t)j] |= 1u << mr;
                std::memcpy(D.data() + j * l, k, sizeof(double) * l);
                next.push_back(j);
            }
        }
        fset.swap(next);
        if (rounds >= budget) break;
    }
    return 0;
}

}  // namespace nmfg
