// Host FP64 FCNNLS on the normal equations (nmfg_fcnnls.cpp).
#pragma once
#include <cstdint>

namespace nmfg {

// K >= 0 minimising ||Y - X K|| given C^T C = X^T X (l x l, row-major) and
// the right-hand sides C^T y_j as rows of cta (p x l); K out as p x l (row j
// = the solution for column j).  eps is the reference's tolerance
// (rust/src/fcnnls.rs).  Returns 0, or -1 for l outside 1..32.
int fcnnls_normal(const double *ctc, const double *cta, int l, int64_t p, double eps, double *K);

}  // namespace nmfg
