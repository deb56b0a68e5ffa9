"""Seeded synthetic inputs for the benchmark workloads.

Every matrix product here is a fixed-order sum of outer products in numpy
elementwise arithmetic (no BLAS), so the generated V is bit-identical on
every machine with the same numpy.

* ``lowrank_uniform``  -- 500 genes x 300 cells, rank 5:
  V = W* H* + |N(0, 0.01)|, W*, H* ~ U(0, 1); W0, H0 ~ U(0, 1) drawn next from
  the same stream.  (0.01 is the normal's standard deviation.)
* ``single_cell``      -- genes x cells log1p-normalised negative-binomial
  counts: W* ~ Gamma(0.3, 1) with unit-sum columns, cell mixtures
  theta_j ~ Dirichlet(0.1), library sizes L_j ~ LogNormal(8, 0.5),
  mu = W* (theta L); counts ~ NB(mean mu, dispersion 0.5) drawn as
  Poisson(Gamma(shape 2, scale mu / 2)); each cell scaled to 1e4 total
  (empty cells stay zero), then log1p.  The unit-sum normalisation of W*'s
  columns is a choice made here to reach single-cell-like sparsity.
"""
from __future__ import annotations

import numpy as np


def outer_sum(A, B):
    """A @ B as a fixed-order sum over the inner index (bit-portable)."""
    out = np.zeros((A.shape[0], B.shape[1]))
    for k in range(A.shape[1]):
        out += A[:, k:k + 1] * B[k:k + 1, :]
    return out


def lowrank_uniform(n=500, p=300, rank=5, seed=0):
    """Returns V, W0, H0."""
    g = np.random.default_rng(seed)
    Ws = g.uniform(0.0, 1.0, (n, rank))
    Hs = g.uniform(0.0, 1.0, (rank, p))
    V = outer_sum(Ws, Hs) + np.abs(g.normal(0.0, 0.01, (n, p)))
    W0 = g.uniform(0.0, 1.0, (n, rank))
    H0 = g.uniform(0.0, 1.0, (rank, p))
    return V, W0, H0


def single_cell(n=2000, p=7750, k=10, seed=1, chunk=65536):
    """log1p-normalised NB counts, n genes x p cells (see module docstring).
    Generated in column chunks to bound host memory."""
    g = np.random.default_rng(seed)
    Ws = g.gamma(0.3, 1.0, (n, k))
    Ws /= Ws.sum(axis=0, keepdims=True)
    V = np.empty((n, p))
    for c0 in range(0, p, chunk):
        c1 = min(p, c0 + chunk)
        m = c1 - c0
        theta = g.dirichlet(np.full(k, 0.1), size=m).T          # k x m
        lib = g.lognormal(8.0, 0.5, size=m)                      # m
        mu = outer_sum(Ws, theta * lib[np.newaxis, :])           # n x m
        lam = g.gamma(2.0, 1.0, size=mu.shape) * (0.5 * mu)
        X = g.poisson(lam).astype(np.float64)
        tot = X.sum(axis=0)
        scale = np.where(tot > 0, 1e4 / np.where(tot > 0, tot, 1.0), 0.0)
        V[:, c0:c1] = np.log1p(X * scale[np.newaxis, :])
    return V
