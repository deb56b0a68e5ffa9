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
  ``capture`` is the fraction of each cell's library that falls on these n
  genes (the rest falls on genes outside the panel): L_j is multiplied by it.
  With every other parameter as stated no reading reaches ~90 % zeros at
  2000 x 7750 (capture 1: 51.6 % with NB size 2, 62 % with size 0.5; W*
  without unit-sum columns: 0.4 %); capture 0.08 gives 89.6 %
  (``SPARSE_CAPTURE``).
* ``single_cell_device`` -- the same distribution generated with torch on a
  device (for the 2000 x 500000 workload, 8 GB in FP64): same steps, torch's
  seeded generator instead of numpy's, so the values differ from
  ``single_cell`` with the same seed.
* ``lowrank_uniform_device`` -- lowrank_uniform's V at any shape, drawn with
  torch on a device (the served stream's requests): V = W* H* + |N(0, 0.01)|
  with W* (n x rank), H* (rank x p) ~ U(0, 1), rank 5 by default.
* ``bf16x2_reconstruct`` / ``fp32_reconstruct`` -- the values the FP32 and
  BF16 hi+lo storage forms of V represent (host reference for tests).
"""
from __future__ import annotations

import numpy as np

# capture fraction at which single_cell(2000, 7750) has about 90 % zeros
SPARSE_CAPTURE = 0.08


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


def single_cell(n=2000, p=7750, k=10, seed=1, chunk=65536, capture=1.0):
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
        lib = g.lognormal(8.0, 0.5, size=m) * capture            # m
        mu = outer_sum(Ws, theta * lib[np.newaxis, :])           # n x m
        lam = g.gamma(2.0, 1.0, size=mu.shape) * (0.5 * mu)
        X = g.poisson(lam).astype(np.float64)
        tot = X.sum(axis=0)
        scale = np.where(tot > 0, 1e4 / np.where(tot > 0, tot, 1.0), 0.0)
        V[:, c0:c1] = np.log1p(X * scale[np.newaxis, :])
    return V


def single_cell_device(n=2000, p=500000, k=20, seed=3, device="cuda", chunk=32768,
                       capture=1.0):
    """single_cell's distribution drawn with torch on `device`; returns an
    n x p float64 tensor there."""
    import torch

    g = torch.Generator(device=device).manual_seed(seed)
    f64 = dict(dtype=torch.float64, device=device)
    Ws = torch._standard_gamma(torch.full((n, k), 0.3, **f64), generator=g)
    Ws /= Ws.sum(dim=0, keepdim=True)
    V = torch.empty((n, p), **f64)
    for c0 in range(0, p, chunk):
        c1 = min(p, c0 + chunk)
        m = c1 - c0
        theta = torch._standard_gamma(torch.full((k, m), 0.1, **f64), generator=g)
        theta /= theta.sum(dim=0, keepdim=True)                     # Dirichlet(0.1)
        lib = torch.empty(m, **f64).log_normal_(8.0, 0.5, generator=g) * capture
        mu = Ws @ (theta * lib[None, :])
        lam = torch._standard_gamma(torch.full((n, m), 2.0, **f64), generator=g) * (0.5 * mu)
        X = torch.poisson(lam, generator=g)
        tot = X.sum(dim=0)
        scale = torch.where(tot > 0, 1e4 / torch.where(tot > 0, tot, torch.ones_like(tot)),
                            torch.zeros_like(tot))
        V[:, c0:c1] = torch.log1p(X * scale[None, :])
    return V


def lowrank_uniform_device(n, p, rank=5, seed=0, device="cuda", chunk=32768):
    """lowrank_uniform's distribution drawn with torch on `device` (columns
    in chunks); returns the n x p float64 tensor V."""
    import torch

    g = torch.Generator(device=device).manual_seed(seed)
    f64 = dict(dtype=torch.float64, device=device)
    Ws = torch.rand((n, rank), generator=g, **f64)
    V = torch.empty((n, p), **f64)
    for c0 in range(0, p, chunk):
        c1 = min(p, c0 + chunk)
        Hs = torch.rand((rank, c1 - c0), generator=g, **f64)
        noise = torch.randn((n, c1 - c0), generator=g, **f64).mul_(0.01).abs_()
        V[:, c0:c1] = torch.addmm(noise, Ws, Hs)
    return V


def _bf16_rne(f32):
    """bfloat16 rounding (nearest, ties to even) of finite float32 values,
    returned as float32."""
    u = np.ascontiguousarray(f32, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32)


def fp32_reconstruct(V):
    """The FP64 values the FP32 storage form of V holds."""
    return np.asarray(V, dtype=np.float64).astype(np.float32).astype(np.float64)


def bf16x2_reconstruct(V):
    """The FP64 values the BF16 hi+lo storage form of V holds:
    f = float32(V), hi = bf16(f), lo = bf16(f - hi), value = hi + lo."""
    f = np.asarray(V, dtype=np.float64).astype(np.float32)
    hi = _bf16_rne(f)
    lo = _bf16_rne(f - hi)
    return hi.astype(np.float64) + lo.astype(np.float64)
