"""Inputs and metrics of the reference's multi-scale benchmark notebook,
restated for the known-answer tests (examples/build_scaling_notebook.py):

* ``make_synthetic``: :89-94 -- W* ~ U(0.1, 1.5), H* ~ U(0.1, 1.5) from
  default_rng(seed), V = max(W* H* + noise N(0, 1), 0);
* ``random_start``: :163-167, 175-176 -- run ``run`` draws W0 then H0 ~
  U(0, sqrt(mean(V) / rank)) from default_rng(2024 + run);
* ``loss``: :183 -- 0.5 ||V - W H||_F^2;
* ``hoyer`` / ``factor_sparsity_w``: :104-117 -- mean Hoyer sparsity of the
  columns of W;
* ``amari_error``: :143-151;
* ``connectivity`` / ``dispersion``: :119-122, 139-141, with the consensus
  over the cells ``consensus_subset`` picks (default_rng(43), :288-289).

The notebook's printed results (examples/benchmark_scaling.ipynb, cells
16-21) are the known answers: tier "medium" is make_synthetic(3000, 50000,
rank=15, seed=1), 100 iterations, K = 3 random starts; hals + NNDSVD runs 25
iterations from nndsvd_init(V, 15, fill='mean', seed=0).
"""
from __future__ import annotations

import numpy as np

# examples/benchmark_scaling.ipynb cell 21 (7 significant digits) and the
# pivot tables of cells 18-19 (3 decimals), tier "medium".
MEDIUM = {
    "n": 3000, "p": 50000, "rank": 15, "seed": 1, "iters": 100, "runs": 3, "nndsvd_iters": 25,
    "ehals_mean_loss": 2.615893e+06,        # cell 21, lowest loss
    "hals_nndsvd_amari": 2.332754e-01,      # cell 21, lowest Amari
    "hals_nndsvd_sparsity_w": 1.934819e-01,  # cell 21, sparsest W
    "sparsity_w_3dp": {"lee": 0.053, "brunet": 0.054, "hals": 0.170, "ehals": 0.157,
                       "dnmf": 0.170},      # cell 19
    "dispersion_3dp": {"lee": 0.820, "brunet": 0.824, "hals": 1.000, "ehals": 0.592,
                       "dnmf": 1.000},      # cell 18
}


def make_synthetic(n_genes, n_cells, rank=10, noise=0.05, seed=0):
    rng = np.random.default_rng(seed)
    W = rng.uniform(0.1, 1.5, (n_genes, rank)).astype(np.float64)
    H = rng.uniform(0.1, 1.5, (rank, n_cells)).astype(np.float64)
    V = W @ H + noise * rng.normal(size=(n_genes, n_cells))
    return np.maximum(V, 0.0), W, H


def random_start(V, rank, run):
    rng = np.random.default_rng(2024 + run)
    scale = float(np.sqrt(V.mean() / rank))
    n, p = V.shape
    W0 = rng.uniform(0, scale, (n, rank))
    H0 = rng.uniform(0, scale, (rank, p))
    return W0, H0


def loss(V, W, H):
    return 0.5 * float(np.linalg.norm(V - W @ H) ** 2)


def hoyer(x):
    x = np.asarray(x).ravel()
    n = x.size
    if n == 0:
        return 0.0
    l1 = float(np.abs(x).sum())
    l2 = float(np.sqrt((x ** 2).sum()))
    if l2 == 0:
        return 0.0
    return (np.sqrt(n) - l1 / l2) / (np.sqrt(n) - 1.0)


def factor_sparsity_w(W):
    return float(np.mean([hoyer(W[:, k]) for k in range(W.shape[1])]))


def amari_error(W_est, W_true):
    P = np.abs(np.linalg.pinv(W_est) @ W_true)
    n = P.shape[0]
    row = sum(P[i, :].sum() / max(P[i, :].max(), 1e-30) - 1 for i in range(n))
    col = sum(P[:, j].sum() / max(P[:, j].max(), 1e-30) - 1 for j in range(n))
    return float((row + col) / (2 * n * (n - 1))) if n > 1 else 0.0


def consensus_subset(n_cells, size=3000, seed=43):
    return np.random.default_rng(seed).choice(n_cells, size, replace=False)


def connectivity(H_sub):
    c = np.argmax(H_sub, axis=0)
    return (c[:, None] == c[None, :]).astype(np.float64)


def dispersion(consensus):
    return float((4.0 * (consensus - 0.5) ** 2).mean())
