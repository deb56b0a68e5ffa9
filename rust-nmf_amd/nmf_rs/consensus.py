"""Rank sweep with consensus clustering (Brunet et al. 2004).

The reference has this layer only as notebook code
(examples/build_scaling_notebook.py:121-141 connectivity / cophenetic /
dispersion, 157-208 the multi-run driver calling nmf_rs.nmf once per run).
Here the ranks x restarts are one device sweep on the resident V (members
grouped by rank into batched solves that advance in the same kernel
launches, the whole sweep one captured graph replay), the consensus
average runs on the GPU, and the average-linkage cophenetic correlation
uses scipy as the notebook does.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib


def cluster_labels(H):
    """Cluster of each column: argmax over the rank axis of H
    (build_scaling_notebook.py:121-124).  H: r x p, or batch x r x p."""
    return np.argmax(np.asarray(H), axis=-2).astype(np.int32)


def consensus_matrix(labels):
    """Mean connectivity over runs: C[i][j] = fraction of runs in which
    columns i and j share a cluster (build_scaling_notebook.py:195-200)."""
    return _lib.consensus_matrix(labels)


def cophenetic_correlation(consensus):
    """Correlation between the distances 1 - C and the cophenetic distances
    of their average-linkage tree, as build_scaling_notebook.py:126-137
    documents it.  That notebook function keeps element [0] of scipy's
    cophenet(Z, Y) -> (c, d), i.e. the scalar c, whose std() is 0, so it
    returns 1.0 for every input; here the correlation is taken against the
    cophenetic distances d (it equals scipy's c)."""
    from scipy.cluster.hierarchy import cophenet, linkage
    from scipy.spatial.distance import squareform

    C = np.array(consensus, dtype=np.float64, copy=True)
    np.fill_diagonal(C, 1.0)
    cond = squareform(1.0 - C, checks=False)
    if cond.size < 2:
        return 1.0
    Z = linkage(cond, method="average")
    _, coph = cophenet(Z, cond)
    if cond.std() == 0 or coph.std() == 0:
        return 1.0
    return float(np.corrcoef(cond, coph)[0, 1])


def dispersion_coefficient(consensus):
    """Kim & Park 2007: mean of 4 (C - 1/2)^2 (build_scaling_notebook.py:139-141)."""
    C = np.asarray(consensus, dtype=np.float64)
    return float((4.0 * (C - 0.5) ** 2).mean())


def uniform_restarts(shape_n, shape_p, rank, restarts, seed_of=lambda r, i: 1000 * r + i):
    """W0, H0 ~ U(0, 1) per restart, seeded by seed_of(rank, restart); returns
    (restarts x n x r, restarts x r x p)."""
    W0 = np.empty((restarts, shape_n, rank))
    H0 = np.empty((restarts, rank, shape_p))
    for i in range(restarts):
        g = np.random.default_rng(seed_of(rank, i))
        W0[i] = g.uniform(0.0, 1.0, (shape_n, rank))
        H0[i] = g.uniform(0.0, 1.0, (rank, shape_p))
    return W0, H0


def rank_sweep(V, ranks, restarts=10, method="brunet", max_iter=200, precision="fp32",
               init=uniform_restarts, cophenetic=True, tune=True):
    """Runs `restarts` solves per rank -- all ranks x restarts as ONE device
    sweep (nmfg_sweep: members grouped by rank into batched solves, one
    captured graph replay; the plan autotuned per shape when ``tune``) --
    and the consensus statistics per rank.  Returns {rank: {...}} plus
    "timing": seconds of the plan set-up (autotuning included), of the
    solves (factors up, the replay, results back) and of the consensus
    statistics; and "plan": the sweep's plan."""
    from . import _METHOD_ALIASES

    algo = _METHOD_ALIASES.get(str(method).lower())
    if algo is None:
        raise ValueError(f"unknown method '{method}'")
    V = np.asarray(V, dtype=np.float64)
    n, p = V.shape
    ranks = [int(r) for r in ranks]
    member_ranks = [r for r in ranks for _ in range(restarts)]
    W0s, H0s = [], []
    for r in ranks:
        W0, H0 = init(n, p, r, int(restarts))
        W0s.extend(W0)
        H0s.extend(H0)
    m = _lib.DeviceMatrix(V, precision=precision)
    try:
        t = time.perf_counter()
        sw = _lib.DeviceSweep(m, member_ranks, _lib.make_params(algo, max_iter), tune=tune)
        t_plan = time.perf_counter() - t
        try:
            t = time.perf_counter()
            sw.set_factors(W0s, H0s)
            sw.launch()
            Ws, Hs = sw.fetch()
            t_solve = time.perf_counter() - t
            plan = sw.plan()
        finally:
            sw.close()
    finally:
        m.close()
    out = {}
    t = time.perf_counter()
    for i, r in enumerate(ranks):
        W = np.stack(Ws[i * restarts:(i + 1) * restarts])
        H = np.stack(Hs[i * restarts:(i + 1) * restarts])
        labels = cluster_labels(H)
        C = consensus_matrix(labels)
        rec = {"W": W, "H": H, "labels": labels, "consensus": C,
               "dispersion": dispersion_coefficient(C)}
        if cophenetic:
            rec["cophenetic"] = cophenetic_correlation(C)
        out[r] = rec
    out["timing"] = {"plan_s": t_plan, "solve_s": t_solve, "consensus_s": time.perf_counter() - t}
    out["plan"] = plan
    return out
