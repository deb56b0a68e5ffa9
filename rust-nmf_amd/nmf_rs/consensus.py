"""Rank sweep with consensus clustering (Brunet et al. 2004).

The reference has this layer only as notebook code
(examples/build_scaling_notebook.py:121-141 connectivity / cophenetic /
dispersion, 157-208 the multi-run driver calling nmf_rs.nmf once per run).
Here every restart of a rank is one member of a batched solve on the
resident V (all restarts advance in the same kernel launches), the
consensus average runs on the GPU, and the average-linkage cophenetic
correlation uses scipy as the notebook does.
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
               init=uniform_restarts, cophenetic=True, workers=4):
    """Runs `restarts` solves per rank (one batched solve per rank) and the
    consensus statistics.  The ranks are spread over `workers` threads, each
    with its own library context (HIP stream) and resident copy of V, so
    batched solves of different ranks overlap on the GPU.  Returns
    {rank: {...}} plus timing under key "timing": wall seconds of the solves
    (all ranks) and of the consensus statistics."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    from . import _METHOD_ALIASES

    algo = _METHOD_ALIASES.get(str(method).lower())
    if algo is None:
        raise ValueError(f"unknown method '{method}'")
    V = np.asarray(V, dtype=np.float64)
    n, p = V.shape
    local = threading.local()
    owned = []
    lock = threading.Lock()

    def matrix():
        m = getattr(local, "m", None)
        if m is None:
            ctx = _lib.Context()
            m = _lib.DeviceMatrix(V, precision=precision, ctx=ctx)
            local.m = m
            with lock:
                owned.append((m, ctx))
        return m

    def solve(r):
        m = matrix()
        W0, H0 = init(n, p, int(r), int(restarts))
        s = _lib.DeviceSolve(m, int(r), _lib.make_params(algo, max_iter), batch=int(restarts))
        try:
            s.set_factors(W0, H0)
            s.launch()
            W, H, _, _ = s.fetch()
        finally:
            s.close()
        return int(r), W, H

    out = {}
    nw = max(1, int(workers))
    gate = threading.Barrier(nw)

    def warm(_):
        matrix()
        gate.wait()  # every worker thread holds its context before the clock starts

    try:
        with ThreadPoolExecutor(nw) as pool:
            list(pool.map(warm, range(nw)))
            t = time.perf_counter()
            solved = list(pool.map(solve, ranks))
            t_solve = time.perf_counter() - t
    finally:
        for m, ctx in owned:
            m.close()
            ctx.close()
    t = time.perf_counter()
    for r, W, H in solved:
        labels = cluster_labels(H)
        C = consensus_matrix(labels)
        rec = {"W": W, "H": H, "labels": labels, "consensus": C,
               "dispersion": dispersion_coefficient(C)}
        if cophenetic:
            rec["cophenetic"] = cophenetic_correlation(C)
        out[r] = rec
    out["timing"] = {"solve_s": t_solve, "consensus_s": time.perf_counter() - t}
    return out
