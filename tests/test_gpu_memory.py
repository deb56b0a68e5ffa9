"""Device memory of a context: blocks released by destroyed matrices and
solves are handed to the context's next allocations, with whatever the last
user left in them.  Every kernel must write a buffer before it reads it, so
results may not depend on the blocks' contents: solves on a context that
fills every block it hands out with 0xff bytes (NaN as doubles,
NMFG_POISON=1) and solves on reused blocks equal, bit for bit, the same
solves on a fresh context."""
import os

import numpy as np
import pytest

from _util import rel
from nmf_rs import _lib, synthetic

pytestmark = pytest.mark.gpu


def context_with(**env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return _lib.Context()   # the variables are read when the context is created
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.fixture(scope="module")
def contexts():
    plain, poison, nocache = (context_with(NMFG_POISON=0), context_with(NMFG_POISON=1),
                              context_with(NMFG_CACHE_MB=0))
    yield plain, poison, nocache
    for c in (plain, poison, nocache):
        c.close()


# algo, n, p, rank, precision, stop, batch
CASES = [
    ("lee", 300, 2000, 10, "fp64", "max_iter", 1),
    ("lee", 300, 2000, 10, "fp32", "rel_change", 1),
    ("lee", 300, 2000, 32, "fp64", "max_iter", 3),
    ("brunet", 600, 16500, 5, "fp64", "rel_change", 1),
    ("brunet", 200, 900, 20, "bf16x2-store", "stationary", 1),
    ("hals", 300, 2000, 16, "fp64", "rel_change", 1),
    ("ehals", 250, 1200, 8, "fp64", "stationary", 1),
    ("dnmf", 300, 2000, 12, "fp32", "max_iter", 1),
    ("offset", 200, 700, 6, "fp64", "stationary", 1),
    ("nsnmf", 200, 700, 6, "fp64", "rel_change", 1),
    ("lsnmf", 200, 700, 6, "fp64", "max_iter", 1),
]


def run_case(ctx, case, V, W0, H0, S):
    algo, n, p, rank, prec, stop, batch = case
    m = _lib.DeviceMatrix(V, precision=prec, ctx=ctx)
    try:
        prm = _lib.make_params(algo, 60, stop=stop, rel_tol=1e-5, min_iter=5,
                               stationary_th=1e-3, check_interval=5, check_niter=3,
                               precision=prec)
        s = _lib.DeviceSolve(m, rank, prm, batch)
        try:
            s.set_factors(W0, H0)
            if algo == "lsnmf":
                s.set_weight(S)
            s.launch()
            first = s.fetch()
            s.launch()          # a relaunch starts again from W0, H0
            second = s.fetch()
        finally:
            s.close()
    finally:
        m.close()
    return first, second


def same(a, b):
    Wa, Ha, ia, da = a
    Wb, Hb, ib, db = b
    return ia == ib and np.array_equal(Wa, Wb) and np.array_equal(Ha, Hb) and np.array_equal(da, db)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[4]}-{c[5]}-r{c[3]}-b{c[6]}")
def test_results_do_not_depend_on_block_contents(contexts, case):
    plain, poison, nocache = contexts
    algo, n, p, rank, prec, stop, batch = case
    V, W0, H0 = synthetic.lowrank_uniform(n, p, max(2, rank // 2), seed=n + p + rank)
    g = np.random.default_rng(rank)
    W0 = g.uniform(0.1, 1.0, ((batch,) if batch > 1 else ()) + (n, rank))
    H0 = g.uniform(0.1, 1.0, ((batch,) if batch > 1 else ()) + (rank, p))
    S = (g.uniform(size=(n, p)) < 0.8).astype(np.float64) if algo == "lsnmf" else None
    ref, again = run_case(plain, case, V, W0, H0, S)
    print(f"{case}: n_iter {ref[2]}", flush=True)
    assert np.all(np.isfinite(ref[0])) and np.all(np.isfinite(ref[1]))
    assert same(ref, again)
    reused = run_case(plain, case, V, W0, H0, S)   # the first run's blocks, as it left them
    assert same(ref, reused[0]) and same(ref, reused[1])
    poisoned = run_case(poison, case, V, W0, H0, S)
    assert same(ref, poisoned[0]) and same(ref, poisoned[1])
    uncached = run_case(nocache, case, V, W0, H0, S)
    assert same(ref, uncached[0])


def test_products_and_nndsvd_on_poisoned_blocks(contexts):
    import nmf_rs

    plain, poison, _ = contexts
    V = synthetic.single_cell(900, 5000, 10, seed=8)
    g = np.random.default_rng(2)
    B = g.normal(size=(5000, 40))
    out = {}
    for name, ctx in (("plain", plain), ("poison", poison)):
        m = _lib.DeviceMatrix(V, precision="fp64", ctx=ctx)
        try:
            C = m.mul(B)
            W, H = nmf_rs.nndsvd_on_device(m, 12, fill="mean", seed=0, v_mean=float(V.mean()))
        finally:
            m.close()
        out[name] = (C, W, H)
    assert rel(out["plain"][0], V @ B) <= 1e-12
    for a, b in zip(out["plain"], out["poison"]):
        assert np.array_equal(a, b)


def test_release_cached_then_reuse(contexts):
    plain, _, _ = contexts
    plain.release_cached()
    case = CASES[0]
    algo, n, p, rank, *_ = case
    V, W0, H0 = synthetic.lowrank_uniform(n, p, rank, seed=3)
    a = run_case(plain, case, V, W0, H0, None)
    plain.release_cached()
    b = run_case(plain, case, V, W0, H0, None)
    assert same(a[0], b[0])
