"""Relative-change stop and the served stream.

The stop rule is checked against the oracle's own per-iteration objective
trace: the GPU must stop at the first iteration t >= min_iter with
|f(t-1) - f(t)| <= rel_tol |f(t-1)|, return the oracle's factors after t
iterations (1e-12) and the same trace.  Served results must equal the same
requests solved one at a time (the stream only changes timing)."""
import numpy as np
import pytest

import oracle
from _util import rel
from nmf_rs import _lib, _rust, serve, synthetic
from test_gpu_parity import NT, TOL

pytestmark = pytest.mark.gpu


def oracle_trace(algo, V, W0, H0, iters, **kw):
    # stationary_th < 0 never stops: one deviance per iteration
    ref = oracle.run(algo, V, W0, H0, iters, stop="stationary", stationary_th=-1.0,
                     num_threads=NT, **kw)
    assert ref["n_iter"] == iters
    return np.array(ref["deviances"])


def pick_tol(f, min_iter):
    """A tolerance that the trace crosses for the first time well inside the
    run, far (in ratio) from every iteration's relative change."""
    rc = np.abs(np.diff(f)) / np.abs(f[:-1])       # rc[t-1]: change at iteration t
    cand = np.sort(rc[min_iter - 1:])
    for lo, hi in zip(cand[:-1], cand[1:]):
        if hi / lo > 1.01 and lo < np.median(cand):
            tol = np.sqrt(lo * hi)
            ts = [t for t in range(min_iter, len(f)) if rc[t - 1] <= tol]
            if ts and ts[0] < len(f) - 2:
                return tol, ts[0]
    pytest.skip("trace has no clear crossing")


@pytest.mark.parametrize("algo", ["lee", "brunet", "hals"])
def test_rel_change_stop_matches_oracle(algo):
    V, W0, H0 = synthetic.lowrank_uniform(200, 150, 4, seed=5)
    iters, min_iter = 60, 5
    f = oracle_trace(algo, V, W0, H0, iters)
    tol, t_stop = pick_tol(f, min_iter)
    out = _rust.py_nmf_run(algo, V, W0, H0, iters, stop="rel_change", rel_tol=tol,
                           min_iter=min_iter)
    print(f"{algo}: tol {tol:.3e} expected stop {t_stop} got {out['n_iter']}", flush=True)
    assert out["n_iter"] == t_stop
    d = np.array(out["deviances"])
    assert len(d) == t_stop + 1
    assert np.max(np.abs(d - f[: t_stop + 1]) / np.abs(f[: t_stop + 1])) <= 1e-12
    ref = oracle.run(algo, V, W0, H0, t_stop, num_threads=NT)
    assert rel(out["W"], ref["W"]) <= TOL and rel(out["H"], ref["H"]) <= TOL


def test_rel_change_runs_at_least_min_iter():
    V, W0, H0 = synthetic.lowrank_uniform(60, 50, 3, seed=6)
    out = _rust.py_nmf_run("lee", V, W0, H0, 40, stop="rel_change", rel_tol=1.0, min_iter=7)
    assert out["n_iter"] == 7 and len(out["deviances"]) == 8


def test_served_stream_equals_one_at_a_time():
    reqs = serve.request_mix(6, seed=3, n_range=(40, 300), p_range=(50, 900), ranks=(2, 9))
    serve.poisson_arrivals(reqs, qps=200.0, seed=4)
    mats = [synthetic.lowrank_uniform(r.n, r.p, 3, seed=r.seed % 1000)[0] for r in reqs]
    srv = serve.Server(workers=3, precision="fp32", max_iter=40, min_iter=5, rel_tol=1e-6)
    try:
        res = srv.run_stream(reqs, mats)
        stats = serve.latency_stats(reqs, res)
        print("latency", stats, flush=True)
        for req, out, V in zip(reqs, res, mats):
            W, H, it, dev, _ = srv.solve(V, req.rank, req.method)
            print(f"request {req.rid} {req.method} {req.n}x{req.p} r={req.rank} it={out.n_iter}",
                  flush=True)
            assert out.n_iter == it and 5 <= it <= 40
            np.testing.assert_array_equal(out.W, W)
            np.testing.assert_array_equal(out.H, H)
            assert out.end >= out.start >= 0.0
        assert stats["p99_s"] >= stats["p50_s"] > 0.0
    finally:
        srv.close()


@pytest.mark.parametrize("stop", ["rel_change", "stationary"])
def test_brunet_fused_objective_equals_separate_pass(stop, monkeypatch):
    """brunet's objective summed inside the next KL column pass (default)
    against the separate objective pass (NMFG_FUSE_OBJ=0): same stop
    iteration, bit-identical W and H (a stop restores the H kept before the
    pass that found it), deviances equal up to summation order.  The
    tolerance is the first of a few at which the separate-pass run stops
    inside the run.  600 x 16500: 258 column tiles, enough work-groups for
    the fused form to be chosen."""
    V, W0, H0 = synthetic.lowrank_uniform(600, 16500, 5, seed=21)

    def kw_of(tol):
        return (dict(rel_tol=tol, min_iter=5) if stop == "rel_change" else
                dict(stationary_th=tol, check_interval=5, check_niter=3))

    # the stationary threshold is absolute (deviance units), the relative one is not
    for tol in (1e-5, 1e-4, 1e-3, 1e-2, 1e-1, 1.0, 10.0, 100.0, 1000.0):
        monkeypatch.setenv("NMFG_FUSE_OBJ", "0")
        b = _rust.py_nmf_run("brunet", V, W0, H0, 200, stop=stop, **kw_of(tol))
        if 5 < b["n_iter"] < 190:
            break
    else:
        pytest.skip("no tolerance stops inside the run")
    monkeypatch.setenv("NMFG_FUSE_OBJ", "1")
    a = _rust.py_nmf_run("brunet", V, W0, H0, 200, stop=stop, **kw_of(tol))
    print(stop, "tol", tol, "n_iter", a["n_iter"], b["n_iter"], flush=True)
    assert a["n_iter"] == b["n_iter"]
    assert np.array_equal(a["W"], b["W"]) and np.array_equal(a["H"], b["H"])
    da, db = np.asarray(a["deviances"]), np.asarray(b["deviances"])
    assert da.shape == db.shape and np.max(np.abs(da - db) / np.abs(db)) <= 1e-13
