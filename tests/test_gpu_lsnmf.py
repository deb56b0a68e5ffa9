"""GPU parity of lsNMF (weighted Frobenius MU, rust/src/lib.rs:946-1043) and
of cv_rank (python/nmf_rs/__init__.py:273-355) against the CPU oracle.
FP64 tolerance 1e-12 on max|X_gpu - X_oracle| / max|X_oracle|."""
import os

import numpy as np
import pytest

import nmf_rs
import oracle
from _util import rel
from nmf_rs import _lib, _rust, synthetic

pytestmark = pytest.mark.gpu

TOL = 1e-12
NT = min(16, os.cpu_count() or 1)


def _inputs(seed, n, p, r):
    g = np.random.default_rng(seed)
    V = synthetic.outer_sum(g.uniform(0, 1, (n, r)), g.uniform(0, 1, (r, p))) + 0.01
    return V, g.uniform(0.1, 1, (n, r)), g.uniform(0.1, 1, (r, p)), g


def _weights(g, shape):
    return {"f64": g.uniform(0.2, 2.0, shape), "mask": (g.random(shape) > 0.1).astype(np.float64)}


def _check(out, ref, what):
    rw, rh = rel(out["W"], ref["W"]), rel(out["H"], ref["H"])
    print(f"{what}: rel W {rw:.2e} H {rh:.2e}", flush=True)
    assert rw <= TOL and rh <= TOL, (what, rw, rh)


@pytest.mark.parametrize("r", [1, 3, 5, 8, 10, 12, 16, 20, 32])
def test_gpu_lsnmf_ranks_and_weight_forms(r):
    """Both weight storages (FP64, one byte for 0/1) and every split shape."""
    V, W0, H0, g = _inputs(500 + r, 70, 203, r)
    for form, S in _weights(g, V.shape).items():
        ref = oracle.run("lsnmf", V, W0, H0, 8, weight=S, num_threads=NT)
        out = _rust.py_nmf_run("lsnmf", V, W0, H0, 8, weight=S)
        _check(out, ref, f"lsnmf r={r} {form}")


@pytest.mark.parametrize("shape", [(1, 1), (1, 300), (300, 1), (17, 129), (129, 17), (1025, 65)])
def test_gpu_lsnmf_edge_shapes(shape):
    n, p = shape
    V, W0, H0, g = _inputs(n * 3 + p, n, p, 2)
    for form, S in _weights(g, V.shape).items():
        ref = oracle.run("lsnmf", V, W0, H0, 5, weight=S)
        out = _rust.py_nmf_run("lsnmf", V, W0, H0, 5, weight=S)
        _check(out, ref, f"lsnmf {shape} {form}")


def test_gpu_lsnmf_masked_entries_do_not_matter():
    V, W0, H0, g = _inputs(41, 90, 140, 4)
    S = (g.random(V.shape) > 0.25).astype(np.float64)
    V2 = np.where(S == 0, g.uniform(0, 50, V.shape), V)
    a = _rust.py_nmf_run("lsnmf", V, W0, H0, 7, weight=S)
    b = _rust.py_nmf_run("lsnmf", V2, W0, H0, 7, weight=S)
    assert np.array_equal(a["W"], b["W"]) and np.array_equal(a["H"], b["H"])


def test_gpu_lsnmf_requires_weight():
    V, W0, H0, _ = _inputs(43, 20, 30, 2)
    with pytest.raises(ValueError):
        _rust.py_nmf_run("lsnmf", V, W0, H0, 3)
    with pytest.raises(ValueError):
        nmf_rs.nmf(V, 2, "ls-nmf", W0=W0, H0=H0, max_iter=3)


def test_gpu_lsnmf_stationary_trace():
    V, W0, H0, g = _inputs(45, 80, 120, 4)
    S = g.uniform(0.5, 1.5, V.shape)
    kw = dict(stop="stationary", stationary_th=1e-6, check_interval=5, check_niter=3)
    ref = oracle.run("lsnmf", V, W0, H0, 300, weight=S, **kw)
    out = _rust.py_nmf_run("lsnmf", V, W0, H0, 300, weight=S, **kw)
    d_ref, d_gpu = np.asarray(ref["deviances"]), np.asarray(out["deviances"])
    print("n_iter", out["n_iter"], ref["n_iter"], flush=True)
    assert out["n_iter"] == ref["n_iter"] and len(d_gpu) == len(d_ref)
    assert np.max(np.abs(d_gpu - d_ref) / np.abs(d_ref)) <= TOL
    _check(out, ref, "lsnmf stationary")


def test_gpu_lsnmf_single_cell_2000x7750_mask():
    V = synthetic.single_cell(2000, 7750, 10, seed=1)
    g = np.random.default_rng(3)
    S = (g.random(V.shape) > 0.05).astype(np.float64)
    W0, H0 = g.uniform(0.1, 1, (2000, 10)), g.uniform(0.1, 1, (10, 7750))
    ref = oracle.run("lsnmf", V * S, W0, H0, 15, weight=S, num_threads=NT)
    out = _rust.py_nmf_run("lsnmf", V * S, W0, H0, 15, weight=S)
    _check(out, ref, "lsnmf single-cell mask")


def test_gpu_lsnmf_batched_members_equal_single_and_fp32():
    V, W0, H0, g = _inputs(47, 90, 150, 5)
    S = (g.random(V.shape) > 0.1).astype(np.float64)
    Ws = np.stack([W0, g.uniform(0.1, 1, W0.shape)])
    Hs = np.stack([H0, g.uniform(0.1, 1, H0.shape)])
    prm = _lib.make_params("lsnmf", 9)
    m = _lib.DeviceMatrix(V)
    try:
        bs = _lib.DeviceSolve(m, 5, prm, batch=2)
        bs.set_factors(Ws, Hs)
        bs.set_weight(S)
        bs.launch()
        Wb, Hb, _, _ = bs.fetch()
        for b in range(2):
            ref = oracle.run("lsnmf", V, Ws[b], Hs[b], 9, weight=S)
            _check({"W": Wb[b], "H": Hb[b]}, ref, f"batched {b}")
            one = _lib.DeviceSolve(m, 5, prm)
            one.set_factors(Ws[b], Hs[b])
            one.set_weight(S)
            one.launch()
            W1, H1, _, _ = one.fetch()
            assert np.array_equal(W1, Wb[b]) and np.array_equal(H1, Hb[b])
            one.close()
        bs.close()
    finally:
        m.close()
    Vr = synthetic.fp32_reconstruct(V)
    ref = oracle.run("lsnmf", Vr, W0, H0, 6, weight=S)
    out = _rust.py_nmf_run("lsnmf", V, W0, H0, 6, weight=S, precision="fp32")
    _check(out, ref, "lsnmf fp32")


def _cv_rank_oracle(V, ranks, n_folds, mask_frac, max_iter, seed):
    """cv_rank restated on the oracle: the reference's draw order
    (__init__.py:322-329), random_init starts, oracle lsNMF fits."""
    n, p = V.shape
    master = np.random.default_rng(seed)
    rows = []
    for k in ranks:
        for fold in range(n_folds):
            fs = int(master.integers(0, 2**31 - 1))
            train = np.random.default_rng(fs).random((n, p)) > mask_frac
            S = train.astype(np.float64)
            W0, H0 = oracle.random_init_ref(V * S, k, seed=fs)
            o = oracle.run("lsnmf", V * S, W0, H0, max_iter, weight=S)
            est = o["W"] @ o["H"]
            rows.append((k, fold, 0.5 * float((((V - est) * S) ** 2).sum()),
                         float(((V[~train] - est[~train]) ** 2).mean())))
    return rows


def test_gpu_cv_rank_matches_oracle_restatement():
    V, _, _, _ = _inputs(49, 60, 45, 3)
    df = nmf_rs.cv_rank(V, [2, 3, 4], n_folds=2, mask_frac=0.1, max_iter=30, init="random",
                        seed=5)
    ref = _cv_rank_oracle(V, [2, 3, 4], 2, 0.1, 30, 5)
    assert list(df.columns) == ["rank", "fold", "train_loss", "test_mse"]
    assert len(df) == len(ref)
    for (_, row), (k, fold, tl, te) in zip(df.iterrows(), ref):
        assert int(row["rank"]) == k and int(row["fold"]) == fold
        assert abs(row["train_loss"] - tl) <= 1e-10 * abs(tl)
        assert abs(row["test_mse"] - te) <= 1e-10 * abs(te)
    # nndsvd starts (the default): the true rank 3 fits the training entries
    # best among 2..4 on this exactly rank-3 + 0.01 matrix and the held-out
    # error is finite
    dn = nmf_rs.cv_rank(V, [2, 3], n_folds=2, mask_frac=0.1, max_iter=60, seed=6)
    assert np.isfinite(dn["test_mse"]).all()
    assert dn[dn["rank"] == 3]["train_loss"].mean() < dn[dn["rank"] == 2]["train_loss"].mean()
