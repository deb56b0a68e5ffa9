"""GPU parity of the offset (lee + per-row offset, rust/src/lib.rs:1649-1759)
and nsNMF (brunet on W S and S H, lib.rs:1761-1833) methods against the CPU
oracle.  FP64 tolerance 1e-12 on max|X_gpu - X_oracle| / max|X_oracle| for W,
H and the offset vector; relative 1e-12 on objective values."""
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


def _inputs(seed, n, p, r, offset_rows=True):
    g = np.random.default_rng(seed)
    V = synthetic.outer_sum(g.uniform(0, 1, (n, r)), g.uniform(0, 1, (r, p))) + 0.01
    if offset_rows:
        V = V + g.uniform(0.0, 2.0, (n, 1))  # per-row baselines for the offset model
    return V, g.uniform(0.1, 1, (n, r)), g.uniform(0.1, 1, (r, p))


def _check(out, ref, what, offset=False):
    rw, rh = rel(out["W"], ref["W"]), rel(out["H"], ref["H"])
    ro = rel(out["offset"], ref["offset"]) if offset else 0.0
    print(f"{what}: rel W {rw:.2e} H {rh:.2e} off {ro:.2e}", flush=True)
    assert rw <= TOL and rh <= TOL and ro <= TOL, (what, rw, rh, ro)


@pytest.mark.parametrize("r", [1, 3, 5, 10, 11, 16, 20, 32])
def test_gpu_offset_ranks(r):
    """Every row-pass flavour: VALU rows (r <= 10) and MFMA rows + rowfin."""
    V, W0, H0 = _inputs(300 + r, 70, 203, r)
    ref = oracle.run("offset", V, W0, H0, 8, num_threads=NT)
    out = _rust.py_nmf_run("offset", V, W0, H0, 8)
    _check(out, ref, f"offset r={r}", offset=True)


@pytest.mark.parametrize("r", [1, 3, 6, 8, 10, 12, 20, 32])
@pytest.mark.parametrize("theta", [0.0, 0.5, 0.9])
def test_gpu_nsnmf_ranks(r, theta):
    V, W0, H0 = _inputs(400 + r, 70, 203, r, offset_rows=False)
    ref = oracle.run("nsnmf", V, W0, H0, 8, theta=theta, num_threads=NT)
    out = _rust.py_nmf_run("nsnmf", V, W0, H0, 8, theta=theta)
    _check(out, ref, f"nsnmf r={r} theta={theta}")
    assert np.allclose(out["W"].sum(axis=0), 1.0, atol=1e-13)


@pytest.mark.parametrize("shape", [(1, 1), (1, 300), (300, 1), (17, 129), (129, 17), (1025, 65)])
def test_gpu_offset_nsnmf_edge_shapes(shape):
    n, p = shape
    g = np.random.default_rng(n * 7 + p)
    r = 2
    V = g.uniform(0, 1, (n, p))
    W0, H0 = g.uniform(0.1, 1, (n, r)), g.uniform(0.1, 1, (r, p))
    for algo in ("offset", "nsnmf"):
        ref = oracle.run(algo, V, W0, H0, 5)
        out = _rust.py_nmf_run(algo, V, W0, H0, 5)
        _check(out, ref, f"{algo} {shape}", offset=algo == "offset")


def test_gpu_offset_given_initial_and_zero_rows():
    """An explicit initial offset (py_nmf_run's `offset`), and rows of V that
    are all zero (rowsum V = 0 -> the eps floor of the offset update)."""
    V, W0, H0 = _inputs(7, 60, 90, 4)
    V[5, :] = 0.0
    V[:, 11] = 0.0
    off0 = np.random.default_rng(8).uniform(0.0, 1.0, 60)
    ref = oracle.run("offset", V, W0, H0, 10, offset=off0)
    out = _rust.py_nmf_run("offset", V, W0, H0, 10, offset=off0)
    _check(out, ref, "offset given", offset=True)
    # the public API returns the offset and uses it in fitted()
    res = nmf_rs.nmf(V, 4, "offset", W0=W0, H0=H0, max_iter=10, offset=off0)
    assert rel(res.offset, ref["offset"]) <= TOL
    assert rel(res.fitted(), ref["W"] @ ref["H"] + ref["offset"][:, None]) <= 1e-12


def test_gpu_offset_nsnmf_stationary_traces():
    """Objective-based stop: the reference's stationary window over
    0.5 ||V - W H - off||^2 (offset) and KL(V, W S H) (nsNMF)."""
    V, W0, H0 = _inputs(9, 80, 120, 4)
    for algo in ("offset", "nsnmf"):
        kw = dict(stop="stationary", stationary_th=1e-6, check_interval=5, check_niter=3)
        ref = oracle.run(algo, V, W0, H0, 300, **kw)
        out = _rust.py_nmf_run(algo, V, W0, H0, 300, **kw)
        d_ref, d_gpu = np.asarray(ref["deviances"]), np.asarray(out["deviances"])
        print(algo, "n_iter", out["n_iter"], ref["n_iter"], flush=True)
        assert out["n_iter"] == ref["n_iter"] and len(d_gpu) == len(d_ref)
        assert np.max(np.abs(d_gpu - d_ref) / np.abs(d_ref)) <= TOL
        _check(out, ref, f"{algo} stationary", offset=algo == "offset")


def test_gpu_offset_nsnmf_single_cell_2000x7750():
    V = synthetic.single_cell(2000, 7750, 10, seed=1)
    g = np.random.default_rng(2)
    W0, H0 = g.uniform(0.1, 1, (2000, 10)), g.uniform(0.1, 1, (10, 7750))
    for algo, iters in (("offset", 20), ("nsnmf", 20)):
        ref = oracle.run(algo, V, W0, H0, iters, num_threads=NT)
        out = _rust.py_nmf_run(algo, V, W0, H0, iters)
        _check(out, ref, f"{algo} single-cell", offset=algo == "offset")


def test_gpu_offset_nsnmf_batched_members_equal_single():
    V, W0, H0 = _inputs(12, 90, 150, 5)
    g = np.random.default_rng(13)
    Ws = np.stack([W0, g.uniform(0.1, 1, W0.shape), g.uniform(0.1, 1, W0.shape)])
    Hs = np.stack([H0, g.uniform(0.1, 1, H0.shape), g.uniform(0.1, 1, H0.shape)])
    for algo in ("offset", "nsnmf"):
        prm = _lib.make_params(algo, 9, theta=0.4)
        m = _lib.DeviceMatrix(V)
        try:
            bs = _lib.DeviceSolve(m, 5, prm, batch=3)
            bs.set_factors(Ws, Hs)
            if algo == "offset":
                offs = np.stack([V.mean(axis=1), V.mean(axis=1) * 0.5, np.zeros(90)])
                bs.set_offset(offs)
            bs.launch()
            Wb, Hb, _, _ = bs.fetch()
            ob = bs.offset() if algo == "offset" else None
            for b in range(3):
                one = _lib.DeviceSolve(m, 5, prm)
                one.set_factors(Ws[b], Hs[b])
                if algo == "offset":
                    one.set_offset(offs[b])
                one.launch()
                W1, H1, _, _ = one.fetch()
                assert np.array_equal(Wb[b], W1) and np.array_equal(Hb[b], H1), (algo, b)
                if algo == "offset":
                    assert np.array_equal(ob[b], one.offset())
                    ref = oracle.run("offset", V, Ws[b], Hs[b], 9, offset=offs[b])
                    _check({"W": W1, "H": H1, "offset": one.offset()}, ref, f"batched {b}",
                           offset=True)
                else:
                    ref = oracle.run("nsnmf", V, Ws[b], Hs[b], 9, theta=0.4)
                    _check({"W": W1, "H": H1}, ref, f"batched nsnmf {b}")
                one.close()
            bs.close()
        finally:
            m.close()


def test_gpu_offset_fp32_form():
    """FP32 storage of V: same arithmetic on the rounded V (row sums included)."""
    V, W0, H0 = _inputs(15, 64, 300, 3)
    Vr = synthetic.fp32_reconstruct(V)
    for algo in ("offset", "nsnmf"):
        ref = oracle.run(algo, Vr, W0, H0, 6)
        out = _rust.py_nmf_run(algo, V, W0, H0, 6, precision="fp32")
        _check(out, ref, f"{algo} fp32", offset=algo == "offset")
