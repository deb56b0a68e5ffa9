"""GPU parity of Kim-Park sparse NMF (snmf/r, snmf/l: rust/src/lib.rs:1198-1334
with FCNNLS, rust/src/fcnnls.rs) against the CPU oracle.  The products with
V run on the GPU; FCNNLS runs in FP64 on the host.  FP64 tolerance 1e-12 on
max|X_gpu - X_oracle| / max|X_oracle| for W and H and relative 1e-12 on every
per-iteration deviance."""
import os

import numpy as np
import pytest

import nmf_rs
import oracle
from _util import rel
from nmf_rs import _rust, synthetic

pytestmark = pytest.mark.gpu

TOL = 1e-12
NT = min(16, os.cpu_count() or 1)
KW = dict(sparsity_h=0.01, sparsity_w=0.01, smoothness_h=-1.0, smoothness_w=-1.0)


def _inputs(seed, n, p, r):
    g = np.random.default_rng(seed)
    V = synthetic.outer_sum(g.uniform(0, 1, (n, r)), g.uniform(0, 1, (r, p))) + 0.01
    return V, g.uniform(0.1, 1, (n, r)), g.uniform(0.1, 1, (r, p))


def _check(out, ref, what):
    rw, rh = rel(out["W"], ref["W"]), rel(out["H"], ref["H"])
    d_gpu, d_ref = np.asarray(out["deviances"]), np.asarray(ref["deviances"])
    rd = float(np.max(np.abs(d_gpu - d_ref) / np.abs(d_ref))) if len(d_ref) else 0.0
    print(f"{what}: rel W {rw:.2e} H {rh:.2e} dev {rd:.2e}", flush=True)
    assert out["n_iter"] == ref["n_iter"] and len(d_gpu) == len(d_ref)
    assert rw <= TOL and rh <= TOL and rd <= TOL, (what, rw, rh, rd)


@pytest.mark.parametrize("algo", ["snmf/r", "snmf/l"])
@pytest.mark.parametrize("r", [1, 2, 4, 7, 10, 16])
def test_gpu_snmf_ranks(algo, r):
    V, W0, H0 = _inputs(600 + r, 70, 203, r)
    ref = oracle.run(algo, V, W0, H0, 8, num_threads=NT, **KW)
    out = _rust.py_nmf_run(algo, V, W0, H0, 8, **KW)
    _check(out, ref, f"{algo} r={r}")


@pytest.mark.parametrize("algo", ["snmf/r", "snmf/l"])
def test_gpu_snmf_params_and_api(algo):
    """Explicit eta / beta, max_iter 0 (normalised start returned), and the
    nmf() routing of sparsity / smoothness (python/nmf_rs/__init__.py:528-538)."""
    V, W0, H0 = _inputs(61, 50, 40, 3)
    kw = dict(sparsity_h=0.2, sparsity_w=0.3, smoothness_h=0.5, smoothness_w=0.7)
    ref = oracle.run(algo, V, W0, H0, 6, **kw)
    out = _rust.py_nmf_run(algo, V, W0, H0, 6, **kw)
    _check(out, ref, f"{algo} eta/beta")
    ref0 = oracle.run(algo, V, W0, H0, 0, **KW)
    out0 = _rust.py_nmf_run(algo, V, W0, H0, 0, **KW)
    _check(out0, ref0, f"{algo} max_iter 0")
    res = nmf_rs.nmf(V, 3, algo, W0=W0, H0=H0, max_iter=5, sparsity=0.05, smoothness=0.4)
    if algo == "snmf/r":
        kw2 = dict(sparsity_h=0.05, sparsity_w=0.0, smoothness_h=0.0, smoothness_w=0.4)
    else:
        kw2 = dict(sparsity_h=0.0, sparsity_w=0.05, smoothness_h=0.4, smoothness_w=0.0)
    ref2 = oracle.run(algo, V, W0, H0, 5, **kw2)
    assert rel(res.W, ref2["W"]) <= TOL and rel(res.H, ref2["H"]) <= TOL
    assert len(res.deviances) == 5


def test_gpu_snmf_single_cell_2000x7750():
    V = synthetic.single_cell(2000, 7750, 10, seed=1)
    g = np.random.default_rng(5)
    W0, H0 = g.uniform(0.1, 1, (2000, 10)), g.uniform(0.1, 1, (10, 7750))
    for algo in ("snmf/r", "snmf/l"):
        ref = oracle.run(algo, V, W0, H0, 3, num_threads=NT, **KW)
        out = _rust.py_nmf_run(algo, V, W0, H0, 3, **KW)
        _check(out, ref, f"{algo} single-cell")
