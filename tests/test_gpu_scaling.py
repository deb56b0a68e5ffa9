"""The GPU path on the reference's multi-scale benchmark
(examples/build_scaling_notebook.py; inputs and metrics restated in
tests/_scaling.py) against the results that notebook printed
(examples/benchmark_scaling.ipynb cells 18-21) and against the oracle's run
of the same tier (tests/golden/scaling_medium.json).

Tier "medium": 3000 x 50000, rank 15, ehals from 3 random starts (100
iterations), hals from NNDSVD (25 iterations).  Tier "large": 5000 x 200000
(8 GB FP64), rank 20, ehals from one random start and hals from NNDSVD (20
iterations), 50 iterations otherwise -- too large for the CPU oracle, so
only the notebook's figures pin it."""
import json
from pathlib import Path

import numpy as np
import pytest

import _scaling as S
import nmf_rs

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "scaling_medium.json"


def sig7(x):
    return float(f"{x:.6e}")


@pytest.fixture(scope="module")
def medium():
    M = S.MEDIUM
    V, W_true, _ = S.make_synthetic(M["n"], M["p"], rank=M["rank"], seed=M["seed"])
    return V, W_true


def test_gpu_medium_tier_ehals_and_hals_nndsvd(medium):
    V, W_true = medium
    M = S.MEDIUM
    ref = json.loads(GOLD.read_text())
    sub = S.consensus_subset(M["p"])
    losses, spars = [], []
    cons = np.zeros((len(sub), len(sub)))
    for run in range(M["runs"]):
        W0, H0 = S.random_start(V, M["rank"], run)
        res = nmf_rs.nmf(V, rank=M["rank"], method="ehals", W0=W0, H0=H0, max_iter=M["iters"])
        f = S.loss(V, res.W, res.H)
        f_ref = float.fromhex(ref["ehals_runs"][run]["loss"])
        print(f"medium ehals run {run}: loss {f:.10e} oracle {f_ref:.10e} "
              f"rel {abs(f - f_ref) / f_ref:.2e}", flush=True)
        assert abs(f - f_ref) <= 1e-9 * f_ref
        losses.append(f)
        spars.append(S.factor_sparsity_w(res.W))
        cons += S.connectivity(res.H[:, sub])
    assert sig7(np.mean(losses)) == M["ehals_mean_loss"]
    assert round(float(np.mean(spars)), 3) == M["sparsity_w_3dp"]["ehals"]
    assert round(S.dispersion(cons / M["runs"]), 3) == M["dispersion_3dp"]["ehals"]

    W0, H0 = nmf_rs.nndsvd_init(V, M["rank"], fill="mean", seed=0)
    res = nmf_rs.nmf(V, rank=M["rank"], method="hals", W0=W0, H0=H0, max_iter=M["nndsvd_iters"])
    amari = S.amari_error(res.W, W_true)
    hs = S.factor_sparsity_w(res.W)
    f = S.loss(V, res.W, res.H)
    f_ref = float.fromhex(ref["hals_nndsvd_loss"])
    print(f"medium hals+NNDSVD: Amari {amari:.9e} sparsity {hs:.9e} loss rel "
          f"{abs(f - f_ref) / f_ref:.2e}", flush=True)
    assert sig7(amari) == M["hals_nndsvd_amari"]
    assert sig7(hs) == M["hals_nndsvd_sparsity_w"]
    assert abs(f - f_ref) <= 1e-9 * f_ref


@pytest.mark.parametrize("method", ["lee", "brunet", "hals", "dnmf"])
def test_gpu_medium_tier_other_methods(medium, method):
    """The notebook's 3-decimal figures for the other methods (random
    starts, 100 iterations, K = 3): Hoyer sparsity of W (cell 19) and
    consensus dispersion (cell 18).  dnmf runs with the nmf() default
    sparsity 0.01 (python/nmf_rs/__init__.py:533-536)."""
    V, _ = medium
    M = S.MEDIUM
    sub = S.consensus_subset(M["p"])
    spars = []
    cons = np.zeros((len(sub), len(sub)))
    for run in range(M["runs"]):
        W0, H0 = S.random_start(V, M["rank"], run)
        res = nmf_rs.nmf(V, rank=M["rank"], method=method, W0=W0, H0=H0, max_iter=M["iters"])
        spars.append(S.factor_sparsity_w(res.W))
        cons += S.connectivity(res.H[:, sub])
    sp, disp = float(np.mean(spars)), S.dispersion(cons / M["runs"])
    print(f"medium {method}: sparsity W {sp:.6f} dispersion {disp:.6f}", flush=True)
    assert round(sp, 3) == M["sparsity_w_3dp"][method]
    assert round(disp, 3) == M["dispersion_3dp"][method]


def test_gpu_large_tier_ehals_and_hals_nndsvd():
    """5000 x 200000 rank 20 (8 GB in FP64): ehals from run 0's random start,
    50 iterations, loss 5.551842e+07; hals from NNDSVD, 20 iterations, Amari
    error 2.146043e-01 and Hoyer sparsity of W 2.158551e-01 (cell 21)."""
    V, W_true, _ = S.make_synthetic(5000, 200000, rank=20, seed=2)
    W0, H0 = S.random_start(V, 20, 0)
    res = nmf_rs.nmf(V, rank=20, method="ehals", W0=W0, H0=H0, max_iter=50)
    f = S.loss(V, res.W, res.H)
    print(f"large ehals loss {f:.9e}", flush=True)
    assert sig7(f) == 5.551842e+07
    W0, H0 = nmf_rs.nndsvd_init(V, 20, fill="mean", seed=0)
    res = nmf_rs.nmf(V, rank=20, method="hals", W0=W0, H0=H0, max_iter=20)
    amari, hs = S.amari_error(res.W, W_true), S.factor_sparsity_w(res.W)
    print(f"large hals+NNDSVD Amari {amari:.9e} sparsity {hs:.9e}", flush=True)
    assert sig7(amari) == 2.146043e-01
    assert sig7(hs) == 2.158551e-01
