"""Storage forms of V (FP32, BF16 hi+lo) on the GPU.

The store forms keep W, H and all arithmetic in FP64, so a form must give
the FP64 algorithm run on the values it stores: the oracle on the rounded V
is the reference (same 1e-12 bar as FP64).  Against the FP64 solve on the
unrounded V the forms are compared by the relative objective (FP32: 1e-6),
on the single-cell workloads.  The FP32 form's own tests (products in FP32)
are in test_gpu_fp32.py."""
import numpy as np
import pytest

import oracle
from _util import rel
from nmf_rs import _lib, _rust, synthetic
from test_gpu_parity import NT, TOL, assert_parity

pytestmark = pytest.mark.gpu

ROUND = {"fp32-store": synthetic.fp32_reconstruct, "bf16x2-store": synthetic.bf16x2_reconstruct}
RUNS = [("lee", 30), ("brunet", 30), ("hals", 8), ("ehals", 8), ("dnmf", 8)]


def kw_for(algo):
    return {"sparsity_h": 0.01, "sparsity_w": 0.01} if algo == "dnmf" else {}


@pytest.mark.parametrize("form", ["fp32-store", "bf16x2-store"])
def test_form_equals_fp64_algorithm_on_stored_values(form):
    V, W0, H0 = synthetic.lowrank_uniform(96, 80, 4, seed=3)
    Vr = ROUND[form](V)
    assert np.max(np.abs(Vr - V)) > 0  # the form really rounds
    for algo, iters in RUNS:
        kw = kw_for(algo)
        print(f"form={form} algo={algo} iters={iters}", flush=True)
        ref = oracle.run(algo, Vr, W0, H0, iters, num_threads=NT, **kw)
        out = _rust.py_nmf_run(algo, V, W0, H0, iters, precision=form, **kw)
        if algo in ("lee", "brunet"):
            assert rel(out["W"], ref["W"]) <= TOL and rel(out["H"], ref["H"]) <= TOL, (form, algo)
        else:
            assert_parity(algo, out, ref, Vr, W0, H0, iters, form, **kw)


def test_form_ragged_shapes():
    # odd p (partial tiles, odd bf16 pair at the end), rows not a multiple of 8
    g = np.random.default_rng(5)
    for n, p in [(13, 257), (130, 33), (1, 5)]:
        V = g.uniform(0.0, 3.0, (n, p))
        W0 = g.uniform(0.1, 1.0, (n, 3))
        H0 = g.uniform(0.1, 1.0, (3, p))
        for form in ("fp32-store", "bf16x2-store"):
            ref = oracle.run("lee", ROUND[form](V), W0, H0, 10, num_threads=NT)
            out = _rust.py_nmf_run("lee", V, W0, H0, 10, precision=form)
            print(f"shape=({n},{p}) form={form}", rel(out["W"], ref["W"]), flush=True)
            assert rel(out["W"], ref["W"]) <= TOL and rel(out["H"], ref["H"]) <= TOL


def frob_obj(V, W, H):
    R = V - W @ H
    return 0.5 * float(np.sum(R * R))


@pytest.fixture(scope="module")
def single_cell_2000x7750():
    import nmf_rs

    V = synthetic.single_cell(2000, 7750, seed=1)
    W0, H0 = nmf_rs.nndsvd_init(V, 10, fill="mean", seed=0)
    return V, W0, H0


@pytest.mark.parametrize("algo", ["lee", "dnmf"])
def test_forms_objective_vs_fp64_single_cell(single_cell_2000x7750, algo):
    V, W0, H0 = single_cell_2000x7750
    kw = kw_for(algo)
    f = {}
    for form in ("fp64", "fp32-store", "bf16x2-store"):
        out = _rust.py_nmf_run(algo, V, W0, H0, 25, precision=form, **kw)
        f[form] = frob_obj(V, out["W"], out["H"])
    d32 = abs(f["fp32-store"] - f["fp64"]) / f["fp64"]
    d16 = abs(f["bf16x2-store"] - f["fp64"]) / f["fp64"]
    print(f"{algo}: objective fp64 {f['fp64']:.10e}; rel fp32 {d32:.3e} bf16x2 {d16:.3e}", flush=True)
    assert d32 <= 1e-6
    assert d16 <= 1e-6


def test_device_tensor_upload_and_forms():
    """V generated on the device (torch) and uploaded device to device; the
    FP32 form of it matches the FP64 solve on the rounded values."""
    import torch

    Vt = synthetic.single_cell_device(300, 4100, k=6, seed=3, chunk=1000)
    V = Vt.cpu().numpy()
    g = np.random.default_rng(1)
    W0 = g.uniform(0.1, 1.0, (300, 6))
    H0 = g.uniform(0.1, 1.0, (6, 4100))
    for form in ("fp64", "fp32-store"):
        m = _lib.DeviceMatrix(Vt, precision=form)
        s = _lib.DeviceSolve(m, 6, _lib.make_params("lee", 20))
        s.set_factors(W0, H0)
        s.launch()
        W, H, n_iter, _ = s.fetch()
        Vr = V if form == "fp64" else synthetic.fp32_reconstruct(V)
        ref = oracle.run("lee", Vr, W0, H0, 20, num_threads=NT)
        print(f"device upload form={form}", rel(W, ref["W"]), flush=True)
        assert rel(W, ref["W"]) <= TOL and rel(H, ref["H"]) <= TOL
        s.close()
        m.close()
    del Vt
    torch.cuda.empty_cache()
