"""The BF16 hi + lo form: lee's W^T V and V H'^T on bf16 MFMA
(v_mfma_f32_16x16x32_bf16) with FP32 accumulation, every operand carried as
two bfloat16 values hi + lo of its float (about 16 significant bits) and a
product formed as hi hi + hi lo + lo hi; Grams, epilogues and the objective
in FP64.  The other methods run FP64 arithmetic on the hi + lo planes, the
same kernels as the bf16x2-store form.

Reference: the oracle (FP64) on the hi + lo values V holds
(bf16x2_reconstruct).  One lee iteration then differs from it by the
factors' hi + lo rounding (<= 2^-18 relative each) and the dropped lo lo
term (<= 2^-18) per product, and by FP32 accumulation (~sqrt(n) 2^-24 of a
sum); W's update uses H' from the same iteration, which carries the first
product's error.  Measured at most 5.1e-6 over the cases here (the rounding
errors are unbiased, so sums average them); the bar is 1e-5.  Over a solve the
figure of merit is the relative objective against the FP64 solve
(north-star bar 1e-6)."""
import numpy as np
import pytest

import oracle
from _util import rel
from nmf_rs import _lib, _rust, synthetic
from test_gpu_parity import NT

pytestmark = pytest.mark.gpu

ONE_IT = 1e-5
DNMF = {"sparsity_h": 0.01, "sparsity_w": 0.01}


def frob(V, W, H):
    R = V - W @ H
    return 0.5 * float(np.sum(R * R))


def check_one(V, W0, H0, what, **kw):
    Vr = synthetic.bf16x2_reconstruct(V)
    ref = oracle.run("lee", Vr, W0, H0, 1, num_threads=NT, **kw)
    out = _rust.py_nmf_run("lee", V, W0, H0, 1, precision="bf16x2", **kw)
    rw, rh = rel(out["W"], ref["W"]), rel(out["H"], ref["H"])
    print(f"bf16x2 1 it lee {what}: rel W {rw:.2e} H {rh:.2e}", flush=True)
    assert rw <= ONE_IT and rh <= ONE_IT
    return rw, rh


@pytest.mark.parametrize("r", [1, 3, 5, 8, 10, 12, 16, 17, 20, 24, 32])
def test_bf16x2_one_iteration_every_rank(r):
    g = np.random.default_rng(900 + r)
    n, p = 70, 203
    V = synthetic.outer_sum(g.uniform(0, 1, (n, r)), g.uniform(0, 1, (r, p))) + 0.01
    W0, H0 = g.uniform(0.1, 1, (n, r)), g.uniform(0.1, 1, (r, p))
    rw, rh = check_one(V, W0, H0, f"r={r}")
    assert rw > 0 or rh > 0  # the products really run in bf16


@pytest.mark.parametrize("shape", [(1, 5), (7, 64), (13, 257), (33, 129), (130, 33), (129, 1000),
                                   (300, 4100)])
def test_bf16x2_ragged_shapes(shape):
    """Rows not a multiple of the 32-row step or of an 8-row block, columns
    not a multiple of the 64-column tile; also without rescaling."""
    n, p = shape
    g = np.random.default_rng(n * 11 + p)
    V = g.uniform(0.0, 3.0, (n, p))
    W0, H0 = g.uniform(0.1, 1.0, (n, 3)), g.uniform(0.1, 1.0, (3, p))
    check_one(V, W0, H0, f"shape {shape}")
    check_one(V, W0, H0, f"shape {shape} no rescale", rescale=False)


def test_bf16x2_row_splits_and_wide_tiles():
    """p >= 32768 (4 row tiles per row-pass work-group) and a tall matrix
    whose few column tiles take row splits in the numerator pass."""
    g = np.random.default_rng(41)
    for n, p, r in [(70, 33000, 20), (3000, 200, 7)]:
        V = g.uniform(0.0, 1.0, (n, p))
        W0, H0 = g.uniform(0.1, 1.0, (n, r)), g.uniform(0.1, 1.0, (r, p))
        check_one(V, W0, H0, f"{n}x{p} r={r}")


def test_bf16x2_other_methods_equal_bf16x2_store():
    """Methods other than lee run the FP64 kernels on the hi + lo planes:
    bit for bit the bf16x2-store form."""
    g = np.random.default_rng(8)
    n, p, r = 150, 700, 6
    V = g.uniform(0.0, 2.0, (n, p))
    W0, H0 = g.uniform(0.1, 1.0, (n, r)), g.uniform(0.1, 1.0, (r, p))
    for algo, kw in [("brunet", {}), ("hals", {}), ("ehals", {}), ("dnmf", DNMF), ("offset", {})]:
        a = _rust.py_nmf_run(algo, V, W0, H0, 10, precision="bf16x2-store", **kw)
        b = _rust.py_nmf_run(algo, V, W0, H0, 10, precision="bf16x2", **kw)
        assert np.array_equal(a["W"], b["W"]) and np.array_equal(a["H"], b["H"]), algo


def test_bf16x2_objective_vs_fp64_single_cell():
    import nmf_rs

    V = synthetic.single_cell(2000, 7750, seed=1)
    W0, H0 = nmf_rs.nndsvd_init(V, 10, fill="mean", seed=0)
    f64 = _rust.py_nmf_run("lee", V, W0, H0, 100)
    b16 = _rust.py_nmf_run("lee", V, W0, H0, 100, precision="bf16x2")
    a, b = frob(V, f64["W"], f64["H"]), frob(V, b16["W"], b16["H"])
    d = abs(a - b) / a
    print(f"bf16x2 lee 2000x7750: objective rel {d:.2e}, rel W {rel(b16['W'], f64['W']):.2e} "
          f"H {rel(b16['H'], f64['H']):.2e}", flush=True)
    assert d <= 1e-6
    assert np.max(np.abs(b16["W"].sum(axis=0) - 1.0)) <= 1e-12


def test_bf16x2_resident_batched_and_deterministic():
    """A batched solve on a resident matrix: members match single solves to
    the form's roundoff; repeated launches give the same bits."""
    g = np.random.default_rng(4)
    n, p, r, B = 400, 3000, 6, 3
    V = g.uniform(0.0, 2.0, (n, p))
    W0, H0 = g.uniform(0.1, 1.0, (B, n, r)), g.uniform(0.1, 1.0, (B, r, p))
    m = _lib.DeviceMatrix(V, precision="bf16x2")
    prm = _lib.make_params("lee", 20, precision="bf16x2")
    sb = _lib.DeviceSolve(m, r, prm, batch=B)
    sb.set_factors(W0, H0)
    sb.launch()
    Wb, Hb, _, _ = sb.fetch()
    sb.launch()
    Wb2, Hb2, _, _ = sb.fetch()
    assert np.array_equal(Wb, Wb2) and np.array_equal(Hb, Hb2)
    sb.close()
    for b in range(B):
        s1 = _lib.DeviceSolve(m, r, prm)
        s1.set_factors(W0[b], H0[b])
        s1.launch()
        W1, H1, _, _ = s1.fetch()
        s1.close()
        print(f"bf16x2 batched member {b}: rel W {rel(Wb[b], W1):.2e}", flush=True)
        assert rel(Wb[b], W1) <= 1e-5 and rel(Hb[b], H1) <= 1e-5
    m.close()


def test_bf16x2_device_tensor_upload_equals_host():
    """V handed over as a device tensor fills both copies as a host upload
    does: the same factors bit for bit."""
    import torch

    g = np.random.default_rng(12)
    n, p, r = 203, 1100, 5
    V = g.uniform(0.0, 2.0, (n, p))
    W0, H0 = g.uniform(0.1, 1.0, (n, r)), g.uniform(0.1, 1.0, (r, p))
    outs = []
    for src in (V, torch.from_numpy(V).cuda()):
        m = _lib.DeviceMatrix(src, precision="bf16x2")
        s = _lib.DeviceSolve(m, r, _lib.make_params("lee", 15, precision="bf16x2"))
        s.set_factors(W0, H0)
        s.launch()
        outs.append(s.fetch())
        s.close()
        m.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
