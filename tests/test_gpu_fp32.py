"""The FP32 form: lee's W^T V and V H'^T on FP32 MFMA with FP32
accumulation over float V and float copies of the streamed factor, brunet's
W H, the ratio V / W H and its products in FP32 on the VALU; Grams,
epilogues and the objective in FP64.  The HALS family (hals, ehals, dnmf)
runs FP64 arithmetic on the float V in this form.

Reference: the oracle (FP64) on the float-rounded V.  For lee and brunet one
iteration differs from it only by the FP32 rounding of the two products
(relative error of a sum of n products ~ sqrt(n) 2^-24, so 2e-5 bounds it at
the sizes here); the HALS family matches it at the FP64 bar.  Over a solve
the figure of merit is the relative objective against the FP64 solve
(north-star bar 1e-6)."""
import numpy as np
import pytest

import oracle
from _util import rel
from nmf_rs import _lib, _rust, synthetic
from test_gpu_parity import NT, assert_parity

pytestmark = pytest.mark.gpu

ONE_IT = 2e-5
DNMF = {"sparsity_h": 0.01, "sparsity_w": 0.01}
PRODUCTS = [("lee", {}), ("brunet", {})]            # FP32 products
HALS_FAMILY = [("hals", {}), ("ehals", {}), ("dnmf", DNMF)]  # FP64 arithmetic on float V
FAMILY = PRODUCTS + HALS_FAMILY


def frob(V, W, H):
    R = V - W @ H
    return 0.5 * float(np.sum(R * R))


def check_one(algo, kw, V, W0, H0, what):
    Vr = synthetic.fp32_reconstruct(V)
    ref = oracle.run(algo, Vr, W0, H0, 1, num_threads=NT, **kw)
    out = _rust.py_nmf_run(algo, V, W0, H0, 1, precision="fp32", **kw)
    rw, rh = rel(out["W"], ref["W"]), rel(out["H"], ref["H"])
    print(f"fp32 1 it {algo} {what}: rel W {rw:.2e} H {rh:.2e}", flush=True)
    if (algo, kw) in PRODUCTS:
        assert rw <= ONE_IT and rh <= ONE_IT
        return rw, rh
    assert_parity(algo, out, ref, Vr, W0, H0, 1, f"fp32 form {what}", **kw)
    return rw, rh


@pytest.mark.parametrize("algo,kw", FAMILY)
@pytest.mark.parametrize("r", [1, 3, 5, 7, 8, 9, 10, 12, 13, 16, 18, 20, 24, 28, 32])
def test_fp32_one_iteration_every_rank(algo, kw, r):
    g = np.random.default_rng(700 + r)
    n, p = 70, 203
    V = synthetic.outer_sum(g.uniform(0, 1, (n, r)), g.uniform(0, 1, (r, p))) + 0.01
    W0, H0 = g.uniform(0.1, 1, (n, r)), g.uniform(0.1, 1, (r, p))
    rw, rh = check_one(algo, kw, V, W0, H0, f"r={r}")
    if (algo, kw) in PRODUCTS:
        assert rw > 0 or rh > 0  # the products really run in FP32


@pytest.mark.parametrize("shape", [(1, 5), (13, 257), (130, 33), (129, 1000), (300, 4100)])
def test_fp32_ragged_shapes(shape):
    n, p = shape
    g = np.random.default_rng(n * 7 + p)
    V = g.uniform(0.0, 3.0, (n, p))
    W0, H0 = g.uniform(0.1, 1.0, (n, 3)), g.uniform(0.1, 1.0, (3, p))
    for algo, kw in FAMILY:
        check_one(algo, kw, V, W0, H0, f"shape {shape}")


def test_fp32_wide_row_tiles():
    """p >= 32768: the row pass blocks 4 row tiles per work-group."""
    g = np.random.default_rng(31)
    n, p, r = 70, 33000, 20
    V = g.uniform(0.0, 1.0, (n, p))
    W0, H0 = g.uniform(0.1, 1.0, (n, r)), g.uniform(0.1, 1.0, (r, p))
    for algo, kw in PRODUCTS + HALS_FAMILY[:1]:
        check_one(algo, kw, V, W0, H0, "70x33000 r=20")


def test_fp32_form_hals_high_rank_stays_with_fp64():
    """2000 x 8000 single-cell, rank 12, hals for 5 iterations: with FP32
    products the iterates left the FP64 solve's (max-relative W difference
    4); with FP64 arithmetic on the float V the objective stays within 1e-6
    of the FP64 solve and W within 1e-4."""
    g = np.random.default_rng(5)
    V = synthetic.single_cell(2000, 8000, seed=2)
    W0, H0 = g.uniform(0.1, 1, (2000, 12)), g.uniform(0.1, 1, (12, 8000))
    a = _rust.py_nmf_run("hals", V, W0, H0, 5)
    b = _rust.py_nmf_run("hals", V, W0, H0, 5, precision="fp32")
    fa, fb = frob(V, a["W"], a["H"]), frob(V, b["W"], b["H"])
    print(f"hals r=12 5 it: rel objective {abs(fa - fb) / fa:.2e} rel W {rel(b['W'], a['W']):.2e}",
          flush=True)
    assert abs(fa - fb) / fa <= 1e-6 and rel(b["W"], a["W"]) <= 1e-4


@pytest.fixture(scope="module")
def sc2():
    import nmf_rs

    V = synthetic.single_cell(2000, 7750, seed=1)
    W0, H0 = nmf_rs.nndsvd_init(V, 10, fill="mean", seed=0)
    return V, W0, H0


@pytest.mark.parametrize("algo,kw,iters", [("lee", {}, 100), ("hals", {}, 25), ("ehals", {}, 25),
                                           ("dnmf", DNMF, 25), ("brunet", {}, 100)])
def test_fp32_objective_vs_fp64_single_cell(sc2, algo, kw, iters):
    V, W0, H0 = sc2
    f64 = _rust.py_nmf_run(algo, V, W0, H0, iters, **kw)
    f32 = _rust.py_nmf_run(algo, V, W0, H0, iters, precision="fp32", **kw)
    obj = (lambda V, W, H: oracle.kl_divergence(V, W, H)) if algo == "brunet" else frob
    a, b = obj(V, f64["W"], f64["H"]), obj(V, f32["W"], f32["H"])
    d = abs(a - b) / a
    print(f"fp32 {algo} 2000x7750: objective rel {d:.2e}, rel W {rel(f32['W'], f64['W']):.2e} "
          f"H {rel(f32['H'], f64['H']):.2e}", flush=True)
    assert d <= 1e-6


@pytest.fixture(scope="module")
def sc2_sparse():
    import nmf_rs

    V = synthetic.single_cell(2000, 7750, seed=1, capture=synthetic.SPARSE_CAPTURE)
    W0, H0 = nmf_rs.nndsvd_init(V, 10, fill="mean", seed=0)
    return V, W0, H0


@pytest.mark.parametrize("algo,kw,iters,form", [("lee", {}, 100, "fp32"), ("lee", {}, 100, "bf16x2"),
                                                ("brunet", {}, 100, "fp32"), ("hals", {}, 25, "fp32"),
                                                ("dnmf", DNMF, 25, "fp32")])
def test_lowprec_objective_vs_fp64_sparse_single_cell(sc2_sparse, algo, kw, iters, form):
    """The 89.6 %-zeros single-cell matrix: objective within 1e-6 of FP64."""
    V, W0, H0 = sc2_sparse
    f64 = _rust.py_nmf_run(algo, V, W0, H0, iters, **kw)
    lo = _rust.py_nmf_run(algo, V, W0, H0, iters, precision=form, **kw)
    obj = (lambda V, W, H: oracle.kl_divergence(V, W, H)) if algo == "brunet" else frob
    a, b = obj(V, f64["W"], f64["H"]), obj(V, lo["W"], lo["H"])
    d = abs(a - b) / a
    print(f"{form} {algo} sparse 2000x7750: objective rel {d:.2e}, rel W {rel(lo['W'], f64['W']):.2e} "
          f"H {rel(lo['H'], f64['H']):.2e}", flush=True)
    assert d <= 1e-6


def test_fp32_resident_batched_and_deterministic():
    """A batched FP32-form solve on a resident matrix: members match single
    solves to FP32 roundoff; repeated launches give the same bits."""
    g = np.random.default_rng(3)
    n, p, r, B = 400, 3000, 6, 3
    V = g.uniform(0.0, 2.0, (n, p))
    W0, H0 = g.uniform(0.1, 1.0, (B, n, r)), g.uniform(0.1, 1.0, (B, r, p))
    m = _lib.DeviceMatrix(V, precision="fp32")
    prm = _lib.make_params("lee", 20, precision="fp32")
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
        print(f"fp32 batched member {b}: rel W {rel(Wb[b], W1):.2e}", flush=True)
        assert rel(Wb[b], W1) <= 1e-5 and rel(Hb[b], H1) <= 1e-5
    m.close()
