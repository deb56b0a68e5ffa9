"""The 2000 x 500000 single-cell workload at rank 20 (8 GB in FP64,
generated on the device), lee and dnmf for 25 iterations in every precision
form, checked at full size by what the updates fix:

* lee with column rescaling: every column of W sums to 1 (col_rescale,
  lib.rs:1460-1470);
* lee without rescaling: the objective never increases (the multiplicative
  updates are monotone);
* dnmf: every column of W has unit norm (extract_diagonal, lib.rs:849-853);
* the 4-byte forms against the FP64 form from the same start: relative
  objective <= 1e-6 (dnmf in the BF16 hi + lo forms: 4e-5, see
  objective_bar).

V comes from the generator with 8 % of each library on these genes (about
90 % zeros).

Oracle parity itself on a 2000 x 40000 rank-20 slice (16-row tiles blocked
four to a work-group in the FP32 row pass, the MFMA row pass in FP64)."""
import numpy as np
import pytest

import oracle
from _util import rel
from nmf_rs import _lib, _rust, synthetic
from test_gpu_parity import NT, TOL, assert_parity

pytestmark = pytest.mark.gpu

N, P, RANK, ITERS = 2000, 500000, 20, 25
FORMS = ["fp64", "fp32-store", "bf16x2-store", "fp32", "bf16x2"]
DNMF = {"sparsity_h": 0.01, "sparsity_w": 0.01}


@pytest.fixture(scope="module")
def big():
    import torch

    V = synthetic.single_cell_device(N, P, k=10, seed=3, capture=synthetic.SPARSE_CAPTURE)
    torch.cuda.synchronize()
    print(f"zero fraction {float((V == 0).double().mean()):.3f}", flush=True)
    g = np.random.default_rng(17)
    W0, H0 = g.uniform(0.1, 1.0, (N, RANK)), g.uniform(0.1, 1.0, (RANK, P))
    yield V, W0, H0
    del V
    torch.cuda.empty_cache()


def frob_device(V, W, H):
    import torch

    Wd, Hd = torch.from_numpy(W).cuda(), torch.from_numpy(H).cuda()
    acc = 0.0
    for c0 in range(0, V.shape[1], 65536):
        R = V[:, c0:c0 + 65536] - Wd @ Hd[:, c0:c0 + 65536]
        acc += float((R * R).sum())
    return 0.5 * acc


def solve(V, algo, form, W0, H0, iters=ITERS, **kw):
    m = _lib.DeviceMatrix(V, precision=form)
    s = _lib.DeviceSolve(m, RANK, _lib.make_params(algo, iters, precision=form, **kw))
    try:
        s.set_factors(W0, H0)
        s.launch()
        return s.fetch()
    finally:
        s.close()
        m.close()


def objective_bar(algo, form):
    """1e-6 relative objective against FP64, except dnmf in the two BF16
    hi + lo forms: dnmf's clamps at zero make its 25-iteration path sensitive
    to V's rounding to about 17 significant bits, and the oracle itself, run
    in FP64 on the BF16 hi + lo values of a 2000 x 20000 matrix of this
    generator (about 90 % zeros, rank 20, the same U(0.1, 1) start), ends 3.1e-5
    from its run on the unrounded V (FP32 values: 1.5e-9)."""
    return 4e-5 if algo == "dnmf" and form.startswith("bf16x2") else 1e-6


@pytest.mark.parametrize("algo", ["lee", "dnmf"])
def test_forms_full_size(big, algo):
    V, W0, H0 = big
    kw = DNMF if algo == "dnmf" else {}
    f64 = None
    for form in FORMS:
        W, H, it, _ = solve(V, algo, form, W0, H0, **kw)
        assert it == ITERS and np.all(np.isfinite(W)) and np.all(np.isfinite(H))
        assert W.min() >= 0.0 and H.min() >= 0.0
        if algo == "lee":
            cs = W.sum(axis=0)
            print(f"{algo} {form}: max |colsum - 1| {np.max(np.abs(cs - 1)):.2e}", flush=True)
            assert np.max(np.abs(cs - 1.0)) <= 1e-12
        else:
            nr = np.sqrt((W * W).sum(axis=0))
            print(f"{algo} {form}: max |colnorm - 1| {np.max(np.abs(nr - 1)):.2e}", flush=True)
            assert np.max(np.abs(nr - 1.0)) <= 1e-12
        f = frob_device(V, W, H)
        if f64 is None:
            f64 = f
        else:
            d = abs(f - f64) / f64
            print(f"{algo} {form}: objective {f:.10e} rel vs FP64 {d:.2e}", flush=True)
            assert d <= objective_bar(algo, form)


@pytest.mark.parametrize("form", ["fp64", "fp32", "bf16x2"])
def test_lee_objective_never_increases(big, form):
    """Per-iteration objective (relative-change stop with tolerance 0, so
    every iteration is evaluated and none stops the run)."""
    V, W0, H0 = big
    _, _, it, dev = solve(V, "lee", form, W0, H0, rescale=False, stop="rel_change",
                          rel_tol=0.0, min_iter=1)
    d = np.asarray(dev)
    print(f"lee {form} without rescaling: {it} iterations, objective {d[0]:.6e} -> {d[-1]:.6e}",
          flush=True)
    assert it == ITERS and len(d) == ITERS + 1
    # FP64: roundoff of the objective's own sum; FP32 and BF16 forms: the
    # products' rounding moves each iterate by ~1e-7 relative
    slack = 1e-13 if form == "fp64" else 1e-6
    assert np.all(np.diff(d) <= slack * np.abs(d[:-1]))


@pytest.fixture(scope="module")
def slice40k():
    V = synthetic.single_cell(N, 40000, seed=11)
    g = np.random.default_rng(19)
    return V, g.uniform(0.1, 1.0, (N, RANK)), g.uniform(0.1, 1.0, (RANK, 40000))


@pytest.mark.parametrize("algo,kw", [("lee", {}), ("dnmf", DNMF)])
def test_slice_fp64_vs_oracle(slice40k, algo, kw):
    V, W0, H0 = slice40k
    out = _rust.py_nmf_run(algo, V, W0, H0, 3, **kw)
    ref = oracle.run(algo, V, W0, H0, 3, num_threads=NT, **kw)
    print(f"2000x40000 r=20 {algo}: rel W {rel(out['W'], ref['W']):.2e} "
          f"H {rel(out['H'], ref['H']):.2e}", flush=True)
    if algo == "lee":
        assert rel(out["W"], ref["W"]) <= TOL and rel(out["H"], ref["H"]) <= TOL
    else:
        assert_parity(algo, out, ref, V, W0, H0, 3, "2000x40000 r=20", **kw)


@pytest.mark.parametrize("algo,kw,form", [("lee", {}, "fp32"), ("dnmf", DNMF, "fp32"),
                                          ("lee", {}, "bf16x2")])
def test_slice_lowprec_one_iteration_vs_oracle(slice40k, algo, kw, form):
    """One iteration of the FP32 / BF16 forms against the oracle on the
    values V holds in the form (bars as in test_gpu_fp32 / test_gpu_bf16)."""
    V, W0, H0 = slice40k
    Vr = synthetic.fp32_reconstruct(V) if form == "fp32" else synthetic.bf16x2_reconstruct(V)
    out = _rust.py_nmf_run(algo, V, W0, H0, 1, precision=form, **kw)
    ref = oracle.run(algo, Vr, W0, H0, 1, num_threads=NT, **kw)
    rw, rh = rel(out["W"], ref["W"]), rel(out["H"], ref["H"])
    print(f"2000x40000 r=20 {form} {algo}: rel W {rw:.2e} H {rh:.2e}", flush=True)
    bar = 2e-5 if form == "fp32" else 1e-5
    assert rw <= bar and rh <= bar
