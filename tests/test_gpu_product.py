"""GPU products with a resident V (nmfg_matrix_mul: V B and V^T B), the
range-finder passes of the device NNDSVD (python/nmf_rs/__init__.py:172-174
calls scikit-learn's randomized_svd, whose products these are), against
numpy in FP64.  Tolerance 1e-12 on max|C_gpu - C_np| / max|C_np|: the pass
kernels sum in another order than BLAS."""
import numpy as np
import pytest

from _util import rel
from nmf_rs import _lib, synthetic

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("shape", [(37, 130), (130, 37), (1, 200), (200, 1), (300, 1000)])
@pytest.mark.parametrize("k", [1, 5, 20, 32, 33, 40, 64, 70])
def test_gpu_product_widths_and_shapes(shape, k):
    """Widths above 32 run as chunks of 32 and a remainder (33 -> 32 + 1,
    70 -> 32 + 32 + 6); every chunk's panel is written whole on the device,
    so the zero padding of the panels is exercised at each width."""
    n, p = shape
    g = np.random.default_rng(n * 100 + p + k)
    V = g.uniform(0.0, 1.0, (n, p))
    m = _lib.DeviceMatrix(V)
    try:
        Bp = g.normal(size=(p, k))
        Bn = g.normal(size=(n, k))
        c1 = m.mul(Bp, trans=False)
        c2 = m.mul(Bn, trans=True)
        assert c1.shape == (n, k) and c2.shape == (p, k)
        r1, r2 = rel(c1, V @ Bp), rel(c2, V.T @ Bn)
        print(f"{shape} k={k}: rel V B {r1:.2e}  V^T B {r2:.2e}", flush=True)
        assert r1 <= TOL and r2 <= TOL
        # the staging buffer is reused: a narrower call after a wider one
        c3 = m.mul(Bp[:, :1], trans=False)
        assert rel(c3, V @ Bp[:, :1]) <= TOL
    finally:
        m.close()


def test_gpu_product_fp32_storage_and_repeat():
    V = synthetic.single_cell(400, 3000, 10, seed=4)
    Vr = synthetic.fp32_reconstruct(V)
    g = np.random.default_rng(9)
    m = _lib.DeviceMatrix(V, precision="fp32")
    try:
        B = g.normal(size=(3000, 40))
        first = m.mul(B)
        assert rel(first, Vr @ B) <= TOL
        again = m.mul(B)
        assert np.array_equal(first, again)  # same chunks, same order
        Bt = g.normal(size=(400, 40))
        assert rel(m.mul(Bt, trans=True), Vr.T @ Bt) <= TOL
    finally:
        m.close()


def test_gpu_product_rejects_bad_shapes():
    m = _lib.DeviceMatrix(np.ones((5, 7)))
    try:
        with pytest.raises(ValueError):
            m.mul(np.ones((5, 2)), trans=False)
        with pytest.raises(ValueError):
            m.mul(np.ones((7, 2)), trans=True)
    finally:
        m.close()


# ---------------------------------------------------------------------------
# Device LU with partial pivoting (P L, scipy.linalg.lu(permute_l=True)[0])
# and the device range finder of the NNDSVD's randomized SVD
# ---------------------------------------------------------------------------
def _lu_structure(PL, X):
    """Partial pivoting's structure: every multiplier |l| <= 1 and one row
    per step j holds 1 at column j and 0 after it."""
    m, w = X.shape
    assert np.all(np.abs(PL) <= 1.0 + 1e-15)
    for j in range(w):
        unit = np.nonzero((PL[:, j] == 1.0) & np.all(PL[:, j + 1:] == 0.0, axis=1))[0]
        assert unit.size >= 1, j


@pytest.mark.parametrize("shape", [(1, 1), (5, 1), (40, 40), (300, 7), (1000, 40), (70000, 33)])
def test_gpu_lu_permute_l_matches_scipy(shape):
    from scipy import linalg

    g = np.random.default_rng(shape[0] + 7 * shape[1])
    X = g.normal(size=shape)
    PL = _lib.lu_permute_l(X)
    ref, _ = linalg.lu(X, permute_l=True, check_finite=False)
    r = rel(PL, ref)
    print(f"lu {shape}: rel {r:.2e}", flush=True)
    assert r <= TOL
    _lu_structure(PL, X)


def test_gpu_lu_ties_and_zero_column():
    """Exact arithmetic: a tie on |x| goes to the first row in the swapped
    order (LAPACK's idamax), a zero column is left unscaled."""
    from scipy import linalg

    X = np.array([[1.0, 2.0, 0.0], [-1.0, 5.0, 0.0], [1.0, 3.0, 0.0], [0.5, 1.0, 0.0]])
    PL = _lib.lu_permute_l(X)
    ref, _ = linalg.lu(X, permute_l=True, check_finite=False)
    assert np.array_equal(PL, ref), (PL, ref)
    Z = np.zeros((6, 3))
    Z[:, 1] = [0.0, 2.0, -2.0, 1.0, 0.0, 0.5]
    PZ = _lib.lu_permute_l(Z)
    refz, _ = linalg.lu(Z, permute_l=True, check_finite=False)
    assert np.array_equal(PZ, refz), (PZ, refz)


def test_gpu_lu_rejects_wide():
    with pytest.raises(ValueError):
        _lib.lu_permute_l(np.ones((3, 5)))


@pytest.mark.parametrize("case", [("tall", 3000, 800, 10, "fp64"), ("wide", 600, 9000, 30, "fp64"),
                                  ("wide32", 900, 5000, 12, "fp32")])
def test_gpu_nndsvd_device_range_finder_matches_sklearn(case):
    """nndsvd_init with the power iterations on the device against the
    scikit-learn sequence (oracle.nndsvd_ref): the device products and LU
    steps round differently, the subspace iteration keeps it at roundoff."""
    import nmf_rs
    import oracle

    _, n, p, rank, prec = case
    V = synthetic.single_cell(n, p, 10, seed=n + p)
    Vr = synthetic.fp32_reconstruct(V) if prec == "fp32" else V
    W0, H0 = oracle.nndsvd_ref(Vr, rank, fill="mean", seed=0)
    m = _lib.DeviceMatrix(V, precision=prec)
    try:
        W, H = nmf_rs.nndsvd_on_device(m, rank, fill="mean", seed=0, v_mean=float(Vr.mean()))
    finally:
        m.close()
    rw, rh = rel(W, W0), rel(H, H0)
    print(f"nndsvd {case}: rel W {rw:.2e} H {rh:.2e}", flush=True)
    assert rw <= 1e-11 and rh <= 1e-11
