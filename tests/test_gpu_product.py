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
    """Widths above 32 run as balanced chunks (33 -> 17 + 16, 70 -> 24 + 23 +
    23); every chunk's panel is written whole on the device, so the zero
    padding of the panels is exercised at each width."""
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
