"""The benchmark generators (CPU): the library-capture fraction of the
single-cell generator and the device form of the low-rank uniform one."""
import numpy as np
import torch

from nmf_rs import synthetic


def test_single_cell_capture_one_is_the_stated_generator():
    a = synthetic.single_cell(300, 700, seed=4)
    b = synthetic.single_cell(300, 700, seed=4, capture=1.0)
    assert np.array_equal(a, b)


def test_single_cell_sparse_capture_zero_fraction():
    dense = synthetic.single_cell(2000, 3000, seed=1)
    sparse = synthetic.single_cell(2000, 3000, seed=1, capture=synthetic.SPARSE_CAPTURE)
    zd, zs = float((dense == 0).mean()), float((sparse == 0).mean())
    print(f"zero fraction: capture 1 {zd:.3f}, capture {synthetic.SPARSE_CAPTURE} {zs:.3f}")
    assert 0.48 <= zd <= 0.56 and 0.88 <= zs <= 0.91
    # no empty cell or gene, and each cell still sums to log1p of 1e4 counts
    assert np.all(sparse.sum(axis=0) > 0) and np.all(sparse.sum(axis=1) > 0)
    assert np.allclose(np.expm1(sparse).sum(axis=0), 1e4)


def test_single_cell_device_capture_on_cpu():
    # the zero fraction depends on the gene count (the library is spread
    # over these genes): 2000 genes as in the benchmark workloads
    V = synthetic.single_cell_device(2000, 3000, k=10, seed=3, device="cpu", chunk=1500,
                                     capture=synthetic.SPARSE_CAPTURE)
    z = float((V == 0).double().mean())
    assert V.shape == (2000, 3000) and V.dtype == torch.float64
    assert 0.88 <= z <= 0.91, z


def test_lowrank_uniform_device_on_cpu():
    V = synthetic.lowrank_uniform_device(300, 2500, seed=7, device="cpu", chunk=1000)
    W = synthetic.lowrank_uniform_device(300, 2500, seed=7, device="cpu", chunk=1000)
    assert torch.equal(V, W)
    X = V.numpy()
    assert X.shape == (300, 2500) and X.min() > 0.0
    # W* H* with W*, H* ~ U(0, 1) (inner rank 5) plus |N(0, 0.01)|: mean
    # 5 / 4 + 0.01 sqrt(2 / pi), and a rank-5 spectrum above the noise
    assert abs(X.mean() - (1.25 + 0.01 * np.sqrt(2.0 / np.pi))) < 0.02
    s = np.linalg.svd(X, compute_uv=False)
    assert s[4] > 20.0 * s[5]
