"""Batched solves (restarts advancing together) and the consensus layer.

A member of a batched solve runs the single solve's kernels on its own grid
slice; where the batch's split counts equal the single solve's (small
shapes) it is bit-for-bit the single solve, and at shapes where the splits
differ it matches the single solve and the oracle at the FP64 bar.  The consensus matrix is
an exact average of 0/1 indicators, checked against numpy."""
import numpy as np
import pytest

import oracle
from _util import rel
from nmf_rs import _lib, consensus, synthetic
from test_gpu_parity import NT, TOL

pytestmark = pytest.mark.gpu

KINDS = [("brunet", {}), ("lee", {}), ("hals", {}), ("ehals", {}),
         ("dnmf", {"sparsity_h": 0.01, "sparsity_w": 0.01})]


@pytest.mark.parametrize("algo,kw", KINDS)
@pytest.mark.parametrize("form", ["fp64", "fp32-store"])
def test_batched_members_equal_single_solves(algo, kw, form):
    g = np.random.default_rng(11)
    n, p, r, B, iters = 130, 257, 4, 3, 12
    V = g.uniform(0.0, 2.0, (n, p))
    W0 = g.uniform(0.1, 1.0, (B, n, r))
    H0 = g.uniform(0.1, 1.0, (B, r, p))
    m = _lib.DeviceMatrix(V, precision=form)
    prm = _lib.make_params(algo, iters, **kw)
    sb = _lib.DeviceSolve(m, r, prm, batch=B)
    sb.set_factors(W0, H0)
    sb.launch()
    Wb, Hb, it, _ = sb.fetch()
    assert it == iters and Wb.shape == (B, n, r) and Hb.shape == (B, r, p)
    for b in range(B):
        s1 = _lib.DeviceSolve(m, r, prm)
        s1.set_factors(W0[b], H0[b])
        s1.launch()
        W1, H1, _, _ = s1.fetch()
        s1.close()
        print(f"algo={algo} form={form} member={b}", flush=True)
        np.testing.assert_array_equal(Wb[b], W1)
        np.testing.assert_array_equal(Hb[b], H1)
    sb.close()
    m.close()


def test_batched_brunet_member_vs_oracle():
    V, _, _ = synthetic.lowrank_uniform(150, 300, 5, seed=0)
    W0, H0 = consensus.uniform_restarts(150, 300, 6, 4)
    m = _lib.DeviceMatrix(V)
    s = _lib.DeviceSolve(m, 6, _lib.make_params("brunet", 40), batch=4)
    s.set_factors(W0, H0)
    s.launch()
    W, H, _, _ = s.fetch()
    for b in (0, 3):
        ref = oracle.run("brunet", V, W0[b], H0[b], 40, num_threads=NT)
        print("member", b, rel(W[b], ref["W"]), rel(H[b], ref["H"]), flush=True)
        assert rel(W[b], ref["W"]) <= TOL and rel(H[b], ref["H"]) <= TOL
    s.close()
    m.close()


def test_consensus_matrix_exact():
    g = np.random.default_rng(2)
    for runs, p, k in [(1, 5, 2), (10, 300, 4), (70, 97, 3)]:
        L = g.integers(0, k, (runs, p)).astype(np.int32)
        ref = np.zeros((p, p))
        for row in L:
            ref += (row[:, None] == row[None, :])
        ref /= runs
        C = consensus.consensus_matrix(L)
        print(f"consensus runs={runs} p={p}", flush=True)
        np.testing.assert_array_equal(C, ref)


def clustered_matrix(n=120, p=210, k=3, seed=4):
    """V = W* H* + noise where every column belongs to one of k programs
    (H* has one dominant entry per column): a clear rank-k clustering."""
    g = np.random.default_rng(seed)
    Ws = g.uniform(0.0, 1.0, (n, k))
    lab = np.arange(p) % k
    Hs = 0.05 * g.uniform(0.0, 1.0, (k, p))
    Hs[lab, np.arange(p)] += 1.0
    return Ws @ Hs + 0.01 * g.uniform(0.0, 1.0, (n, p)), lab


def test_rank_sweep_small():
    V, lab = clustered_matrix()
    res = consensus.rank_sweep(V, [2, 3, 4], restarts=5, method="brunet", max_iter=100,
                               precision="fp64")
    for r in (2, 3, 4):
        rec = res[r]
        L = rec["labels"]
        assert L.shape == (5, V.shape[1])
        ref = np.zeros((V.shape[1],) * 2)
        for row in L:
            ref += (row[:, None] == row[None, :])
        ref /= 5
        np.testing.assert_array_equal(rec["consensus"], ref)
        assert -1.0 <= rec["cophenetic"] <= 1.0 and 0.0 <= rec["dispersion"] <= 1.0
        print(f"rank {r}: cophenetic {rec['cophenetic']:.4f} dispersion {rec['dispersion']:.4f}",
              flush=True)
    # rank 3 recovers the planted clustering in every restart
    C3 = res[3]["consensus"]
    same = lab[:, None] == lab[None, :]
    assert np.all(C3[same] == 1.0) and np.all(C3[~same] == 0.0)
    assert res[3]["dispersion"] == 1.0 and abs(res[3]["cophenetic"] - 1.0) <= 1e-12


def test_batched_members_with_batch_dependent_splits():
    """1500 x 3000 brunet rank 5, 10 restarts: the single solve splits its
    column pass's rows (few tiles, idle CUs) more than the batch does, so
    the partial sums are added in another order; members still match the
    single solve and the oracle at 1e-12."""
    V, _, _ = synthetic.lowrank_uniform(1500, 3000, 5, seed=0)
    W0, H0 = consensus.uniform_restarts(1500, 3000, 5, 10)
    m = _lib.DeviceMatrix(V)
    prm = _lib.make_params("brunet", 40)
    s = _lib.DeviceSolve(m, 5, prm, batch=10)
    s.set_factors(W0, H0)
    s.launch()
    W, H, _, _ = s.fetch()
    s.close()
    for b in (0, 9):
        s1 = _lib.DeviceSolve(m, 5, prm)
        s1.set_factors(W0[b], H0[b])
        s1.launch()
        W1, H1, _, _ = s1.fetch()
        s1.close()
        ref = oracle.run("brunet", V, W0[b], H0[b], 40, num_threads=NT)
        print("member", b, rel(W[b], W1), rel(H[b], H1), rel(W[b], ref["W"]),
              rel(H[b], ref["H"]), flush=True)
        assert rel(W[b], W1) <= TOL and rel(H[b], H1) <= TOL
        assert rel(W[b], ref["W"]) <= TOL and rel(H[b], ref["H"]) <= TOL
    m.close()
