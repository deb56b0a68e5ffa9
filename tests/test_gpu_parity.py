"""GPU parity: the gfx950 path against the CPU oracle and the reference's
known answers.  FP64 tolerance 1e-12 on the scale-free metric
max|X_gpu - X_oracle| / max|X_oracle| for W and H, and relative 1e-12 on
objectives (0.5 ||V - WH||^2 and the KL divergence, rust/src/lib.rs:1406-1445)."""
import os

import numpy as np
import pytest

import nmf_rs
import oracle
from _util import golden_cases, quickstart_inputs, rel, unhex
from nmf_rs import _rust, synthetic

pytestmark = pytest.mark.gpu

TOL = 1e-12
NT = min(16, os.cpu_count() or 1)


def gpu_run(algo, V, W0, H0, iters, **kw):
    return _rust.py_nmf_run(algo, V, W0, H0, iters, **kw)


def sensitivity(algo, V, W0, H0, iters, ref, **kw):
    """How far the oracle itself moves when its inputs move by one ulp: the
    roundoff floor of this (method, input).  HALS-type sweeps on noise-like
    inputs amplify roundoff by ~1e8 (clamps at eps flip), so no summation
    order other than the oracle's own can get closer than this."""
    a = oracle.run(algo, V, np.nextafter(W0, np.inf), H0, iters, **kw)
    b = oracle.run(algo, V, W0, np.nextafter(H0, np.inf), iters, **kw)
    return max(rel(a["W"], ref["W"]), rel(a["H"], ref["H"]), rel(b["W"], ref["W"]),
               rel(b["H"], ref["H"]))


FLOOR_FACTOR = 3.0


def assert_parity(algo, out, ref, V, W0, H0, iters, what="", **kw):
    """FP64 parity: max-relative error <= max(1e-12, 3 x the input's own
    roundoff floor (sensitivity())), i.e. 1e-12 or inside the spread of the
    oracle's own 1-ulp-perturbed runs.  For
    the HALS family the first iteration, where roundoff has had little room
    to grow, is checked the same way on its own (1e-12, else 3x its own
    floor)."""
    if algo in ("hals", "ehals", "dnmf") and iters > 1:
        o1 = gpu_run(algo, V, W0, H0, 1, **kw)
        r1 = oracle.run(algo, V, W0, H0, 1, **kw)
        assert_parity(algo, o1, r1, V, W0, H0, 1, what + " (first iteration)", **kw)
    rw, rh = rel(out["W"], ref["W"]), rel(out["H"], ref["H"])
    if rw <= TOL and rh <= TOL:
        return
    sens = sensitivity(algo, V, W0, H0, iters, ref, **kw)
    print(f"{what} {algo}: rel W {rw:.3e} H {rh:.3e} vs 1-ulp floor {sens:.3e} "
          f"(ratio {max(rw, rh) / sens:.2f})", flush=True)
    bar = max(TOL, FLOOR_FACTOR * sens)
    assert rw <= bar and rh <= bar, (what, algo, rw, rh, sens)


def test_gpu_golden_tiny_runs():
    for c in golden_cases():
        V, W0, H0 = unhex(c["V"]), unhex(c["W0"]), unhex(c["H0"])
        for key, rr in c["runs"].items():
            algo = rr.get("algo", key)
            kw = {"sparsity_h": 0.01, "sparsity_w": 0.01} if algo == "dnmf" else {}
            if algo in ("snmf/r", "snmf/l"):
                kw = {"sparsity_h": 0.01, "sparsity_w": 0.01, "smoothness_h": -1.0,
                      "smoothness_w": -1.0}
            if "theta" in rr:
                kw["theta"] = rr["theta"]
            if "weight" in rr:
                kw["weight"] = unhex(rr["weight"])
            print(f"case={c['name']} run={key} shape={V.shape}", flush=True)
            out = gpu_run(algo, V, W0, H0, rr["iters"], **kw)
            ref = {"W": unhex(rr["W"]), "H": unhex(rr["H"])}
            if "off" in rr:
                assert rel(out["offset"], unhex(rr["off"])) <= TOL, (c["name"], key)
            if algo in ("lee", "brunet", "offset", "nsnmf", "lsnmf", "snmf/r", "snmf/l"):
                assert rel(out["W"], ref["W"]) <= TOL, (c["name"], key)
                assert rel(out["H"], ref["H"]) <= TOL, (c["name"], key)
            else:
                assert_parity(algo, out, ref, V, W0, H0, rr["iters"], c["name"], **kw)


def test_gpu_single_steps_golden():
    for c in golden_cases():
        if not c["steps"]:
            continue
        V, W0, H0 = unhex(c["V"]), unhex(c["W0"]), unhex(c["H0"])
        st = c["steps"]
        assert rel(nmf_rs.update_h_lee(V, W0, H0), unhex(st["lee_h"])) <= TOL
        assert rel(nmf_rs.update_w_lee(V, W0, H0), unhex(st["lee_w"])) <= TOL
        assert rel(nmf_rs.update_h_brunet(V, W0, H0), unhex(st["brunet_h"])) <= TOL
        assert rel(nmf_rs.update_w_brunet(V, W0, H0), unhex(st["brunet_w"])) <= TOL


def test_gpu_lowrank_500x300_lee():
    V, W0, H0 = synthetic.lowrank_uniform()
    ref = oracle.run("lee", V, W0, H0, 100, num_threads=NT)
    out = gpu_run("lee", V, W0, H0, 100)
    print("lowrank lee rel W", rel(out["W"], ref["W"]), "H", rel(out["H"], ref["H"]))
    assert rel(out["W"], ref["W"]) <= TOL and rel(out["H"], ref["H"]) <= TOL
    f_gpu = oracle.euclidean_distance(V, out["W"], out["H"])
    f_ref = oracle.euclidean_distance(V, ref["W"], ref["H"])
    assert abs(f_gpu - f_ref) / f_ref <= TOL


@pytest.mark.parametrize("r", [1, 2, 3, 7, 9, 11, 13, 17, 25, 31, 32])
def test_gpu_every_rank_bucket(r):
    g = np.random.default_rng(100 + r)
    n, p = 70, 203
    V = synthetic.outer_sum(g.uniform(0, 1, (n, r)), g.uniform(0, 1, (r, p))) + 0.01
    W0, H0 = g.uniform(0.1, 1, (n, r)), g.uniform(0.1, 1, (r, p))
    for algo in ("lee", "brunet", "hals", "ehals", "dnmf"):
        kw = {"sparsity_h": 0.01, "sparsity_w": 0.01} if algo == "dnmf" else {}
        ref = oracle.run(algo, V, W0, H0, 6, **kw)
        out = gpu_run(algo, V, W0, H0, 6, **kw)
        assert_parity(algo, out, ref, V, W0, H0, 6, f"rank{r}", **kw)


def test_gpu_rank_beyond_max_raises():
    V = np.ones((8, 6))
    r = nmf_rs._lib.lib().nmfg_max_rank() + 1
    with pytest.raises(RuntimeError):
        gpu_run("lee", V, np.ones((8, r)), np.ones((r, 6)), 1)


@pytest.mark.parametrize("shape", [(1, 1), (1, 300), (300, 1), (17, 129), (129, 17)])
def test_gpu_edge_shapes(shape):
    n, p = shape
    g = np.random.default_rng(n * 1000 + p)
    r = 2
    V = g.uniform(0, 1, (n, p))
    W0, H0 = g.uniform(0.1, 1, (n, r)), g.uniform(0.1, 1, (r, p))
    for algo in ("lee", "brunet", "hals"):
        ref = oracle.run(algo, V, W0, H0, 4)
        out = gpu_run(algo, V, W0, H0, 4)
        assert_parity(algo, out, ref, V, W0, H0, 4, str(shape))


def test_gpu_zero_rows_cols_and_zero_factor_column():
    g = np.random.default_rng(4)
    V = g.uniform(0, 1, (40, 33))
    V[5, :] = 0.0
    V[:, 7] = 0.0
    W0, H0 = g.uniform(0.1, 1, (40, 3)), g.uniform(0.1, 1, (3, 33))
    W0[9, :] = 0.0  # a WH == 0 row: brunet's q = 0 branch
    for algo in ("brunet", "lee", "hals", "dnmf"):
        kw = {"sparsity_h": 0.01, "sparsity_w": 0.01} if algo == "dnmf" else {}
        ref = oracle.run(algo, V, W0, H0, 12, **kw)
        out = gpu_run(algo, V, W0, H0, 12, **kw)
        assert_parity(algo, out, ref, V, W0, H0, 12, "zero rows/cols", **kw)


def test_gpu_deterministic():
    V, W0, H0 = synthetic.lowrank_uniform(300, 200, 4, seed=8)
    a = gpu_run("brunet", V, W0, H0, 15)
    b = gpu_run("brunet", V, W0, H0, 15)
    assert np.array_equal(a["W"], b["W"]) and np.array_equal(a["H"], b["H"])


# --------------------------------------------------------------------------
# Reference-held known answers (executed quickstart notebook) on the GPU
# --------------------------------------------------------------------------
def test_gpu_quickstart_brunet_and_lee():
    V, W0, H0 = quickstart_inputs()
    res = nmf_rs.nmf(V, rank=5, method="brunet", W0=W0, H0=H0, max_iter=200)
    err = np.linalg.norm(V - res.fitted())
    assert abs(err - 9.11162723892749) / 9.11162723892749 < 1e-12  # cell 5
    res = nmf_rs.nmf(V, rank=5, method="lee", W0=W0, H0=H0, max_iter=200)
    assert f"{np.linalg.norm(V - res.fitted()):.4f}" == "8.0928"  # cell 7


def test_gpu_quickstart_stationary():
    V, W0, H0 = quickstart_inputs()
    res = nmf_rs.nmf(V, rank=5, method="brunet", W0=W0, H0=H0, max_iter=2000, stop="stationary",
                     stationary_th=1e-6, check_interval=10, check_niter=10)
    assert res.n_iter == 2000 and len(res.deviances) == 2001  # cell 9
    assert abs(res.deviances[-1] - 0.5964519986834893) / 0.5964519986834893 < 1e-12


def test_gpu_stationary_matches_oracle_trace():
    V, W0, H0 = synthetic.lowrank_uniform(120, 90, 3, seed=12)
    for algo in ("lee", "brunet", "hals", "ehals", "dnmf"):
        kw = dict(stop="stationary", stationary_th=1e-3, check_interval=5, check_niter=3)
        ref = oracle.run(algo, V, W0, H0, 60, **kw)
        out = gpu_run(algo, V, W0, H0, 60, **kw)
        assert out["n_iter"] == ref["n_iter"], algo
        d = np.asarray(out["deviances"])
        assert d.shape == ref["deviances"].shape
        assert np.max(np.abs(d - ref["deviances"]) / np.abs(ref["deviances"])) < 1e-11, algo
        assert_parity(algo, out, ref, V, W0, H0, 60, "stationary", **kw)


# --------------------------------------------------------------------------
# Single-cell-shaped 2000 x 7750, rank 10 (the benchmark workload)
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def sc2():
    V = synthetic.single_cell(2000, 7750, seed=1)
    W0, H0 = oracle.nndsvd_ref(V, 10, fill="mean", seed=0)
    return V, W0, H0


@pytest.fixture(scope="module")
def sc2_sparse():
    """The same generator with 8 % of each library on these genes: 89.6 % zeros."""
    V = synthetic.single_cell(2000, 7750, seed=1, capture=synthetic.SPARSE_CAPTURE)
    assert 0.88 <= float((V == 0).mean()) <= 0.91
    W0, H0 = oracle.nndsvd_ref(V, 10, fill="mean", seed=0)
    return V, W0, H0


@pytest.mark.parametrize("algo,iters", [("lee", 100), ("brunet", 100), ("hals", 25),
                                        ("dnmf", 25), ("ehals", 25)])
def test_gpu_single_cell_2000x7750(sc2, algo, iters):
    _single_cell_case(sc2, algo, iters)


@pytest.mark.parametrize("algo,iters", [("lee", 100), ("brunet", 100), ("hals", 25),
                                        ("dnmf", 25)])
def test_gpu_single_cell_sparse_2000x7750(sc2_sparse, algo, iters):
    _single_cell_case(sc2_sparse, algo, iters)


def _single_cell_case(data, algo, iters):
    V, W0, H0 = data
    kw = {"sparsity_h": 0.01, "sparsity_w": 0.01} if algo == "dnmf" else {}
    print(f"single_cell algo={algo} iters={iters}", flush=True)
    ref = oracle.run(algo, V, W0, H0, iters, num_threads=NT, **kw)
    out = gpu_run(algo, V, W0, H0, iters, **kw)
    rw, rh = rel(out["W"], ref["W"]), rel(out["H"], ref["H"])
    print(f"  rel W {rw:.3e} H {rh:.3e}", flush=True)
    if algo in ("lee", "brunet"):
        assert rw <= TOL and rh <= TOL
    else:
        assert_parity(algo, out, ref, V, W0, H0, iters, "single_cell", num_threads=NT, **kw)
    obj = oracle.kl_divergence if algo == "brunet" else oracle.euclidean_distance
    f_gpu, f_ref = obj(V, out["W"], out["H"]), obj(V, ref["W"], ref["H"])
    assert abs(f_gpu - f_ref) / abs(f_ref) <= TOL


def test_gpu_nndsvd_matches_sklearn_path(sc2):
    V, W0, H0 = sc2
    W, H = nmf_rs.nndsvd_init(V, 10, fill="mean", seed=0)
    # Range-finder products run in another summation order than BLAS; the
    # subspace iteration keeps the difference at roundoff scale.
    print("nndsvd rel W", rel(W, W0), "H", rel(H, H0))
    assert rel(W, W0) <= 1e-11 and rel(H, H0) <= 1e-11


def test_gpu_resident_solve_graph_replay(sc2):
    V, W0, H0 = sc2
    m = nmf_rs._lib.DeviceMatrix(V)
    s = nmf_rs._lib.DeviceSolve(m, 10, nmf_rs._lib.make_params("lee", 20))
    s.set_factors(W0, H0)
    outs = []
    for _ in range(3):
        s.launch()
        outs.append(s.fetch())
    ref = oracle.run("lee", V, W0, H0, 20, num_threads=NT)
    for W, H, it, _ in outs:
        assert it == 20
        assert np.array_equal(W, outs[0][0]) and np.array_equal(H, outs[0][1])
        assert rel(W, ref["W"]) <= TOL and rel(H, ref["H"]) <= TOL
    s.close()
    m.close()


@pytest.mark.parametrize("algo", ["lee", "hals", "ehals", "dnmf"])
def test_gpu_wide_matrix_external_gram(algo):
    """p > 32768: the column pass takes the Gram of W from its own kernel
    (more than 1024 column tiles)."""
    g = np.random.default_rng(21)
    n, p, r = 40, 40000, 5
    V = g.uniform(0.0, 1.0, (n, p))
    W0 = g.uniform(0.1, 1.0, (n, r))
    H0 = g.uniform(0.1, 1.0, (r, p))
    kw = {"sparsity_h": 0.01, "sparsity_w": 0.01} if algo == "dnmf" else {}
    ref = oracle.run(algo, V, W0, H0, 6, num_threads=NT, **kw)
    out = gpu_run(algo, V, W0, H0, 6, **kw)
    print(f"wide {algo}: rel W {rel(out['W'], ref['W']):.3e} H {rel(out['H'], ref['H']):.3e}",
          flush=True)
    assert_parity(algo, out, ref, V, W0, H0, 6, "wide", **kw)


@pytest.mark.parametrize("shape", [(1, 40), (17, 129), (33, 1), (130, 257)])
@pytest.mark.parametrize("rank", [11, 20, 30])
def test_gpu_mfma_row_pass_ragged(shape, rank):
    """Ranks >= 11 take the MFMA row pass (16-row blocks, column splits,
    partial sums): ragged row counts and narrow matrices."""
    n, p = shape
    g = np.random.default_rng(n * 1000 + p + rank)
    V = g.uniform(0.0, 1.0, (n, p))
    W0 = g.uniform(0.1, 1.0, (n, rank))
    H0 = g.uniform(0.1, 1.0, (rank, p))
    for algo in ("lee", "hals"):
        ref = oracle.run(algo, V, W0, H0, 5, num_threads=NT)
        out = gpu_run(algo, V, W0, H0, 5)
        print(f"mfma row {shape} r={rank} {algo}: rel W {rel(out['W'], ref['W']):.3e}", flush=True)
        assert_parity(algo, out, ref, V, W0, H0, 5, f"mfma-row {shape}")


@pytest.mark.parametrize("rank", [12, 20, 32])
def test_gpu_mfma_row_pass_row_tiles_wide(rank):
    """p >= 32768 at rank >= 11: the MFMA row pass blocks 4 row tiles (64
    rows) per work-group; ragged row count (70 = 64 + 6) and the external
    Gram column pass."""
    g = np.random.default_rng(300 + rank)
    n, p = 70, 33000
    V = g.uniform(0.0, 1.0, (n, p))
    W0 = g.uniform(0.1, 1.0, (n, rank))
    H0 = g.uniform(0.1, 1.0, (rank, p))
    for algo in ("lee", "hals"):
        ref = oracle.run(algo, V, W0, H0, 4, num_threads=NT)
        out = gpu_run(algo, V, W0, H0, 4)
        print(f"row tiles r={rank} {algo}: rel W {rel(out['W'], ref['W']):.3e}", flush=True)
        assert_parity(algo, out, ref, V, W0, H0, 4, f"row-tiles r={rank}")
