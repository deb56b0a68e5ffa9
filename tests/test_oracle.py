"""Pins the CPU oracle (oracle/nmf_oracle.c) before it is trusted as the
parity checker:

* bit-for-bit against golden vectors written by the independent pure-Python
  restatement (tests/golden/make_golden.py, oracle/pyref.py);
* against the answers the reference's executed quickstart notebook holds
  (examples/tutorial_quickstart.ipynb cells 5, 7, 9, 13);
* against invariants fixed by the update algebra;
* a planted mistake in the Python restatement must break the bitwise match.
"""
import numpy as np
import pytest

import oracle
import pyref
from _util import golden_cases, quickstart_inputs, unhex


@pytest.fixture(scope="module")
def cases():
    return golden_cases()


def test_oracle_matches_golden_runs_bitwise(cases):
    n_checked = 0
    for c in cases:
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
            for nt in (1, 3):
                out = oracle.run(algo, V, W0, H0, rr["iters"], num_threads=nt, **kw)
                assert np.array_equal(out["W"], unhex(rr["W"])), (c["name"], key, nt)
                assert np.array_equal(out["H"], unhex(rr["H"])), (c["name"], key, nt)
                if "off" in rr:
                    assert np.array_equal(out["offset"], unhex(rr["off"])), (c["name"], key, nt)
                n_checked += 1
    assert n_checked >= 2 * (3 * 12 + 10)


def test_oracle_single_steps_and_objectives_bitwise(cases):
    for c in cases:
        if not c["steps"]:
            continue
        V, W0, H0 = unhex(c["V"]), unhex(c["W0"]), unhex(c["H0"])
        st = c["steps"]
        assert np.array_equal(oracle.update_h_lee(V, W0, H0), unhex(st["lee_h"]))
        assert np.array_equal(oracle.update_w_lee(V, W0, H0), unhex(st["lee_w"]))
        assert np.array_equal(oracle.update_h_brunet(V, W0, H0), unhex(st["brunet_h"]))
        assert np.array_equal(oracle.update_w_brunet(V, W0, H0), unhex(st["brunet_w"]))
        assert oracle.euclidean_distance(V, W0, H0) == float.fromhex(st["euclid"])
        assert oracle.kl_divergence(V, W0, H0) == float.fromhex(st["kl"])


def test_planted_mistake_is_caught(cases, monkeypatch):
    """A forward instead of reverse W^T V sum in lee's H update (a plausible
    slip) must change the result bits."""
    c = cases[0]
    V, W0, H0 = unhex(c["V"]), unhex(c["W0"]), unhex(c["H0"])

    def wrong_lee_h(Vl, W, H, eps):
        n, r, p = len(W), len(W[0]), len(H[0])
        G = [[sum(W[u][i] * W[u][j] for u in reversed(range(n))) for j in range(r)]
             for i in range(r)]
        out = [[0.0] * p for _ in range(r)]
        for j in range(p):
            for i in range(r):
                num = 0.0
                for u in range(n):  # planted: forward order
                    num += W[u][i] * Vl[u][j]
                den = 0.0
                for l in reversed(range(r)):
                    den += G[i][l] * H[l][j]
                t = H[i][j] * num
                out[i][j] = (t if t > eps else eps) / (den + eps)
        return out

    monkeypatch.setattr(pyref, "lee_update_h", wrong_lee_h)
    Wl = [list(map(float, row)) for row in W0]
    Hl = [list(map(float, row)) for row in H0]
    Vl = [list(map(float, row)) for row in V]
    W, H = pyref.run("lee", Vl, Wl, Hl, 12)
    assert not (np.array_equal(np.array(W), unhex(c["runs"]["lee"]["W"]))
                and np.array_equal(np.array(H), unhex(c["runs"]["lee"]["H"])))


# --------------------------------------------------------------------------
# Known answers held by the reference (executed quickstart notebook)
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def qs():
    return quickstart_inputs()


def test_quickstart_inputs_reproduce(qs):
    V, _, _ = qs
    # cell 1 output
    assert V.shape == (200, 60)
    assert V.min() == 0.6152416663367443
    assert V.max() == 8.061756004056358


def test_quickstart_brunet_200(qs):
    V, W0, H0 = qs
    out = oracle.run("brunet", V, W0, H0, 200)
    err = np.linalg.norm(V - out["W"] @ out["H"])
    # cell 5 prints 9.11162723892749 (its W@H and norm use that machine's BLAS)
    assert abs(err - 9.11162723892749) / 9.11162723892749 < 1e-14


def test_quickstart_lee_200(qs):
    V, W0, H0 = qs
    out = oracle.run("lee", V, W0, H0, 200)
    assert f"{np.linalg.norm(V - out['W'] @ out['H']):.4f}" == "8.0928"  # cell 7


def test_quickstart_stationary_brunet(qs):
    V, W0, H0 = qs
    out = oracle.run("brunet", V, W0, H0, 2000, stop="stationary", stationary_th=1e-6,
                     check_interval=10, check_niter=10)
    # cell 9: stopped after 2000 iters; final KL = 0.5964519986834893; 2001 values
    assert out["n_iter"] == 2000
    assert len(out["deviances"]) == 2001
    assert out["deviances"][-1] == 0.5964519986834893


def test_quickstart_one_brunet_step_delta_kl(qs):
    V, W0, H0 = qs
    H1 = oracle.update_h_brunet(V, W0, H0)
    W1 = oracle.update_w_brunet(V, W0, H1)

    def kl(W, H):
        wh = W @ H
        return float(((V * np.log(np.where(V > 0, V / (wh + 1e-30), 1))) - V + wh).sum())

    assert kl(W0, H0) - kl(W1, H1) == 808788.7003988085  # cell 13


# --------------------------------------------------------------------------
# Invariants
# --------------------------------------------------------------------------
def _small(seed=5, n=30, p=20, r=3):
    g = np.random.default_rng(seed)
    V = g.uniform(0.0, 2.0, (n, r)) @ g.uniform(0.0, 2.0, (r, p)) + g.uniform(0, 0.1, (n, p))
    return V, g.uniform(0.1, 1.0, (n, r)), g.uniform(0.1, 1.0, (r, p))


def test_lee_rescale_columns_sum_to_one():
    V, W0, H0 = _small()
    out = oracle.run("lee", V, W0, H0, 7)
    assert np.allclose(out["W"].sum(axis=0), 1.0, rtol=0, atol=1e-14)


def test_brunet_marginal_identities():
    """After the H update colsum(W H') = colsum(V); after the W update
    rowsum(W' H') = rowsum(V) (no WH == 0 entries)."""
    V, W0, H0 = _small(7)
    H1 = oracle.update_h_brunet(V, W0, H0)
    assert np.allclose((W0 @ H1).sum(axis=0), V.sum(axis=0), rtol=1e-13)
    W1 = oracle.update_w_brunet(V, W0, H1)
    assert np.allclose((W1 @ H1).sum(axis=1), V.sum(axis=1), rtol=1e-13)


def test_brunet_reduces_kl_toy():
    """The reference's own Rust test (lib.rs:2208-2227)."""
    V = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]])
    out = oracle.run("brunet", V, np.full((3, 2), 0.5), np.ones((2, 2)), 50, stop="stationary",
                     stationary_th=1e-12, check_interval=50, check_niter=10)
    assert (out["W"] >= 0).all() and (out["H"] >= 0).all()
    assert out["deviances"][-1] <= out["deviances"][0]


def test_kl_trace_non_increasing():
    V, W0, H0 = _small(9)
    out = oracle.run("brunet", V, W0, H0, 60, stop="stationary", check_interval=1000)
    d = out["deviances"]
    assert np.all(np.diff(d) <= 1e-12 * d[0])


def test_exact_factorisation_is_fixed_point():
    g = np.random.default_rng(3)
    Ws, Hs = g.uniform(0.5, 1.5, (12, 2)), g.uniform(0.5, 1.5, (2, 9))
    V = Ws @ Hs
    for algo in ("brunet", "lee"):
        out = oracle.run(algo, V, Ws, Hs, 3, rescale=False)
        assert np.allclose(out["W"] @ out["H"], V, rtol=1e-12)


def test_iterates_respect_floor():
    V, W0, H0 = _small(11)
    out = oracle.run("lee", V, W0, H0, 5, eps=1e-9)
    assert out["H"].min() >= 1e-9 * 0.999
    for algo in ("hals", "ehals", "dnmf"):
        o = oracle.run(algo, V, W0, H0, 5)
        assert o["W"].min() > 0 and o["H"].min() > 0


def test_dnmf_unit_columns():
    V, W0, H0 = _small(13)
    out = oracle.run("dnmf", V, W0, H0, 4, sparsity_h=0.01, sparsity_w=0.01)
    assert np.allclose(np.linalg.norm(out["W"], axis=0), 1.0, atol=1e-14)


def test_hals_close_to_sklearn_cd():
    """Loose library cross-check: sklearn's coordinate-descent NMF (the same
    Gauss-Seidel family, floor 0 instead of eps) reaches a similar loss."""
    from sklearn.decomposition import NMF

    V, W0, H0 = _small(17, 40, 30, 3)
    out = oracle.run("hals", V, W0, H0, 200)
    m = NMF(3, init="custom", solver="cd", max_iter=200, tol=0, alpha_W=0.0)
    Wt = m.fit_transform(V.T, W=H0.T.copy(), H=W0.T.copy())
    ours = np.linalg.norm(V - out["W"] @ out["H"])
    theirs = np.linalg.norm(V.T - Wt @ m.components_)
    assert abs(ours - theirs) / theirs < 1e-3


def test_smoothing_matrix():
    S = oracle.smoothing_matrix(3, 0.5)  # lib.rs:2194-2206
    off = 0.5 / 3.0
    assert np.allclose(S, np.full((3, 3), off) + np.eye(3) * 0.5, atol=1e-15)


def test_nndsvd_ref_closed_form():
    V, _, _ = _small(19, 40, 25, 3)
    from sklearn.utils.extmath import randomized_svd

    U, S, Vt = randomized_svd(V, n_components=3, n_oversamples=10, random_state=0)
    W, H = oracle.nndsvd_ref(V, 3, fill="zero", seed=0)
    assert np.allclose(W[:, 0], np.sqrt(S[0]) * np.abs(U[:, 0]))
    assert np.allclose(H[0], np.sqrt(S[0]) * np.abs(Vt[0]))
    assert (W >= 0).all() and (H >= 0).all()
    Wm, Hm = oracle.nndsvd_ref(V, 3, fill="mean", seed=0)
    assert (Wm > 0).all() and (Hm > 0).all()
    assert np.all(Wm[W > 0] == W[W > 0])


# --------------------------------------------------------------------------
# offset and nsNMF
# --------------------------------------------------------------------------
def _col_rescale_loop(W):
    """col_rescale (lib.rs:1460-1470) with a forward scalar column sum."""
    W = W.copy()
    for k in range(W.shape[1]):
        s = 0.0
        for x in W[:, k]:
            s += x
        if s > 0.0:
            W[:, k] = W[:, k] / s
    return W


def test_offset_with_zero_offset_is_lee_without_rescale():
    """off = 0 stays 0 (off * ... = 0), den_addon = W^T 0 = 0 and
    off[i] * rowsum(H) = 0: offset reduces to lee without the rescale, bit
    for bit."""
    V, W0, H0 = _small(21)
    a = oracle.run("offset", V, W0, H0, 9, offset=np.zeros(V.shape[0]))
    b = oracle.run("lee", V, W0, H0, 9, rescale=False)
    assert np.array_equal(a["W"], b["W"]) and np.array_equal(a["H"], b["H"])
    assert np.all(a["offset"] == 0.0)


def test_offset_exact_model_is_fixed_point():
    """V = W H + off 1^T exactly: W^T V = W^T (W H) + W^T off, V H^T = W H H^T
    + off rowsum(H)^T and rowsum(V) = rowsum(W H + off), so every update
    returns its input up to the eps terms.  A dropped or misplaced offset
    term (den_addon, off[i] * rowsum(H)_k, the offset update) moves it."""
    g = np.random.default_rng(4)
    Ws, Hs = g.uniform(0.5, 1.5, (14, 3)), g.uniform(0.5, 1.5, (3, 11))
    off = g.uniform(0.2, 0.8, 14)
    V = Ws @ Hs + off[:, None]
    out = oracle.run("offset", V, Ws, Hs, 4, offset=off)
    assert np.allclose(out["W"], Ws, rtol=1e-7)
    assert np.allclose(out["H"], Hs, rtol=1e-7)
    assert np.allclose(out["offset"], off, rtol=1e-7)
    # away from the model the iterate moves
    bad = oracle.run("offset", V, Ws, Hs, 1, offset=off * 1.5)
    assert not np.allclose(bad["offset"], off * 1.5, rtol=1e-3)


def test_offset_default_is_row_means_and_objective():
    V, W0, H0 = _small(23)
    a = oracle.run("offset", V, W0, H0, 3)
    b = oracle.run("offset", V, W0, H0, 3, offset=V.mean(axis=1))
    assert np.allclose(a["W"], b["W"], rtol=1e-13) and np.allclose(a["offset"], b["offset"], rtol=1e-13)
    s = oracle.run("offset", V, W0, H0, 3, stop="stationary", check_interval=1000)
    f0 = 0.5 * np.sum((V - W0 @ H0 - V.mean(axis=1)[:, None]) ** 2)
    assert abs(s["deviances"][0] - f0) <= 1e-12 * f0
    f3 = 0.5 * np.sum((V - s["W"] @ s["H"] - s["offset"][:, None]) ** 2)
    assert abs(s["deviances"][-1] - f3) <= 1e-12 * f3


def test_nsnmf_theta0_is_brunet_steps_and_rescale():
    """theta = 0 gives S = I exactly: one iteration is the brunet H and W
    steps followed by col_rescale, bit for bit."""
    V, W0, H0 = _small(25)
    W, H = W0, H0
    for _ in range(3):
        H = oracle.update_h_brunet(V, W, H)
        W = _col_rescale_loop(oracle.update_w_brunet(V, W, H))
    out = oracle.run("nsnmf", V, W0, H0, 3, theta=0.0)
    assert np.array_equal(out["W"], W) and np.array_equal(out["H"], H)


def test_nsnmf_step_uses_ws_for_h_and_sh_for_w():
    """One iteration at theta = 0.4 equals the brunet H step on W S and the W
    step on S H' (products by numpy, so to rounding), then col_rescale; the
    columns of W sum to 1 and the objective is KL(V, W S H)."""
    V, W0, H0 = _small(27, 25, 18, 4)
    S = oracle.smoothing_matrix(4, 0.4)
    H1 = oracle.update_h_brunet(V, W0 @ S, H0)
    W1 = _col_rescale_loop(oracle.update_w_brunet(V, W0, S @ H1))
    out = oracle.run("nsnmf", V, W0, H0, 1, theta=0.4, stop="stationary", check_interval=1000)
    assert np.allclose(out["H"], H1, rtol=1e-13, atol=0)
    assert np.allclose(out["W"], W1, rtol=1e-13, atol=0)
    assert np.allclose(out["W"].sum(axis=0), 1.0, atol=1e-14)
    est = out["W"] @ S @ out["H"]
    kl = np.sum(V * np.log(V / est) - V + est)
    assert abs(out["deviances"][-1] - kl) <= 1e-12 * abs(kl)
    # the swapped placement (S H in the H step) is a different iterate
    Hx = oracle.update_h_brunet(V, W0, S @ H0)
    assert not np.allclose(Hx, H1, rtol=1e-6)


# --------------------------------------------------------------------------
# lsNMF
# --------------------------------------------------------------------------
def test_lsnmf_unit_weight_is_lee_without_rescale():
    """S = 1: W^T (W H) = (W^T W) H and (W H) H^T = W (H H^T), so lsNMF is
    lee without the rescale up to summation order."""
    V, W0, H0 = _small(31)
    a = oracle.run("lsnmf", V, W0, H0, 6, weight=np.ones_like(V))
    b = oracle.run("lee", V, W0, H0, 6, rescale=False)
    assert np.allclose(a["W"], b["W"], rtol=1e-12, atol=0)
    assert np.allclose(a["H"], b["H"], rtol=1e-12, atol=0)


def test_lsnmf_masked_entries_do_not_matter():
    """Entries with weight 0 drop out of both products and of the loss: any
    value there gives the same bits (cv_rank's held-out entries)."""
    V, W0, H0 = _small(33)
    g = np.random.default_rng(34)
    S = (g.random(V.shape) > 0.3).astype(np.float64)
    V2 = np.where(S == 0, g.uniform(0, 50, V.shape), V)
    a = oracle.run("lsnmf", V, W0, H0, 5, weight=S, stop="stationary", check_interval=1000)
    b = oracle.run("lsnmf", V2, W0, H0, 5, weight=S, stop="stationary", check_interval=1000)
    assert np.array_equal(a["W"], b["W"]) and np.array_equal(a["H"], b["H"])
    assert np.array_equal(a["deviances"], b["deviances"])
    f = 0.5 * np.sum(((V - a["W"] @ a["H"]) * S) ** 2)
    assert abs(a["deviances"][-1] - f) <= 1e-12 * f
    # and they do matter without the mask
    c = oracle.run("lsnmf", V2, W0, H0, 5, weight=np.ones_like(V))
    assert not np.allclose(c["W"], a["W"], rtol=1e-6)


# --------------------------------------------------------------------------
# FCNNLS and Kim-Park snmf
# --------------------------------------------------------------------------
def test_fcnnls_reference_known_answer():
    """The reference's own test (fcnnls.rs:519-534): K >= 0 and a residual
    below 1 on its 5 x 3 problem; here also the exact NNLS optimum, which
    for this problem is the unconstrained least-squares solution."""
    X = np.array([[1.0, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1]])
    Y = np.array([[2.0, 1.0], [3.0, 2.0], [4.0, 3.0], [4.5, 2.5], [6.5, 4.5]])
    K, P = oracle.fcnnls(X, Y, 0.0)
    assert (K >= 0).all() and ((X @ K - Y) ** 2).sum() < 1.0
    ls = np.linalg.lstsq(X, Y, rcond=None)[0]
    assert np.allclose(K, ls, rtol=1e-13, atol=1e-13) and P.all()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fcnnls_matches_scipy_nnls(seed):
    """eps = 0: the active-set solution is the NNLS optimum (scipy's
    Lawson-Hanson nnls, column by column)."""
    from scipy.optimize import nnls

    g = np.random.default_rng(seed)
    X = g.uniform(-1, 1, (40, 7))
    Y = g.uniform(-1, 2, (40, 60))
    K, P = oracle.fcnnls(X, Y, 0.0)
    Ks = np.stack([nnls(X, Y[:, j])[0] for j in range(Y.shape[1])], axis=1)
    assert np.allclose(K, Ks, rtol=0, atol=1e-12)
    assert np.array_equal(P, K > 0)


def test_fcnnls_positive_eps_drops_variable_zero_quirk():
    """With eps > 0 every column with an inactive variable sits in the inner
    loop's column set; when some column needs a drop, the others take the
    default argmin (row 0) and lose variable 0, and the shared 3 l round cap
    can stop columns short of feasibility (fcnnls.rs:318-392).  Both
    restatements agree bit for bit, and the result is not the eps = 0 NNLS
    optimum (it even keeps negative entries here)."""
    g = np.random.default_rng(0)
    X = g.uniform(-1, 1, (30, 6))
    Y = g.uniform(-1, 2, (30, 25))
    K9, _ = oracle.fcnnls(X, Y, 1e-9)
    Kp = np.array(pyref.fcnnls(X.tolist(), Y.tolist(), 1e-9))
    assert np.array_equal(K9, Kp)
    K0, _ = oracle.fcnnls(X, Y, 0.0)
    assert np.abs(K9 - K0).max() > 0.1 and K9.min() < 0.0 <= K0.min()


def test_snmf_deviances_and_transpose_symmetry():
    """snmf/l on V equals snmf/r on V^T with the factors swapped, bit for
    bit (lib.rs:1209-1226, 1325-1332); one deviance per iteration equal to
    0.5 ||V - W H||^2."""
    V, W0, H0 = _small(35, 24, 18, 3)
    kw = dict(sparsity_h=0.02, sparsity_w=0.02, smoothness_h=-1.0, smoothness_w=-1.0)
    a = oracle.run("snmf/l", V, W0, H0, 6, **kw)
    b = oracle.run("snmf/r", V.T.copy(), H0.T.copy(), W0.T.copy(), 6, **kw)
    assert np.array_equal(a["W"], b["H"].T) and np.array_equal(a["H"], b["W"].T)
    assert a["n_iter"] == 6 and len(a["deviances"]) == 6
    f = 0.5 * np.sum((V - a["W"] @ a["H"]) ** 2)
    assert abs(a["deviances"][-1] - f) <= 1e-12 * f
    assert (a["W"] >= 0).all() and (a["H"] >= 0).all()
