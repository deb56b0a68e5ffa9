"""Pins of the oracle's HALS family and NNDSVD that do not reuse its own
arithmetic: what the mathematics fixes (each Gauss-Seidel coordinate update
is the exact minimiser of the objective in that coordinate, clamped at eps),
compositions the reference's drivers define (ehals, dnmf), a library that
carries the same closed form (scikit-learn's NNDSVD), and the reference
notebook's published results (examples/benchmark_scaling.ipynb, through the
fixture tests/golden/scaling_medium.json that make_scaling_pins.py writes
after asserting them)."""
import json
from pathlib import Path

import numpy as np
import pytest
from scipy.optimize import minimize_scalar

import _scaling as S
import oracle

EPS = 1e-9
GOLD = Path(__file__).resolve().parent / "golden"


def _problem(seed, n=9, p=7, r=3, low_cols=()):
    """Positive low-rank + noise data; columns in low_cols are scaled down
    so that some coordinates of H hit the eps floor."""
    g = np.random.default_rng(seed)
    V = g.uniform(0.2, 1.5, (n, r)) @ g.uniform(0.2, 1.5, (r, p)) + g.uniform(0, 0.3, (n, p))
    for j in low_cols:
        V[:, j] *= 1e-3
    W0 = g.uniform(0.1, 1.0, (n, r))
    H0 = g.uniform(0.1, 1.0, (r, p))
    return V, W0, H0


def _coord_lstsq(resid, w):
    """argmin_t 0.5 ||resid - w t||^2 over t >= eps, by a least-squares
    solve of the one-column system (not the Gram-based closed form)."""
    t = np.linalg.lstsq(w[:, None], resid, rcond=None)[0][0] if np.any(w) else 0.0
    return max(t, EPS)


def _coord_scalar(resid, w, lam, gam):
    """argmin_t 0.5 ||resid - w t||^2 + lam t + 0.5 gam t^2 over t >= eps,
    by bounded scalar minimisation of the objective itself."""
    f = lambda t: 0.5 * float(np.sum((resid - w * t) ** 2)) + lam * t + 0.5 * gam * t * t
    hi = 10.0 * (abs(float(w @ resid)) / max(float(w @ w), 1e-300) + 1.0)
    res = minimize_scalar(f, bounds=(EPS, hi), method="bounded",
                          options={"xatol": 1e-14, "maxiter": 2000})
    return max(res.x, EPS)


def _gs_sweep_h(V, W, H, coord):
    """H step as coordinate descent: columns of H independently, k = 0..r-1
    in order, each coordinate minimised with the current others (already
    updated ones included)."""
    H = H.copy()
    for j in range(V.shape[1]):
        for k in range(H.shape[0]):
            others = W @ H[:, j] - W[:, k] * H[k, j]
            H[k, j] = coord(V[:, j] - others, W[:, k])
    return H


def _gs_sweep_w(V, W, H, coord):
    """W step: the H step on the transposed problem V^T = H^T W^T."""
    return _gs_sweep_h(V.T, H.T, W.T, coord).T


def test_hals_iteration_is_exact_coordinate_descent():
    """One hals iteration (lib.rs:664-677) equals Gauss-Seidel coordinate
    minimisation of 0.5 ||V - WH||^2 clamped at eps: the H sweep with W0,
    then the W sweep with the new H.  Columns 2 and 5 are scaled down so
    that clamps are active."""
    V, W0, H0 = _problem(3, low_cols=(2, 5))
    out = oracle.run("hals", V, W0, H0, 1)
    H1 = _gs_sweep_h(V, W0, H0, _coord_lstsq)
    assert (H1 == EPS).any(), "no clamp exercised"
    W1 = _gs_sweep_w(V, W0, H1, _coord_lstsq)
    assert np.max(np.abs(out["H"] - H1)) <= 1e-12 * np.max(np.abs(H1))
    assert np.max(np.abs(out["W"] - W1)) <= 1e-12 * np.max(np.abs(W1))
    # the same check separates Gauss-Seidel from a Jacobi sweep (every
    # coordinate from the old column), the likeliest restatement slip
    jac = H0.copy()
    for j in range(V.shape[1]):
        for k in range(H0.shape[0]):
            others = W0 @ H0[:, j] - W0[:, k] * H0[k, j]
            jac[k, j] = _coord_lstsq(V[:, j] - others, W0[:, k])
    assert np.max(np.abs(jac - H1)) > 1e-3 * np.max(np.abs(H1))


def test_hals_zero_column_of_w():
    """A zero column of W (Gram diagonal 0, the max(G_kk, eps) guard): the
    coordinate's minimiser is any t; the sweep leaves eps, and the other
    coordinates are still exact minimisers."""
    V, W0, H0 = _problem(5)
    W0[:, 1] = 0.0
    out = oracle.run("hals", V, W0, H0, 1)
    H1 = _gs_sweep_h(V, W0, H0, _coord_lstsq)
    assert np.all(H1[1] == EPS)
    assert np.max(np.abs(out["H"] - H1)) <= 1e-12 * np.max(np.abs(H1))


def test_dnmf_iteration_is_regularised_coordinate_descent_then_normalised():
    """One dnmf iteration (lib.rs:862-884): the H and W sweeps minimise
    0.5 ||V - WH||^2 + lam * sum(x) + 0.5 * gam * ||x||^2 coordinate-wise
    (clamped at eps), then W's columns are scaled to unit norm and H's rows
    carry the scales, W' H' = W H.  lam and gam differ between H and W so a
    swapped or sign-flipped term fails."""
    V, W0, H0 = _problem(7, low_cols=(4,))
    lh, lw, gh, gw = 0.3, 0.05, 0.2, 0.7
    out = oracle.run("dnmf", V, W0, H0, 1, sparsity_h=lh, sparsity_w=lw, smoothness_h=gh,
                     smoothness_w=gw)
    H1 = _gs_sweep_h(V, W0, H0, lambda r, w: _coord_scalar(r, w, lh, gh))
    W1 = _gs_sweep_w(V, W0, H1, lambda r, w: _coord_scalar(r, w, lw, gw))
    nw = np.linalg.norm(W1, axis=0)
    assert np.allclose(np.linalg.norm(out["W"], axis=0), 1.0, atol=1e-14)
    assert np.max(np.abs(out["W"] - W1 / nw)) <= 1e-6 * np.max(np.abs(W1 / nw))
    assert np.max(np.abs(out["H"] - H1 * nw[:, None])) <= 1e-6 * np.max(np.abs(H1 * nw[:, None]))
    # the regularisation moves the result (so the check above sees lam, gam)
    plain = oracle.run("dnmf", V, W0, H0, 1)
    assert np.max(np.abs(plain["H"] - out["H"])) > 1e-3 * np.max(np.abs(out["H"]))


def test_dnmf_without_regularisation_is_hals_up_to_scaling():
    """With lam = gam = 0 the dnmf sweeps are the HALS sweeps, and HALS is
    equivariant under W -> W D^-1, H -> D H up to the eps clamps; so on data
    whose iterates stay above eps, dnmf's W H after several iterations equals
    hals's."""
    g = np.random.default_rng(11)
    Ws, Hs = g.uniform(0.5, 1.5, (12, 3)), g.uniform(0.5, 1.5, (3, 10))
    V = Ws @ Hs + g.uniform(0, 0.05, (12, 10))
    W0 = Ws * g.uniform(0.8, 1.2, Ws.shape)  # a start near the generating factors
    H0 = Hs * g.uniform(0.8, 1.2, Hs.shape)
    for it in (1, 4):
        d = oracle.run("dnmf", V, W0, H0, it)
        h = oracle.run("hals", V, W0, H0, it)
        assert min(h["W"].min(), h["H"].min()) > 1e3 * EPS
        assert np.allclose(d["W"] @ d["H"], h["W"] @ h["H"], rtol=1e-11, atol=0)


def test_ehals_is_hals_from_the_extrapolated_point():
    """ehals (lib.rs:737-778): iteration t runs one HALS iteration from
    max(eps, X_t + 0.5 (X_t - X_{t-1})) for X = W and H, X_0 = X_{-1}."""
    V, W0, H0 = _problem(13, low_cols=(1,))
    e1 = oracle.run("ehals", V, W0, H0, 1)
    h1 = oracle.run("hals", V, np.maximum(W0, EPS), np.maximum(H0, EPS), 1)
    assert np.array_equal(e1["W"], h1["W"]) and np.array_equal(e1["H"], h1["H"])
    e2 = oracle.run("ehals", V, W0, H0, 2)
    ext = lambda x, prev: np.where(x + 0.5 * (x - prev) > EPS, x + 0.5 * (x - prev), EPS)
    h2 = oracle.run("hals", V, ext(h1["W"], W0), ext(h1["H"], H0), 1)
    assert np.array_equal(e2["W"], h2["W"]) and np.array_equal(e2["H"], h2["H"])


def test_nndsvd_matches_scikit_learn_initialisation():
    """nndsvd_init (python/nmf_rs/__init__.py:168-217) with fill='zero' is
    scikit-learn's NNDSVD (_initialize_nmf, init='nndsvd': same randomized
    SVD arguments and closed form, entries below 1e-6 zeroed, ties of the
    positive / negative mass to the negative part), wherever sklearn's entry
    is >= 1e-6 and no component's masses tie; with fill='mean' and sklearn's
    'nndsvda' (zeros -> mean(V)) likewise."""
    from sklearn.decomposition._nmf import _initialize_nmf

    for seed, (n, p, r) in enumerate([(60, 45, 5), (40, 90, 7), (120, 30, 4)]):
        g = np.random.default_rng(50 + seed)
        V = g.uniform(0, 1, (n, r)) @ g.uniform(0, 1, (r, p)) + g.uniform(0, 0.2, (n, p))
        W, H = oracle.nndsvd_ref(V, r, fill="zero", seed=0)
        Ws, Hs = _initialize_nmf(V, r, init="nndsvd", random_state=0)
        for ours, theirs in ((W, Ws), (H, Hs)):
            keep = theirs >= 1e-6
            assert np.allclose(ours[keep], theirs[keep], rtol=1e-12, atol=0)
            assert np.all(ours[~keep] < 1e-6)
        Wm, Hm = oracle.nndsvd_ref(V, r, fill="mean", seed=0)
        Wa, Ha = _initialize_nmf(V, r, init="nndsvda", random_state=0)
        for ours, theirs, z in ((Wm, Wa, W), (Hm, Ha, H)):
            keep = z >= 1e-6
            assert np.allclose(ours[keep], theirs[keep], rtol=1e-12, atol=0)
            assert np.all(ours[z == 0] == V.mean())


def test_scaling_notebook_medium_tier_fixture():
    """The oracle on the reference notebook's medium tier reproduced its
    printed results (asserted by make_scaling_pins.py when it wrote the
    fixture; re-checked here on the stored figures)."""
    f = GOLD / "scaling_medium.json"
    if not f.exists():
        pytest.skip("fixture not generated")
    d = json.loads(f.read_text())
    M = S.MEDIUM
    losses = [float.fromhex(r["loss"]) for r in d["ehals_runs"]]
    assert len(losses) == M["runs"]
    assert float(f"{np.mean(losses):.6e}") == M["ehals_mean_loss"]
    assert round(d["ehals_dispersion"], 3) == M["dispersion_3dp"]["ehals"]
    assert round(d["ehals_sparsity_w"], 3) == M["sparsity_w_3dp"]["ehals"]
    assert float(f"{d['hals_nndsvd_amari']:.6e}") == M["hals_nndsvd_amari"]
    assert float(f"{d['hals_nndsvd_sparsity_w']:.6e}") == M["hals_nndsvd_sparsity_w"]


def test_scaling_metrics_known_values():
    """The notebook metrics as restated: Amari error 0 for a scaled
    permutation, Hoyer sparsity 1 for a one-hot and 0 for a constant vector,
    dispersion 1 for a 0/1 consensus."""
    W = np.random.default_rng(0).uniform(0.1, 1, (20, 4))
    P = np.eye(4)[[2, 0, 3, 1]] * np.array([2.0, 0.5, 3.0, 1.5])
    assert S.amari_error(W @ P, W) < 1e-12
    assert abs(S.hoyer(np.eye(9)[3]) - 1.0) < 1e-15 and S.hoyer(np.ones(9)) == 0.0
    C = S.connectivity(np.array([[1.0, 0, 1, 0], [0, 1, 0, 1]]))
    assert S.dispersion(C) == 1.0
