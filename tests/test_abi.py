"""CPU-side checks of the C ABI library and the Python mirror of the
reference API: the library loads, exports every symbol include/nmfgpu.h
declares, host-only entry points behave as the reference's, and argument
validation matches python/nmf_rs/__init__.py.  No GPU compute is called."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import nmf_rs
from nmf_rs import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "nmfgpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nmfg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == names


def test_params_defaults_match_reference():
    """py_nmf_run defaults (rust/src/lib.rs:1912-1933)."""
    p = _lib.Params()
    _lib.lib().nmfg_params_init(ctypes.byref(p))
    assert p.max_iter == 2000 and p.eps == 1e-9 and p.rescale == 1 and p.theta == 0.5
    assert p.stationary_th == 2.220446049250313e-16
    assert p.check_interval == 50 and p.check_niter == 10 and p.stop == 0
    assert p.precision == 0


def test_smoothing_matrix_host():
    S = nmf_rs.smoothing_matrix(4, 0.3)
    off = 0.3 / 4
    assert np.array_equal(S, np.where(np.eye(4) > 0, (1 - 0.3) + off, off))


def test_max_rank():
    assert _lib.lib().nmfg_max_rank() >= 30


def test_no_cpu_fallback_without_device(monkeypatch):
    """Without a visible gfx950 GPU the context cannot be created and the
    compute entry points fail loudly."""
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, %r); import numpy as np, nmf_rs\n"
            "try:\n"
            "    nmf_rs.nmf(np.ones((4, 3)), 2, 'lee', max_iter=2)\n"
            "except RuntimeError as e:\n"
            "    print('RuntimeError', e)\n" % str(ROOT / "rust-nmf_amd"))
    env = dict(__import__("os").environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                         timeout=120)
    assert "RuntimeError" in out.stdout, out.stdout + out.stderr


def test_method_aliases_match_reference():
    # python/nmf_rs/__init__.py:362-387
    ref = {"brunet": "brunet", "kl": "brunet", "lee": "lee", "frobenius": "lee",
           "euclidean": "lee", "offset": "offset", "nsnmf": "nsnmf", "ns": "nsnmf",
           "ns_nmf": "nsnmf", "hals": "hals", "ehals": "ehals", "e-hals": "ehals",
           "extrap_hals": "ehals", "dnmf": "dnmf", "diag_nmf": "dnmf", "rcppml": "dnmf",
           "ls-nmf": "lsnmf", "lsnmf": "lsnmf", "snmf/r": "snmf/r", "snmf_r": "snmf/r",
           "snmfr": "snmf/r", "snmf/l": "snmf/l", "snmf_l": "snmf/l", "snmfl": "snmf/l"}
    assert nmf_rs._METHOD_ALIASES == ref


@pytest.mark.parametrize("kw,err", [
    (dict(method="nope"), "unknown method"),
    (dict(V=-np.ones((3, 2))), "non-negative"),
    (dict(W0=np.ones((4, 2))), "W0.shape"),
    (dict(H0=np.ones((2, 5))), "H0.shape"),
    (dict(offset=np.ones(2)), "offset.shape"),
    (dict(weight=np.ones((2, 2))), "weight.shape"),
])
def test_nmf_validation_errors(kw, err):
    args = dict(V=np.ones((3, 2)), rank=2, method="lee", W0=np.ones((3, 2)), H0=np.ones((2, 2)))
    args.update(kw)
    V = args.pop("V")
    with pytest.raises(ValueError, match=re.escape(err)):
        nmf_rs.nmf(V, **args)


def test_one_dimensional_input_rejected():
    with pytest.raises(ValueError, match="must be 2-D"):
        nmf_rs.update_h_lee(np.ones(3), np.ones((3, 1)), np.ones((1, 1)))


def test_random_init_matches_reference_restatement():
    import oracle

    V = np.abs(np.random.default_rng(0).normal(size=(7, 5)))
    a = nmf_rs.random_init(V, 3, seed=11)
    b = oracle.random_init_ref(V, 3, seed=11)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_nmfresult_fitted():
    r = nmf_rs.NMFResult(W=np.ones((2, 1)), H=np.full((1, 3), 2.0), n_iter=1, method="lee")
    assert np.array_equal(r.fitted(), np.full((2, 3), 2.0))
    assert repr(r) == "NMFResult(method='lee', n=2, p=3, rank=1, n_iter=1)"
    r.offset = np.array([1.0, 2.0])
    assert np.array_equal(r.fitted()[:, 0], [3.0, 4.0])


def test_consensus_statistics_known_answers():
    """Block consensus (two perfectly stable clusters) has cophenetic and
    dispersion 1; a constant 1/2 consensus has dispersion 0; the seed rule of
    the restarts is 1000 * rank + restart."""
    from nmf_rs import consensus

    C = np.kron(np.eye(2), np.ones((3, 3)))
    assert abs(consensus.cophenetic_correlation(C) - 1.0) <= 1e-12
    assert consensus.dispersion_coefficient(C) == 1.0
    assert consensus.dispersion_coefficient(np.full((4, 4), 0.5)) == 0.0
    W0, H0 = consensus.uniform_restarts(5, 7, 3, 2)
    g = np.random.default_rng(3001)
    np.testing.assert_array_equal(W0[1], g.uniform(0.0, 1.0, (5, 3)))
    np.testing.assert_array_equal(H0[1], g.uniform(0.0, 1.0, (3, 7)))
    np.testing.assert_array_equal(consensus.cluster_labels(np.array([[1., 0., 2.], [0., 3., 2.]])),
                                  [0, 1, 0])


def test_cophenetic_is_the_correlation():
    """On unstructured labels the statistic is below 1 and equals scipy's own
    cophenetic correlation coefficient c of the average-linkage tree."""
    from scipy.cluster.hierarchy import cophenet, linkage
    from scipy.spatial.distance import squareform

    from nmf_rs import consensus

    g = np.random.default_rng(0)
    L = g.integers(0, 3, (5, 60))
    C = np.zeros((60, 60))
    for row in L:
        C += row[:, None] == row[None, :]
    C /= 5
    D = 1.0 - C
    np.fill_diagonal(D, 0.0)
    cond = squareform(D, checks=False)
    c_scipy, _ = cophenet(linkage(cond, method="average"), cond)
    got = consensus.cophenetic_correlation(C)
    assert got < 0.99
    assert abs(got - c_scipy) <= 1e-12


def _fwd_normal_equations(X, Y):
    """X^T X and (X^T Y)^T with forward sums over the observations (the
    oracle's order), as Python floats."""
    nobs, l = X.shape
    ctc = np.zeros((l, l))
    for a in range(l):
        for b in range(l):
            s = 0.0
            for o in range(nobs):
                s += float(X[o, a]) * float(X[o, b])
            ctc[a, b] = s
    cta = np.zeros((Y.shape[1], l))
    for j in range(Y.shape[1]):
        for a in range(l):
            s = 0.0
            for o in range(nobs):
                s += float(X[o, a]) * float(Y[o, j])
            cta[j, a] = s
    return ctc, cta


@pytest.mark.parametrize("seed,eps,p", [(0, 0.0, 30), (0, 1e-9, 30), (1, 1e-9, 30), (2, 0.0, 30),
                                        (3, 1e-9, 3000), (4, 0.0, 3000)])
def test_host_fcnnls_matches_oracle_bitwise(seed, eps, p):
    """The library's host FCNNLS (used by snmf; its rounds spread over host
    threads at 3000 columns) against the oracle's on the same normal
    equations: same bits, including the eps > 0 behaviour."""
    import oracle
    from nmf_rs import _lib

    g = np.random.default_rng(seed)
    X = g.uniform(-1, 1, (25, 6)) if seed != 2 else g.uniform(0, 1, (25, 6))
    Y = g.uniform(-1, 2, (25, p))
    ctc, cta = _fwd_normal_equations(X, Y)
    K = np.empty((p, 6))
    rc = _lib.lib().nmfg_fcnnls(_lib.ptr(np.ascontiguousarray(ctc)), _lib.ptr(np.ascontiguousarray(cta)),
                                6, p, float(eps), _lib.ptr(K))
    assert rc == 0
    Ko, _ = oracle.fcnnls(X, Y, eps)
    assert np.array_equal(K.T, Ko)


def test_nndsvd_closed_form_matches_oracle():
    """The package's NNDSVD closed form (u parts of all components at once,
    v parts row by row) against the oracle's per-component restatement: random SVD factors, a component with
    equal positive and negative mass (the positive side wins), components of
    zero mass on both sides, every fill mode.  The norms are summed in
    another order, hence 1e-15 rather than bit-exact."""
    import nmf_rs
    import oracle

    g = np.random.default_rng(11)
    n, p, rank = 120, 170, 9
    V = g.uniform(0.0, 1.0, (n, p))
    U = g.normal(size=(n, rank))
    Vt = g.normal(size=(rank, p))
    S = np.sort(g.uniform(0.5, 4.0, rank))[::-1]
    U[:, 2] = 0.0                      # zero mass on both sides
    U[:, 4] = np.abs(U[:, 4])          # u >= 0, v <= 0: both products 0
    Vt[4] = -np.abs(Vt[4])
    U[:, 6] = np.r_[np.ones(60), -np.ones(60)]   # equal masses: positive side
    Vt[6] = np.r_[np.ones(85), -np.ones(85)]
    for fill in ("zero", "eps", "mean"):
        W, H = nmf_rs._nndsvd_from_svd(U, S, Vt, rank, fill, float(V.mean()))
        W0, H0 = oracle.nndsvd_from_svd(V, U, S, Vt, rank, fill)
        assert np.max(np.abs(W - W0)) <= 1e-15 * np.max(np.abs(W0)), fill
        assert np.max(np.abs(H - H0)) <= 1e-15 * np.max(np.abs(H0)), fill
        assert np.array_equal(W <= 0.0, W0 <= 0.0) and np.array_equal(H <= 0.0, H0 <= 0.0)


@pytest.mark.parametrize("m,w,cond", [(5000, 34, 1e2), (3000, 24, 1e5), (40, 40, 1e3), (2000, 34, 1e9),
                                      (20, 30, 1e1)])
def test_orthonormal_basis_is_householder_up_to_signs(m, w, cond):
    """The randomized SVD's QR step (CholeskyQR2, Householder when Y is ill
    conditioned or wide): orthonormal columns spanning Y, each column equal
    to LAPACK's Householder Q's up to its sign, to within eps cond(Y)."""
    from scipy import linalg

    g = np.random.default_rng(w)
    U, _ = np.linalg.qr(g.normal(size=(m, min(m, w))))
    V, _ = np.linalg.qr(g.normal(size=(w, w)))
    s = np.logspace(0, -np.log10(cond), min(m, w))
    Y = (U * s) @ V[:, : min(m, w)].T * 1e3
    Q = nmf_rs._orthonormal_basis(Y)
    Qh, _ = linalg.qr(Y, mode="economic", check_finite=False)
    assert Q.shape == Qh.shape
    assert np.max(np.abs(Q.T @ Q - np.eye(Q.shape[1]))) <= 1e-14
    sg = np.sign(np.sum(Q * Qh, axis=0))
    assert np.all(sg != 0)
    assert np.max(np.abs(Q * sg - Qh)) <= 1e-15 * max(cond, 10.0) * 100
