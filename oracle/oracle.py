"""CPU oracle for the dense NMF update path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this module, and only as the checker (or the timed CPU baseline);
the product library never calls it.

* ``run`` / ``update_*`` / ``euclidean_distance`` / ``kl_divergence`` wrap the
  C restatement in ``nmf_oracle.c`` (rust/src/lib.rs arithmetic, bit-exact for
  lee and brunet).
* ``nndsvd_ref`` restates ``nndsvd_init`` (python/nmf_rs/__init__.py:129-219)
  on top of scikit-learn's ``randomized_svd`` -- the same library routine the
  reference calls, with the same arguments.
* ``random_init_ref`` restates ``random_init`` (__init__.py:102-126).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libnmf_oracle.so"

ALGO = {"brunet": 0, "lee": 1, "offset": 2, "nsnmf": 3, "hals": 4, "ehals": 5, "dnmf": 6,
        "lsnmf": 7, "snmf/r": 8, "snmf/l": 9}
STOP = {"max_iter": 0, "stationary": 1}


class _Params(ctypes.Structure):
    _fields_ = [
        ("algo", ctypes.c_int32),
        ("max_iter", ctypes.c_int64),
        ("eps", ctypes.c_double),
        ("rescale", ctypes.c_int32),
        ("sparsity_h", ctypes.c_double),
        ("sparsity_w", ctypes.c_double),
        ("smoothness_h", ctypes.c_double),
        ("smoothness_w", ctypes.c_double),
        ("stop", ctypes.c_int32),
        ("stationary_th", ctypes.c_double),
        ("check_interval", ctypes.c_int64),
        ("check_niter", ctypes.c_int64),
        ("num_threads", ctypes.c_int32),
        ("theta", ctypes.c_double),
    ]


_lib = None


def build() -> Path:
    """Compile the C restatement with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        dp = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.c_int64
        i32 = ctypes.c_int32
        L.orc_run.argtypes = [ctypes.POINTER(_Params), dp, i64, i64, dp, dp, i32,
                              ctypes.POINTER(i64), dp, i64, ctypes.POINTER(i64)]
        L.orc_run.restype = ctypes.c_int
        L.orc_run_offset.argtypes = [ctypes.POINTER(_Params), dp, i64, i64, dp, dp, i32, dp, i32,
                                     ctypes.POINTER(i64), dp, i64, ctypes.POINTER(i64)]
        L.orc_run_offset.restype = ctypes.c_int
        L.orc_run_lsnmf.argtypes = [ctypes.POINTER(_Params), dp, dp, i64, i64, dp, dp, i32,
                                    ctypes.POINTER(i64), dp, i64, ctypes.POINTER(i64)]
        L.orc_run_lsnmf.restype = ctypes.c_int
        for nm in ("orc_euclidean_update_h", "orc_euclidean_update_w"):
            f = getattr(L, nm)
            f.argtypes = [dp, i64, i64, dp, dp, i32, ctypes.c_double, dp, i32]
            f.restype = None
        for nm in ("orc_divergence_update_h", "orc_divergence_update_w"):
            f = getattr(L, nm)
            f.argtypes = [dp, i64, i64, dp, dp, i32, dp, i32]
            f.restype = None
        for nm in ("orc_euclidean_distance", "orc_kl_divergence"):
            f = getattr(L, nm)
            f.argtypes = [dp, i64, i64, dp, dp, i32]
            f.restype = ctypes.c_double
        L.orc_fcnnls.argtypes = [dp, dp, i64, i32, i64, ctypes.c_double, dp,
                                 ctypes.POINTER(ctypes.c_uint8)]
        L.orc_fcnnls.restype = ctypes.c_int
        L.orc_smoothing_matrix.argtypes = [i32, ctypes.c_double, dp]
        L.orc_smoothing_matrix.restype = None
        _lib = L
    return _lib


def _c(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def run(algo, V, W0, H0, max_iter, *, eps=1e-9, rescale=True, theta=0.5, offset=None,
        weight=None, sparsity_h=0.0, sparsity_w=0.0, smoothness_h=0.0, smoothness_w=0.0,
        stop="max_iter", stationary_th=2.220446049250313e-16, check_interval=50,
        check_niter=10, num_threads=1):
    """Full driver; same keyword meaning as ``py_nmf_run`` (lib.rs:1912-1933).
    For "offset" the result also holds "offset" (initial: ``offset`` or
    rowMeans(V))."""
    V = _c(V)
    W = _c(W0).copy()
    H = _c(H0).copy()
    n, p = V.shape
    r = W.shape[1]
    prm = _Params(ALGO[algo], int(max_iter), float(eps), int(bool(rescale)),
                  float(sparsity_h), float(sparsity_w), float(smoothness_h),
                  float(smoothness_w), STOP[stop], float(stationary_th),
                  int(check_interval), int(check_niter), int(num_threads or 1), float(theta))
    cap = int(max_iter) + 1 if stop == "stationary" else 0
    if algo in ("snmf/r", "snmf/l"):
        cap = int(max_iter)
    dev = np.zeros(max(cap, 1))
    n_iter = ctypes.c_int64(0)
    n_dev = ctypes.c_int64(0)
    if algo == "lsnmf":
        S = _c(weight)
        rc = lib().orc_run_lsnmf(ctypes.byref(prm), _p(V), _p(S), n, p, _p(W), _p(H), r,
                                 ctypes.byref(n_iter), _p(dev), cap, ctypes.byref(n_dev))
        if rc != 0:
            raise ValueError("oracle: bad lsnmf arguments")
        return {"W": W, "H": H, "n_iter": int(n_iter.value),
                "deviances": dev[: n_dev.value].copy()}
    off = _c(offset).copy() if offset is not None else np.zeros(n)
    rc = lib().orc_run_offset(ctypes.byref(prm), _p(V), n, p, _p(W), _p(H), r, _p(off),
                              int(offset is not None), ctypes.byref(n_iter), _p(dev), cap,
                              ctypes.byref(n_dev))
    if rc != 0:
        raise ValueError(f"oracle: unsupported algo {algo!r}")
    out = {"W": W, "H": H, "n_iter": int(n_iter.value), "deviances": dev[: n_dev.value].copy()}
    if algo == "offset":
        out["offset"] = off
    return out


def update_h_lee(V, W, H, eps=1e-9, num_threads=1):
    V, W, H = _c(V), _c(W), _c(H)
    out = np.empty_like(H)
    lib().orc_euclidean_update_h(_p(V), V.shape[0], V.shape[1], _p(W), _p(H), W.shape[1],
                                 float(eps), _p(out), int(num_threads))
    return out


def update_w_lee(V, W, H, eps=1e-9, num_threads=1):
    V, W, H = _c(V), _c(W), _c(H)
    out = np.empty_like(W)
    lib().orc_euclidean_update_w(_p(V), V.shape[0], V.shape[1], _p(W), _p(H), W.shape[1],
                                 float(eps), _p(out), int(num_threads))
    return out


def update_h_brunet(V, W, H, num_threads=1):
    V, W, H = _c(V), _c(W), _c(H)
    out = np.empty_like(H)
    lib().orc_divergence_update_h(_p(V), V.shape[0], V.shape[1], _p(W), _p(H), W.shape[1],
                                  _p(out), int(num_threads))
    return out


def update_w_brunet(V, W, H, num_threads=1):
    V, W, H = _c(V), _c(W), _c(H)
    out = np.empty_like(W)
    lib().orc_divergence_update_w(_p(V), V.shape[0], V.shape[1], _p(W), _p(H), W.shape[1],
                                  _p(out), int(num_threads))
    return out


def euclidean_distance(V, W, H):
    V, W, H = _c(V), _c(W), _c(H)
    return lib().orc_euclidean_distance(_p(V), V.shape[0], V.shape[1], _p(W), _p(H), W.shape[1])


def kl_divergence(V, W, H):
    V, W, H = _c(V), _c(W), _c(H)
    return lib().orc_kl_divergence(_p(V), V.shape[0], V.shape[1], _p(W), _p(H), W.shape[1])


def fcnnls(X, Y, eps=0.0):
    """FCNNLS (rust/src/fcnnls.rs:270-501): K >= 0 minimising ||Y - X K||;
    returns (K, Pset)."""
    X, Y = _c(X), _c(Y)
    nobs, l = X.shape
    p = Y.shape[1]
    K = np.empty((l, p))
    P = np.empty((l, p), dtype=np.uint8)
    rc = lib().orc_fcnnls(_p(X), _p(Y), nobs, l, p, float(eps), _p(K),
                          P.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    if rc != 0:
        raise ValueError("oracle: fcnnls supports at most 64 variables")
    return K, P.astype(bool)


def smoothing_matrix(r, theta):
    out = np.empty((r, r))
    lib().orc_smoothing_matrix(int(r), float(theta), _p(out))
    return out


# --------------------------------------------------------------------------
# Python-side reference restatements
# --------------------------------------------------------------------------
def random_init_ref(V, rank, seed=None, *, max_value=None):
    """random_init (python/nmf_rs/__init__.py:102-126): default_rng(seed),
    W0 then H0 ~ U(0, max V)."""
    V = np.asarray(V)
    n, p = V.shape
    rng = np.random.default_rng(seed)
    hi = float(max_value if max_value is not None else (np.max(V) if V.size else 1.0))
    return rng.uniform(0.0, hi, size=(n, rank)), rng.uniform(0.0, hi, size=(rank, p))


def nndsvd_from_svd(V, U, S, Vt, rank, fill="mean"):
    """NNDSVD closed form + fill (python/nmf_rs/__init__.py:176-217)."""
    n, p = V.shape
    W = np.zeros((n, rank))
    H = np.zeros((rank, p))
    r0 = np.sqrt(S[0])
    W[:, 0] = r0 * np.abs(U[:, 0])
    H[0, :] = r0 * np.abs(Vt[0, :])
    for k in range(1, rank):
        u, v = U[:, k], Vt[k, :]
        up, un = np.maximum(u, 0.0), np.maximum(-u, 0.0)
        vp, vn = np.maximum(v, 0.0), np.maximum(-v, 0.0)
        nup, nun = np.linalg.norm(up), np.linalg.norm(un)
        nvp, nvn = np.linalg.norm(vp), np.linalg.norm(vn)
        mp, mn = nup * nvp, nun * nvn
        if mp >= mn:
            sig = np.sqrt(S[k] * mp)
            if mp > 0:
                W[:, k] = sig * up / nup
                H[k, :] = sig * vp / nvp
        else:
            sig = np.sqrt(S[k] * mn)
            if mn > 0:
                W[:, k] = sig * un / nun
                H[k, :] = sig * vn / nvn
    if fill == "zero":
        pass
    elif fill == "eps":
        W[W < 1e-9] = 1e-9
        H[H < 1e-9] = 1e-9
    elif fill == "mean":
        m = float(V.mean()) if V.size else 0.0
        W[W <= 0.0] = m
        H[H <= 0.0] = m
    else:
        raise ValueError(f"unknown fill mode '{fill}': supported zero / mean / eps")
    return W, H


def nndsvd_ref(V, rank, *, fill="mean", seed=0, n_oversamples=10):
    """nndsvd_init (python/nmf_rs/__init__.py:129-219): scikit-learn's
    randomized_svd (installed 1.7.2; the reference pins only >=1.2,
    pyproject.toml:63) with the reference's arguments, then the closed form."""
    from sklearn.utils.extmath import randomized_svd

    V = np.asarray(V, dtype=np.float64)
    if (V < 0).any():
        raise ValueError("V must be non-negative for NNDSVD initialisation.")
    U, S, Vt = randomized_svd(V, n_components=rank, n_oversamples=n_oversamples,
                              random_state=seed)
    return nndsvd_from_svd(V, U, S, Vt, rank, fill)
