"""nmf_rs on MI355X: the omicverse/rust-NMF Python API over gfx950 kernels.

Drop-in counterpart of the reference package ``nmf_rs``
(python/nmf_rs/__init__.py).  Same public names, arguments, defaults,
aliases, validation and exceptions; the native module ``nmf_rs._rust`` is
replaced by ``._rust``, a ctypes binding of the C ABI in include/nmfgpu.h
whose compute runs on the GPU (no CPU fallback).

GPU path: every method of the reference -- brunet, lee, offset, nsNMF,
lsNMF, hals, ehals, dnmf, snmf/r, snmf/l (FP64 arithmetic; snmf's
FCNNLS normal-equation solves in FP64 on the host, its products with V on
the GPU) -- plus ``nndsvd_init`` (its 16 passes over V run on the GPU) and
``cv_rank`` (lsNMF fits).

Additions with no reference counterpart (keyword-only, defaults keep the
reference behaviour): ``precision=`` on ``nmf``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._rust import (
    py_divergence_update_h as _div_h,
    py_divergence_update_w as _div_w,
    py_euclidean_update_h as _euc_h,
    py_euclidean_update_w as _euc_w,
    py_nmf_run as _nmf_run,
    py_smoothing_matrix as _smoothing_matrix,
    set_num_threads as _set_num_threads,
)

__all__ = [
    "NMFResult",
    "nmf",
    "update_h_brunet",
    "update_w_brunet",
    "update_h_lee",
    "update_w_lee",
    "smoothing_matrix",
    "random_init",
    "nndsvd_init",
    "cv_rank",
    "set_num_threads",
]


@dataclass
class NMFResult:
    """Outcome of an NMF run (python/nmf_rs/__init__.py:59-99)."""
    W: np.ndarray
    H: np.ndarray
    n_iter: int
    deviances: np.ndarray = field(default_factory=lambda: np.empty(0, dtype=np.float64))
    method: str = "brunet"
    offset: Optional[np.ndarray] = None

    @property
    def basis(self) -> np.ndarray:
        return self.W

    @property
    def coef(self) -> np.ndarray:
        return self.H

    def fitted(self) -> np.ndarray:
        """The model estimate W H (plus the offset for the offset method)."""
        est = self.W @ self.H
        if self.offset is not None:
            est = est + self.offset[:, None]
        return est

    def __repr__(self) -> str:
        return (f"NMFResult(method='{self.method}', n={self.W.shape[0]}, "
                f"p={self.H.shape[1]}, rank={self.W.shape[1]}, n_iter={self.n_iter})")


def random_init(V, rank, seed=None, *, max_value=None):
    """W0 (n x rank) then H0 (rank x p) ~ U(0, max V) from
    ``numpy.random.default_rng(seed)`` (python/nmf_rs/__init__.py:102-126)."""
    n, p = V.shape
    gen = np.random.default_rng(seed)
    top = float(max_value if max_value is not None else (np.max(V) if V.size else 1.0))
    return gen.uniform(0.0, top, size=(n, rank)), gen.uniform(0.0, top, size=(rank, p))


# ---------------------------------------------------------------------------
# NNDSVD (python/nmf_rs/__init__.py:129-219)
# ---------------------------------------------------------------------------
def _sign_fix(U, Vt, u_based):
    """Deterministic SVD signs: the entry of largest magnitude in each
    column of U (u_based) or row of Vt is made positive."""
    if u_based:
        idx = np.argmax(np.abs(U), axis=0)
        sg = np.sign(U[idx, np.arange(U.shape[1])])
    else:
        idx = np.argmax(np.abs(Vt), axis=1)
        sg = np.sign(Vt[np.arange(Vt.shape[0]), idx])
    return U * sg[np.newaxis, :], Vt * sg[:, np.newaxis]


def _orthonormal_basis(Y, max_cond=1e6):
    """Q of the economic QR of the tall matrix Y, up to the signs of its
    columns: CholeskyQR2 (Fukaya, Nakatsukasa, Yanagisawa, Yamamoto 2014:
    Q = Y R^-1 with R the Cholesky factor of Y^T Y, twice), three BLAS-3
    calls per pass instead of LAPACK's Householder panels.  Its Q is
    Householder's Q times a diagonal of signs, to within eps cond(Y); the
    randomized SVD's sign fix removes the signs.  When the first pass's R
    shows a condition above max_cond (diagonal ratio) or Y^T Y is not
    numerically positive definite, Householder QR is used instead."""
    from scipy import linalg

    Q = Y
    for it in range(2):
        try:
            R = linalg.cholesky(Q.T @ Q, lower=False, check_finite=False)
        except linalg.LinAlgError:
            break
        d = np.abs(np.diag(R))
        if it == 0 and not (d.min() > 0.0 and d.max() <= max_cond * d.min()):
            break
        Q = linalg.solve_triangular(R, Q.T, trans="T", lower=False, check_finite=False).T
    else:
        return Q
    return linalg.qr(Y, mode="economic", check_finite=False)[0]


def _randomized_svd(dev: "_lib.DeviceMatrix", rank: int, n_oversamples: int, seed: int):
    """Rank-``rank`` randomized SVD of the resident V with the sequence of
    scikit-learn's ``randomized_svd`` as the reference calls it
    (__init__.py:172-174): Gaussian sketch from ``RandomState(seed)``,
    n_iter = 7 (4 when rank >= 0.1 min(n, p)) power iterations normalised by
    LU with permuted L, economic QR (by CholeskyQR2 where Y is well
    conditioned, _orthonormal_basis), SVD of the projected matrix (gesdd),
    sign fix.  The power iterations run on the device (products and the LU
    steps, nmfg_matrix_range_finder) when both dimensions of V are at least
    rank + n_oversamples; otherwise (and for n_iter <= 2, which scikit-learn
    runs without a normaliser) the LU steps run on the host between GPU
    products.  The final QR and the small SVD run on the host."""
    from scipy import linalg

    n, p = dev.shape
    width = rank + n_oversamples
    n_iter = 7 if rank < 0.1 * min(n, p) else 4
    flip = n < p  # work on V^T when V is wide
    gen = np.random.RandomState(seed)
    if flip:
        Q = gen.normal(size=(n, width))
        fwd = lambda X: dev.mul(X, trans=True)    # (V^T) X
        back = lambda X: dev.mul(X, trans=False)  # (V^T)^T X
    else:
        Q = gen.normal(size=(p, width))
        fwd = lambda X: dev.mul(X, trans=False)
        back = lambda X: dev.mul(X, trans=True)
    if n_iter > 2 and min(n, p) >= width:
        Y = dev.range_finder(Q, n_iter, flip)
    else:
        for _ in range(n_iter):
            if n_iter > 2:
                Q, _ = linalg.lu(fwd(Q), permute_l=True, check_finite=False)
                Q, _ = linalg.lu(back(Q), permute_l=True, check_finite=False)
            else:
                Q = back(fwd(Q))
        Y = fwd(Q)
    Q = _orthonormal_basis(Y)
    B = back(Q).T
    Uh, s, Vt = linalg.svd(B, full_matrices=False, lapack_driver="gesdd")
    U = Q @ Uh
    U, Vt = _sign_fix(U, Vt, u_based=not flip)
    if flip:
        return Vt[:rank].T, s[:rank], U[:, :rank].T
    return U[:, :rank], s[:rank], Vt[:rank]


def nndsvd_init(V, rank, *, fill="mean", seed=0, n_oversamples=10):
    """NNDSVD initialisation (Boutsidis & Gallopoulos 2008) with the
    reference's arguments, fill modes and tie rule
    (python/nmf_rs/__init__.py:129-219).  V may also be a float64 torch
    tensor on the GPU (no host copy)."""
    if _lib._is_device_tensor(V):
        if bool((V < 0).any()):
            raise ValueError("V must be non-negative for NNDSVD initialisation.")
        v_mean = float(V.mean()) if V.numel() else 0.0
    else:
        V = np.asarray(V, dtype=np.float64)
        if (V < 0).any():
            raise ValueError("V must be non-negative for NNDSVD initialisation.")
        v_mean = float(V.mean()) if V.size else 0.0
    if fill not in ("zero", "mean", "eps"):
        raise ValueError(f"unknown fill mode '{fill}': supported zero / mean / eps")
    dev = _lib.DeviceMatrix(V)
    try:
        return nndsvd_on_device(dev, rank, fill=fill, seed=seed, n_oversamples=n_oversamples,
                                v_mean=v_mean)
    finally:
        dev.close()


def _nndsvd_from_svd(U, S, Vt, rank, fill, v_mean):
    """NNDSVD closed form from a rank-``rank`` SVD and the fill
    (python/nmf_rs/__init__.py:176-217): component 0 from |u|, |v|; each
    later one from the sign side (positive on ties) of larger mass."""
    n, p = U.shape[0], Vt.shape[1]
    W = np.empty((n, rank))
    H = np.empty((rank, p))
    root = np.sqrt(S[0])
    np.multiply(root, np.abs(U[:, 0]), out=W[:, 0])
    np.multiply(root, np.abs(Vt[0, :]), out=H[0, :])
    if rank > 1:
        # components 1..: the positive and negative parts of u_k (all at
        # once) and of v_k (row by row, in the row of H that the chosen side
        # then fills: no p-long temporaries, and Vt may be a transposed
        # view), their norms, the side of larger mass (the positive side on
        # ties), then scale * part / norm where the mass is positive.  The
        # sums of squares use einsum, not BLAS dot: with many host BLAS
        # threads the per-row calls cost more in thread start-up than work
        # (1508 x 82612 rank 24 with 16 threads: 96 against 20 ms).
        Uk = U[:, 1:rank]
        up = np.maximum(Uk, 0.0)
        un = up - Uk                                   # max(-u, 0), exact
        nup = np.sqrt(np.einsum("ik,ik->k", up, up))
        nun = np.sqrt(np.einsum("ik,ik->k", un, un))
        nvp, nvn = np.empty(rank - 1), np.empty(rank - 1)
        for k in range(rank - 1):
            hk = H[1 + k]
            np.maximum(Vt[1 + k], 0.0, out=hk)
            nvp[k] = np.sqrt(np.einsum("j,j->", hk, hk))
            np.subtract(hk, Vt[1 + k], out=hk)         # max(-v, 0), exact
            nvn[k] = np.sqrt(np.einsum("j,j->", hk, hk))
        mp, mn = nup * nvp, nun * nvn
        pos = mp >= mn
        mass = np.where(pos, mp, mn)
        ok = mass > 0
        na = np.where(ok, np.where(pos, nup, nun), 1.0)
        nb = np.where(ok, np.where(pos, nvp, nvn), 1.0)
        scale = np.where(ok, np.sqrt(S[1:rank] * np.where(ok, mass, 1.0)), 0.0)
        a = W[:, 1:rank]
        np.copyto(a, np.where(pos[np.newaxis, :], up, un))
        a *= scale[np.newaxis, :]
        a /= na[np.newaxis, :]
        a[:, ~ok] = 0.0
        for k in range(rank - 1):
            hk = H[1 + k]                              # holds max(-v, 0)
            if not ok[k]:
                hk[:] = 0.0
                continue
            if pos[k]:
                np.maximum(Vt[1 + k], 0.0, out=hk)
            hk *= scale[k]
            hk /= nb[k]
    if fill == "eps":
        np.copyto(W, 1e-9, where=W < 1e-9)
        np.copyto(H, 1e-9, where=H < 1e-9)
    elif fill == "mean":
        np.copyto(W, v_mean, where=W <= 0.0)
        np.copyto(H, v_mean, where=H <= 0.0)
    return W, H


def nndsvd_on_device(dev, rank, *, fill="mean", seed=0, n_oversamples=10, v_mean=None):
    """nndsvd_init on a matrix already resident on the GPU
    (``_lib.DeviceMatrix``, any storage form); ``v_mean`` is mean(V) for
    fill='mean'.  No reference counterpart (the served path uploads V once)."""
    if fill not in ("zero", "mean", "eps"):
        raise ValueError(f"unknown fill mode '{fill}': supported zero / mean / eps")
    if fill == "mean" and v_mean is None:
        raise ValueError("fill='mean' needs v_mean")
    U, S, Vt = _randomized_svd(dev, int(rank), int(n_oversamples), seed)
    return _nndsvd_from_svd(U, S, Vt, int(rank), fill, v_mean)


# ---------------------------------------------------------------------------
# Single steps (python/nmf_rs/__init__.py:222-264)
# ---------------------------------------------------------------------------
def _f64_c(arr, name):
    """float64 C-contiguous 2-D (python/nmf_rs/__init__.py:235-240)."""
    a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64))
    if a.ndim != 2:
        raise ValueError(f"{name} must be 2-D, got shape {a.shape}")
    return a


def update_h_brunet(V, W, H):
    """One KL update of H (R `std.divergence.update.h`)."""
    return _div_h(_f64_c(V, "V"), _f64_c(W, "W"), _f64_c(H, "H"))


def update_w_brunet(V, W, H):
    """One KL update of W (R `std.divergence.update.w`)."""
    return _div_w(_f64_c(V, "V"), _f64_c(W, "W"), _f64_c(H, "H"))


def update_h_lee(V, W, H, eps=1e-9):
    """One Frobenius update of H (R `std.euclidean.update.h`)."""
    return _euc_h(_f64_c(V, "V"), _f64_c(W, "W"), _f64_c(H, "H"), float(eps))


def update_w_lee(V, W, H, eps=1e-9):
    """One Frobenius update of W (R `std.euclidean.update.w`)."""
    return _euc_w(_f64_c(V, "V"), _f64_c(W, "W"), _f64_c(H, "H"), float(eps))


def smoothing_matrix(rank, theta):
    """S = (1 - theta) I + (theta / r) 1 1^T."""
    return _smoothing_matrix(int(rank), float(theta))


def _cv_folds(n, p, ranks, n_folds, mask_frac, seed):
    """(rank, fold, fold seed, training mask) in the reference's draw order
    (python/nmf_rs/__init__.py:322-329): one default_rng(seed) draws a seed
    per (rank, fold); that seed's generator draws the mask (True = training
    entry, held out with probability mask_frac)."""
    master = np.random.default_rng(seed)
    for k in ranks:
        for fold in range(n_folds):
            fs = int(master.integers(0, 2**31 - 1))
            yield int(k), fold, fs, np.random.default_rng(fs).random((n, p)) > mask_frac


def cv_rank(V, ranks: Sequence[int], *, method="hals", n_folds=4, mask_frac=0.05, max_iter=100,
            init="nndsvd", seed=0, num_threads=None, **nmf_kwargs):
    """Cross-validated rank selection (python/nmf_rs/__init__.py:273-355):
    per (rank, fold), ls-nmf on V with the fold's held-out entries weighted
    0 (a 0/1 weight, held as one byte per entry on the GPU), started from
    nndsvd_init (or random_init) of the training matrix; reports the
    weighted training loss and the held-out MSE.  ``method`` is accepted as
    in the reference (which does not use it)."""
    try:
        import pandas as pd
    except ImportError as e:
        raise ImportError("cv_rank() requires pandas") from e
    V = _f64_c(V, "V")
    if (V < 0).any():
        raise ValueError("V must be non-negative")
    n, p = V.shape
    rows = []
    for k, fold, fs, train in _cv_folds(n, p, ranks, n_folds, mask_frac, seed):
        weight = train.astype(np.float64)
        V_train = V * weight
        if init == "nndsvd":
            W0, H0 = nndsvd_init(V_train, k, fill="mean", seed=fs)
        else:
            W0, H0 = random_init(V_train, k, seed=fs)
        res = nmf(V_train, rank=k, method="ls-nmf", W0=W0, H0=H0, weight=weight,
                  max_iter=max_iter, num_threads=num_threads, **nmf_kwargs)
        est = res.fitted()
        resid = (V - est) * weight
        held = ~train
        rows.append({"rank": k, "fold": fold, "train_loss": 0.5 * float((resid ** 2).sum()),
                     "test_mse": float(((V[held] - est[held]) ** 2).mean())})
    return pd.DataFrame(rows)


_METHOD_ALIASES = {
    # python/nmf_rs/__init__.py:362-387
    "brunet": "brunet", "kl": "brunet",
    "lee": "lee", "frobenius": "lee", "euclidean": "lee",
    "offset": "offset",
    "nsnmf": "nsnmf", "ns": "nsnmf", "ns_nmf": "nsnmf",
    "hals": "hals",
    "ehals": "ehals", "e-hals": "ehals", "extrap_hals": "ehals",
    "dnmf": "dnmf", "diag_nmf": "dnmf", "rcppml": "dnmf",
    "ls-nmf": "lsnmf", "lsnmf": "lsnmf",
    "snmf/r": "snmf/r", "snmf_r": "snmf/r", "snmfr": "snmf/r",
    "snmf/l": "snmf/l", "snmf_l": "snmf/l", "snmfl": "snmf/l",
}


def set_num_threads(n):
    """rayon pool size in the reference; accepted and ignored on the GPU
    (True on the first call, False afterwards, as build_global)."""
    return _set_num_threads(int(n))


def nmf(V, rank, method="brunet", *, W0=None, H0=None, seed=None, max_iter=2000, eps=1e-9,
        rescale=True, theta=0.5, offset=None, weight=None, sparsity=0.01, smoothness=1.0,
        stop="max_iter", stationary_th=float(np.finfo(np.float64).eps), check_interval=50,
        check_niter=10, num_threads=None, precision="fp64") -> NMFResult:
    """V ~ W H with the reference's signature and defaults
    (python/nmf_rs/__init__.py:396-568); computed on the GPU."""
    method_norm = _METHOD_ALIASES.get(method.lower())
    if method_norm is None:
        raise ValueError(f"unknown method '{method}'. supported: brunet, lee, offset, nsNMF "
                         "(plus aliases KL/Frobenius/ns).")
    V_arr = _f64_c(V, "V")
    if (V_arr < 0).any():
        raise ValueError("V must be non-negative.")
    n, p = V_arr.shape
    if W0 is None or H0 is None:
        w_init, h_init = random_init(V_arr, rank, seed=seed)
        W0 = w_init if W0 is None else W0
        H0 = h_init if H0 is None else H0
    W0 = _f64_c(W0, "W0")
    H0 = _f64_c(H0, "H0")
    if W0.shape != (n, rank):
        raise ValueError(f"W0.shape {W0.shape} != ({n}, {rank})")
    if H0.shape != (rank, p):
        raise ValueError(f"H0.shape {H0.shape} != ({rank}, {p})")
    offset_arr = None
    if offset is not None:
        offset_arr = np.ascontiguousarray(np.asarray(offset, dtype=np.float64))
        if offset_arr.shape != (n,):
            raise ValueError(f"offset.shape {offset_arr.shape} != ({n},)")
    weight_arr = None
    if weight is not None:
        weight_arr = np.ascontiguousarray(np.asarray(weight, dtype=np.float64))
        if weight_arr.shape != V_arr.shape:
            raise ValueError(f"weight.shape {weight_arr.shape} != V.shape {V_arr.shape}")
    # sparsity / smoothness routing (python/nmf_rs/__init__.py:528-538)
    if method_norm == "snmf/r":
        sh, sw, gh, gw = float(sparsity), 0.0, 0.0, float(smoothness)
    elif method_norm == "snmf/l":
        sh, sw, gh, gw = 0.0, float(sparsity), float(smoothness), 0.0
    elif method_norm == "dnmf":
        sh = sw = float(sparsity)
        gh = gw = 0.0
    else:
        sh = sw = gh = gw = 0.0
    out = _nmf_run(method_norm, V_arr, W0, H0, int(max_iter), eps=float(eps),
                   rescale=bool(rescale), theta=float(theta), offset=offset_arr,
                   weight=weight_arr, sparsity_h=sh, sparsity_w=sw, smoothness_h=gh,
                   smoothness_w=gw, stop=stop, stationary_th=float(stationary_th),
                   check_interval=int(check_interval), check_niter=int(check_niter),
                   num_threads=int(num_threads) if num_threads is not None else None,
                   precision=precision)
    return NMFResult(W=np.asarray(out["W"]), H=np.asarray(out["H"]), n_iter=int(out["n_iter"]),
                     deviances=np.asarray(out["deviances"], dtype=np.float64),
                     method=method_norm,
                     offset=np.asarray(out["offset"]) if "offset" in out else None)
