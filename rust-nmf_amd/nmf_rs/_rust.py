"""GPU counterpart of the reference's native module ``nmf_rs._rust``
(rust/src/lib.rs:1835-2187): the same seven entry points with the same
arguments, return values and exceptions, backed by the gfx950 C ABI
(include/nmfgpu.h) instead of rayon CPU loops.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, context, f64c, lib, ptr

_ALGO_ALIASES = {
    # Algo::from_str (lib.rs:1472-1505)
    "brunet": "brunet", "kl": "brunet",
    "lee": "lee", "frobenius": "lee", "euclidean": "lee",
    "offset": "offset",
    "nsnmf": "nsnmf", "ns": "nsnmf", "ns_nmf": "nsnmf",
    "hals": "hals",
    "ehals": "ehals", "e-hals": "ehals", "extrap_hals": "ehals",
    "dnmf": "dnmf", "diag_nmf": "dnmf", "diagonalised_nmf": "dnmf", "rcppml": "dnmf",
    "ls-nmf": "lsnmf", "lsnmf": "lsnmf",
    "snmf/r": "snmf/r", "snmf_r": "snmf/r", "snmfr": "snmf/r",
    "snmf/l": "snmf/l", "snmf_l": "snmf/l", "snmfl": "snmf/l",
}


def _algo(name: str) -> str:
    a = _ALGO_ALIASES.get(str(name).lower())
    if a is None:
        raise ValueError(
            f"unknown algorithm '{name}': supported: brunet, lee, offset, nsNMF, hals, ehals, "
            "dnmf, lsNMF, snmf/r, snmf/l")
    return a


def _shapes(v, w, h):
    v, w, h = f64c(v), f64c(w), f64c(h)
    if v.ndim != 2 or w.ndim != 2 or h.ndim != 2:
        raise ValueError("v, w and h must be 2-D")
    n, p = v.shape
    r = w.shape[1]
    if w.shape[0] != n or h.shape != (r, p):
        raise ValueError(f"shape mismatch: V {v.shape}, W {w.shape}, H {h.shape}")
    return v, w, h, n, p, r


def set_num_threads(n):
    """lib.rs:1841-1847: True on the first call only (the GPU ignores n)."""
    return bool(lib().nmfg_set_num_threads(int(n)))


def py_divergence_update_h(v, w, h):
    """lib.rs:1849-1862 -> divergence_update_h_impl (lib.rs:117-144)."""
    v, w, h, n, p, r = _shapes(v, w, h)
    out = np.empty((r, p))
    check(lib().nmfg_divergence_update_h(context(), ptr(v), n, p, ptr(w), ptr(h), r, ptr(out)))
    return out


def py_divergence_update_w(v, w, h):
    """lib.rs:1864-1877 -> divergence_update_w_impl (lib.rs:146-204)."""
    v, w, h, n, p, r = _shapes(v, w, h)
    out = np.empty((n, r))
    check(lib().nmfg_divergence_update_w(context(), ptr(v), n, p, ptr(w), ptr(h), r, ptr(out)))
    return out


def py_euclidean_update_h(v, w, h, eps):
    """lib.rs:1879-1892 -> euclidean_update_h_impl (lib.rs:247-276)."""
    v, w, h, n, p, r = _shapes(v, w, h)
    out = np.empty((r, p))
    check(lib().nmfg_euclidean_update_h(context(), ptr(v), n, p, ptr(w), ptr(h), r,
                                        float(eps), ptr(out)))
    return out


def py_euclidean_update_w(v, w, h, eps):
    """lib.rs:1894-1907 -> euclidean_update_w_impl (lib.rs:278-322)."""
    v, w, h, n, p, r = _shapes(v, w, h)
    out = np.empty((n, r))
    check(lib().nmfg_euclidean_update_w(context(), ptr(v), n, p, ptr(w), ptr(h), r,
                                        float(eps), ptr(out)))
    return out


def py_smoothing_matrix(r, theta):
    """lib.rs:2171-2175."""
    out = np.empty((int(r), int(r)))
    check(lib().nmfg_smoothing_matrix(int(r), float(theta), ptr(out)))
    return out


def py_nmf_run(algo, v, w0, h0, max_iter, *, eps=1e-9, rescale=True, theta=0.5, offset=None,
               weight=None, sparsity_h=0.0, sparsity_w=0.0, smoothness_h=0.0, smoothness_w=0.0,
               stop="max_iter", stationary_th=2.220446049250313e-16, check_interval=50,
               check_niter=10, num_threads=None, precision="fp64", rel_tol=1e-6, min_iter=1,
               ctx=None):
    """lib.rs:1909-2169.  Returns {"W", "H", "n_iter", "deviances"}, plus
    "offset" (1-d) for the offset method (lib.rs:2160-2164).

    ``precision``, ``stop="rel_change"`` with ``rel_tol`` / ``min_iter``, and
    ``ctx`` (a ``_lib.Context``; default the process-wide one) are
    keyword-only and have no reference counterpart; the defaults keep the
    reference's behaviour.
    """
    a = _algo(algo)
    if a == "lsnmf" and weight is None:
        raise ValueError("method='lsnmf' requires a `weight` matrix of the same shape as V")
    if int(max_iter) < 0:
        raise OverflowError("can't convert negative int to unsigned")
    prm = _lib.make_params(a, max_iter, eps=eps, rescale=rescale, theta=theta,
                           sparsity_h=sparsity_h, sparsity_w=sparsity_w,
                           smoothness_h=smoothness_h, smoothness_w=smoothness_w, stop=stop,
                           stationary_th=stationary_th, check_interval=check_interval,
                           check_niter=check_niter, num_threads=num_threads,
                           precision=precision, rel_tol=rel_tol, min_iter=min_iter)
    v, w0, h0, n, p, r = _shapes(v, w0, h0)
    W = w0.copy()
    H = h0.copy()
    cap = int(max_iter) + 1 if prm.stop != _lib.STOP["max_iter"] else 0
    if a in ("snmf/r", "snmf/l"):
        cap = int(max_iter)  # Kim-Park records a deviance every iteration (lib.rs:1313-1321)
    dev = np.empty(max(cap, 1))
    n_iter = ctypes.c_int64(0)
    n_dev = ctypes.c_int64(0)
    handle = ctx.handle if ctx is not None else context()
    if a == "lsnmf":
        wgt = f64c(weight)
        if wgt.shape != v.shape:
            raise ValueError(f"weight.shape {wgt.shape} != V.shape {v.shape}")
        check(lib().nmfg_run_lsnmf(handle, ctypes.byref(prm), ptr(v), n, p, p, ptr(wgt), p,
                                   ptr(W), ptr(H), r, ctypes.byref(n_iter), ptr(dev), cap,
                                   ctypes.byref(n_dev)))
        return {"W": W, "H": H, "n_iter": int(n_iter.value),
                "deviances": [float(x) for x in dev[: n_dev.value]]}
    off = None
    if a == "offset":
        # py_nmf_run reads `offset` as a length-n float64 array (lib.rs:1944, 1963)
        off = np.zeros(n) if offset is None else f64c(offset).copy()
        if offset is not None and off.shape != (n,):
            raise ValueError(f"offset.shape {off.shape} != ({n},)")
    check(lib().nmfg_run_offset(handle, ctypes.byref(prm), ptr(v), n, p, p, ptr(W), ptr(H), r,
                                ptr(off) if off is not None else None,
                                int(offset is not None and off is not None),
                                ctypes.byref(n_iter), ptr(dev), cap, ctypes.byref(n_dev)))
    out = {"W": W, "H": H, "n_iter": int(n_iter.value),
           "deviances": [float(x) for x in dev[: n_dev.value]]}
    if off is not None:
        out["offset"] = off
    return out
