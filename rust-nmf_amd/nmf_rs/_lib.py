"""ctypes binding of the C ABI in include/nmfgpu.h (libnmfgpu.so).

Loads the in-tree gfx950 library and maps its status codes onto Python
exceptions the way the reference's PyO3 layer does: invalid arguments raise
ValueError (rust/src/lib.rs:1499-1502, 1520-1523, 1967-1971), everything else
RuntimeError (lib.rs:1981-1986).  There is no CPU fallback: a missing
library or device raises RuntimeError.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libnmfgpu.so"

NMFG_OK = 0
NMFG_EINVAL = -1
NMFG_ENOMEM = -2
NMFG_EDEVICE = -3
NMFG_EUNSUPPORTED = -4
NMFG_EBUDGET = -5

ALGO = {"brunet": 0, "lee": 1, "offset": 2, "nsnmf": 3, "hals": 4, "ehals": 5, "dnmf": 6,
        "lsnmf": 7, "snmf/r": 8, "snmf/l": 9}
STOP = {"max_iter": 0, "stationary": 1, "rel_change": 2}
# precision forms (enum nmfg_precision): "fp32-store" / "bf16x2-store" keep
# FP64 arithmetic on the rounded V; "fp32" runs the products in FP32,
# "bf16x2" lee's products on bf16 MFMA over hi + lo operands
PRECISION = {"fp64": 0, "fp32-store": 1, "bf16x2-store": 2, "fp32": 3, "bf16x2": 4}

# Exported symbols, one per declaration in include/nmfgpu.h.
EXPORTS = (
    "nmfg_params_init", "nmfg_ctx_create", "nmfg_ctx_destroy", "nmfg_ctx_stream",
    "nmfg_ctx_release_cached", "nmfg_run",
    "nmfg_euclidean_update_h", "nmfg_euclidean_update_w", "nmfg_divergence_update_h",
    "nmfg_divergence_update_w", "nmfg_smoothing_matrix", "nmfg_set_num_threads",
    "nmfg_last_error", "nmfg_max_rank", "nmfg_matrix_create", "nmfg_matrix_upload",
    "nmfg_matrix_destroy", "nmfg_matrix_mul", "nmfg_solve_create", "nmfg_solve_set_factors",
    "nmfg_solve_launch", "nmfg_solve_fetch", "nmfg_solve_profile", "nmfg_solve_destroy",
    "nmfg_solve_create_batched", "nmfg_solve_batch", "nmfg_consensus", "nmfg_run_offset",
    "nmfg_matrix_range_finder", "nmfg_lu_permute_l",
    "nmfg_solve_set_offset", "nmfg_solve_get_offset", "nmfg_run_lsnmf", "nmfg_solve_set_weight",
    "nmfg_fcnnls", "nmfg_run_device_noalloc", "nmfg_sweep_create", "nmfg_sweep_set_factors",
    "nmfg_sweep_launch", "nmfg_sweep_fetch", "nmfg_sweep_plan", "nmfg_sweep_destroy",
    "nmfg_budget_create", "nmfg_budget_destroy", "nmfg_ctx_set_budget", "nmfg_budget_query",
    "nmfg_footprint", "nmfg_device_memory", "nmfg_solve_profile_evicted",
    "nmfg_group_create", "nmfg_group_add", "nmfg_group_step", "nmfg_group_size",
    "nmfg_group_stats", "nmfg_group_destroy",
)


class Params(ctypes.Structure):
    """struct nmfg_params (py_nmf_run keyword arguments, lib.rs:1912-1933)."""
    _fields_ = [
        ("algo", ctypes.c_int32),
        ("max_iter", ctypes.c_int64),
        ("eps", ctypes.c_double),
        ("rescale", ctypes.c_int32),
        ("theta", ctypes.c_double),
        ("sparsity_h", ctypes.c_double),
        ("sparsity_w", ctypes.c_double),
        ("smoothness_h", ctypes.c_double),
        ("smoothness_w", ctypes.c_double),
        ("stop", ctypes.c_int32),
        ("stationary_th", ctypes.c_double),
        ("check_interval", ctypes.c_int64),
        ("check_niter", ctypes.c_int64),
        ("num_threads", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("rel_tol", ctypes.c_double),
        ("min_iter", ctypes.c_int64),
    ]


_lib = None
_ctx = None
_lock = threading.Lock()

_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_vp = ctypes.c_void_p


def _declare(L):
    sig = {
        "nmfg_params_init": (None, [ctypes.POINTER(Params)]),
        "nmfg_ctx_create": (ctypes.c_int, [_i32, ctypes.POINTER(_vp)]),
        "nmfg_ctx_destroy": (ctypes.c_int, [_vp]),
        "nmfg_ctx_stream": (_vp, [_vp]),
        "nmfg_ctx_release_cached": (ctypes.c_int, [_vp]),
        "nmfg_run": (ctypes.c_int, [_vp, ctypes.POINTER(Params), _dp, _i64, _i64, _i64, _dp, _dp,
                                    _i32, ctypes.POINTER(_i64), _dp, _i64, ctypes.POINTER(_i64)]),
        "nmfg_euclidean_update_h": (ctypes.c_int, [_vp, _dp, _i64, _i64, _dp, _dp, _i32,
                                                   ctypes.c_double, _dp]),
        "nmfg_euclidean_update_w": (ctypes.c_int, [_vp, _dp, _i64, _i64, _dp, _dp, _i32,
                                                   ctypes.c_double, _dp]),
        "nmfg_divergence_update_h": (ctypes.c_int, [_vp, _dp, _i64, _i64, _dp, _dp, _i32, _dp]),
        "nmfg_divergence_update_w": (ctypes.c_int, [_vp, _dp, _i64, _i64, _dp, _dp, _i32, _dp]),
        "nmfg_smoothing_matrix": (ctypes.c_int, [_i32, ctypes.c_double, _dp]),
        "nmfg_set_num_threads": (ctypes.c_int, [_i64]),
        "nmfg_last_error": (ctypes.c_char_p, []),
        "nmfg_max_rank": (_i32, []),
        "nmfg_matrix_create": (ctypes.c_int, [_vp, _i64, _i64, _i32, ctypes.POINTER(_vp)]),
        "nmfg_matrix_upload": (ctypes.c_int, [_vp, _dp, _i64]),
        "nmfg_matrix_destroy": (ctypes.c_int, [_vp]),
        "nmfg_matrix_mul": (ctypes.c_int, [_vp, _i32, _dp, _i32, _dp]),
        "nmfg_solve_create": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(Params),
                                             ctypes.POINTER(_vp)]),
        "nmfg_solve_set_factors": (ctypes.c_int, [_vp, _dp, _dp]),
        "nmfg_solve_launch": (ctypes.c_int, [_vp]),
        "nmfg_solve_fetch": (ctypes.c_int, [_vp, _dp, _dp, ctypes.POINTER(_i64), _dp, _i64,
                                            ctypes.POINTER(_i64)]),
        "nmfg_solve_profile": (ctypes.c_int, [_vp, _dp, ctypes.POINTER(_i64), _dp,
                                              ctypes.POINTER(_i64)]),
        "nmfg_solve_destroy": (ctypes.c_int, [_vp]),
        "nmfg_solve_create_batched": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(Params), _i32,
                                                     ctypes.POINTER(_vp)]),
        "nmfg_solve_batch": (_i32, [_vp]),
        "nmfg_consensus": (ctypes.c_int, [_vp, ctypes.POINTER(_i32), _i32, _i64, _dp]),
        "nmfg_run_offset": (ctypes.c_int, [_vp, ctypes.POINTER(Params), _dp, _i64, _i64, _i64,
                                           _dp, _dp, _i32, _dp, _i32, ctypes.POINTER(_i64), _dp,
                                           _i64, ctypes.POINTER(_i64)]),
        "nmfg_solve_set_offset": (ctypes.c_int, [_vp, _dp]),
        "nmfg_run_lsnmf": (ctypes.c_int, [_vp, ctypes.POINTER(Params), _dp, _i64, _i64, _i64, _dp,
                                          _i64, _dp, _dp, _i32, ctypes.POINTER(_i64), _dp, _i64,
                                          ctypes.POINTER(_i64)]),
        "nmfg_solve_set_weight": (ctypes.c_int, [_vp, _dp, _i64]),
        "nmfg_fcnnls": (ctypes.c_int, [_dp, _dp, _i32, _i64, ctypes.c_double, _dp]),
        "nmfg_solve_get_offset": (ctypes.c_int, [_vp, _dp]),
        "nmfg_sweep_create": (ctypes.c_int, [_vp, ctypes.POINTER(_i32), _i32, ctypes.POINTER(Params),
                                             _i32, ctypes.POINTER(_vp)]),
        "nmfg_sweep_set_factors": (ctypes.c_int, [_vp, ctypes.POINTER(_dp), ctypes.POINTER(_dp)]),
        "nmfg_sweep_launch": (ctypes.c_int, [_vp]),
        "nmfg_sweep_fetch": (ctypes.c_int, [_vp, ctypes.POINTER(_dp), ctypes.POINTER(_dp)]),
        "nmfg_sweep_plan": (ctypes.c_int, [_vp, ctypes.c_char_p, _i64, ctypes.POINTER(ctypes.c_double)]),
        "nmfg_sweep_destroy": (ctypes.c_int, [_vp]),
        "nmfg_budget_create": (ctypes.c_int, [_i64, ctypes.POINTER(_vp)]),
        "nmfg_budget_destroy": (ctypes.c_int, [_vp]),
        "nmfg_ctx_set_budget": (ctypes.c_int, [_vp, _vp]),
        "nmfg_budget_query": (ctypes.c_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                             ctypes.POINTER(_i64)]),
        "nmfg_solve_profile_evicted": (ctypes.c_int, [_vp, _i64, _dp, ctypes.POINTER(_i64), _dp,
                                                      ctypes.POINTER(_i64)]),
        "nmfg_device_memory": (ctypes.c_int, [_i32, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
        "nmfg_footprint": (ctypes.c_int, [ctypes.POINTER(Params), _i64, _i64, _i32, _i32,
                                          ctypes.POINTER(_i64)]),
        "nmfg_run_device_noalloc": (ctypes.c_int, [_vp, ctypes.POINTER(Params), _vp, _i64, _i64,
                                                   _i64, _vp, _vp, _i64, _i32,
                                                   ctypes.POINTER(_i64), _dp, _i64,
                                                   ctypes.POINTER(_i64)]),
        "nmfg_matrix_range_finder": (ctypes.c_int, [_vp, _i32, _dp, _i32, _i32, _dp]),
        "nmfg_lu_permute_l": (ctypes.c_int, [_vp, _dp, _i64, _i32, _dp]),
        "nmfg_group_create": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_vp)]),
        "nmfg_group_add": (ctypes.c_int, [_vp, _vp]),
        "nmfg_group_step": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), _i32, ctypes.POINTER(_i32)]),
        "nmfg_group_size": (ctypes.c_int, [_vp, ctypes.POINTER(_i32)]),
        "nmfg_group_stats": (ctypes.c_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                            ctypes.POINTER(_i32)]),
        "nmfg_group_destroy": (ctypes.c_int, [_vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _share_hip_runtime():
    """One HIP runtime per process.  torch's wheel ships its own
    libamdhip64 / libhsa-runtime64 under a different file name; loaded after
    this library it would start a second runtime that cannot open the GPU.
    Loading torch first (when installed) makes this library bind to that same
    runtime instead."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """The loaded library (raises RuntimeError when it was not built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not LIB_PATH.exists():
                    raise RuntimeError(
                        f"GPU backend not built: {LIB_PATH} is missing. Run "
                        "`make -C rust-nmf_amd` (or __graft_entry__.build()).")
                _share_hip_runtime()
                L = ctypes.CDLL(str(LIB_PATH))
                _declare(L)
                _lib = L
    return _lib


class BudgetError(RuntimeError):
    """A device allocation or a request that does not fit the memory budget
    (NMFG_EBUDGET)."""


def check(rc: int):
    if rc == NMFG_OK:
        return
    msg = lib().nmfg_last_error().decode(errors="replace")
    if rc == NMFG_EINVAL:
        raise ValueError(msg)
    if rc == NMFG_EBUDGET:
        raise BudgetError(f"nmfgpu error {rc}: {msg}")
    raise RuntimeError(f"nmfgpu error {rc}: {msg}")


def context():
    """The process-wide context on device 0 (created on first use)."""
    global _ctx
    if _ctx is None:
        with _lock:
            if _ctx is None:
                h = _vp()
                check(lib().nmfg_ctx_create(0, ctypes.byref(h)))
                _ctx = h
    return _ctx


class Budget:
    """A device-memory budget (nmfg_budget) that contexts charge their
    blocks to; allocations beyond it raise BudgetError."""

    def __init__(self, nbytes: int):
        h = _vp()
        check(lib().nmfg_budget_create(int(nbytes), ctypes.byref(h)))
        self.handle = h

    def query(self):
        """(cap, used, peak) in bytes."""
        c, u, pk = _i64(0), _i64(0), _i64(0)
        check(lib().nmfg_budget_query(self.handle, ctypes.byref(c), ctypes.byref(u),
                                      ctypes.byref(pk)))
        return c.value, u.value, pk.value

    def close(self):
        if getattr(self, "handle", None):
            check(lib().nmfg_budget_destroy(self.handle))
            self.handle = None


def footprint(params, n, p, r, nndsvd_width=0):
    """Device bytes of an n x p matrix in params' form plus one solve of rank
    r (plus NNDSVD's range finder of the given width), nmfg_footprint."""
    b = _i64(0)
    check(lib().nmfg_footprint(ctypes.byref(params), int(n), int(p), int(r), int(nndsvd_width),
                               ctypes.byref(b)))
    return b.value


def device_memory(device=0):
    """(free, total) bytes of the device (hipMemGetInfo)."""
    f, t = _i64(0), _i64(0)
    check(lib().nmfg_device_memory(int(device), ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


class Context:
    """A library context of its own (nmfg_ctx: device 0 and a HIP stream of
    its own), for work that should overlap with other contexts' work, e.g.
    one per serving worker.  The default process-wide one is context()."""

    def __init__(self, device: int = 0):
        h = _vp()
        check(lib().nmfg_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h

    def stream_handle(self) -> int:
        return int(lib().nmfg_ctx_stream(self.handle) or 0)

    def set_budget(self, budget: "Budget | None"):
        """Charges the context's device blocks to ``budget`` (None detaches)."""
        check(lib().nmfg_ctx_set_budget(self.handle, budget.handle if budget else None))

    def release_cached(self):
        """Frees the device memory the context keeps from destroyed matrices
        and solves (nmfg_ctx_release_cached)."""
        check(lib().nmfg_ctx_release_cached(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            lib().nmfg_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def f64c(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def make_params(algo="brunet", max_iter=2000, *, eps=1e-9, rescale=True, theta=0.5,
                sparsity_h=0.0, sparsity_w=0.0, smoothness_h=0.0, smoothness_w=0.0,
                stop="max_iter", stationary_th=2.220446049250313e-16, check_interval=50,
                check_niter=10, num_threads=None, precision="fp64", rel_tol=1e-6,
                min_iter=1) -> Params:
    a = str(algo).lower()
    if a not in ALGO:
        raise ValueError(f"unknown algorithm '{algo}'")
    s = str(stop).lower()
    s = {"fixed": "max_iter", "none": "max_iter"}.get(s, s)
    if s not in STOP:
        raise ValueError(f"unknown stopping criterion '{stop}': supported: max_iter, stationary, "
                         "rel_change")
    if precision not in PRECISION:
        raise ValueError(f"unknown precision '{precision}': supported "
                         f"{', '.join(PRECISION)}")
    p = Params()
    lib().nmfg_params_init(ctypes.byref(p))
    p.algo = ALGO[a]
    p.max_iter = int(max_iter)
    p.eps = float(eps)
    p.rescale = int(bool(rescale))
    p.theta = float(theta)
    p.sparsity_h = float(sparsity_h)
    p.sparsity_w = float(sparsity_w)
    p.smoothness_h = float(smoothness_h)
    p.smoothness_w = float(smoothness_w)
    p.stop = STOP[s]
    p.stationary_th = float(stationary_th)
    p.check_interval = int(check_interval)
    p.check_niter = int(check_niter)
    p.num_threads = int(num_threads) if num_threads else 0
    p.precision = PRECISION[precision]
    p.rel_tol = float(rel_tol)
    p.min_iter = int(min_iter)
    return p


def _is_device_tensor(x):
    return hasattr(x, "data_ptr") and getattr(x, "is_cuda", False)


class DeviceMatrix:
    """V resident on the GPU (nmfg_matrix)."""

    def __init__(self, V, precision="fp64", ctx: "Context | None" = None):
        if precision not in PRECISION:
            raise ValueError(f"unknown precision '{precision}': supported "
                         f"{', '.join(PRECISION)}")
        dev = _is_device_tensor(V)
        if dev:
            # an FP64 CUDA tensor: uploaded device to device
            import torch

            if V.dtype != torch.float64 or V.dim() != 2:
                raise ValueError("a device V must be a 2-D float64 tensor")
            V = V.contiguous()
            # the upload runs on the library context's stream, which is not
            # ordered after torch's: wait until V (and the copy above) is
            # written
            torch.cuda.current_stream(V.device).synchronize()
        else:
            V = f64c(V)
            if V.ndim != 2:
                raise ValueError(f"V must be 2-D, got shape {V.shape}")
        self.shape = tuple(V.shape)
        self.precision = precision
        h = _vp()
        self.ctx = ctx
        check(lib().nmfg_matrix_create(ctx.handle if ctx is not None else context(),
                                       self.shape[0], self.shape[1], PRECISION[precision],
                                       ctypes.byref(h)))
        self._h = h
        src = ctypes.cast(V.data_ptr(), _dp) if dev else ptr(V)
        check(lib().nmfg_matrix_upload(h, src, self.shape[1]))

    def mul(self, B, trans=False):
        """V @ B (trans=False, B p x k) or V.T @ B (trans=True, B n x k)."""
        B = f64c(B)
        n, p = self.shape
        k = B.shape[1]
        if B.shape[0] != (n if trans else p):
            raise ValueError(f"B.shape {B.shape} does not match V.shape {self.shape}")
        C = np.empty((p if trans else n, k))
        check(lib().nmfg_matrix_mul(self._h, 1 if trans else 0, ptr(B), k, ptr(C)))
        return C

    def range_finder(self, Q0, n_iter, flip):
        """randomized_range_finder's LU power iterations on the device
        (nmfg_matrix_range_finder): A = V^T when flip else V, Q0 the
        b x width sketch; returns A Q (a x width) for the final QR."""
        Q0 = f64c(Q0)
        n, p = self.shape
        a, b = (p, n) if flip else (n, p)
        if Q0.ndim != 2 or Q0.shape[0] != b:
            raise ValueError(f"Q0.shape {Q0.shape} does not match V.shape {self.shape}")
        width = Q0.shape[1]
        Y = np.empty((a, width))
        check(lib().nmfg_matrix_range_finder(self._h, 1 if flip else 0, ptr(Q0), width,
                                             int(n_iter), ptr(Y)))
        return Y

    def close(self):
        if getattr(self, "_h", None):
            # a matrix of a context already destroyed is left alone (the
            # library would touch the freed context)
            if self.ctx is None or getattr(self.ctx, "handle", None) is not None:
                lib().nmfg_matrix_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceSolve:
    """One solve on a resident V (nmfg_solve): W0/H0 stay on the device and
    every launch restarts from them (a captured HIP graph for max_iter).
    With ``batch`` > 1, that many independent solves of the same rank
    advance together (nmfg_solve_create_batched); factors then carry a
    leading batch axis."""

    def __init__(self, matrix: DeviceMatrix, rank: int, params: Params, batch: int = 1):
        self.matrix = matrix
        self.rank = int(rank)
        self.batch = int(batch)
        h = _vp()
        if self.batch == 1:
            check(lib().nmfg_solve_create(matrix._h, self.rank, ctypes.byref(params),
                                          ctypes.byref(h)))
        else:
            check(lib().nmfg_solve_create_batched(matrix._h, self.rank, ctypes.byref(params),
                                                  self.batch, ctypes.byref(h)))
        self._h = h
        self.max_iter = params.max_iter

    def _shapes(self):
        n, p = self.matrix.shape
        lead = () if self.batch == 1 else (self.batch,)
        return lead + (n, self.rank), lead + (self.rank, p)

    def set_factors(self, W0, H0):
        W0, H0 = f64c(W0), f64c(H0)
        ws, hs = self._shapes()
        if W0.shape != ws or H0.shape != hs:
            raise ValueError(f"W0/H0 shapes {W0.shape}/{H0.shape} do not match {ws}/{hs}")
        check(lib().nmfg_solve_set_factors(self._h, ptr(W0), ptr(H0)))

    def set_weight(self, S):
        """Weight matrix of an lsNMF solve (n x p, shared by a batch)."""
        S = f64c(S)
        if S.shape != self.matrix.shape:
            raise ValueError(f"weight shape {S.shape} != {self.matrix.shape}")
        check(lib().nmfg_solve_set_weight(self._h, ptr(S), S.shape[1]))

    def set_offset(self, off):
        """Initial offset of an offset solve (n, or batch x n); the default is
        rowMeans(V)."""
        n = self.matrix.shape[0]
        off = f64c(off)
        want = (n,) if self.batch == 1 else (self.batch, n)
        if off.shape != want:
            raise ValueError(f"offset shape {off.shape} != {want}")
        check(lib().nmfg_solve_set_offset(self._h, ptr(off)))

    def offset(self):
        """Current offset of an offset solve (after fetch)."""
        n = self.matrix.shape[0]
        out = np.empty((n,) if self.batch == 1 else (self.batch, n))
        check(lib().nmfg_solve_get_offset(self._h, ptr(out)))
        return out

    def launch(self):
        check(lib().nmfg_solve_launch(self._h))

    def fetch(self):
        ws, hs = self._shapes()
        W = np.empty(ws)
        H = np.empty(hs)
        it = _i64(0)
        nd = _i64(0)
        cap = int(self.max_iter) + 1
        dev = np.empty(max(cap, 1))
        check(lib().nmfg_solve_fetch(self._h, ptr(W), ptr(H), ctypes.byref(it), ptr(dev), cap,
                                     ctypes.byref(nd)))
        return W, H, int(it.value), dev[: nd.value].copy()

    def profile(self, evict_bytes=0):
        """(column-pass ms, launches, row-pass ms, launches) of max_iter
        back-to-back launches each; with evict_bytes > 0 a read of that many
        bytes precedes every pass (cold caches) and is not counted."""
        cm, rm = ctypes.c_double(0), ctypes.c_double(0)
        cn, rn = _i64(0), _i64(0)
        if evict_bytes:
            check(lib().nmfg_solve_profile_evicted(self._h, int(evict_bytes), ctypes.byref(cm),
                                                   ctypes.byref(cn), ctypes.byref(rm),
                                                   ctypes.byref(rn)))
        else:
            check(lib().nmfg_solve_profile(self._h, ctypes.byref(cm), ctypes.byref(cn),
                                           ctypes.byref(rm), ctypes.byref(rn)))
        return cm.value, int(cn.value), rm.value, int(rn.value)

    def close(self):
        if getattr(self, "_h", None):
            m = self.matrix
            # a solve whose matrix or context is already destroyed is left
            # alone (the library would touch freed objects)
            if getattr(m, "_h", None) is not None and (
                    m.ctx is None or getattr(m.ctx, "handle", None) is not None):
                lib().nmfg_solve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceSweep:
    """Solves of uneven ranks on one resident V as one graph replay
    (nmfg_sweep): ``ranks`` lists every member's rank (e.g. ranks x
    restarts of a consensus rank sweep); ``tune`` autotunes the plan."""

    def __init__(self, matrix: DeviceMatrix, ranks, params: Params, tune: bool = True):
        self.matrix = matrix
        self.ranks = [int(r) for r in ranks]
        arr = (_i32 * len(self.ranks))(*self.ranks)
        h = _vp()
        check(lib().nmfg_sweep_create(matrix._h, arr, len(self.ranks), ctypes.byref(params),
                                      1 if tune else 0, ctypes.byref(h)))
        self._h = h

    def _ptrs(self, mats):
        arr = (_dp * len(mats))()
        for i, a in enumerate(mats):
            arr[i] = ptr(a)
        return arr

    def set_factors(self, W0s, H0s):
        n, p = self.matrix.shape
        W0s = [f64c(w) for w in W0s]
        H0s = [f64c(h) for h in H0s]
        for r, w, h in zip(self.ranks, W0s, H0s):
            if w.shape != (n, r) or h.shape != (r, p):
                raise ValueError(f"member factors {w.shape}/{h.shape} do not match rank {r}")
        self._keep = (W0s, H0s)
        check(lib().nmfg_sweep_set_factors(self._h, self._ptrs(W0s), self._ptrs(H0s)))

    def launch(self):
        check(lib().nmfg_sweep_launch(self._h))

    def fetch(self):
        n, p = self.matrix.shape
        Ws = [np.empty((n, r)) for r in self.ranks]
        Hs = [np.empty((r, p)) for r in self.ranks]
        check(lib().nmfg_sweep_fetch(self._h, self._ptrs(Ws), self._ptrs(Hs)))
        return Ws, Hs

    def plan(self):
        import json

        buf = ctypes.create_string_buffer(8192)
        ms = ctypes.c_double(0)
        check(lib().nmfg_sweep_plan(self._h, buf, 8192, ctypes.byref(ms)))
        out = json.loads(buf.value.decode())
        out["tuned_ms"] = ms.value
        return out

    def close(self):
        if getattr(self, "_h", None):
            m = self.matrix
            if getattr(m, "_h", None) is not None and (
                    m.ctx is None or getattr(m.ctx, "handle", None) is not None):
                lib().nmfg_sweep_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StepGroup:
    """Concurrent solves advanced together (nmfg_group) on ``branches``
    streams (1..8): each member stays on one branch; a step starts a round
    on every idle branch (10-iteration chunks of its members, one stop-flag
    read-back for the branch) and returns once some round has ended.
    ``add`` may be called from any thread; a solve belongs to the group from
    ``add`` until ``step`` returns it, and then holds the results its own
    launch would give."""

    GROUP_MAX = 64

    def __init__(self, ctx: "Context | None" = None, branches: int = 4):
        self.ctx = ctx
        h = _vp()
        check(lib().nmfg_group_create(ctx.handle if ctx is not None else context(),
                                      int(branches), ctypes.byref(h)))
        self._h = h
        self._members = {}
        self._lock = threading.Lock()

    def add(self, solve: DeviceSolve):
        key = solve._h.value
        with self._lock:
            fresh = key not in self._members
            self._members[key] = solve
        try:
            check(lib().nmfg_group_add(self._h, solve._h))
        except Exception:
            if fresh:
                with self._lock:
                    self._members.pop(key, None)
            raise

    def size(self) -> int:
        n = _i32(0)
        check(lib().nmfg_group_size(self._h, ctypes.byref(n)))
        return int(n.value)

    def step(self):
        """One step; returns the DeviceSolves that finished in it."""
        out = (_vp * self.GROUP_MAX)()
        nf = _i32(0)
        check(lib().nmfg_group_step(self._h, out, self.GROUP_MAX, ctypes.byref(nf)))
        with self._lock:
            return [self._members.pop(out[i]) for i in range(nf.value)]

    def stats(self):
        steps, ms = _i64(0), _i64(0)
        mx = _i32(0)
        check(lib().nmfg_group_stats(self._h, ctypes.byref(steps), ctypes.byref(ms), ctypes.byref(mx)))
        return {"steps": int(steps.value), "member_steps": int(ms.value),
                "max_members": int(mx.value)}

    def close(self):
        if getattr(self, "_h", None):
            lib().nmfg_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def consensus_matrix(labels):
    """C[i][j] = fraction of the rows of ``labels`` (runs x p cluster labels)
    in which items i and j share a label, computed on the GPU."""
    L = np.ascontiguousarray(np.asarray(labels, dtype=np.int32))
    if L.ndim != 2:
        raise ValueError("labels must be runs x p")
    runs, p = L.shape
    C = np.empty((p, p))
    check(lib().nmfg_consensus(context(), L.ctypes.data_as(ctypes.POINTER(_i32)), runs, p,
                               ptr(C)))
    return C


def lu_permute_l(X):
    """P L of the LU with partial pivoting of X (m x w, m >= w) on the GPU:
    scipy.linalg.lu(X, permute_l=True)[0] (nmfg_lu_permute_l)."""
    X = f64c(X)
    if X.ndim != 2:
        raise ValueError("X must be 2-D")
    m, w = X.shape
    PL = np.empty((m, w))
    check(lib().nmfg_lu_permute_l(context(), ptr(X), m, w, ptr(PL)))
    return PL


def stream_handle() -> int:
    """The hipStream_t (as an int) every call on the context runs on."""
    return int(lib().nmfg_ctx_stream(context()) or 0)
