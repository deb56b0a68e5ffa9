"""nmfg_run_device_noalloc: the no-allocation variant of nmfg_run, through
the C ABI, on device buffers (torch tensors), against nmfg_run on the same
inputs -- two calls in a row, the second on a smaller problem (the scratch
array keeps the first call's contents), for every method it runs, with the
fixed-iteration and the objective-based stops."""
import ctypes

import numpy as np
import pytest

from _util import rel
from nmf_rs import _lib, _rust

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _problem(n, p, r, seed):
    g = np.random.default_rng(seed)
    V = g.uniform(0, 1, (n, r)) @ g.uniform(0, 1, (r, p)) + g.uniform(0, 0.05, (n, p))
    return V, g.uniform(0.1, 1, (n, r)), g.uniform(0.1, 1, (r, p))


def noalloc_run(algo, V, W0, H0, iters, ldh_pad=3, **kw):
    import torch

    n, p = V.shape
    r = W0.shape[1]
    ldv = (p + 127) // 128 * 128
    dV = torch.zeros((n, ldv), dtype=torch.float64, device="cuda")
    dV[:, :p] = torch.from_numpy(V)
    dW = torch.from_numpy(np.ascontiguousarray(W0)).cuda()
    dH = torch.zeros((r, p + ldh_pad), dtype=torch.float64, device="cuda")
    dH[:, :p] = torch.from_numpy(H0)
    torch.cuda.synchronize()
    prm = _lib.make_params(algo, iters, **kw)
    it, nd = ctypes.c_int64(0), ctypes.c_int64(0)
    cap = iters + 1
    dev = np.zeros(cap)
    _lib.check(_lib.lib().nmfg_run_device_noalloc(
        _lib.context(), ctypes.byref(prm), dV.data_ptr(), n, p, ldv, dW.data_ptr(),
        dH.data_ptr(), p + ldh_pad, r, ctypes.byref(it), _lib.ptr(dev), cap, ctypes.byref(nd)))
    return {"W": dW.cpu().numpy(), "H": dH[:, :p].cpu().numpy(), "n_iter": it.value,
            "deviances": dev[: nd.value]}


CASES = [("lee", {}), ("brunet", {}), ("hals", {}), ("ehals", {}),
         ("dnmf", {"sparsity_h": 0.01, "sparsity_w": 0.01}), ("nsnmf", {"theta": 0.5})]


@pytest.mark.parametrize("algo,kw", CASES)
def test_noalloc_equals_nmfg_run_twice(algo, kw):
    for (n, p, r, seed) in ((300, 1000, 6, 1), (120, 300, 4, 2)):
        V, W0, H0 = _problem(n, p, r, seed)
        print(f"noalloc algo={algo} shape={n}x{p} r={r}", flush=True)
        out = noalloc_run(algo, V, W0, H0, 15, **kw)
        ref = _rust.py_nmf_run(algo, V, W0, H0, 15, **kw)
        print(f"  rel W {rel(out['W'], ref['W']):.2e} H {rel(out['H'], ref['H']):.2e} "
              f"bitwise {np.array_equal(out['W'], ref['W']) and np.array_equal(out['H'], ref['H'])}",
              flush=True)
        assert out["n_iter"] == 15
        assert rel(out["W"], ref["W"]) <= TOL and rel(out["H"], ref["H"]) <= TOL


@pytest.mark.parametrize("algo,stop", [("lee", "rel_change"), ("brunet", "stationary"),
                                       ("hals", "rel_change")])
def test_noalloc_objective_stops(algo, stop):
    kw = {"stop": stop}
    if stop == "rel_change":
        kw.update(rel_tol=1e-5, min_iter=5)
    else:
        kw.update(stationary_th=1e-4, check_interval=5, check_niter=3)
    for (n, p, r, seed) in ((200, 900, 5, 3), (90, 260, 3, 4)):
        V, W0, H0 = _problem(n, p, r, seed)
        print(f"noalloc stop={stop} algo={algo} shape={n}x{p}", flush=True)
        out = noalloc_run(algo, V, W0, H0, 80, **kw)
        ref = _rust.py_nmf_run(algo, V, W0, H0, 80, **kw)
        assert out["n_iter"] == ref["n_iter"]
        d, dr = np.asarray(out["deviances"]), np.asarray(ref["deviances"])
        assert d.shape == dr.shape and np.max(np.abs(d - dr) / np.abs(dr)) <= TOL
        assert rel(out["W"], ref["W"]) <= TOL and rel(out["H"], ref["H"]) <= TOL


def test_noalloc_rejects_what_it_cannot_hold():
    import torch

    # rank 32 over 131072 columns: H0 and H alone need 64 MiB of scratch
    p, r = 131072, 32
    dV = torch.zeros((1, p), dtype=torch.float64, device="cuda")
    dW = torch.ones((1, r), dtype=torch.float64, device="cuda")
    dH = torch.ones((r, p), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    prm = _lib.make_params("lee", 2)
    it, nd = ctypes.c_int64(0), ctypes.c_int64(0)
    L = _lib.lib()
    rc = L.nmfg_run_device_noalloc(_lib.context(), ctypes.byref(prm), dV.data_ptr(), 1, p, p,
                                   dW.data_ptr(), dH.data_ptr(), p, r, ctypes.byref(it), None, 0,
                                   ctypes.byref(nd))
    assert rc == _lib.NMFG_ENOMEM, rc
    assert b"scratch" in L.nmfg_last_error()
    # a row stride that is not a multiple of 128
    rc = L.nmfg_run_device_noalloc(_lib.context(), ctypes.byref(prm), dV.data_ptr(), 1, 100, 100,
                                   dW.data_ptr(), dH.data_ptr(), p, 2, ctypes.byref(it), None, 0,
                                   ctypes.byref(nd))
    assert rc == _lib.NMFG_EINVAL, rc
    # a method it does not run
    prm = _lib.make_params("offset", 2)
    rc = L.nmfg_run_device_noalloc(_lib.context(), ctypes.byref(prm), dV.data_ptr(), 1, 128, 128,
                                   dW.data_ptr(), dH.data_ptr(), p, 2, ctypes.byref(it), None, 0,
                                   ctypes.byref(nd))
    assert rc == _lib.NMFG_EUNSUPPORTED, rc
