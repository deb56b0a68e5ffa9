"""Writes tests/golden/golden_vectors.json.

Inputs come from numpy's default_rng (PCG64; stable across numpy versions),
and expected outputs from the independent pure-Python scalar restatement
``oracle/pyref.py``.  Every float is stored as ``float.hex`` so the fixture
is exact.  Shapes are tiny so the pure-Python loops finish in seconds.

The ``ci_smoke`` case reuses the reference's CI smoke-test inputs
(.github/workflows/ci.yml:58-63: V = |N(0,1)| 80x30 from default_rng(0),
random_init(V, 4, seed=1), 20 iterations).

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
import pyref  # noqa: E402

OUT = Path(__file__).resolve().parent / "golden_vectors.json"


def hexm(a):
    a = np.asarray(a, dtype=np.float64)
    return {"shape": list(a.shape), "hex": [float(x).hex() for x in a.ravel()]}


def lists(a):
    return [list(map(float, row)) for row in np.asarray(a, dtype=np.float64)]


def random_init(V, rank, seed):
    # random_init (python/nmf_rs/__init__.py:102-126)
    n, p = V.shape
    rng = np.random.default_rng(seed)
    hi = float(np.max(V))
    return rng.uniform(0.0, hi, size=(n, rank)), rng.uniform(0.0, hi, size=(rank, p))


def case_inputs(seed, n, p, r, zero_row=None, zero_col=None):
    rng = np.random.default_rng(seed)
    Ws = rng.uniform(0.1, 1.5, (n, r))
    Hs = rng.uniform(0.1, 1.5, (r, p))
    V = np.zeros((n, p))
    for k in range(r):  # fixed-order sum: bit-portable, no BLAS
        V = V + Ws[:, k:k + 1] * Hs[k:k + 1, :]
    V = V + rng.uniform(0.0, 0.05, (n, p))
    if zero_row is not None:
        V[zero_row, :] = 0.0
    if zero_col is not None:
        V[:, zero_col] = 0.0
    W0 = rng.uniform(0.0, 1.0, (n, r))
    H0 = rng.uniform(0.0, 1.0, (r, p))
    return V, W0, H0


def add_lsnmf(c, Vl, Wl, Hl, iters, seed):
    # a general positive weight and a 0/1 mask (cv_rank's held-out pattern)
    g = np.random.default_rng(seed)
    n, p = len(Vl), len(Vl[0])
    Sw = g.uniform(0.2, 2.0, (n, p))
    Sm = (g.random((n, p)) > 0.2).astype(np.float64)
    for name, S in (("lsnmf", Sw), ("lsnmf_mask", Sm)):
        W, H = pyref.run("lsnmf", Vl, Wl, Hl, iters, eps=1e-9, weight=lists(S))
        c["runs"][name] = {"algo": "lsnmf", "weight": hexm(S), "iters": iters, "W": hexm(W),
                           "H": hexm(H)}


def add_snmf(c, Vl, Wl, Hl, iters):
    # Kim-Park snmf/r and snmf/l with the Python defaults (sparsity 0.01,
    # smoothness -1 -> eta = max V) and FCNNLS eps 1e-9
    for name in ("snmf/r", "snmf/l"):
        W, H = pyref.run(name, Vl, Wl, Hl, iters, eps=1e-9, sparsity=0.01, smoothness=-1.0)
        c["runs"][name] = {"iters": iters, "W": hexm(W), "H": hexm(H)}


def add_offset_nsnmf(c, Vl, Wl, Hl, iters):
    # offset from rowMeans(V) (the reference default); nsNMF at theta 0.5
    # (the reference default) and 0.3
    W, H, off = pyref.run("offset", Vl, Wl, Hl, iters, eps=1e-9)
    c["runs"]["offset"] = {"iters": iters, "W": hexm(W), "H": hexm(H), "off": hexm(off)}
    for name, theta in (("nsnmf", 0.5), ("nsnmf_theta03", 0.3)):
        W, H = pyref.run("nsnmf", Vl, Wl, Hl, iters, theta=theta)
        c["runs"][name] = {"algo": "nsnmf", "theta": theta, "iters": iters, "W": hexm(W),
                           "H": hexm(H)}


def main():
    cases = []
    specs = [
        ("small_a", 11, 12, 9, 3, None, None),
        ("small_zero_row_col", 12, 10, 7, 2, 3, 4),
        ("rank1", 13, 6, 5, 1, None, None),
    ]
    for name, seed, n, p, r, zr, zc in specs:
        V, W0, H0 = case_inputs(seed, n, p, r, zr, zc)
        Vl, Wl, Hl = lists(V), lists(W0), lists(H0)
        c = {"name": name, "V": hexm(V), "W0": hexm(W0), "H0": hexm(H0), "runs": {}, "steps": {}}
        # single steps (py_euclidean_update_{h,w}, py_divergence_update_{h,w})
        c["steps"]["lee_h"] = hexm(pyref.lee_update_h(Vl, Wl, Hl, 1e-9))
        c["steps"]["lee_w"] = hexm(pyref.lee_update_w(Vl, Wl, Hl, 1e-9))
        c["steps"]["brunet_h"] = hexm(pyref.brunet_update_h(Vl, Wl, Hl))
        c["steps"]["brunet_w"] = hexm(pyref.brunet_update_w(Vl, Wl, Hl))
        c["steps"]["euclid"] = float(pyref.euclidean_distance(Vl, Wl, Hl)).hex()
        c["steps"]["kl"] = float(pyref.kl_divergence(Vl, Wl, Hl)).hex()
        for algo, iters in (("lee", 12), ("brunet", 12), ("hals", 6), ("ehals", 6), ("dnmf", 6)):
            W, H = pyref.run(algo, Vl, Wl, Hl, iters, eps=1e-9, sparsity=0.01)
            c["runs"][algo] = {"iters": iters, "W": hexm(W), "H": hexm(H)}
        add_offset_nsnmf(c, Vl, Wl, Hl, 8)
        add_lsnmf(c, Vl, Wl, Hl, 8, seed + 100)
        add_snmf(c, Vl, Wl, Hl, 5)
        cases.append(c)

    # The reference CI smoke inputs (ci.yml:58-63), 20 iterations.
    rng = np.random.default_rng(0)
    V = np.abs(rng.normal(size=(80, 30)))
    W0, H0 = random_init(V, 4, seed=1)
    Vl, Wl, Hl = lists(V), lists(W0), lists(H0)
    c = {"name": "ci_smoke", "V": hexm(V), "W0": hexm(W0), "H0": hexm(H0), "runs": {}, "steps": {}}
    for algo in ("lee", "brunet", "hals"):
        W, H = pyref.run(algo, Vl, Wl, Hl, 20, eps=1e-9)
        c["runs"][algo] = {"iters": 20, "W": hexm(W), "H": hexm(H)}
    add_offset_nsnmf(c, Vl, Wl, Hl, 20)
    add_lsnmf(c, Vl, Wl, Hl, 20, 7)
    add_snmf(c, Vl, Wl, Hl, 10)
    cases.append(c)

    meta = {"generator": "tests/golden/make_golden.py", "numpy": np.__version__,
            "eps": 1e-9, "dnmf_sparsity": 0.01}
    OUT.write_text(json.dumps({"meta": meta, "cases": cases}, separators=(",", ":")))
    print("wrote", OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
