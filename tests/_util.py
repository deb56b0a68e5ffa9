"""Shared helpers for the test suite."""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def unhex(d):
    return np.array([float.fromhex(x) for x in d["hex"]]).reshape(d["shape"])


def golden_cases():
    return json.loads((GOLDEN / "golden_vectors.json").read_text())["cases"]


def rel(a, b):
    """max|a - b| / max|b| (scale-free parity metric)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def quickstart_inputs():
    """Inputs of the reference quickstart notebook
    (examples/build_notebooks.py:65-69, 81)."""
    rng = np.random.default_rng(0)
    n, p, rank = 200, 60, 5
    W_true = rng.uniform(0.1, 1.5, (n, rank))
    H_true = rng.uniform(0.1, 1.5, (rank, p))
    V = W_true @ H_true + rng.uniform(0, 0.05, (n, p))
    g = np.random.default_rng(42)
    hi = float(np.max(V))
    W0 = g.uniform(0.0, hi, size=(n, rank))
    H0 = g.uniform(0.0, hi, size=(rank, p))
    return V, W0, H0
