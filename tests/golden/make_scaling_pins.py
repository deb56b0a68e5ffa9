"""Writes tests/golden/scaling_medium.json: the C oracle on the reference's
multi-scale benchmark, tier "medium" (examples/build_scaling_notebook.py,
inputs restated in tests/_scaling.py), checked against the numbers that
notebook printed (examples/benchmark_scaling.ipynb cells 18-21):

* ehals from K = 3 random starts, 100 iterations: mean 0.5 ||V - WH||^2 =
  2.615893e+06 (7 significant digits), mean Hoyer sparsity of W 0.157,
  consensus dispersion 0.592 (3 decimals);
* hals from nndsvd_init(V, 15, fill='mean', seed=0), 25 iterations: Amari
  error against the generating W* 2.332754e-01 and Hoyer sparsity of W
  1.934819e-01 (7 significant digits).

The script asserts every figure before writing the fixture, whose per-run
losses (float.hex) the GPU test compares with.  About 12 minutes on 8 cores
(300 ehals iterations of the C oracle at 3000 x 50000, rank 15).

Run:  python tests/golden/make_scaling_pins.py [--threads N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(HERE.parent))
import oracle  # noqa: E402
import _scaling as S  # noqa: E402

OUT = HERE / "scaling_medium.json"


def sig7(x):
    return float(f"{x:.6e}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    M = S.MEDIUM
    t0 = time.perf_counter()
    V, W_true, _ = S.make_synthetic(M["n"], M["p"], rank=M["rank"], seed=M["seed"])
    sub = S.consensus_subset(M["p"])
    rank = M["rank"]
    runs = []
    cons = np.zeros((len(sub), len(sub)))
    for run in range(M["runs"]):
        W0, H0 = S.random_start(V, rank, run)
        t = time.perf_counter()
        out = oracle.run("ehals", V, W0, H0, M["iters"], num_threads=args.threads)
        f = S.loss(V, out["W"], out["H"])
        runs.append({"loss": f.hex(), "sparsity_w": S.factor_sparsity_w(out["W"])})
        cons += S.connectivity(out["H"][:, sub])
        print(f"ehals run {run}: loss {f:.9e} ({time.perf_counter() - t:.0f} s)", flush=True)
    mean_loss = float(np.mean([float.fromhex(r["loss"]) for r in runs]))
    disp = S.dispersion(cons / M["runs"])
    spars = float(np.mean([r["sparsity_w"] for r in runs]))
    W0n, H0n = oracle.nndsvd_ref(V, rank, fill="mean", seed=0)
    out = oracle.run("hals", V, W0n, H0n, M["nndsvd_iters"], num_threads=args.threads)
    amari = S.amari_error(out["W"], W_true)
    hspars = S.factor_sparsity_w(out["W"])
    hloss = S.loss(V, out["W"], out["H"])
    print(f"ehals mean loss {mean_loss:.9e} (notebook {M['ehals_mean_loss']:.6e}), "
          f"dispersion {disp:.6f}, sparsity W {spars:.6f}; hals+NNDSVD Amari {amari:.9e} "
          f"(notebook {M['hals_nndsvd_amari']:.6e}), sparsity W {hspars:.9e}", flush=True)
    assert sig7(mean_loss) == M["ehals_mean_loss"], mean_loss
    assert round(disp, 3) == M["dispersion_3dp"]["ehals"], disp
    assert round(spars, 3) == M["sparsity_w_3dp"]["ehals"], spars
    assert sig7(amari) == M["hals_nndsvd_amari"], amari
    assert sig7(hspars) == M["hals_nndsvd_sparsity_w"], hspars
    OUT.write_text(json.dumps({
        "meta": {"generator": "tests/golden/make_scaling_pins.py", "numpy": np.__version__,
                 "tier": "medium", "seconds": round(time.perf_counter() - t0)},
        "ehals_runs": runs, "ehals_mean_loss": mean_loss, "ehals_dispersion": disp,
        "ehals_sparsity_w": spars, "hals_nndsvd_amari": amari, "hals_nndsvd_sparsity_w": hspars,
        "hals_nndsvd_loss": hloss.hex(),
    }, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
