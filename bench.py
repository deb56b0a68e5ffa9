"""Benchmark: FP64 lee on the single-cell-shaped 2000 genes x 7750 cells
matrix, rank 10, NNDSVD (fill='mean') start, 100 iterations per solve.

A step is one full solve (100 lee iterations) from the same start, replayed
as one captured HIP graph with V, W0 and H0 resident in HBM.  Prints one JSON
line: iterations/s (whole job), the roofline of the dominant kernel measured
with HIP events (back to back, V cached in the MALL, and with V evicted
before every pass), lee at 2000 x 500000 rank 20 in the FP64 and FP32 forms
(V 32 times the MALL), the CPU oracle timed on this host, the batched rank sweep
(140 brunet solves with consensus statistics, solves/s), the served
stream (Poisson arrivals, p50 / p99 latency at the 1e-6 relative stop) and
the other update rules on the same matrix (GPU and CPU iterations/s).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
import argparse
import json

import numpy as np
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "rust-nmf_amd"))

METRIC = "iters/s + %HBM peak, lee k=10 2k×7.75k; batched solves/s + p99 at 1e-6 rel"
HBM_PEAK_GBS = 8000.0
N, P, R, ITERS = 2000, 7750, 10, 100


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(n, p, r, sv=8, sf=8):
    """Per iteration: one V pass per half-iteration plus factor traffic
    (H read twice and written in the column pass, W likewise in the row pass)."""
    col = n * p * sv + sf * (n * r + 2 * r * p)
    row = n * p * sv + sf * (r * p + 2 * n * r)
    return col, row


def traffic_from_profiles(key, name="lee_fp64_2000x7750_r10_traffic.json"):
    """HBM bytes per launch of a pass kernel from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel scaled), if present."""
    f = ROOT / "profiles" / name
    if f.exists():
        try:
            return json.loads(f.read_text()).get(key)
        except (OSError, ValueError):
            return None
    return None


def copy_calibration():
    """A plain device copy of a known size (the 2000x7750 FP64 matrix), so
    that a rocprofv3 --pmc run of this script reads FETCH_SIZE / WRITE_SIZE
    of a kernel whose true HBM traffic is known."""
    import torch

    x = torch.rand(N * P, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"copy calibration: 3 copies of {x.numel() * 8} bytes")


def precision_forms(n=2000, p=500000, rank=20, iters=25):
    """The 2000 x 500000 single-cell workload (generated on the device, seed
    3; 8 GB in FP64) in each storage form of V: lee and dnmf, 25 iterations
    from the same NNDSVD start.  Reports the iteration rate, the objective
    0.5||V - WH||^2 (FP64, against the unrounded V) and its relative
    difference to the FP64 form, and max|dW|/max|W|, max|dH|/max|H|."""
    import torch

    import nmf_rs
    from nmf_rs import _lib, synthetic

    t = time.perf_counter()
    Vt = synthetic.single_cell_device(n, p, k=10, seed=3, capture=synthetic.SPARSE_CAPTURE)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t
    t = time.perf_counter()
    W0, H0 = nmf_rs.nndsvd_init(Vt, rank, fill="mean", seed=0)
    nndsvd_s = time.perf_counter() - t
    log(f"forms: generated {n}x{p} in {gen_s:.1f} s, NNDSVD {nndsvd_s:.1f} s")

    def objective(W, H):
        Wd = torch.from_numpy(W).cuda()
        Hd = torch.from_numpy(H).cuda()
        acc = 0.0
        for c0 in range(0, p, 65536):
            R = Vt[:, c0:c0 + 65536] - Wd @ Hd[:, c0:c0 + 65536]
            acc += float((R * R).sum())
        return 0.5 * acc

    ext = torch.cuda.ExternalStream(_lib.stream_handle())
    out = {"workload": f"{n} x {p} single-cell (device generator, seed 3), rank {rank}, "
                       f"NNDSVD mean init, {iters} iterations", "generate_s": round(gen_s, 2),
           "nndsvd_s": round(nndsvd_s, 2)}
    for algo in ("lee", "dnmf"):
        kw = {"sparsity": 0.01} if algo == "dnmf" else {}
        base = None
        for form in ("fp64", "fp32-store", "bf16x2-store", "fp32", "bf16x2"):
            m = _lib.DeviceMatrix(Vt, precision=form)
            prm = _lib.make_params(algo, iters, sparsity_h=kw.get("sparsity", 0.0),
                                   sparsity_w=kw.get("sparsity", 0.0), precision=form)
            s = _lib.DeviceSolve(m, rank, prm)
            s.set_factors(W0, H0)
            s.launch()
            ext.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(ext)
            s.launch()
            e1.record(ext)
            ext.synchronize()
            ms = e0.elapsed_time(e1)
            W, H, _, _ = s.fetch()
            s.close()
            m.close()
            f = objective(W, H)
            sv = 8 if form == "fp64" else 4
            b_iter = 2 * n * p * sv + 8 * (3 * rank * p + 3 * n * rank)
            rec = {"iters_s": round(iters / (ms / 1e3), 1),
                   "hbm_frac_algorithmic": round(b_iter * iters / (ms / 1e3) / (HBM_PEAK_GBS * 1e9), 4),
                   "objective": f}
            if base is None:
                base = (f, W, H)
            else:
                rec["rel_objective_vs_fp64"] = abs(f - base[0]) / base[0]
                rec["rel_W_vs_fp64"] = float(np.max(np.abs(W - base[1])) / np.max(np.abs(base[1])))
                rec["rel_H_vs_fp64"] = float(np.max(np.abs(H - base[2])) / np.max(np.abs(base[2])))
            log(f"forms: {algo} {form}: {json.dumps(rec)}")
            out[f"{algo}_{form}"] = rec
    del Vt
    torch.cuda.empty_cache()
    return out


def large_bench(n=2000, p=500000, rank=20, iters=25):
    """The 2000 x 500000 single-cell workload at rank 20 (generated on the
    device, seed 3; 8 GB in FP64, 32 times the 256 MB MALL, so every pass
    reads V from HBM): lee in the FP64, FP32 and BF16 hi + lo forms from one
    NNDSVD start.
    Each form: a warm-up solve, then one 25-iteration solve timed with HIP
    events on the library's stream; pass averages from nmfg_solve_profile;
    a roofline for the slower pass (bytes of V in the form's storage plus
    the factor traffic at 8 B per entry, per launch, over its average
    time)."""
    import torch

    import nmf_rs
    from nmf_rs import _lib, synthetic

    Vt = synthetic.single_cell_device(n, p, k=10, seed=3, capture=synthetic.SPARSE_CAPTURE)
    torch.cuda.synchronize()
    log(f"large: zero fraction {float((Vt == 0).double().mean()):.3f}")
    W0, H0 = nmf_rs.nndsvd_init(Vt, rank, fill="mean", seed=0)
    ext = torch.cuda.ExternalStream(_lib.stream_handle())

    def objective(W, H):
        Wd, Hd = torch.from_numpy(W).cuda(), torch.from_numpy(H).cuda()
        acc = 0.0
        for c0 in range(0, p, 65536):
            R = Vt[:, c0:c0 + 65536] - Wd @ Hd[:, c0:c0 + 65536]
            acc += float((R * R).sum())
        return 0.5 * acc

    out = {"workload": f"lee, {n} x {p} single-cell (device generator, seed 3), rank {rank}, "
                       f"NNDSVD mean init, {iters} iterations per solve, V resident in HBM"}
    f64 = None
    for form in ("fp64", "fp32", "bf16x2"):
        m = _lib.DeviceMatrix(Vt, precision=form)
        s = _lib.DeviceSolve(m, rank, _lib.make_params("lee", iters, precision=form))
        s.set_factors(W0, H0)
        s.launch()
        ext.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(ext)
        s.launch()
        e1.record(ext)
        ext.synchronize()
        ms = e0.elapsed_time(e1)
        W, H, _, _ = s.fetch()
        cm, cn, rm, rn = s.profile()
        s.close()
        m.close()
        sv = 8 if form == "fp64" else 4
        col_b, row_b = algorithmic_bytes(n, p, rank, sv=sv)
        col_us, row_us = cm / cn * 1e3, rm / rn * 1e3
        if col_us >= row_us:
            kname, us, nb, tk = "column pass (W^T V)", col_us, col_b, f"{form}_col_pass_bytes_per_launch"
        else:
            kname, us, nb, tk = "row pass (V H^T)", row_us, row_b, f"{form}_row_pass_bytes_per_launch"
        ach = nb / (us * 1e-6) / 1e9
        iters_s = iters / (ms / 1e3)
        f = objective(W, H)
        rec = {"iters_s": round(iters_s, 1), "ms_per_solve": round(ms, 3),
               "hbm_frac_algorithmic": round((col_b + row_b) * iters_s / (HBM_PEAK_GBS * 1e9), 4),
               "col_avg_us": round(col_us, 1), "row_avg_us": round(row_us, 1),
               "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(ach, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(ach / HBM_PEAK_GBS, 4),
                            "traffic": traffic_from_profiles(tk, "lee_2000x500000_r20_traffic.json"),
                            "avg_us": round(us, 1)},
               "objective": f}
        if f64 is None:
            f64 = (f, iters_s, W, H)
        else:
            rec["speedup_vs_fp64"] = round(iters_s / f64[1], 3)
            rec["rel_objective_vs_fp64"] = abs(f - f64[0]) / f64[0]
            rec["max_rel_dW_vs_fp64"] = float(np.max(np.abs(W - f64[2])) / np.max(np.abs(f64[2])))
            rec["max_rel_dH_vs_fp64"] = float(np.max(np.abs(H - f64[3])) / np.max(np.abs(f64[3])))
        log(f"large: lee {form}: {json.dumps(rec)}")
        out[form] = rec
    del Vt
    torch.cuda.empty_cache()
    return out


def kl_objective_device(V, W, H):
    """KL objective (lib.rs:1406-1445 terms) of host factors, in FP64 on the
    GPU: sum of v log(v / wh) - v + wh over v > 0, wh > 0, plus wh where v = 0."""
    import torch

    Vd = V if hasattr(V, "is_cuda") else torch.from_numpy(V).cuda()
    WH = torch.from_numpy(W).cuda() @ torch.from_numpy(H).cuda()
    pos = (Vd > 0) & (WH > 0)
    t = torch.where(pos, Vd * torch.log(torch.where(pos, Vd / WH, torch.ones_like(WH))) - Vd + WH,
                    torch.zeros_like(WH))
    t = t + torch.where(Vd > 0, torch.zeros_like(WH), WH)
    return float(t.sum())


def kl_objective_host(V, W, H):
    """The same KL objective as kl_objective_device, in FP64 on the host."""
    WH = W @ H
    pos = (V > 0) & (WH > 0)
    t = np.where(pos, V * np.log(np.where(pos, V, 1.0) / np.where(pos, WH, 1.0)) - V + WH, 0.0)
    return float(np.sum(t) + np.sum(np.where(V > 0, 0.0, WH)))


def rank_sweep_bench(baseline=True):
    """Rank sweep: the 1500 x 3000 low-rank + noise matrix (rank-5 uniform
    factors, |N(0, 0.01)|, seed 0), brunet 200 iterations, ranks 2..15 x 10
    restarts (W0, H0 ~ U(0, 1), seed 1000 r + restart): 140 solves.

    * FP32 form, batched: the whole sweep as one device sweep (members
      grouped by rank into batched solves, one captured graph replay, plan
      autotuned on the first call and cached); solves/s = 140 / wall time
      of the solves (factors up, the replay, results back); the consensus
      matrices (GPU) and cophenetic correlations (scipy) are timed apart.
    * FP64, unbatched: the plain form, one solve at a time on one resident
      matrix (set factors, one graph replay, fetch), as a caller looping
      over nmf() calls would run it on the GPU.
    * Accuracy of every FP32 member against its FP64 solve (which the
      parity tests pin to the oracle at 1e-12): relative KL objective, max
      relative |dW|, |dH|.
    * The same device sweep in the storage-only form (V in FP32, FP64
      arithmetic: "fp32_store"), with its solves/s and accuracy.
    baseline=False leaves out the FP64 solves, the storage-only sweep and the
    accuracy figures (for kernel traces of the sweep alone)."""
    from nmf_rs import _lib, consensus, synthetic

    V, _, _ = synthetic.lowrank_uniform(1500, 3000, 5, seed=0)
    ranks = list(range(2, 16))
    # first call: code objects, allocator and the plan's autotuning
    t = time.perf_counter()
    consensus.rank_sweep(V, ranks, restarts=10, method="brunet", max_iter=200,
                         precision="fp32", cophenetic=False)
    first_s = time.perf_counter() - t
    t = time.perf_counter()
    res = consensus.rank_sweep(V, ranks, restarts=10, method="brunet", max_iter=200,
                               precision="fp32")
    wall = time.perf_counter() - t
    tim = res["timing"]
    solves = 10 * len(ranks)
    if not baseline:
        out = {"workload": "brunet 1500x3000, ranks 2..15 x 10 restarts, 200 iterations (140 solves)",
               "solves": solves, "solves_per_s": round(solves / tim["solve_s"], 2),
               "solve_s": round(tim["solve_s"], 3), "plan": res["plan"], "wall_s": round(wall, 3)}
        log(f"rank sweep: {json.dumps(out)}")
        return out
    # FP64 unbatched baseline
    m = _lib.DeviceMatrix(V)
    f64 = {}
    try:
        prm = {r: _lib.make_params("brunet", 200) for r in ranks}
        s = _lib.DeviceSolve(m, 2, prm[2])  # warm-up
        s.set_factors(*[x[0] for x in consensus.uniform_restarts(1500, 3000, 2, 1)])
        s.launch()
        s.fetch()
        s.close()
        t = time.perf_counter()
        for r in ranks:
            W0, H0 = consensus.uniform_restarts(1500, 3000, r, 10)
            for b in range(10):
                s = _lib.DeviceSolve(m, r, prm[r])
                s.set_factors(W0[b], H0[b])
                s.launch()
                W, H, _, _ = s.fetch()
                s.close()
                f64[(r, b)] = (W, H)
        f64_s = time.perf_counter() - t
    finally:
        m.close()
    # the members' objectives on the host (numpy's BLAS): the products are
    # small, and this keeps the figures independent of the device libraries
    f64_obj = {k: kl_objective_host(V, W, H) for k, (W, H) in f64.items()}

    def accuracy(res):
        dobj, dw, dh = [], [], []
        for r in ranks:
            for b in range(10):
                W64, H64 = f64[(r, b)]
                W32, H32 = res[r]["W"][b], res[r]["H"][b]
                a, c = f64_obj[(r, b)], kl_objective_host(V, W32, H32)
                dobj.append(abs(c - a) / abs(a))
                dw.append(float(np.max(np.abs(W32 - W64)) / np.max(np.abs(W64))))
                dh.append(float(np.max(np.abs(H32 - H64)) / np.max(np.abs(H64))))
        return {"members": len(dobj), "max_rel_objective": max(dobj),
                "median_rel_objective": float(np.median(dobj)),
                "p99_rel_objective": float(np.percentile(dobj, 99)),
                "members_above_1e-6": int(np.sum(np.asarray(dobj) > 1e-6)),
                "max_rel_dW": max(dw), "max_rel_dH": max(dh)}

    # the same sweep with V stored in FP32 and FP64 arithmetic (the
    # storage-only form): fewer solves/s, the FP64 algorithm on float(V)
    consensus.rank_sweep(V, ranks, restarts=10, method="brunet", max_iter=200,
                         precision="fp32-store", cophenetic=False)
    res_st = consensus.rank_sweep(V, ranks, restarts=10, method="brunet", max_iter=200,
                                  precision="fp32-store", cophenetic=False)
    sweep_rate = solves / tim["solve_s"]
    base_rate = solves / f64_s
    store_rate = solves / res_st["timing"]["solve_s"]
    out = {"workload": "brunet 1500x3000, ranks 2..15 x 10 restarts, 200 iterations (140 solves)",
           "solves": solves,
           "solves_per_s": round(sweep_rate, 2), "form": "fp32, one device sweep (one graph replay)",
           "solve_s": round(tim["solve_s"], 3), "consensus_s": round(tim["consensus_s"], 3),
           "plan_s_first_call_with_tuning": round(first_s, 3), "plan": res["plan"],
           "wall_s": round(wall, 3),
           "fp64_unbatched_solves_per_s": round(base_rate, 2),
           "speedup_vs_fp64_unbatched": round(sweep_rate / base_rate, 2),
           "accuracy_vs_fp64": accuracy(res),
           "fp32_store": {"form": "V stored in FP32, FP64 arithmetic, one device sweep",
                          "solves_per_s": round(store_rate, 2),
                          "speedup_vs_fp64_unbatched": round(store_rate / base_rate, 2),
                          "plan": res_st["plan"], "accuracy_vs_fp64": accuracy(res_st)},
           "cophenetic": {r: round(res[r]["cophenetic"], 6) for r in ranks},
           "dispersion": {r: round(res[r]["dispersion"], 6) for r in ranks}}
    log(f"rank sweep: {json.dumps(out)}")
    return out


def serve_accuracy(reqs, res, mats, means, sample=4):
    """Served FP32-form results against FP64 solves of the same requests:
    the request of largest rank, of largest p, the first brunet and the first
    hals request; each FP64 solve runs exactly as many iterations as the
    served solve, from the served solve's own NNDSVD start (the form's
    arithmetic) and from NNDSVD on the FP64 V (end to end), and is compared
    with the served factors by the relative objective (Frobenius, or KL for
    brunet), in FP64 on the GPU."""
    import torch

    from nmf_rs import serve

    pick = [max(range(len(reqs)), key=lambda i: reqs[i].rank),
            max(range(len(reqs)), key=lambda i: reqs[i].p)]
    for m in ("brunet", "hals"):
        pick += [i for i, r in enumerate(reqs) if r.method == m][:1]
    pick = list(dict.fromkeys(pick))[:sample]
    rows = []
    for i in pick:
        r, x = reqs[i], res[i]
        if x.error:
            continue
        ref = serve.Server(workers=1, precision="fp64", max_iter=x.n_iter, min_iter=x.n_iter,
                           rel_tol=1e-6)
        try:
            W, H, it, _, _ = ref.solve(mats[i], r.rank, r.method, means[i], init=(x.W0, x.H0))
            We, He, _, _, _ = ref.solve(mats[i], r.rank, r.method, means[i])
        finally:
            ref.close()

        def obj(W, H):
            if r.method == "brunet":
                return kl_objective_device(mats[i], W, H)
            R = mats[i] - torch.from_numpy(W).cuda() @ torch.from_numpy(H).cuda()
            return 0.5 * float(torch.sum(R * R))

        fa, fb, fe = obj(W, H), obj(x.W, x.H), obj(We, He)
        rows.append({"method": r.method, "n": r.n, "p": r.p, "rank": r.rank, "n_iter": x.n_iter,
                     "fp64_n_iter": it, "rel_objective": float(abs(fb - fa) / abs(fa)),
                     "max_rel_dW": float(np.max(np.abs(x.W - W)) / np.max(np.abs(W))),
                     "max_rel_dH": float(np.max(np.abs(x.H - H)) / np.max(np.abs(H))),
                     "rel_objective_fp64_own_nndsvd": float(abs(fb - fe) / abs(fe))})
    return {"what": "served FP32 form vs FP64 from the same NNDSVD start at the same iteration "
                    "count: relative objective, max relative |dW|, |dH| (beside them: the "
                    "objective vs FP64 with NNDSVD on the FP64 V)",
            "max_rel_objective": max((q["rel_objective"] for q in rows), default=None),
            "max_rel_dW": max((q["max_rel_dW"] for q in rows), default=None),
            "max_rel_dH": max((q["max_rel_dH"] for q in rows), default=None),
            "requests": rows}


def serve_bench(count=48, workers=8, load=0.7):
    """Served stream: `count` requests with n ~ logU(500, 5000), p ~
    logU(1000, 1e5), rank ~ U{5..30}, method lee 50 % / hals 30 % / brunet
    20 %, each V from the low-rank uniform generator (W* H* + |N(0, 0.01)|,
    inner rank 5) generated on the device (seed per request).  Each request uploads V in FP32 storage, runs NNDSVD on it and
    solves for 25..100 iterations with the 1e-6 relative-change stop.
    Open-loop Poisson arrivals at `load` times the closed-loop capacity,
    measured on all the requests submitted at once.  Latency =
    completion - arrival.  Two modes, each at its own capacity: continuous
    batching (the workers prepare requests, one step group advances every
    solve in service together; the headline) and one solve per worker
    stream."""
    import torch

    from nmf_rs import serve, synthetic

    reqs = serve.request_mix(count, seed=5)
    mats, means = [], []
    for r in reqs:
        V = synthetic.lowrank_uniform_device(r.n, r.p, seed=r.seed % 100000, chunk=16384)
        mats.append(V)
        means.append(float(V.mean()))
    torch.cuda.synchronize()

    def run(cobatch):
        # budget: 90 % of the device memory free once the requests' matrices
        # are resident (hipMemGetInfo), requests admitted first in, first out
        srv = serve.Server(workers=workers, precision="fp32", max_iter=100, min_iter=25,
                           rel_tol=1e-6, budget_fraction=0.9, cobatch=cobatch, keep_init=True)
        try:
            cap = srv.capacity(reqs, mats, means)
            qps = load * cap
            serve.poisson_arrivals(reqs, qps, seed=6)
            res = srv.run_stream(reqs, mats, means)
            metrics = srv.metrics()
        finally:
            srv.close()
        span = max(x.end for x in res) - reqs[0].arrival
        lat = serve.latency_stats(reqs, res)
        its = [x.n_iter for x in res]
        rec = {"workers": workers, "qps": round(qps, 3), "target_load": load,
               "capacity_solves_per_s": round(cap, 3),
               "solves_per_s": round(count / span, 3),
               "latency_p50_s": round(lat["p50_s"], 4), "latency_p99_s": round(lat["p99_s"], 4),
               "latency_mean_s": round(lat["mean_s"], 4),
               "iterations_mean": round(float(np.mean(its)), 2),
               "iterations_min": int(min(its)), "iterations_max": int(max(its)),
               "nndsvd_s_mean": round(float(np.mean([x.timings["nndsvd_s"] for x in res])), 4),
               "solve_s_mean": round(float(np.mean([x.timings["solve_s"] for x in res])), 4),
               "slowest": [{"method": reqs[i].method, "n": reqs[i].n, "p": reqs[i].p,
                            "rank": reqs[i].rank, "n_iter": res[i].n_iter,
                            "service_s": round(res[i].end - res[i].start, 4),
                            **{k: round(v, 4) for k, v in res[i].timings.items()}}
                           for i in sorted(range(len(res)),
                                           key=lambda i: res[i].start - res[i].end)[:3]],
               "budget": metrics}
        return rec, res

    per_worker, _ = run(False)
    log(f"serve (one solve per worker stream): {json.dumps(per_worker)}")
    # the plain path: FP64, one request at a time (one worker), same requests
    plain = serve.Server(workers=1, precision="fp64", max_iter=100, min_iter=25, rel_tol=1e-6)
    try:
        plain_cap = plain.capacity(reqs, mats, means)
    finally:
        plain.close()
    out, res = run(True)
    out["fp64_one_at_a_time_solves_per_s"] = round(plain_cap, 3)
    out["capacity_over_fp64_one_at_a_time"] = round(out["capacity_solves_per_s"] / plain_cap, 3)
    out["mode"] = "continuous batching: one step group advances every solve in service"
    accuracy = serve_accuracy(reqs, res, mats, means)
    out = {"workload": f"{count} requests, n logU(500,5000) x p logU(1e3,1e5), rank U{{5..30}}, "
                       "lee/hals/brunet 50/30/20 %, V = W*H* + |N(0, 0.01)| (inner rank 5), "
                       "FP32-stored V, NNDSVD, 25..100 iterations, "
                       "stop at 1e-6 relative objective change", **out,
           "accuracy_vs_fp64": accuracy, "per_worker_streams": per_worker}
    log(f"serve: {json.dumps(out)}")
    del mats
    torch.cuda.empty_cache()
    return out


def variants_bench(V, W0, H0, steps=5, iters=100, cpu_iters=3):
    """The other update rules on the same resident 2000 x 7750 matrix, rank
    10, same start: GPU iterations/s over `steps` graph-replayed solves of
    `iters` iterations, and the C oracle's iterations/s over `cpu_iters`
    iterations on this host's cores."""
    import torch
    from nmf_rs import _lib

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    from nmf_rs import _rust

    nt = min(16, os.cpu_count() or 1)
    m = _lib.DeviceMatrix(V)
    ext = torch.cuda.ExternalStream(_lib.stream_handle())
    out = {}
    # lsNMF with cv_rank's kind of weight: 5 % of the entries held out (0/1)
    mask = (np.random.default_rng(11).random(V.shape) > 0.05).astype(np.float64)
    try:
        for algo in ("brunet", "offset", "nsnmf", "lsnmf", "hals", "ehals", "dnmf"):
            kw = {"sparsity_h": 0.01, "sparsity_w": 0.01} if algo == "dnmf" else {}
            s = _lib.DeviceSolve(m, R, _lib.make_params(algo, iters, **kw))
            s.set_factors(W0, H0)
            if algo == "lsnmf":
                s.set_weight(mask)
                kw = {"weight": mask}
            s.launch()
            ext.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(ext)
            for _ in range(steps):
                s.launch()
            e1.record(ext)
            ext.synchronize()
            gpu = steps * iters / (e0.elapsed_time(e1) / 1e3)
            s.close()
            t = time.perf_counter()
            oracle.run(algo, V, W0, H0, cpu_iters, num_threads=nt, **kw)
            cpu = cpu_iters / (time.perf_counter() - t)
            out[algo] = {"gpu_iters_s": round(gpu, 1), "cpu_iters_s": round(cpu, 2),
                         "gpu_over_cpu": round(gpu / cpu, 1)}
            log(f"variant {algo}: {json.dumps(out[algo])}")
        # Kim-Park snmf: one host call each (V up over PCIe, FCNNLS on the
        # host between the GPU products), so the rate includes the copies
        kw = {"sparsity_h": 0.01, "sparsity_w": 0.01, "smoothness_h": -1.0, "smoothness_w": -1.0}
        for algo in ("snmf/r", "snmf/l"):
            si = 10
            _rust.py_nmf_run(algo, V, W0, H0, 1, **kw)
            t = time.perf_counter()
            _rust.py_nmf_run(algo, V, W0, H0, si, **kw)
            gpu = si / (time.perf_counter() - t)
            t = time.perf_counter()
            oracle.run(algo, V, W0, H0, 2, num_threads=nt, **kw)
            cpu = 2 / (time.perf_counter() - t)
            out[algo] = {"gpu_iters_s": round(gpu, 1), "cpu_iters_s": round(cpu, 2),
                         "gpu_over_cpu": round(gpu / cpu, 1),
                         "note": "host call per solve: V upload and FCNNLS on the host included"}
            log(f"variant {algo}: {json.dumps(out[algo])}")
    finally:
        m.close()
    return {"workload": f"2000x7750 single-cell, rank 10, FP64, {iters} iterations per solve "
                        f"(snmf: 10 per host call); lsNMF weight: 5 % held out (0/1); "
                        f"CPU: C oracle, {cpu_iters} iterations (snmf 2), {nt} threads",
            "methods": out}


def cpu_baseline(V, W0, H0, budget_s=15.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    nt = min(16, os.cpu_count() or 1)
    oracle.lib()
    t = time.perf_counter()
    oracle.run("lee", V, W0, H0, 2, num_threads=nt)
    per = (time.perf_counter() - t) / 2
    k = int(max(2, min(100, budget_s / max(per, 1e-6))))
    t = time.perf_counter()
    oracle.run("lee", V, W0, H0, k, num_threads=nt)
    dt = time.perf_counter() - t
    return {"value": k / dt, "unit": "iters/s", "cores": nt, "kind": "port",
            "sample": f"C oracle lee, {k} iterations of the same 2000x7750 r=10 solve "
                      f"(one-time V^T build included), {nt} threads"}


def objective_bench(V, W0, H0):
    """Cost of the objective-based stop: lee and brunet at 2000 x 7750 rank
    10 for 100 iterations with no stop (one graph replay) and with the
    relative-change stop at tolerance 0 (the objective pass and the stop
    test run after every iteration, on the device)."""
    from nmf_rs import _lib

    m = _lib.DeviceMatrix(V)
    out = {}
    try:
        for algo in ("lee", "brunet"):
            rec = {}
            for stop in ("max_iter", "rel_change"):
                prm = _lib.make_params(algo, ITERS, stop=stop, rel_tol=0.0, min_iter=1)
                s = _lib.DeviceSolve(m, R, prm)
                s.set_factors(W0, H0)
                s.launch()
                s.fetch()
                t = time.perf_counter()
                for _ in range(3):
                    s.launch()
                    _, _, it, _ = s.fetch()
                rec[stop + "_us_per_iter"] = round((time.perf_counter() - t) / (3 * it) * 1e6, 2)
                s.close()
            rec["objective_us_per_iter"] = round(rec["rel_change_us_per_iter"] -
                                                 rec["max_iter_us_per_iter"], 2)
            out[algo] = rec
            log(f"objective {algo}: {rec}")
    finally:
        m.close()
    return {"workload": "2000x7750 single-cell, rank 10, FP64, 100 iterations, relative-change "
                        "stop at tolerance 0 vs no stop", "methods": out}


def range_finder_bench():
    """NNDSVD on a resident 5000 x 100000 FP32-stored single-cell matrix at
    rank 30 (width 40, 7 power iterations): the device range finder alone
    and the whole nndsvd_on_device, each timed on its second call (NNDSVD
    with the default host BLAS threads and with 8)."""
    import torch

    import nmf_rs
    from nmf_rs import _lib, synthetic

    n, p, rank = 5000, 100000, 30
    Vd = synthetic.single_cell_device(n, p, k=10, seed=3)
    torch.cuda.synchronize()
    m = _lib.DeviceMatrix(Vd, precision="fp32")
    v_mean = float(Vd.mean())
    del Vd
    torch.cuda.empty_cache()
    try:
        Q = np.random.RandomState(0).normal(size=(n, rank + 10))
        ts = []
        for _ in range(2):
            t = time.perf_counter()
            m.range_finder(Q, 7, True)
            ts.append(time.perf_counter() - t)
        from threadpoolctl import threadpool_limits

        tn = {}
        for threads in (None, 8):
            # host BLAS threads of the QR, the small SVD and U = Q Uh: the
            # default (one per core) and 8, as a serving worker would have
            lim = threadpool_limits(limits=threads) if threads else None
            try:
                for _ in range(2):
                    t = time.perf_counter()
                    nmf_rs.nndsvd_on_device(m, rank, fill="mean", seed=0, v_mean=v_mean)
                    tn[threads] = time.perf_counter() - t
            finally:
                if lim is not None:
                    lim.restore_original_limits()
    finally:
        m.close()
    out = {"workload": f"NNDSVD of a {n}x{p} FP32-stored single-cell matrix, rank {rank}",
           "range_finder_ms": round(ts[1] * 1e3, 1), "nndsvd_ms": round(tn[None] * 1e3, 1),
           "nndsvd_ms_8_host_threads": round(tn[8] * 1e3, 1)}
    log(f"range finder: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain copy of known size (for counter scaling)")
    ap.add_argument("--no-large", action="store_true",
                    help="skip lee at 2000x500000 rank 20 in the FP64 and FP32 forms")
    ap.add_argument("--sweep-no-baseline", action="store_true",
                    help="rank sweep without its FP64 baseline and accuracy figures")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the 140-solve batched rank sweep with consensus")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the served stream (Poisson arrivals, p50 / p99 latency)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the other update rules (brunet, offset, nsNMF, lsNMF, HALS "
                         "family, snmf) on the same matrix")
    ap.add_argument("--forms", action="store_true",
                    help="also run the 2000x500000 rank-20 workload in the FP64 / FP32 / "
                         "BF16 hi+lo storage forms of V")
    ap.add_argument("--objective", action="store_true",
                    help="also time the objective-based stop (lee, brunet) per iteration")
    ap.add_argument("--range-finder", action="store_true",
                    help="also time NNDSVD's device range finder at 5000x100000 rank 30")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard V's columns across GPUs; see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import nmf_rs
    from nmf_rs import _lib, synthetic

    log("generating 2000x7750 single-cell matrix")
    V = synthetic.single_cell(N, P, seed=1, capture=synthetic.SPARSE_CAPTURE)
    zero_frac = float((V == 0).mean())
    log(f"zero fraction {zero_frac:.3f}; NNDSVD init on the GPU")
    t = time.perf_counter()
    W0, H0 = nmf_rs.nndsvd_init(V, R, fill="mean", seed=0)
    nndsvd_s = time.perf_counter() - t  # first call: library and context start-up included
    t = time.perf_counter()
    nmf_rs.nndsvd_init(V, R, fill="mean", seed=0)
    nndsvd_warm_s = time.perf_counter() - t

    m = _lib.DeviceMatrix(V)
    s = _lib.DeviceSolve(m, R, _lib.make_params("lee", ITERS))
    s.set_factors(W0, H0)
    ext = torch.cuda.ExternalStream(_lib.stream_handle())
    for _ in range(args.warmup):
        s.launch()
    ext.synchronize()
    torch.cuda.synchronize()

    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(ext)
    for _ in range(args.steps):
        s.launch()
    e1.record(ext)
    ext.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = gpu_ms / args.steps
    iters_s = args.steps * ITERS / (gpu_ms / 1e3)
    W, H, n_iter, _ = s.fetch()
    log(f"{args.steps} solves: {gpu_ms:.3f} ms GPU ({wall * 1e3:.1f} ms wall), "
        f"{iters_s:.0f} it/s")

    # The one-call host entry (V in and W, H out over PCIe, context and
    # buffers set up per call), for reference only.
    from nmf_rs import _rust
    _rust.py_nmf_run("lee", V, W0, H0, ITERS)
    t = time.perf_counter()
    _rust.py_nmf_run("lee", V, W0, H0, ITERS)
    host_call_s = time.perf_counter() - t

    col_ms, col_n, row_ms, row_n = s.profile()
    # the same passes with a 512 MiB read before each (V out of L2 and MALL)
    ccol_ms, ccol_n, crow_ms, crow_n = s.profile(evict_bytes=512 << 20)
    col_b, row_b = algorithmic_bytes(N, P, R)
    col_avg, row_avg = col_ms / col_n, row_ms / row_n
    if col_ms >= row_ms:
        name, avg, nbytes, tkey = "column pass (k_col_mfma, W^T V)", col_avg, col_b, "col_pass_bytes_per_launch"
    else:
        name, avg, nbytes, tkey = "row pass (k_row, V H^T)", row_avg, row_b, "row_pass_bytes_per_launch"
    achieved = nbytes / (avg * 1e-3) / 1e9
    ccol_avg, crow_avg = ccol_ms / ccol_n, crow_ms / crow_n
    if ccol_avg >= crow_avg:
        cname, cavg, cbytes = "column pass (k_col_mfma, W^T V)", ccol_avg, col_b
    else:
        cname, cavg, cbytes = "row pass (k_row, V H^T)", crow_avg, row_b
    cold = cbytes / (cavg * 1e-3) / 1e9
    log(f"cold caches: col pass avg {ccol_avg * 1e3:.2f} us, row pass avg {crow_avg * 1e3:.2f} us")
    b_iter = col_b + row_b
    log(f"col pass avg {col_avg * 1e3:.2f} us, row pass avg {row_avg * 1e3:.2f} us")

    out = {
        "metric": METRIC,
        "value": round(iters_s, 2),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded NB single-cell generator, log1p-normalised)",
        "config": {"workload": "lee FP64, 2000 genes x 7750 cells, rank 10, NNDSVD mean init, "
                               "100 iterations per step, V resident in HBM",
                   "n": N, "p": P, "rank": R, "iters_per_step": ITERS,
                   "library_capture": synthetic.SPARSE_CAPTURE, "zero_fraction": round(zero_frac, 4)},
        "hbm_frac_algorithmic": round(b_iter * iters_s / (HBM_PEAK_GBS * 1e9), 4),
        "bytes_per_iter": b_iter,
        "nndsvd_s": round(nndsvd_s, 4),
        "nndsvd_warm_s": round(nndsvd_warm_s, 4),
        "pcie_inclusive_iters_s": round(ITERS / host_call_s, 1),
        "roofline": {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(tkey), "avg_us": round(avg * 1e3, 3),
                     "col_avg_us": round(col_avg * 1e3, 3), "row_avg_us": round(row_avg * 1e3, 3),
                     "v_cache": "back-to-back passes over the same 124 MB V: V stays in the "
                                "256 MB MALL between launches"},
        "roofline_cold": {"bound": "hbm", "kernel": cname, "achieved": round(cold, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(cold / HBM_PEAK_GBS, 4),
                          "avg_us": round(cavg * 1e3, 3), "col_avg_us": round(ccol_avg * 1e3, 3),
                          "row_avg_us": round(crow_avg * 1e3, 3),
                          "v_cache": "a 512 MiB read before every pass evicts V from L2 and "
                                     "MALL; the read's own time is subtracted"},
    }
    if args.copy_calib:
        copy_calibration()
    if not args.no_large:
        out["large"] = large_bench()
    if not args.no_sweep:
        out["rank_sweep"] = rank_sweep_bench(baseline=not args.sweep_no_baseline)
    if not args.no_serve:
        out["serve"] = serve_bench()
    if not args.no_variants:
        out["variants"] = variants_bench(V, W0, H0)
    if args.forms:
        out["precision_forms"] = precision_forms()
    if args.objective:
        out["objective"] = objective_bench(V, W0, H0)
    if args.range_finder:
        out["range_finder"] = range_finder_bench()
    if not args.no_cpu:
        log("timing the CPU oracle")
        out["cpu_baseline"] = cpu_baseline(V, W0, H0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
