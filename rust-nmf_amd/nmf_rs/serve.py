"""Served solves: a stream of independent NMF requests on one GPU.

No reference counterpart (the reference runs one solve per Python call).
Each worker thread owns a library context (its own HIP stream), so requests
overlap on the device; every request uploads its V once in the served
storage form, runs NNDSVD on the resident matrix, then the solve with the
relative-change stop (min_iter .. max_iter iterations, stop when the
objective changes by less than rel_tol between iterations).  The library
calls release the GIL, so the workers' device work overlaps with each
other's host work (request set-up, NNDSVD's small factorisations).
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class Request:
    rid: int
    n: int
    p: int
    rank: int
    method: str
    arrival: float = 0.0          # seconds after the stream starts
    seed: int = 0


@dataclass
class Result:
    rid: int
    W: np.ndarray
    H: np.ndarray
    n_iter: int
    deviances: list
    start: float = 0.0            # service start (s, stream clock)
    end: float = 0.0
    timings: dict = field(default_factory=dict)


def request_mix(count, seed=0, n_range=(500, 5000), p_range=(1000, 100000), ranks=(5, 30),
                methods=(("lee", 0.5), ("hals", 0.3), ("brunet", 0.2))):
    """Request shapes: n, p log-uniform over the ranges, rank uniform over
    the inclusive range, method by the given probabilities."""
    g = np.random.default_rng(seed)
    names = [m for m, _ in methods]
    prob = np.array([w for _, w in methods], dtype=np.float64)
    prob /= prob.sum()
    out = []
    for i in range(count):
        n = int(np.floor(np.exp(g.uniform(np.log(n_range[0]), np.log(n_range[1])))))
        p = int(np.floor(np.exp(g.uniform(np.log(p_range[0]), np.log(p_range[1])))))
        r = int(g.integers(ranks[0], ranks[1] + 1))
        m = str(g.choice(names, p=prob))
        out.append(Request(i, n, p, r, m, seed=int(g.integers(0, 2**31 - 1))))
    return out


def poisson_arrivals(requests, qps, seed=1):
    """Assigns open-loop Poisson arrival times at rate qps (requests / s)."""
    g = np.random.default_rng(seed)
    t = 0.0
    for r in requests:
        t += float(g.exponential(1.0 / qps))
        r.arrival = t
    return requests


class Server:
    """A pool of workers, each with its own library context."""

    def __init__(self, workers=4, precision="fp32", max_iter=100, min_iter=25, rel_tol=1e-6,
                 host_threads=None):
        import os

        self.workers = int(workers)
        # host BLAS threads per worker (NNDSVD's small factorisations): the
        # workers share the cores instead of each starting a full pool
        cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        self.host_threads = int(host_threads or max(1, (cores or 1) // self.workers))
        self.precision = precision
        self.max_iter = int(max_iter)
        self.min_iter = int(min_iter)
        self.rel_tol = float(rel_tol)
        self._local = threading.local()
        self._ctxs = []
        self._lock = threading.Lock()
        # The host BLAS thread count is process-global state: set it once,
        # before any worker runs, and restore it on close.  Changing it per
        # request from the workers would change it while other workers are
        # inside BLAS / LAPACK calls, which OpenBLAS does not support.
        from threadpoolctl import threadpool_limits

        self._blas_limits = threadpool_limits(limits=self.host_threads)
        self._pool = ThreadPoolExecutor(self.workers, thread_name_prefix="nmf-serve")

    def _ctx(self):
        c = getattr(self._local, "ctx", None)
        if c is None:
            c = _lib.Context()
            self._local.ctx = c
            with self._lock:
                self._ctxs.append(c)
        return c

    def solve(self, V, rank, method, v_mean=None):
        """One request, synchronously on the calling thread's context.  V:
        numpy array or float64 device tensor."""
        return self._solve(V, rank, method, v_mean)

    def _solve(self, V, rank, method, v_mean):
        from . import _METHOD_ALIASES, nndsvd_on_device

        algo = _METHOD_ALIASES[method]
        ctx = self._ctx()
        t0 = time.perf_counter()
        if v_mean is None:
            v_mean = float(V.mean())
        m = _lib.DeviceMatrix(V, precision=self.precision, ctx=ctx)
        t1 = time.perf_counter()
        try:
            W0, H0 = nndsvd_on_device(m, int(rank), fill="mean", seed=0, v_mean=v_mean)
            t2 = time.perf_counter()
            prm = _lib.make_params(algo, self.max_iter, stop="rel_change", rel_tol=self.rel_tol,
                                   min_iter=self.min_iter, precision=self.precision)
            s = _lib.DeviceSolve(m, int(rank), prm)
            try:
                s.set_factors(W0, H0)
                s.launch()
                W, H, n_iter, dev = s.fetch()
            finally:
                s.close()
            t3 = time.perf_counter()
        finally:
            m.close()
        return W, H, n_iter, list(dev), {"upload_s": t1 - t0, "nndsvd_s": t2 - t1,
                                         "solve_s": t3 - t2}

    def capacity(self, requests, matrices, means=None):
        """Closed-loop throughput (requests / s) with every worker busy: the
        given requests are all submitted at once."""
        t = time.perf_counter()
        futs = [self._pool.submit(self.solve, matrices[i], r.rank, r.method,
                                  None if means is None else means[i])
                for i, r in enumerate(requests)]
        for f in futs:
            f.result()
        return len(requests) / (time.perf_counter() - t)

    def run_stream(self, requests, matrices, means=None):
        """Submits each request at its arrival time (open loop) and waits for
        all; returns Results in request order.  matrices[i] is request i's V."""
        results = [None] * len(requests)
        futs = []
        t_start = time.perf_counter()

        def job(i):
            r = requests[i]
            st = time.perf_counter() - t_start
            W, H, it, dev, tm = self.solve(matrices[i], r.rank, r.method,
                                           None if means is None else means[i])
            results[i] = Result(r.rid, W, H, it, dev, st, time.perf_counter() - t_start, tm)

        for i, r in enumerate(requests):
            delay = r.arrival - (time.perf_counter() - t_start)
            if delay > 0:
                time.sleep(delay)
            futs.append(self._pool.submit(job, i))
        for f in futs:
            f.result()
        return results

    def close(self):
        self._pool.shutdown(wait=True)
        with self._lock:
            for c in self._ctxs:
                c.close()
            self._ctxs.clear()
        if self._blas_limits is not None:
            self._blas_limits.restore_original_limits()
            self._blas_limits = None


def latency_stats(requests, results):
    """Latency (completion - arrival) percentiles in seconds."""
    lat = np.array([res.end - req.arrival for req, res in zip(requests, results)])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
