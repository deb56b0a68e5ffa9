"""Served solves: a stream of independent NMF requests on one GPU.

No reference counterpart (the reference runs one solve per Python call).
Each worker thread owns a library context (its own HIP stream), so requests
overlap on the device; every request uploads its V once in the served
storage form, runs NNDSVD on the resident matrix, then the solve with the
relative-change stop (min_iter .. max_iter iterations, stop when the
objective changes by less than rel_tol between iterations).  The library
calls release the GIL, so the workers' device work overlaps with each
other's host work (request set-up, NNDSVD's small factorisations).

With a device-memory budget (budget_bytes, or budget_fraction of the free
device memory that hipMemGetInfo reports) requests are admitted first in,
first out: each reserves its footprint (nmfg_footprint: V in the served
form, NNDSVD's range finder and the solve) before it allocates, waits while
the reservations of the requests in service leave too little room, and is
refused with BudgetError when it exceeds the whole budget.  The workers'
contexts are attached to a library budget of the same size, which frees
idle cached blocks of any context to make room and refuses (NMFG_EBUDGET)
an allocation beyond it.  Requests on different matrices are not fused
into shared launches: each request's kernels read its own V, so fusing
saves launches but no HBM traffic, and the contexts' streams already run
the requests' kernels side by side.
"""
from __future__ import annotations

import threading
import time
from collections import deque
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
    error: str = ""               # set when the request was refused (BudgetError)
    W0: np.ndarray = None         # the solve's start (NNDSVD), with keep_init
    H0: np.ndarray = None


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


class Admission:
    """First-in, first-out admission against a device-memory budget: a
    request reserves its footprint in bytes before it allocates and returns
    it when it is done.  A request waits while it is not at the head of the
    queue or its footprint does not fit beside the reservations in service
    (a large request is not overtaken by smaller ones behind it); one larger
    than the whole budget is refused at once."""

    def __init__(self, cap):
        self.cap = int(cap)
        self.reserved = 0
        self.reserved_peak = 0
        self.max_depth = 0
        self.waited = 0
        self.wait_s = 0.0
        self.rejected = 0
        self.admitted = 0
        self._cv = threading.Condition()
        self._queue = deque()

    def acquire(self, nbytes):
        """Blocks until `nbytes` are reserved; returns the seconds waited."""
        nbytes = int(nbytes)
        t0 = time.perf_counter()
        with self._cv:
            if nbytes > self.cap:
                self.rejected += 1
                raise _lib.BudgetError(f"request needs {nbytes} bytes of device memory; "
                                       f"the budget is {self.cap} bytes")
            ticket = object()
            self._queue.append(ticket)
            self.max_depth = max(self.max_depth, len(self._queue))
            blocked = False
            while self._queue[0] is not ticket or self.reserved + nbytes > self.cap:
                blocked = True
                self._cv.wait()
            self._queue.popleft()
            self.reserved += nbytes
            self.reserved_peak = max(self.reserved_peak, self.reserved)
            self.admitted += 1
            waited = time.perf_counter() - t0
            if blocked:
                self.waited += 1
                self.wait_s += waited
            self._cv.notify_all()
        return waited

    def release(self, nbytes):
        with self._cv:
            self.reserved -= int(nbytes)
            self._cv.notify_all()

    def depth(self):
        """Requests waiting for admission now."""
        with self._cv:
            return len(self._queue)


class Server:
    """A pool of workers, each with its own library context; optionally a
    device-memory budget with first-in, first-out admission.

    With ``cobatch`` (continuous batching) the workers only prepare requests
    (upload, NNDSVD, the solve's buffers and start) and hand the solves to
    one device loop, which advances every solve in service together: one
    step group (nmfg_group) runs rounds of 10-iteration chunks of its
    members (more for members cheaper than the costliest) on ``branches``
    streams that advance independently, with one stop-flag read-back per
    branch round; solves prepared meanwhile join at the next step and finished
    ones leave.  Without it each worker runs its own solve on its own
    stream.  With ``keep_init`` each Result of a stream also holds the
    solve's NNDSVD start (W0, H0)."""

    def __init__(self, workers=4, precision="fp32", max_iter=100, min_iter=25, rel_tol=1e-6,
                 host_threads=None, budget_bytes=None, budget_fraction=None, device=0,
                 cobatch=False, branches=4, keep_init=False):
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
        self.keep_init = bool(keep_init)
        self._local = threading.local()
        self._ctxs = []
        self._lock = threading.Lock()
        # The host BLAS thread count is process-global state: set it once,
        # before any worker runs, and restore it on close.  Changing it per
        # request from the workers would change it while other workers are
        # inside BLAS / LAPACK calls, which OpenBLAS does not support.
        # The limit reaches only the BLAS libraries loaded when it is set:
        # load scipy.linalg's (NNDSVD's QR, Cholesky and SVD) first.  Loaded
        # later by the first request, it ran unlimited in every worker, and
        # NNDSVD took 430-480 ms per request instead of 55-62 (8 workers,
        # 16 cores) for as long as that first server lived.
        import scipy.linalg  # noqa: F401
        from threadpoolctl import threadpool_limits

        self._blas_limits = threadpool_limits(limits=self.host_threads)
        self.budget = None
        self.admission = None
        if budget_bytes is not None or budget_fraction is not None:
            if budget_bytes is None:
                free, _ = _lib.device_memory(device)
                budget_bytes = int(float(budget_fraction) * free)
            self.budget = _lib.Budget(int(budget_bytes))
            self.admission = Admission(int(budget_bytes))
        self.cobatch = bool(cobatch)
        self._group = None
        if self.cobatch:
            self._gctx = _lib.Context(device)
            if self.budget is not None:
                self._gctx.set_budget(self.budget)
            self._group = _lib.StepGroup(self._gctx, branches=branches)
            self._gcv = threading.Condition()
            self._gwait = {}      # solve handle -> [event, error]
            self._gactive = 0     # solves added and not yet returned
            self._gstop = False
            self._gthread = threading.Thread(target=self._device_loop, name="nmf-serve-device",
                                             daemon=True)
            self._gthread.start()
        self._pool = ThreadPoolExecutor(self.workers, thread_name_prefix="nmf-serve")

    def _device_loop(self):
        """Steps the group while it has members; wakes the workers whose
        solves finished."""
        while True:
            with self._gcv:
                while self._gactive == 0 and not self._gstop:
                    self._gcv.wait()
                if self._gactive == 0 and self._gstop:
                    return
            try:
                done = self._group.step()
                err = None
            except Exception as e:  # every waiting solve gets the error
                done, err = None, e
            with self._gcv:
                if err is not None:
                    for w in self._gwait.values():
                        w[1] = err
                        w[0].set()
                    self._gactive -= len(self._gwait)
                    self._gwait.clear()
                    continue
                for sv in done:
                    w = self._gwait.pop(sv._h.value)
                    w[0].set()
                self._gactive -= len(done)

    def _run_grouped(self, solve):
        """Runs a prepared solve in the step group; returns when it finished."""
        w = [threading.Event(), None]
        with self._gcv:
            self._gwait[solve._h.value] = w
            self._gactive += 1
            try:
                self._group.add(solve)
            except Exception:
                self._gwait.pop(solve._h.value, None)
                self._gactive -= 1
                raise
            self._gcv.notify_all()
        w[0].wait()
        if w[1] is not None:
            raise w[1]

    def _ctx(self):
        c = getattr(self._local, "ctx", None)
        if c is None:
            c = _lib.Context()
            if self.budget is not None:
                c.set_budget(self.budget)
            self._local.ctx = c
            with self._lock:
                self._ctxs.append(c)
        return c

    def solve(self, V, rank, method, v_mean=None, init=None):
        """One request, synchronously on the calling thread's context.  V:
        numpy array or float64 device tensor.  init=(W0, H0) replaces the
        NNDSVD start.  Returns W, H, iterations, deviances, timings."""
        return self._solve(V, rank, method, v_mean, init)[:5]

    def footprint(self, n, p, rank, method):
        """Device bytes one request allocates (nmfg_footprint): V in the
        served form, NNDSVD's range finder of width rank + 10, the solve."""
        from . import _METHOD_ALIASES

        prm = _lib.make_params(_METHOD_ALIASES[method], self.max_iter, stop="rel_change",
                               rel_tol=self.rel_tol, min_iter=self.min_iter,
                               precision=self.precision)
        return _lib.footprint(prm, n, p, int(rank), int(rank) + 10)

    def _solve(self, V, rank, method, v_mean, init=None):
        from . import _METHOD_ALIASES, nndsvd_on_device

        algo = _METHOD_ALIASES[method]
        n, p = (int(d) for d in V.shape)
        need = 0
        wait_s = 0.0
        if self.admission is not None:
            need = self.footprint(n, p, rank, method)
            wait_s = self.admission.acquire(need)
        try:
            ctx = self._ctx()
            t0 = time.perf_counter()
            if v_mean is None:
                v_mean = float(V.mean())
            m = _lib.DeviceMatrix(V, precision=self.precision, ctx=ctx)
            t1 = time.perf_counter()
            try:
                if init is None:
                    W0, H0 = nndsvd_on_device(m, int(rank), fill="mean", seed=0, v_mean=v_mean)
                else:
                    W0, H0 = (np.asarray(a, dtype=np.float64) for a in init)
                t2 = time.perf_counter()
                prm = _lib.make_params(algo, self.max_iter, stop="rel_change",
                                       rel_tol=self.rel_tol, min_iter=self.min_iter,
                                       precision=self.precision)
                s = _lib.DeviceSolve(m, int(rank), prm)
                try:
                    s.set_factors(W0, H0)
                    if self.cobatch:
                        self._run_grouped(s)
                    else:
                        s.launch()
                    W, H, n_iter, dev = s.fetch()
                finally:
                    s.close()
                t3 = time.perf_counter()
            finally:
                m.close()
        finally:
            if self.admission is not None:
                self.admission.release(need)
        return W, H, n_iter, list(dev), {"admission_wait_s": wait_s, "upload_s": t1 - t0,
                                         "nndsvd_s": t2 - t1, "solve_s": t3 - t2,
                                         "footprint_bytes": need}, (W0, H0)

    def metrics(self):
        """Budget and queue figures: the admission queue (requests waiting
        now, its largest depth, how many waited and for how long in all,
        how many were refused), the bytes reserved (now and at the peak) and
        the bytes the library holds for the workers under the budget (now,
        at the peak: blocks in use and idle cached blocks)."""
        out = {}
        if self.cobatch:
            g = self._group.stats()
            g["mean_members_per_step"] = (round(g["member_steps"] / g["steps"], 3)
                                          if g["steps"] else 0.0)
            out["cobatch"] = g
        if self.admission is None:
            out["budget_bytes"] = None
            return out
        a = self.admission
        cap, used, peak = self.budget.query()
        return {**out, "budget_bytes": cap, "queue_depth": a.depth(), "max_queue_depth": a.max_depth,
                "admitted": a.admitted, "waited": a.waited, "wait_s_total": round(a.wait_s, 6),
                "rejected": a.rejected, "reserved_bytes": a.reserved,
                "reserved_peak_bytes": a.reserved_peak, "device_held_bytes": used,
                "device_held_peak_bytes": peak}

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
            try:
                W, H, it, dev, tm, (W0, H0) = self._solve(matrices[i], r.rank, r.method,
                                                          None if means is None else means[i])
            except _lib.BudgetError as e:
                now = time.perf_counter() - t_start
                results[i] = Result(r.rid, None, None, 0, [], st, now, {}, str(e))
                return
            results[i] = Result(r.rid, W, H, it, dev, st, time.perf_counter() - t_start, tm)
            if self.keep_init:
                results[i].W0, results[i].H0 = W0, H0

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
        if self._group is not None:
            with self._gcv:
                self._gstop = True
                self._gcv.notify_all()
            self._gthread.join()
            self._group.close()
            self._group = None
            self._gctx.close()
        with self._lock:
            for c in self._ctxs:
                c.close()
            self._ctxs.clear()
        if self.budget is not None:
            self.budget.close()
            self.budget = None
        if self._blas_limits is not None:
            self._blas_limits.restore_original_limits()
            self._blas_limits = None


def latency_stats(requests, results):
    """Latency (completion - arrival) percentiles in seconds, over the
    requests that were served."""
    lat = np.array([res.end - req.arrival for req, res in zip(requests, results)
                    if not res.error])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
