"""Step groups (continuous batching): solves of different methods, ranks,
shapes, stops, precision forms and contexts advanced together, one
10-iteration chunk of every member per step on parallel streams.  A member
runs exactly the kernels of its solve run alone, so W, H, the iteration
count and the deviances must be the same bits as nmfg_solve_launch gives;
solves that join while others are running and members that stop at
different steps included.  Served streams in the continuous-batching mode
must equal the same requests solved one at a time."""
import numpy as np
import pytest

from nmf_rs import _lib, serve, synthetic

pytestmark = pytest.mark.gpu

DNMF = {"sparsity_h": 0.01, "sparsity_w": 0.01}

# (method, n, p, rank, stop, form, extra params)
CASES = [
    ("lee", 300, 900, 6, "rel_change", "fp64", {}),
    ("brunet", 250, 1100, 8, "rel_change", "fp64", {}),
    ("hals", 120, 700, 5, "stationary", "fp64", {}),
    ("dnmf", 200, 640, 12, "rel_change", "fp64", DNMF),
    ("ehals", 90, 333, 3, "max_iter", "fp64", {}),
    ("nsNMF", 160, 500, 7, "rel_change", "fp64", {}),
    ("offset", 140, 610, 4, "rel_change", "fp64", {}),
    ("lee", 500, 2100, 20, "rel_change", "fp32", {}),
    ("brunet", 410, 1500, 10, "rel_change", "fp32", {}),
    ("lee", 350, 1300, 9, "rel_change", "bf16x2", {}),
    ("lsnmf", 130, 420, 5, "rel_change", "fp64", {}),
    ("brunet", 2000, 7750, 10, "rel_change", "fp64", {}),  # fused objective in the column pass
]


def make_case(i, case, ctx):
    method, n, p, r, stop, form, kw = case
    V = synthetic.lowrank_uniform(n, p, max(2, r // 2), seed=100 + i)[0]
    g = np.random.default_rng(200 + i)
    W0, H0 = g.uniform(0.1, 1.0, (n, r)), g.uniform(0.1, 1.0, (r, p))
    # tolerances that stop the members at different iterations
    prm = _lib.make_params(method, 60, precision=form, stop=stop,
                           rel_tol=[3e-2, 1e-2, 3e-3, 1e-3][i % 4], min_iter=3,
                           stationary_th=1e-2, check_interval=10, check_niter=5, **kw)
    m = _lib.DeviceMatrix(V, precision=form, ctx=ctx)
    s = _lib.DeviceSolve(m, r, prm)
    s.set_factors(W0, H0)
    if method == "lsnmf":
        S = (g.random((n, p)) > 0.1).astype(np.float64)
        s.set_weight(S)
    return m, s


def test_group_members_equal_alone():
    ctxs = [_lib.Context(), _lib.Context()]
    gctx = _lib.Context()
    alone, grouped = [], []
    try:
        for i, case in enumerate(CASES):
            m, s = make_case(i, case, ctxs[i % 2])
            s.launch()
            alone.append(s.fetch())
            grouped.append((m, s))
        group = _lib.StepGroup(gctx, branches=4)
        try:
            # half join at once, the rest two steps later (while others run)
            for m, s in grouped[: len(grouped) // 2]:
                group.add(s)
            out = {}
            steps = 0
            while len(out) < len(grouped):
                if steps == 2:
                    for m, s in grouped[len(grouped) // 2:]:
                        group.add(s)
                for s in group.step():
                    out[s._h.value] = s.fetch()
                steps += 1
                assert steps < 200
            st = group.stats()
            print("group stats", st, flush=True)
            assert st["max_members"] == len(grouped) or st["max_members"] >= len(grouped) // 2
            assert group.size() == 0
        finally:
            group.close()
        for case, (m, s), ref in zip(CASES, grouped, alone):
            W, H, it, dev = out[s._h.value]
            print(f"{case[0]} {case[1]}x{case[2]} r={case[3]} {case[4]} {case[5]}: "
                  f"n_iter {it} (alone {ref[2]})", flush=True)
            assert it == ref[2]
            np.testing.assert_array_equal(W, ref[0])
            np.testing.assert_array_equal(H, ref[1])
            np.testing.assert_array_equal(dev, ref[3])
        # the members stopped at different iterations: the group ran them
        # past each other's stops
        assert len({ref[2] for ref in alone}) > 1
    finally:
        for m, s in grouped:
            s.close()
            m.close()
        for c in ctxs + [gctx]:
            c.close()


def test_group_reuses_a_solve_and_rejects_misuse():
    """A solve returned by a step can join again (same results); adding a
    member twice or with no factors is refused."""
    ctx = _lib.Context()
    try:
        m, s = make_case(0, CASES[0], ctx)
        s.launch()
        ref = s.fetch()
        group = _lib.StepGroup(ctx, branches=2)
        try:
            for _ in range(2):
                group.add(s)
                with pytest.raises(ValueError):
                    group.add(s)
                got = []
                while not got:
                    got = group.step()
                assert got == [s]
                W, H, it, dev = s.fetch()
                assert it == ref[2]
                np.testing.assert_array_equal(W, ref[0])
                np.testing.assert_array_equal(dev, ref[3])
            assert group.step() == []
            s2 = _lib.DeviceSolve(m, 6, _lib.make_params("lee", 10))
            with pytest.raises(ValueError):
                group.add(s2)
            s2.close()
        finally:
            group.close()
        s.close()
        m.close()
    finally:
        ctx.close()


def test_cobatched_stream_equals_one_at_a_time():
    reqs = serve.request_mix(10, seed=5, n_range=(60, 400), p_range=(100, 1500), ranks=(2, 14))
    serve.poisson_arrivals(reqs, qps=400.0, seed=6)
    mats = [synthetic.lowrank_uniform(r.n, r.p, 3, seed=r.seed % 1000)[0] for r in reqs]
    srv = serve.Server(workers=3, precision="fp32", max_iter=50, min_iter=5, rel_tol=1e-6,
                       cobatch=True)
    try:
        res = srv.run_stream(reqs, mats)
        met = srv.metrics()
    finally:
        srv.close()
    print("metrics", met, flush=True)
    alone = serve.Server(workers=1, precision="fp32", max_iter=50, min_iter=5, rel_tol=1e-6)
    try:
        for req, out, V in zip(reqs, res, mats):
            assert not out.error, out.error
            W, H, it, dev, _ = alone.solve(V, req.rank, req.method)
            print(f"request {req.rid} {req.method} {req.n}x{req.p} r={req.rank} it={out.n_iter}",
                  flush=True)
            assert out.n_iter == it
            np.testing.assert_array_equal(out.W, W)
            np.testing.assert_array_equal(out.H, H)
            np.testing.assert_array_equal(out.deviances, dev)
    finally:
        alone.close()
    assert met["cobatch"]["steps"] > 0 and met["cobatch"]["max_members"] >= 2
