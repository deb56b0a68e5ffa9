"""First-in, first-out admission against a device-memory budget (host
logic, no GPU): requests wait in arrival order while their footprint does
not fit beside the reservations in service, a smaller request is not let
past a waiting larger one, and a request above the whole budget is refused.
The footprint query (nmfg_footprint) is host arithmetic over the solve's
buffer layout."""
import threading
import time

import pytest

from nmf_rs import _lib
from nmf_rs.serve import Admission


def _start(adm, nbytes, log, name):
    def run():
        adm.acquire(nbytes)
        log.append(name)

    t = threading.Thread(target=run)
    t.start()
    return t


def _wait_depth(adm, d):
    for _ in range(2000):
        if adm.depth() >= d:
            return
        time.sleep(0.001)
    raise AssertionError("queue never reached depth %d" % d)


def test_fifo_waits_and_no_overtaking():
    adm = Admission(100)
    log = []
    adm.acquire(60)                      # a request in service
    tb = _start(adm, 50, log, "b")       # does not fit: waits
    _wait_depth(adm, 1)
    tc = _start(adm, 10, log, "c")       # would fit, but b is ahead of it
    _wait_depth(adm, 2)
    time.sleep(0.02)
    assert log == [] and adm.reserved == 60 and adm.max_depth == 2
    adm.release(60)
    tb.join(5)
    tc.join(5)
    assert log == ["b", "c"]
    assert adm.reserved == 60 and adm.reserved_peak == 60
    assert adm.waited == 2 and adm.admitted == 3 and adm.depth() == 0
    adm.release(50)
    adm.release(10)
    assert adm.reserved == 0


def test_request_above_budget_is_refused():
    adm = Admission(100)
    with pytest.raises(_lib.BudgetError):
        adm.acquire(101)
    assert adm.rejected == 1 and adm.reserved == 0
    assert adm.acquire(100) == pytest.approx(0.0, abs=0.1)
    adm.release(100)


def test_reservations_never_exceed_budget():
    adm = Admission(1000)
    sizes = [300, 700, 200, 900, 100, 500, 400, 1000, 50, 600] * 3
    in_service = []
    lock = threading.Lock()
    worst = [0]

    def req(nb):
        adm.acquire(nb)
        with lock:
            in_service.append(nb)
            worst[0] = max(worst[0], sum(in_service))
        time.sleep(0.002)
        with lock:
            in_service.remove(nb)
        adm.release(nb)

    ts = [threading.Thread(target=req, args=(nb,)) for nb in sizes]
    for t in ts:
        t.start()
    for t in ts:
        t.join(20)
    assert worst[0] <= 1000 and adm.reserved_peak <= 1000
    assert adm.admitted == len(sizes) and adm.reserved == 0


def test_footprint_grows_with_the_request():
    prm = _lib.make_params("lee", 100, precision="fp32")
    small = _lib.footprint(prm, 300, 5000, 6, 16)
    assert small >= 4 * 300 * 5000  # at least V in FP32
    assert _lib.footprint(prm, 600, 5000, 6, 16) > small
    assert _lib.footprint(prm, 300, 9000, 6, 16) > small
    assert _lib.footprint(prm, 300, 5000, 24, 34) > small
    # the FP64 form stores V in 8 bytes and needs no staging rows
    f64 = _lib.footprint(_lib.make_params("lee", 100), 300, 5000, 6, 16)
    assert f64 >= 8 * 300 * 5000
    with pytest.raises(ValueError):
        _lib.footprint(prm, 0, 5000, 6, 16)


def test_grid_coverage_every_method_form_rank_and_width():
    """The library's set-up checks on the host that every column-tiled pass
    covers V's p columns and ends inside its padded rows; footprint runs that
    set-up without a device.  Every method, precision form, template rank
    and p around the tile widths (32, 64, 128) must pass it."""
    methods = ["lee", "brunet", "hals", "ehals", "dnmf", "offset", "nsnmf", "lsnmf"]
    forms = ["fp64", "fp32-store", "bf16x2-store", "fp32", "bf16x2"]
    widths = [1, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 7750, 32769]
    n = 0
    for m in methods:
        kw = {"sparsity_h": 0.01, "sparsity_w": 0.01} if m == "dnmf" else {}
        for f in forms:
            prm = _lib.make_params(m, 10, precision=f, **kw)
            for r in (1, 5, 10, 16, 17, 20, 24, 32):
                for p in widths:
                    assert _lib.footprint(prm, 40, p, r) > 0
                    n += 1
    assert n == len(methods) * len(forms) * 8 * len(widths)


def test_budget_handle():
    b = _lib.Budget(1 << 20)
    assert b.query() == (1 << 20, 0, 0)
    b.close()
    with pytest.raises(ValueError):
        _lib.Budget(-1)
