"""CPU tests of the request loop's admission, continuous batching and failure
handling (fake executor, no GPU)."""

import threading
import time

import pytest

from jaxframes_amd.serving import FrameStore, Server, poisson_arrivals


class _Q:
    def __init__(self, qid, fp, dur=0.02, fail=False):
        self.qid, self.fp, self.dur, self.fail = qid, fp, dur, fail


def _make(slots=2, budget=100, log=None):
    lock = threading.Lock()
    state = {"running": 0, "peak": 0}

    def ex(store, q):
        with lock:
            state["running"] += 1
            state["peak"] = max(state["peak"], state["running"])
            if log is not None:
                log.append(("start", q.qid, time.perf_counter()))
        time.sleep(q.dur)
        with lock:
            state["running"] -= 1
        if q.fail:
            raise RuntimeError(f"query {q.qid} failed")
        return q.qid

    srv = Server(FrameStore(budget), slots=slots, hbm_budget_bytes=budget, executor=ex,
                 use_streams=False, footprint_fn=lambda st, q: q.fp)
    return srv, state


def test_slots_bound_concurrency():
    srv, st = _make(slots=3, budget=10**9)
    futs = [srv.submit(_Q(i, 1)) for i in range(12)]
    assert [f.result()[0] for f in futs] == list(range(12))
    srv.close()
    assert st["peak"] == 3 and srv.peak_running == 3


def test_memory_budget_limits_admission():
    srv, st = _make(slots=8, budget=100)
    futs = [srv.submit(_Q(i, 40)) for i in range(6)]   # at most 2 fit at once
    for f in futs:
        f.result()
    srv.close()
    assert srv.peak_reserved <= 100
    assert st["peak"] == 2


def test_oversize_request_rejected_without_blocking_others():
    srv, _ = _make(slots=2, budget=100)
    big = srv.submit(_Q(0, 1000))
    small = srv.submit(_Q(1, 10))
    with pytest.raises(MemoryError):
        big.result()
    assert small.result()[0] == 1
    srv.close()


def test_errors_reach_the_caller_and_free_the_slot():
    srv, _ = _make(slots=1, budget=100)
    bad = srv.submit(_Q(0, 50, fail=True))
    good = srv.submit(_Q(1, 50))
    with pytest.raises(RuntimeError):
        bad.result()
    assert good.result()[0] == 1
    assert srv.running == 0 and srv.reserved == 0
    srv.close()


def test_continuous_admission_starts_waiting_request_when_a_slot_frees():
    log = []
    srv, _ = _make(slots=2, budget=10**9, log=log)
    a = srv.submit(_Q(0, 1, dur=0.05))
    b = srv.submit(_Q(1, 1, dur=0.30))
    c = srv.submit(_Q(2, 1, dur=0.01))       # waits for a slot
    for f in (a, b, c):
        f.result()
    srv.close()
    start = {qid: t for _, qid, t in log}
    # c starts when a finishes, long before b does
    assert start[2] - start[0] < 0.2
    assert srv.admitted_order == [1, 2, 3]


def test_fifo_head_of_line_respects_budget():
    srv, _ = _make(slots=4, budget=100)
    futs = [srv.submit(_Q(0, 60, dur=0.05)), srv.submit(_Q(1, 60, dur=0.01)),
            srv.submit(_Q(2, 10, dur=0.01))]
    for f in futs:
        f.result()
    srv.close()
    assert srv.admitted_order == [1, 2, 3]   # request 3 does not overtake request 2


def test_poisson_arrivals_rate():
    a = poisson_arrivals(20000, qps=500.0, seed=1)
    assert all(x < y for x, y in zip(a, a[1:]))
    assert abs(len(a) / a[-1] - 500.0) / 500.0 < 0.05
