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


def _make_batching(slots=1, budget=10**9, max_batch=32):
    runs = []

    def ex(store, q):
        time.sleep(q.dur)
        runs.append([q.qid])
        return q.qid

    def bex(store, qs):
        runs.append([q.qid for q in qs])
        return [q.qid for q in qs]

    srv = Server(FrameStore(budget), slots=slots, hbm_budget_bytes=budget, executor=ex,
                 use_streams=False, footprint_fn=lambda st, q: q.fp, batch_executor=bex,
                 batchable_fn=lambda st, q: getattr(q, "batch", False), max_batch=max_batch)
    return srv, runs


def _bq(qid, fp=1):
    q = _Q(qid, fp, dur=0.0)
    q.batch = True
    return q


def test_queued_batchable_requests_join_the_head_batch():
    srv, runs = _make_batching(slots=1)
    blocker = srv.submit(_Q(0, 1, dur=0.1))
    time.sleep(0.02)                        # the blocker holds the only slot
    futs = [srv.submit(_bq(1)), srv.submit(_Q(2, 1, dur=0.0)), srv.submit(_bq(3)),
            srv.submit(_bq(4))]
    assert [f.result()[0] for f in [blocker] + futs] == [0, 1, 2, 3, 4]
    srv.close()
    # the head batch took 1, 3, 4 past the non-batchable 2, which ran next
    assert runs == [[0], [1, 3, 4], [2]]
    assert srv.batches == [1, 3, 1]


def test_batch_respects_budget_and_max_batch():
    srv, runs = _make_batching(slots=1, budget=100, max_batch=3)
    blocker = srv.submit(_Q(0, 10, dur=0.1))
    time.sleep(0.02)
    futs = [srv.submit(_bq(i, fp=40)) for i in range(1, 6)]
    for f in [blocker] + futs:
        f.result()
    srv.close()
    assert srv.peak_reserved <= 100
    assert all(len(r) <= 3 for r in runs)
    assert [i for r in runs for i in r] == [0, 1, 2, 3, 4, 5]   # arrival order kept


def test_batch_failure_reaches_every_member():
    def bex(store, qs):
        raise RuntimeError("batch failed")
    srv = Server(FrameStore(100), slots=1, hbm_budget_bytes=100, executor=lambda st, q: time.sleep(0.1),
                 use_streams=False, footprint_fn=lambda st, q: 1, batch_executor=bex,
                 batchable_fn=lambda st, q: getattr(q, "batch", False))
    blocker = srv.submit(_Q(0, 1))
    time.sleep(0.02)
    futs = [srv.submit(_bq(i)) for i in (1, 2)]
    blocker.result()
    for f in futs:
        with pytest.raises(RuntimeError):
            f.result()
    assert srv.running == 0 and srv.reserved == 0
    srv.close()


def test_switch_interval_untouched_by_default_set_only_when_asked_and_restored():
    import sys
    before = sys.getswitchinterval()
    srv, _ = _make(slots=1)
    assert sys.getswitchinterval() == before
    srv.close()
    srv = Server(FrameStore(100), slots=1, hbm_budget_bytes=100, executor=lambda st, q: None,
                 use_streams=False, footprint_fn=lambda st, q: 1, switch_interval=1e-4)
    assert abs(sys.getswitchinterval() - 1e-4) < 1e-9
    srv.close()
    assert sys.getswitchinterval() == before


def test_admission_timeout_fails_only_the_late_request():
    srv, _ = _make(slots=1, budget=100)
    blocker = srv.submit(_Q(0, 10, dur=0.25))
    time.sleep(0.02)
    late = srv.submit(_Q(1, 10, dur=0.0), timeout=0.05)
    patient = srv.submit(_Q(2, 10, dur=0.0))
    with pytest.raises(TimeoutError):
        late.result(timeout=5)
    assert blocker.result()[0] == 0 and patient.result()[0] == 2
    assert 2 not in srv.admitted_order
    srv.close()
    assert srv.running == 0 and srv.reserved == 0


def test_cancelled_request_never_runs():
    log = []
    srv, _ = _make(slots=1, budget=100, log=log)
    blocker = srv.submit(_Q(0, 10, dur=0.15))
    time.sleep(0.02)
    gone = srv.submit(_Q(1, 10, dur=0.0))
    kept = srv.submit(_Q(2, 10, dur=0.0))
    assert gone.cancel()
    assert blocker.result()[0] == 0 and kept.result()[0] == 2
    srv.close()
    assert [q for _, q, _ in log] == [0, 2]
    assert not blocker.cancel()     # admitted requests cannot be withdrawn


def test_injected_library_error_fails_one_request_without_wedging():
    from jaxframes_amd import _lib

    import ctypes

    def ex(store, q):
        nb = ctypes.c_uint64(0)          # a real C-ABI call that needs no GPU
        _lib.call("jfx_compact_workspace_size", 0, ctypes.byref(nb))
        return q.qid

    srv = Server(FrameStore(100), slots=2, hbm_budget_bytes=100, executor=ex, use_streams=False,
                 footprint_fn=lambda st, q: 1, batch_executor=None)
    _lib.inject_failure("jfx_compact_workspace_size", status=2, count=1)
    try:
        futs = [srv.submit(_Q(i, 1)) for i in range(6)]
        outcomes = []
        for f in futs:
            try:
                outcomes.append(f.result(timeout=5)[0])
            except _lib.JfxError as e:
                assert "injected" in str(e)
                outcomes.append("error")
    finally:
        _lib.inject_failure("jfx_compact_workspace_size", count=0)
        srv.close()
    assert outcomes.count("error") == 1
    assert sorted(o for o in outcomes if o != "error") == sorted(set(range(6)) - {outcomes.index("error")})
    assert srv.running == 0 and srv.reserved == 0
