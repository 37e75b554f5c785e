"""Serving: a request loop over device-resident frames.

- FrameStore keeps frames resident in HBM and, when a frame is loaded, builds
  its key order once (stable radix sort with the row index: sorted keys, the
  permutation, values in key order). Queries then never sort the resident side.
- Queries: FilterGroupBy (filter(value <op> t) -> groupby(key).agg(sum, count,
  mean)) and Join (inner join of a resident frame with a small dimension
  table). Results equal the plain path's: the reference semantics of
  groupby_aggregate (parallel_algorithms.py:733-803) on the filtered rows and
  of sort_merge_join (:938-1075).
- Server does continuous batching: requests queue on arrival and are admitted
  whenever a slot is free AND their HBM footprint fits the budget next to the
  requests already running; each slot is a worker thread with its own HIP
  stream (default priority), so admitted requests overlap on the device and a
  request that arrives while others run starts as soon as one finishes.
  Latency is arrival -> result ready on the host.
"""

from __future__ import annotations

import ctypes
import sys
import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass, field

import torch

from . import _lib
from ._tensors import key_code, to_device, widen_key


# ---- resident frames ---------------------------------------------------------------

class Frame:
    """A resident frame (int keys, value column) and its indexes:

    - key order (stable): sorted keys and the row permutation, reused by joins
      (pairs then come in the reference's (key, left row, right row) order);
    - (key, value) order for float values: the values ascending inside each
      key group (NaN last), float64 prefix sums of them restarted at every
      group (segmented_prefix), group starts, the
      first NaN row of each group and the group keys, which answer
      filter -> groupby by binary search (jfx_served_groups).
    """

    def __init__(self, name: str, keys, values):
        from .multikey import multikey_heads
        from .sorting import radix_sort
        self.name = name
        self.keys, _ = widen_key(to_device(keys))
        self.values = to_device(values, self.keys.device)
        self.n = self.keys.shape[0]
        dev = self.keys.device
        sk, _, perm = radix_sort(self.keys, None, True, return_index=True, index_dtype=torch.int32)
        self.sorted_keys = sk
        self.perm = perm                                  # uint32 row index (int32 storage)
        self.sorted_values = self.values[perm.long()] if self.n else self.values
        parts = [self.keys, self.values, self.sorted_keys, self.perm, self.sorted_values]
        self.ngroups = 0
        if self.n:
            # key groups of the key order: joins run over them
            gid, starts = multikey_heads([sk])
            G = starts.shape[0]
            self.ngroups = G
            self.group_keys = sk[starts]
            self.group_starts = torch.empty(G + 1, dtype=torch.int32, device=dev)
            self.group_starts[:G] = starts.to(torch.int32)
            self.group_starts[G] = self.n
            parts += [self.group_keys, self.group_starts]
        self.indexed = self.values.dtype in (torch.float32, torch.float64) and self.n > 0
        if self.indexed:
            # (key, value) order: stable by value, then stable by key
            _, _, pv = radix_sort(self.values, None, True, want_keys=False, return_index=True)
            _, _, p2 = radix_sort(self.keys[pv], None, True, want_keys=False, return_index=True)
            kv = pv[p2]
            self.kv_values = self.values[kv]
            nan = torch.isnan(self.kv_values)
            nn = torch.bincount(gid[nan], minlength=G)
            npos = torch.bincount(gid[torch.isposinf(self.kv_values)], minlength=G)
            nneg = torch.bincount(gid[torch.isneginf(self.kv_values)], minlength=G)
            ends = self.group_starts[1:].long()
            self.group_nan_start = (ends - nn).to(torch.int32)
            self.group_pinf_start = (ends - nn - npos).to(torch.int32)
            self.group_ninf_end = (self.group_starts[:G].long() + nneg).to(torch.int32)
            # finite values only: infinities and NaN are counted per group
            fin = torch.where(torch.isfinite(self.kv_values), self.kv_values,
                              torch.zeros_like(self.kv_values))
            self.prefix = torch.zeros(self.n + 1, dtype=torch.float64, device=dev)
            self.prefix[1:] = segmented_prefix(fin.double(), self.group_starts)
            # the index columns in jfx_served_query field order
            self.index_ptrs = (self.kv_values.data_ptr(), self.prefix.data_ptr(),
                               self.group_starts.data_ptr(), self.group_nan_start.data_ptr(),
                               self.group_ninf_end.data_ptr(), self.group_pinf_start.data_ptr(),
                               self.group_keys.data_ptr())
            parts += [self.kv_values, self.group_nan_start, self.group_pinf_start,
                      self.group_ninf_end, self.prefix]
        self.nbytes = sum(t.numel() * t.element_size() for t in parts)


def segmented_prefix(x: torch.Tensor, starts: torch.Tensor) -> torch.Tensor:
    """Inclusive prefix sums of x restarted at every group (groups are the
    row ranges [starts[g], starts[g + 1])): out[i] = sum of x[starts[g] .. i]
    for the group g of row i. A log-step (Hillis-Steele) scan whose adds
    only ever combine partial sums of one group, so each result's rounding
    error is bounded by the group's own values (about log2(group rows) *
    2^-53 * sum |x| over the group), whatever the other groups hold; a
    global cumsum differenced at the group start would carry the rounding of
    the running total of every earlier group."""
    n = x.shape[0]
    if n == 0:
        return x.clone()
    st = starts.long()
    sizes = st[1:] - st[:-1]
    row_start = torch.repeat_interleave(st[:-1], sizes)
    pos = torch.arange(n, device=x.device) - row_start     # row's offset in its group
    out = x.clone()
    d = 1
    longest = int(sizes.max())
    while d < longest:
        add = torch.zeros_like(out)
        add[d:] = out[:-d]
        out = out + torch.where(pos >= d, add, torch.zeros_like(add))
        d *= 2
    return out


class FrameStore:
    """Resident frames with HBM accounting against a budget."""

    def __init__(self, hbm_budget_bytes: int):
        self.budget = int(hbm_budget_bytes)
        self.frames: dict[str, Frame] = {}
        self.lock = threading.Lock()

    def load(self, name: str, keys, values) -> Frame:
        fr = Frame(name, keys, values)
        with self.lock:
            if self.resident_bytes() + fr.nbytes > self.budget:
                raise MemoryError(f"frame {name} ({fr.nbytes} B) does not fit the HBM budget")
            self.frames[name] = fr
        return fr

    def resident_bytes(self) -> int:
        return sum(f.nbytes for f in self.frames.values())

    def __getitem__(self, name: str) -> Frame:
        return self.frames[name]


# ---- queries -----------------------------------------------------------------------

@dataclass
class FilterGroupBy:
    frame: str
    op: str
    threshold: float
    aggs: tuple = ("sum", "count", "mean")


@dataclass
class Join:
    frame: str
    dim_keys: torch.Tensor
    dim_values: torch.Tensor
    how: str = "inner"


@dataclass
class Request:
    query: object
    arrival: float
    future: Future = field(default_factory=Future)
    rid: int = 0
    footprint: int = 0
    t_admit: float = 0.0     # perf_counter() when a slot took it
    t_done: float = 0.0      # when its result was ready on the host
    deadline: float = float("inf")   # perf_counter() by which it must be admitted


def footprint(store: FrameStore, q) -> int:
    """Upper bound of the device bytes a query allocates while it runs."""
    fr = store[q.frame]
    kb, vb = fr.keys.element_size(), fr.values.element_size()
    if isinstance(q, FilterGroupBy):
        if fr.indexed:
            # per-group key, count, sum and mean outputs, lookback scratch
            return fr.ngroups * (kb + 8 + 2 * vb + 1) + (1 << 20)
        # positions, gathered keys + values, group outputs, segreduce scratch
        return fr.n * (4 + kb + vb) + fr.n * (kb + 3 * 8) + fr.n * 16 + (1 << 22)
    nd = q.dim_keys.shape[0]
    # dimension codes, their sort and per-group runs and offsets; the pairs
    # (key + 2 payloads) are bounded by n when the dimension keys are unique
    return nd * (3 * 4 + 2 * 8) + fr.ngroups * (8 + 8) + fr.n * (kb + 2 * vb) + (1 << 22)


# ---- executors -----------------------------------------------------------------------

def _segreduce(sk, sv, aggs):
    """Head-flag segmented reduction of key-ordered rows, every requested
    aggregation in one pass."""
    from .grouping import _VAL_CODE, _alloc_out, _segreduce_ws
    n = sk.shape[0]
    dev = sk.device
    kc = key_code(sk)
    vcode = _VAL_CODE[sv.dtype]
    out_keys = torch.empty(max(n, 1), dtype=sk.dtype, device=dev)
    outs = {fn: _alloc_out(fn, sv, max(n, 1), dev) for fn in aggs}
    G = 0
    if n:
        ws = _segreduce_ws(vcode, n, dev)
        g = ctypes.c_uint64(0)
        ptrs = [_lib.ptr(outs.get(f)) for f in ("sum", "count", "min", "max", "mean")]
        _lib.call("jfx_segreduce", kc, _lib.ptr(sk), vcode, _lib.ptr(sv), None, n, _lib.ptr(out_keys),
                  *ptrs, ctypes.byref(g), _lib.ptr(ws), ws.numel(),
                  ctypes.c_void_p(_lib.stream_handle(dev)))
        G = g.value
    return out_keys[:G].clone(), {fn: o[:G].clone() for fn, o in outs.items()}


_INDEXED_AGGS = {"sum", "count", "mean"}
_SERVED_QMAX = 1024


class ServedQuery(ctypes.Structure):
    """jfx_served_query: one indexed filter -> groupby query of a batch."""
    _fields_ = [("values", ctypes.c_void_p), ("prefix", ctypes.c_void_p),
                ("group_starts", ctypes.c_void_p), ("nan_start", ctypes.c_void_p),
                ("ninf_end", ctypes.c_void_p), ("pinf_start", ctypes.c_void_p),
                ("group_keys", ctypes.c_void_p), ("ngroups", ctypes.c_uint32),
                ("cmp", ctypes.c_int32), ("threshold", ctypes.c_double)]


def _served_batch(store: FrameStore, qs: list) -> list:
    """Indexed filter -> groupby queries on frames with one (key type, value
    type), all in one jfx_served_batch launch; [(keys, {agg: column})]."""
    from .compaction import OPS
    frs = [store[q.frame] for q in qs]
    dev = frs[0].keys.device
    kdt, vdt = frs[0].group_keys.dtype, frs[0].kv_values.dtype
    Q = len(qs)
    table = (ServedQuery * Q)()
    tot = 0
    for i, (q, fr) in enumerate(zip(qs, frs)):
        if q.op not in OPS:
            raise ValueError(f"unsupported comparison {q.op!r}")
        table[i] = ServedQuery(*fr.index_ptrs, fr.ngroups, OPS[q.op], float(q.threshold))
        tot += fr.ngroups
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_served_batch_workspace_size", Q, tot, ctypes.byref(nb))
    ws = _lib.workspace(nb.value, dev)
    cap = max(tot, 1)
    keys = torch.empty(cap, dtype=kdt, device=dev)
    cnt = torch.empty(cap, dtype=torch.int64, device=dev)
    sm = torch.empty(cap, dtype=vdt, device=dev)
    mean = torch.empty(cap, dtype=vdt, device=dev)
    starts = (ctypes.c_uint64 * (Q + 1))()
    vt = _lib.JFX_F32 if vdt == torch.float32 else _lib.JFX_F64
    _lib.call("jfx_served_batch", vt, keys.element_size(), table, Q, _lib.ptr(keys), _lib.ptr(cnt),
              _lib.ptr(sm), _lib.ptr(mean), starts, _lib.ptr(ws), ws.numel(),
              ctypes.c_void_p(_lib.stream_handle(dev)))
    full = {"sum": sm, "count": cnt, "mean": mean}
    out = []
    for i, q in enumerate(qs):
        a, b = starts[i], starts[i + 1]
        out.append((keys[a:b], {agg: full[agg][a:b] for agg in q.aggs}))
    return out


def run_filter_groupby(store: FrameStore, q: FilterGroupBy):
    """filter -> groupby. On an indexed frame: one launch that evaluates every
    group by binary search and prefix sums and writes the groups that keep
    rows compacted (jfx_served_batch with one query).
    Otherwise on the frame's key order: the rows passing the filter are
    already grouped, so a segmented reduction over them finishes the query
    (no sort)."""
    from .compaction import filter_rows, take
    fr = store[q.frame]
    if fr.indexed and set(q.aggs) <= _INDEXED_AGGS:
        return _served_batch(store, [q])[0]
    pos = filter_rows(fr.sorted_values, q.op, q.threshold)
    sk = take(fr.sorted_keys, pos)
    sv = take(fr.sorted_values, pos)
    return _segreduce(sk, sv, q.aggs)


def run_join(store: FrameStore, q: Join):
    """Inner join of a resident frame with a dimension table: only the
    dimension side is sorted; the frame's key order is reused."""
    from .joining import fill_values, payload_source
    from .sorting import radix_sort
    if q.how != "inner":
        raise ValueError("served joins are inner joins")
    fr = store[q.frame]
    dev = fr.keys.device
    dk, _ = widen_key(to_device(q.dim_keys, dev))
    if dk.dtype != fr.keys.dtype:
        dk = dk.to(fr.keys.dtype)
    kc = key_code(fr.keys)
    if fr.ngroups and kc in _INT_KEYS:
        return _join_by_groups(fr, dk, q.dim_values, kc)
    sd, _, dperm = radix_sort(dk, None, True, return_index=True, index_dtype=torch.int32)
    L, R = fr.n, dk.shape[0]
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_join_workspace_size", L, ctypes.byref(nb))
    ws = _lib.workspace(nb.value, dev)
    stream = ctypes.c_void_p(_lib.stream_handle(dev))
    m = ctypes.c_uint64(0)
    _lib.call("jfx_join_count", kc, _lib.ptr(fr.sorted_keys), L, _lib.ptr(sd), R, ctypes.byref(m),
              _lib.ptr(ws), ws.numel(), stream)
    M = m.value
    # the frame's payload is already in key order: the fill reads it by sorted
    # position (nearly sequential); the dimension payload is read by row
    keys, lv, rv, _, _ = fill_values(kc, fr.sorted_keys, None, L, sd, dperm, R, M,
                                     payload_source(fr.sorted_values, dev), False,
                                     payload_source(q.dim_values, dev), True, ws,
                                     keys_dtype=fr.keys.dtype)
    return keys, {"left_value": lv, "right_value": rv}


_INT_KEYS = (_lib.JFX_U32, _lib.JFX_I32, _lib.JFX_U64, _lib.JFX_I64)


def _join_by_groups(fr: Frame, dk: torch.Tensor, dim_values, kc: int):
    """Inner join over the frame's key groups (jfx_served_join_count / _fill):
    dimension rows are mapped to groups and sorted by group with a fixed
    plan; each group's rows are contiguous in key order, so the fill streams
    the frame's key-ordered payload."""
    from .joining import payload_source
    dev = dk.device
    R, G = dk.shape[0], fr.ngroups
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_served_join_workspace_size", R, G, ctypes.byref(nb))
    ws = _lib.workspace(nb.value, dev)
    stream = ctypes.c_void_p(_lib.stream_handle(dev))
    m = ctypes.c_uint64(0)
    _lib.call("jfx_served_join_count", kc, _lib.ptr(fr.group_keys), _lib.ptr(fr.group_starts), G,
              _lib.ptr(dk), R, ctypes.byref(m), _lib.ptr(ws), ws.numel(), stream)
    M = m.value
    lsrc = payload_source(fr.sorted_values, dev)
    rsrc = payload_source(dim_values, dev)
    if rsrc.shape[0] != R:
        raise ValueError("dimension keys and values must have the same length")
    keys = torch.empty(M, dtype=fr.keys.dtype, device=dev)
    lv = torch.empty(M, dtype=lsrc.dtype, device=dev)
    rv = torch.empty(M, dtype=rsrc.dtype, device=dev)
    if M:
        _lib.call("jfx_served_join_fill", kc, _lib.ptr(fr.group_keys), _lib.ptr(fr.group_starts), G,
                  lsrc.element_size(), _lib.ptr(lsrc), R, rsrc.element_size(), _lib.ptr(rsrc), M,
                  _lib.ptr(keys), _lib.ptr(lv), _lib.ptr(rv), _lib.ptr(ws), ws.numel(), stream)
    return keys, {"left_value": lv, "right_value": rv}


def execute_filter_batch(store: FrameStore, qs: list) -> list:
    """Several indexed filter -> groupby queries as one batch: one
    jfx_served_batch launch evaluates every group of every query and writes
    the groups that keep rows compacted, with one host synchronisation
    however many queries the batch holds. Results equal run_filter_groupby's
    query by query."""
    frs = [store[q.frame] for q in qs]
    kinds = {(fr.group_keys.dtype, fr.kv_values.dtype) for fr in frs}
    out = [None] * len(qs)
    for kind in kinds:   # one launch per (key type, value type) and per 1024 queries
        idx = [i for i, fr in enumerate(frs) if (fr.group_keys.dtype, fr.kv_values.dtype) == kind]
        for a in range(0, len(idx), _SERVED_QMAX):
            part = idx[a:a + _SERVED_QMAX]
            for i, r in zip(part, _served_batch(store, [qs[i] for i in part])):
                out[i] = r
    return out


def batchable(store: FrameStore, q) -> bool:
    """Queries execute_filter_batch can take."""
    return (isinstance(q, FilterGroupBy) and store[q.frame].indexed
            and set(q.aggs) <= _INDEXED_AGGS)


def execute(store: FrameStore, q):
    if isinstance(q, FilterGroupBy):
        return run_filter_groupby(store, q)
    if isinstance(q, Join):
        return run_join(store, q)
    raise TypeError(f"unknown query {type(q).__name__}")


def run_plain(store: FrameStore, q):
    """The plain per-query path the server is measured against: the same
    operators on the frame as stored, one query at a time, nothing reused
    (filter, gather, full groupby with its sort; join with both sides sorted)."""
    from .compaction import filter_rows, take
    from .grouping import groupby_aggregate
    from .joining import sort_merge_join
    fr = store[q.frame]
    if isinstance(q, FilterGroupBy):
        pos = filter_rows(fr.values, q.op, q.threshold)
        k = take(fr.keys, pos)
        v = take(fr.values, pos)
        return groupby_aggregate(k, {a: v for a in q.aggs}, {a: a for a in q.aggs})
    keys, out = sort_merge_join(fr.keys, {"value": fr.values}, q.dim_keys, {"value": q.dim_values},
                                q.how)
    return keys, out


# ---- the request loop ------------------------------------------------------------------

class Server:
    """Continuous-batching request loop.

    submit() queues a request (FIFO); a request is admitted when a slot is
    free and its footprint fits budget - (footprints of running requests).
    Each slot is a worker thread with its own stream; executor(store, query)
    runs the query there and the result is ready when the stream has
    drained. A request whose footprint exceeds the whole budget fails at
    once with MemoryError.
    """

    def __init__(self, store: FrameStore, slots: int = 4, hbm_budget_bytes: int | None = None,
                 executor=execute, use_streams: bool = True, footprint_fn=footprint,
                 batch_executor=execute_filter_batch, batchable_fn=batchable,
                 max_batch: int = _SERVED_QMAX,
                 switch_interval: float | None = None):
        # Slot threads spend most of their time in GPU waits, which release
        # the GIL; a thread whose wait has ended must get the GIL back before
        # it can launch the next step. With the interpreter's default 5 ms
        # switch interval that hand-off, not the GPU, can bound throughput
        # (measured 4.0-5.0x the plain path instead of 5.6-5.8x). A caller
        # that owns the process may pass switch_interval (e.g. 1e-4): the
        # interpreter-wide setting is changed while the server is open and
        # restored by close(); by default it is left alone.
        self._prev_switch = sys.getswitchinterval()
        if switch_interval is not None:
            sys.setswitchinterval(switch_interval)
        self.store = store
        self.slots = slots
        self.budget = int(hbm_budget_bytes if hbm_budget_bytes is not None else store.budget)
        self.executor = executor
        self.batch_executor = batch_executor
        self.batchable_fn = batchable_fn
        self.max_batch = max_batch
        self.batches = []   # sizes of the batches run
        self.requests: list[Request] = []   # every request submitted, for timing reports
        self.footprint_fn = footprint_fn
        self.cv = threading.Condition()
        self.pending: deque[Request] = deque()
        self.running = 0
        self.reserved = 0
        self.peak_running = 0
        self.peak_reserved = 0
        self.admitted_order: list[int] = []
        self.closed = False
        self._timed = False   # some request was given an admission timeout
        self._rid = 0
        # one stream per slot, made here so warmup() can prime each stream's
        # allocator pool before requests arrive
        self.streams = [torch.cuda.Stream() if use_streams else None for _ in range(slots)]
        self.workers = [threading.Thread(target=self._worker, args=(i,), daemon=True)
                        for i in range(slots)]
        for w in self.workers:
            w.start()

    def submit(self, query, arrival: float | None = None, timeout: float | None = None) -> Future:
        """Queue a query; the future resolves to (result, latency_s). timeout:
        seconds the request may wait for admission before it fails with
        TimeoutError. future.cancel() withdraws a request not yet admitted."""
        fp = self.footprint_fn(self.store, query)
        with self.cv:
            self._rid += 1
            now = time.perf_counter()
            req = Request(query, now if arrival is None else arrival, rid=self._rid, footprint=fp,
                          deadline=now + timeout if timeout is not None else float("inf"))
            if timeout is not None:
                self._timed = True
            if fp > self.budget:
                req.future.set_exception(MemoryError(
                    f"request {req.rid} needs {fp} B, more than the {self.budget} B budget"))
                return req.future
            self.pending.append(req)
            self.requests.append(req)
            self.cv.notify_all()
        return req.future

    def _expire(self):
        # called with the lock held: fail the requests whose admission
        # deadline has passed (they never reach a slot) and drop cancelled
        # ones. Only servers that were given a timeout pay for the sweep;
        # otherwise cancelled requests are dropped when they are popped.
        if not self._timed:
            return
        now = time.perf_counter()
        keep = deque()
        for r in self.pending:
            if r.future.cancelled():
                continue
            if now > r.deadline:
                r.future.set_exception(TimeoutError(
                    f"request {r.rid} was not admitted within its timeout"))
                continue
            keep.append(r)
        self.pending = keep

    def _admit(self) -> Request | None:
        # called with the lock held: FIFO head, if a slot and the memory allow
        self._expire()
        if not self.pending or self.running >= self.slots:
            return None
        head = self.pending[0]
        if self.reserved + head.footprint > self.budget:
            return None
        self.pending.popleft()
        if not head.future.set_running_or_notify_cancel():
            return self._admit()      # cancelled between the sweep and now
        self.running += 1
        self.reserved += head.footprint
        self.peak_running = max(self.peak_running, self.running)
        self.peak_reserved = max(self.peak_reserved, self.reserved)
        self.admitted_order.append(head.rid)
        return head

    def _take_batch(self, head: Request) -> list:
        # called with the lock held after head (the FIFO head) was admitted:
        # queued batchable requests join its batch, in arrival order, while
        # they fit the budget. Requests they pass over keep their places; the
        # FIFO head is always admitted first, so none of them starves.
        batch = [head]
        if self.batch_executor is None or not self.batchable_fn(self.store, head.query):
            return batch
        keep = deque()
        while self.pending and len(batch) < self.max_batch:
            r = self.pending.popleft()
            if r.future.cancelled():
                continue
            if (self.batchable_fn(self.store, r.query)
                    and self.reserved + r.footprint <= self.budget
                    and r.future.set_running_or_notify_cancel()):
                self.reserved += r.footprint
                self.peak_reserved = max(self.peak_reserved, self.reserved)
                self.admitted_order.append(r.rid)
                batch.append(r)
            else:
                keep.append(r)
        keep.extend(self.pending)
        self.pending = keep
        return batch

    def _run(self, batch: list):
        if len(batch) == 1:
            return [self.executor(self.store, batch[0].query)]
        return self.batch_executor(self.store, [r.query for r in batch])

    def warmup(self, queries: list):
        """Run queries once on every slot's stream (single queries and, for
        the batchable ones, one batch), so the first requests do not pay for
        device allocations; results are discarded."""
        batchable_qs = [q for q in queries if self.batchable_fn(self.store, q)]
        for s in self.streams:
            if s is None:
                continue
            with torch.cuda.stream(s):
                for q in queries:
                    self.executor(self.store, q)
                if self.batch_executor is not None and len(batchable_qs) > 1:
                    self.batch_executor(self.store, batchable_qs)
            s.synchronize()

    def _worker(self, i: int):
        stream = self.streams[i]
        while True:
            with self.cv:
                req = self._admit()
                while req is None:
                    if self.closed and not self.pending:
                        return
                    wait = None
                    if self._timed:   # wake up for the nearest admission deadline
                        nxt = min((r.deadline for r in self.pending), default=float("inf"))
                        if nxt != float("inf"):
                            wait = max(nxt - time.perf_counter(), 0.0) + 1e-4
                    self.cv.wait(wait)
                    req = self._admit()
                batch = self._take_batch(req)
                self.batches.append(len(batch))
            t_admit = time.perf_counter()
            for r in batch:
                r.t_admit = t_admit
            try:
                if stream is not None:
                    with torch.cuda.stream(stream):
                        results = self._run(batch)
                    stream.synchronize()
                else:
                    results = self._run(batch)
                done = time.perf_counter()
                for r, res in zip(batch, results):
                    r.t_done = done
                    r.future.set_result((res, done - r.arrival))
            except BaseException as e:  # delivered to the callers through the futures
                for r in batch:
                    if not r.future.done():
                        r.future.set_exception(e)
            finally:
                with self.cv:
                    self.running -= 1
                    self.reserved -= sum(r.footprint for r in batch)
                    self.cv.notify_all()

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for w in self.workers:
            w.join()
        sys.setswitchinterval(self._prev_switch)


def poisson_arrivals(n: int, qps: float, seed: int = 0):
    """n arrival offsets (seconds) of a Poisson process of rate qps."""
    g = torch.Generator()
    g.manual_seed(seed)
    gaps = torch.empty(n, dtype=torch.float64).exponential_(qps, generator=g)
    return torch.cumsum(gaps, 0).tolist()


def replay(server: Server, queries: list, arrivals: list):
    """Submit queries at their arrival offsets; returns [(result, latency_s)]."""
    t0 = time.perf_counter()
    futs = []
    for q, at in zip(queries, arrivals):
        delay = t0 + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        futs.append(server.submit(q, arrival=t0 + at))
    return [f.result() for f in futs]
