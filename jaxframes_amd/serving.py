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
      key group (NaN last), float64 prefix sums of them, group starts, the
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
        # identity rows: a join fill then reports sorted positions, and the left
        # payload comes from sorted_values in (nearly) sequential order
        self.iota = torch.arange(self.n, dtype=torch.int32, device=dev)
        parts = [self.keys, self.values, self.sorted_keys, self.perm, self.sorted_values, self.iota]
        self.indexed = self.values.dtype in (torch.float32, torch.float64) and self.n > 0
        if self.indexed:
            # (key, value) order: stable by value, then stable by key
            _, _, pv = radix_sort(self.values, None, True, want_keys=False, return_index=True)
            _, _, p2 = radix_sort(self.keys[pv], None, True, want_keys=False, return_index=True)
            kv = pv[p2]
            self.kv_values = self.values[kv]
            gid, starts = multikey_heads([sk])
            G = starts.shape[0]
            self.group_keys = sk[starts]
            self.group_starts = torch.empty(G + 1, dtype=torch.int32, device=dev)
            self.group_starts[:G] = starts.to(torch.int32)
            self.group_starts[G] = self.n
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
            self.prefix[1:] = torch.cumsum(fin.double(), 0)
            self.ngroups = G
            parts += [self.kv_values, self.group_keys, self.group_starts, self.group_nan_start,
                      self.group_pinf_start, self.group_ninf_end, self.prefix]
        self.nbytes = sum(t.numel() * t.element_size() for t in parts)


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


def footprint(store: FrameStore, q) -> int:
    """Upper bound of the device bytes a query allocates while it runs."""
    fr = store[q.frame]
    kb, vb = fr.keys.element_size(), fr.values.element_size()
    if isinstance(q, FilterGroupBy):
        if fr.indexed:
            # per-group count / sum / mean / mask, positions, the gathered groups
            return fr.ngroups * (8 + 2 * vb + 1 + 4 + kb + 8 + 2 * vb) + (1 << 20)
        # positions, gathered keys + values, group outputs, segreduce scratch
        return fr.n * (4 + kb + vb) + fr.n * (kb + 3 * 8) + fr.n * 16 + (1 << 22)
    nd = q.dim_keys.shape[0]
    # sorted dimension side, join scratch; pairs bounded by n * duplicates of
    # the dimension keys (unique dimension keys: at most one pair per row)
    return nd * (2 * kb + 8 + 4) + fr.n * (4 + 4 + kb + 2 * vb) + (1 << 22)


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


def run_filter_groupby(store: FrameStore, q: FilterGroupBy):
    """filter -> groupby. On an indexed frame: one binary-search pass over the
    groups (jfx_served_groups) and a compaction of the groups that keep rows.
    Otherwise on the frame's key order: the rows passing the filter are
    already grouped, so a segmented reduction over them finishes the query
    (no sort)."""
    from .compaction import OPS, compact_rows, filter_rows, take
    fr = store[q.frame]
    if fr.indexed and set(q.aggs) <= _INDEXED_AGGS:
        if q.op not in OPS:
            raise ValueError(f"unsupported comparison {q.op!r}")
        dev = fr.keys.device
        G = fr.ngroups
        vt = _lib.JFX_F32 if fr.kv_values.dtype == torch.float32 else _lib.JFX_F64
        cnt = torch.empty(G, dtype=torch.int64, device=dev)
        sm = torch.empty(G, dtype=fr.kv_values.dtype, device=dev)
        mean = torch.empty(G, dtype=fr.kv_values.dtype, device=dev)
        mask = torch.empty(G, dtype=torch.uint8, device=dev)
        val = (ctypes.c_float if vt == _lib.JFX_F32 else ctypes.c_double)(q.threshold)
        _lib.call("jfx_served_groups", vt, _lib.ptr(fr.kv_values), _lib.ptr(fr.prefix),
                  _lib.ptr(fr.group_starts), _lib.ptr(fr.group_nan_start),
                  _lib.ptr(fr.group_ninf_end), _lib.ptr(fr.group_pinf_start), G, OPS[q.op],
                  ctypes.byref(val), _lib.ptr(cnt), _lib.ptr(sm), _lib.ptr(mean), _lib.ptr(mask),
                  ctypes.c_void_p(_lib.stream_handle(dev)))
        pos = compact_rows(mask)
        full = {"sum": sm, "count": cnt, "mean": mean}
        return take(fr.group_keys, pos), {a: take(full[a], pos) for a in q.aggs}
    pos = filter_rows(fr.sorted_values, q.op, q.threshold)
    sk = take(fr.sorted_keys, pos)
    sv = take(fr.sorted_values, pos)
    return _segreduce(sk, sv, q.aggs)


def run_join(store: FrameStore, q: Join):
    """Inner join of a resident frame with a dimension table: only the
    dimension side is sorted; the frame's key order is reused."""
    from .joining import gather
    from .sorting import radix_sort
    if q.how != "inner":
        raise ValueError("served joins are inner joins")
    fr = store[q.frame]
    dev = fr.keys.device
    dk, _ = widen_key(to_device(q.dim_keys, dev))
    if dk.dtype != fr.keys.dtype:
        dk = dk.to(fr.keys.dtype)
    sd, _, dperm = radix_sort(dk, None, True, return_index=True, index_dtype=torch.int32)
    kc = key_code(fr.keys)
    L, R = fr.n, dk.shape[0]
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_join_workspace_size", L, ctypes.byref(nb))
    ws = _lib.workspace(nb.value, dev)
    stream = ctypes.c_void_p(_lib.stream_handle(dev))
    m = ctypes.c_uint64(0)
    _lib.call("jfx_join_count", kc, _lib.ptr(fr.sorted_keys), L, _lib.ptr(sd), R, ctypes.byref(m),
              _lib.ptr(ws), ws.numel(), stream)
    M = m.value
    li = torch.empty(M, dtype=torch.int32, device=dev)
    ri = torch.empty(M, dtype=torch.int32, device=dev)
    keys = torch.empty(M, dtype=fr.keys.dtype, device=dev)
    if M:
        # left rows as sorted positions (iota): pairs keep the reference's
        # (key, left row, right row) order because the key order is stable
        _lib.call("jfx_join_fill", kc, _lib.ptr(fr.sorted_keys), _lib.ptr(fr.iota), L, _lib.ptr(sd),
                  _lib.ptr(dperm), R, _lib.ptr(li), _lib.ptr(ri), _lib.ptr(keys), _lib.ptr(ws),
                  ws.numel(), stream)
    return keys, {"left_value": gather(fr.sorted_values, li),
                  "right_value": gather(to_device(q.dim_values, dev), ri)}


def execute_filter_batch(store: FrameStore, qs: list) -> list:
    """Several indexed filter -> groupby queries as one batch: per-query
    group passes into slices of shared buffers, ONE compaction of all the
    groups that keep rows and one gather per output column, so the batch
    costs a fixed number of host synchronisations however many queries it
    holds. Results equal run_filter_groupby's query by query."""
    from .compaction import OPS, compact_rows, take
    frs = [store[q.frame] for q in qs]
    kinds = {(fr.group_keys.dtype, fr.kv_values.dtype) for fr in frs}
    if len(kinds) > 1:   # one sub-batch per (key type, value type)
        out = [None] * len(qs)
        for kind in kinds:
            idx = [i for i, fr in enumerate(frs) if (fr.group_keys.dtype, fr.kv_values.dtype) == kind]
            for i, r in zip(idx, execute_filter_batch(store, [qs[i] for i in idx])):
                out[i] = r
        return out
    dev = frs[0].keys.device
    vdt = frs[0].kv_values.dtype
    Gs = [fr.ngroups for fr in frs]
    offs = [0]
    for G in Gs:
        offs.append(offs[-1] + G)
    tot = offs[-1]
    cnt = torch.empty(tot, dtype=torch.int64, device=dev)
    sm = torch.empty(tot, dtype=vdt, device=dev)
    mean = torch.empty(tot, dtype=vdt, device=dev)
    mask = torch.empty(tot, dtype=torch.uint8, device=dev)
    vt = _lib.JFX_F32 if vdt == torch.float32 else _lib.JFX_F64
    ctype = ctypes.c_float if vt == _lib.JFX_F32 else ctypes.c_double
    stream = ctypes.c_void_p(_lib.stream_handle(dev))
    for q, fr, o, G in zip(qs, frs, offs, Gs):
        if q.op not in OPS:
            raise ValueError(f"unsupported comparison {q.op!r}")
        val = ctype(q.threshold)
        _lib.call("jfx_served_groups", vt, _lib.ptr(fr.kv_values), _lib.ptr(fr.prefix),
                  _lib.ptr(fr.group_starts), _lib.ptr(fr.group_nan_start),
                  _lib.ptr(fr.group_ninf_end), _lib.ptr(fr.group_pinf_start), G, OPS[q.op],
                  ctypes.byref(val), _lib.ptr(cnt[o:]), _lib.ptr(sm[o:]), _lib.ptr(mean[o:]),
                  _lib.ptr(mask[o:]), stream)
    pos = compact_rows(mask)
    keys_all = torch.cat([fr.group_keys for fr in frs])
    cuts = torch.searchsorted(pos, torch.tensor(offs, dtype=torch.int64, device=dev)).tolist()
    k, c, s_, m = take(keys_all, pos), take(cnt, pos), take(sm, pos), take(mean, pos)
    full = {"sum": s_, "count": c, "mean": m}
    out = []
    for i, q in enumerate(qs):
        a, b = cuts[i], cuts[i + 1]
        out.append((k[a:b], {agg: full[agg][a:b] for agg in q.aggs}))
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
                 batch_executor=execute_filter_batch, batchable_fn=batchable, max_batch: int = 32):
        self.store = store
        self.slots = slots
        self.budget = int(hbm_budget_bytes if hbm_budget_bytes is not None else store.budget)
        self.executor = executor
        self.batch_executor = batch_executor
        self.batchable_fn = batchable_fn
        self.max_batch = max_batch
        self.batches = []   # sizes of the batches run
        self.footprint_fn = footprint_fn
        self.cv = threading.Condition()
        self.pending: deque[Request] = deque()
        self.running = 0
        self.reserved = 0
        self.peak_running = 0
        self.peak_reserved = 0
        self.admitted_order: list[int] = []
        self.closed = False
        self._rid = 0
        self.workers = [threading.Thread(target=self._worker, args=(i, use_streams), daemon=True)
                        for i in range(slots)]
        for w in self.workers:
            w.start()

    def submit(self, query, arrival: float | None = None) -> Future:
        fp = self.footprint_fn(self.store, query)
        with self.cv:
            self._rid += 1
            req = Request(query, time.perf_counter() if arrival is None else arrival, rid=self._rid,
                          footprint=fp)
            if fp > self.budget:
                req.future.set_exception(MemoryError(
                    f"request {req.rid} needs {fp} B, more than the {self.budget} B budget"))
                return req.future
            self.pending.append(req)
            self.cv.notify_all()
        return req.future

    def _admit(self) -> Request | None:
        # called with the lock held: FIFO head, if a slot and the memory allow
        if not self.pending or self.running >= self.slots:
            return None
        head = self.pending[0]
        if self.reserved + head.footprint > self.budget:
            return None
        self.pending.popleft()
        self.running += 1
        self.reserved += head.footprint
        self.peak_running = max(self.peak_running, self.running)
        self.peak_reserved = max(self.peak_reserved, self.reserved)
        self.admitted_order.append(head.rid)
        return head

    def _take_batch(self, head: Request) -> list:
        # called with the lock held after head was admitted: the batchable
        # requests queued right behind it join its batch while they fit
        batch = [head]
        if self.batch_executor is None or not self.batchable_fn(self.store, head.query):
            return batch
        while (self.pending and len(batch) < self.max_batch
               and self.batchable_fn(self.store, self.pending[0].query)
               and self.reserved + self.pending[0].footprint <= self.budget):
            r = self.pending.popleft()
            self.reserved += r.footprint
            self.peak_reserved = max(self.peak_reserved, self.reserved)
            self.admitted_order.append(r.rid)
            batch.append(r)
        return batch

    def _run(self, batch: list):
        if len(batch) == 1:
            return [self.executor(self.store, batch[0].query)]
        return self.batch_executor(self.store, [r.query for r in batch])

    def _worker(self, i: int, use_streams: bool):
        stream = torch.cuda.Stream() if use_streams else None
        while True:
            with self.cv:
                req = self._admit()
                while req is None:
                    if self.closed and not self.pending:
                        return
                    self.cv.wait()
                    req = self._admit()
                batch = self._take_batch(req)
                self.batches.append(len(batch))
            try:
                if stream is not None:
                    with torch.cuda.stream(stream):
                        results = self._run(batch)
                    stream.synchronize()
                else:
                    results = self._run(batch)
                done = time.perf_counter()
                for r, res in zip(batch, results):
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
