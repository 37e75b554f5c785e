"""Batched groupby: thousands of independent groupbys in one device pipeline.

Input: one concatenated table (keys, values) with frame offsets; frame f is
rows offsets[f]:offsets[f+1]. Output: every frame's groups, frame-major, as
SortBasedGroupBy.groupby_aggregate (parallel_algorithms.py:733-803) would give
frame by frame, with a CSR offset array per frame.

The device work is jfx_batched_groupby (composite (frame, key) keys, one radix
sort, one segmented reduction) and has no host synchronisation, so a plan can
be captured once in a HIP graph and replayed; the digit width and the tile
shape of the sort are picked per size class by a small autotuner that times
the candidates on the first call. When enough rows belong to frames whose keys
span fewer than 1024 values (measured once when the plan is made), the plan
runs jfx_batched_groupby_routed instead: those frames are reduced in one-pass
LDS tables and only the other frames' rows are sorted.
"""

from __future__ import annotations

import ctypes
import json
import os
import threading
import time

import numpy as np
import torch

from . import _lib
from ._tensors import key_code, to_device, widen_key
from .grouping import _VAL_CODE, _alloc_out, _value_tensor

AGGS = ("sum", "count", "min", "max", "mean")


# onesweep tile shapes of jfx_batched_groupby (sort_shape): (waves, keys per
# lane) for 32-bit and for 64-bit composite keys
SORT_SHAPES = {0: ((16, 8), (16, 16)), 1: ((8, 16), (16, 8)), 2: ((8, 8), (4, 16)),
               3: ((4, 16), (8, 8)), 4: ((16, 16), (16, 16))}


class Autotuner:
    """Best sort configuration -- digit width (8 or 11 bits per radix pass)
    and onesweep tile shape (waves per workgroup x keys per lane, which sets
    the tile size) -- per (row-count class, composite bit width), measured
    on first use by timing every candidate. Thread-safe; optionally persisted
    as JSON (path from JFX_TUNE_FILE)."""

    # shape 4 (16384-key tiles) exists for 8-bit digits only
    CANDIDATES = tuple((d, sh) for d in (8, 11) for sh in sorted(SORT_SHAPES)
                       if not (d == 11 and sh == 4))

    def __init__(self, path: str | None = None):
        self._lock = threading.Lock()
        self.path = path if path is not None else os.environ.get("JFX_TUNE_FILE")
        self.table: dict[str, tuple] = {}
        self.timings: dict[str, dict] = {}
        if self.path and os.path.exists(self.path):
            with open(self.path) as f:
                for k, v in json.load(f).items():
                    self.table[k] = (int(v), 0) if isinstance(v, int) else (int(v[0]), int(v[1]))

    @staticmethod
    def key(n: int, total_bits: int) -> str:
        return f"n2^{max(n, 1).bit_length()}:bits{total_bits}"

    def lookup(self, n: int, total_bits: int):
        with self._lock:
            return self.table.get(self.key(n, total_bits))

    def tune(self, n: int, total_bits: int, run, reps: int = 2) -> tuple:
        """run(digit_bits, sort_shape) launches one batch on the current
        stream; returns the fastest (digit_bits, sort_shape)."""
        k = self.key(n, total_bits)
        with self._lock:
            if k in self.table:
                return self.table[k]
        best, best_t, times = None, None, {}
        for cand in self.CANDIDATES:
            run(*cand)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                run(*cand)
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / reps
            times[f"d{cand[0]}/shape{cand[1]}"] = round(t * 1e3, 3)
            if best_t is None or t < best_t:
                best, best_t = cand, t
        with self._lock:
            self.table[k] = best
            self.timings[k] = times
            if self.path:
                with open(self.path, "w") as f:
                    json.dump({kk: list(v) for kk, v in self.table.items()}, f)
        return best


TUNER = Autotuner()


class BatchResult:
    """Groups of every frame: offsets (int64[F+1]) into keys and the
    aggregate columns."""

    def __init__(self, offsets, keys, values: dict):
        self.offsets = offsets
        self.keys = keys
        self.values = values

    def frame(self, f: int):
        a, b = int(self.offsets[f]), int(self.offsets[f + 1])
        return self.keys[a:b], {c: v[a:b] for c, v in self.values.items()}


# rows one device pipeline takes (its row indices and lookback counts are 30-bit)
MAX_ROWS = (1 << 30) - 1

# Narrow frames take the one-pass tables when they hold at least this share of
# the rows: routing saves their sort passes (about 80 bytes of traffic per row
# at 4 passes) but copies the wide rows' values once more and places the wide
# groups after the reduction (about 8 bytes per wide row and 32 per wide group).
ROUTE_MIN_SHARE = 0.15
_OUT_BITS = {"sum": 1, "count": 2, "min": 4, "max": 8, "mean": 16}


def _key_max(kbytes: int, signed: bool) -> int:
    return (1 << (8 * kbytes - 1)) - 1 if signed else (1 << (8 * kbytes)) - 1


def _ordered_image(x: int, kbytes: int, signed: bool) -> int:
    """Order-preserving unsigned image of an integer key (the kernels' ordered())."""
    w = 8 * kbytes
    return ((x + (1 << (w - 1))) if signed else x) & ((1 << w) - 1)


class BatchedGroupBy:
    """A reusable plan for one table shape: frame offsets, key range, value
    type and aggregations fixed at construction.

    keys / values: the concatenated columns (device tensors). frame_offsets:
    int64 offsets, length nframes + 1. key_range: (min, max) of the keys over
    all frames if known, else measured once here. aggs: subset of AGGS.
    A table of more than max_rows rows is cut at frame boundaries into
    several pipelines, launched back to back (each frame must fit one).
    """

    def __init__(self, keys, values, frame_offsets, *, key_range=None, aggs=("sum",),
                 digit_bits: int = 0, sort_shape: int = 0, tuner: Autotuner | None = TUNER,
                 max_rows: int = MAX_ROWS, route: bool | None = None):
        _lib.require_gpu()
        for a in aggs:
            if a not in AGGS:
                raise ValueError(f"Unsupported aggregation function: {a}")
        self.keys, self.korig = widen_key(to_device(keys))
        self.dev = self.keys.device
        self.vals, self.vorig = _value_tensor(to_device(values, self.dev))
        self.offsets = to_device(torch.as_tensor(frame_offsets, dtype=torch.int64), self.dev)
        self.n = self.keys.shape[0]
        self.nframes = self.offsets.shape[0] - 1
        if self.vals.shape[0] != self.n or int(self.offsets[-1]) != self.n or int(self.offsets[0]) != 0:
            raise ValueError("frame offsets must run from 0 to the number of rows")
        self.kc = key_code(self.keys)
        kb = self.keys.element_size()
        signed = self.keys.dtype in (torch.int32, torch.int64)
        # per-chunk key ranges (one read of the keys): the key range when none
        # is given, and the routing plan below
        ranges = None
        if self.n > 0 and (key_range is None or (route is not False and self.n <= max_rows)):
            ranges = self._chunk_ranges()
        if key_range is None:
            if self.n == 0:
                key_range = (0, 0)
            else:
                off = (1 << (8 * kb - 1)) if signed else 0
                key_range = (int(ranges[3][:, 0].min()) - off, int(ranges[3][:, 1].max()) - off)
        self.parts = None
        self.aggs = tuple(aggs)
        self.route = None
        if self.n > max_rows:
            self._split(key_range, digit_bits, sort_shape, tuner, max_rows, route)
            return
        self.kmin, kmax = int(key_range[0]), int(key_range[1])
        span = kmax - self.kmin
        self.key_bits = max(1, span.bit_length())
        self.frame_bits = max(0, (self.nframes - 1).bit_length())
        self.total_bits = self.key_bits + self.frame_bits
        if self.total_bits > 64:
            raise ValueError("frame and key bits exceed 64")
        self.comp_dtype = torch.int32 if self.total_bits <= 32 else torch.int64
        self.vcode = _VAL_CODE[self.vals.dtype]
        self.aggs = tuple(aggs)
        cap = max(self.n, 1)
        self.out_keys = torch.empty(cap, dtype=self.comp_dtype, device=self.dev)
        self.outs = {a: _alloc_out(a, self.vals, cap, self.dev) for a in self.aggs}
        self.ngroups = torch.zeros(1, dtype=torch.int32, device=self.dev)
        if route is not False and self.vals.dtype != torch.float64 and ranges is not None:
            self._plan_route(ranges, force=route is True)
        nb = ctypes.c_uint64(0)
        if self.route is not None:
            outs = sum(_OUT_BITS[a] for a in self.aggs)
            _lib.call("jfx_batched_routed_workspace_size", self.total_bits, self.vcode,
                      self.nframes, ctypes.byref(self.route), outs, ctypes.byref(nb))
        else:
            _lib.call("jfx_batched_groupby_workspace_size", self.total_bits, self.vcode, self.n,
                      ctypes.byref(nb))
        self.ws = torch.empty(nb.value, dtype=torch.uint8, device=self.dev)
        self.graph = None
        if digit_bits:
            self.digit_bits, self.sort_shape = digit_bits, sort_shape
        elif tuner is not None:
            self.digit_bits, self.sort_shape = tuner.tune(self.n, self.total_bits, self._launch)
        else:
            self.digit_bits, self.sort_shape = 8, sort_shape

    def _chunk_ranges(self):
        """Row chunks of every frame (at most BATCH_CHUNK_ROWS rows each) and
        the min / max ordered key image of each: (frame of each chunk, first
        rows, row counts, uint64[chunks, 2])."""
        offs = self.offsets.cpu().numpy().astype(np.int64)
        rows = np.diff(offs)
        CH = _lib.BATCH_CHUNK_ROWS
        nch = (rows + CH - 1) // CH
        nc = int(nch.sum())
        fr = np.repeat(np.arange(self.nframes, dtype=np.int64), nch)
        first = np.cumsum(nch) - nch
        kin = np.arange(nc, dtype=np.int64) - np.repeat(first, nch)
        c_row0 = offs[fr] + kin * CH
        c_rows = np.minimum(CH, offs[fr + 1] - c_row0)
        d_row0 = torch.from_numpy(c_row0).to(self.dev)
        d_rows = torch.from_numpy(c_rows.astype(np.int32)).to(self.dev)
        mm = torch.empty(2 * nc, dtype=torch.int64, device=self.dev)
        s = torch.cuda.current_stream(self.dev)
        _lib.call("jfx_batched_key_ranges", self.kc, _lib.ptr(self.keys), _lib.ptr(d_row0),
                  _lib.ptr(d_rows), nc, _lib.ptr(mm), ctypes.c_void_p(s.cuda_stream))
        return fr, c_row0, c_rows, mm.cpu().numpy().view(np.uint64).reshape(-1, 2)

    def _plan_route(self, ranges, force: bool):
        """From the per-chunk key ranges: when frames spanning fewer than
        1024 values hold at least ROUTE_MIN_SHARE of the rows (or force),
        build the jfx_batched_routing arrays: those frames' row chunks and
        table offsets, the wide frames' compacted row offsets and each
        frame's narrow index."""
        fr, c_row0, c_rows, mmh = ranges
        offs = self.offsets.cpu().numpy().astype(np.int64)
        rows = np.diff(offs)
        F = self.nframes
        CH = _lib.BATCH_CHUNK_ROWS
        nch = (rows + CH - 1) // CH
        fmin = np.full(F, np.iinfo(np.uint64).max, dtype=np.uint64)
        fmax = np.zeros(F, dtype=np.uint64)
        np.minimum.at(fmin, fr, mmh[:, 0])
        np.maximum.at(fmax, fr, mmh[:, 1])
        kb = self.keys.element_size()
        signed = self.keys.dtype in (torch.int32, torch.int64)
        lo = np.uint64(_ordered_image(self.kmin, kb, signed))
        hi = np.uint64(_ordered_image(self.kmin + (1 << self.key_bits) - 1, kb, signed)
                       if self.kmin + (1 << self.key_bits) - 1 <= _key_max(kb, signed)
                       else _ordered_image(_key_max(kb, signed), kb, signed))
        has = rows > 0
        span = np.where(has, fmax - np.where(has, fmin, 0), 0)
        narrow = has & (span < _lib.BATCH_NARROW_SPAN) & (fmin >= lo) & (fmax <= hi)
        nrows = int(rows[narrow].sum())
        if nrows == 0 or (not force and nrows < ROUTE_MIN_SHARE * self.n):
            return
        nidx = np.flatnonzero(narrow)
        widx = np.flatnonzero(~narrow)
        nspan = span[nidx].astype(np.uint64)
        csel = narrow[fr]
        chunk_narrow = np.searchsorted(nidx, fr[csel])
        cent = nspan[chunk_narrow] + 1
        chunk_part = np.cumsum(cent) - cent
        mbase = np.cumsum(nspan + 1) - (nspan + 1)
        chunk0 = np.zeros(nidx.size + 1, dtype=np.int64)
        chunk0[1:] = np.cumsum(nch[nidx])
        woffs = np.zeros(widx.size + 1, dtype=np.int64)
        woffs[1:] = np.cumsum(rows[widx])
        frame_narrow = np.full(F, 0xFFFFFFFF, dtype=np.uint32)
        frame_narrow[nidx] = np.arange(nidx.size, dtype=np.uint32)
        arrays = {
            "narrow_frame": nidx.astype(np.uint32), "narrow_omin": fmin[nidx],
            "narrow_span": nspan.astype(np.uint32), "narrow_chunk0": chunk0.astype(np.uint32),
            "narrow_mbase": mbase.astype(np.uint64), "chunk_row0": c_row0[csel].astype(np.uint64),
            "chunk_rows": c_rows[csel].astype(np.uint32),
            "chunk_narrow": chunk_narrow.astype(np.uint32),
            "chunk_part": chunk_part.astype(np.uint64), "wide_frame": widx.astype(np.uint32),
            "wide_offsets": woffs, "frame_narrow": frame_narrow,
        }
        # one device buffer, every array at an 8-byte boundary
        spans, pos = {}, 0
        for name, a in arrays.items():
            spans[name] = (pos, a.nbytes)
            pos += (a.nbytes + 7) // 8 * 8
        host = np.zeros(max(pos, 8), dtype=np.uint8)
        for name, a in arrays.items():
            o, nb = spans[name]
            host[o:o + nb] = np.ascontiguousarray(a).view(np.uint8)
        self._route_buf = torch.from_numpy(host).to(self.dev)
        base = self._route_buf.data_ptr()
        r = _lib.BatchedRouting()
        r.n_narrow, r.n_chunks, r.n_wide = nidx.size, int(csel.sum()), widx.size
        r.n_wide_rows = int(woffs[-1])
        r.part_entries = int(cent.sum())
        r.merged_entries = int((nspan + 1).sum())
        for name in arrays:
            setattr(r, name, base + spans[name][0])
        self.route = r
        self.narrow_frames, self.narrow_rows = int(nidx.size), nrows
        self.narrow_ids = nidx

    def _split(self, key_range, digit_bits, sort_shape, tuner, max_rows, route=None):
        offs = self.offsets.cpu().tolist()
        cuts, start = [0], 0
        for f in range(self.nframes):
            if offs[f + 1] - offs[f] > max_rows:
                raise ValueError(f"frame {f} has {offs[f + 1] - offs[f]} rows; at most {max_rows}")
            if offs[f + 1] - offs[start] > max_rows:
                cuts.append(f)
                start = f
        cuts.append(self.nframes)
        self.parts = []
        for a, b in zip(cuts, cuts[1:]):
            r0, r1 = offs[a], offs[b]
            sub = torch.tensor([o - r0 for o in offs[a:b + 1]], dtype=torch.int64)
            self.parts.append(BatchedGroupBy(self.keys[r0:r1], self.vals[r0:r1], sub,
                                             key_range=key_range, aggs=self.aggs,
                                             digit_bits=digit_bits, sort_shape=sort_shape,
                                             tuner=tuner, max_rows=max_rows, route=route))
        self.digit_bits = self.parts[0].digit_bits
        self.sort_shape = self.parts[0].sort_shape

    def _launch(self, digit_bits: int | None = None, sort_shape: int | None = None, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream(self.dev)
        ptr = lambda a: _lib.ptr(self.outs.get(a))
        if self.route is not None:
            _lib.call("jfx_batched_groupby_routed", self.kc, _lib.ptr(self.keys), self.vcode,
                      _lib.ptr(self.vals), self.n, _lib.ptr(self.offsets), self.nframes,
                      self.kmin & ((1 << 64) - 1), self.key_bits, digit_bits or self.digit_bits,
                      self.sort_shape if sort_shape is None else sort_shape,
                      ctypes.byref(self.route), _lib.ptr(self.out_keys), ptr("sum"), ptr("count"),
                      ptr("min"), ptr("max"), ptr("mean"), _lib.ptr(self.ngroups),
                      _lib.ptr(self.ws), self.ws.numel(), ctypes.c_void_p(s.cuda_stream))
            return
        _lib.call("jfx_batched_groupby", self.kc, _lib.ptr(self.keys), self.vcode, _lib.ptr(self.vals),
                  self.n, _lib.ptr(self.offsets), self.nframes, self.kmin & ((1 << 64) - 1),
                  self.key_bits, digit_bits or self.digit_bits,
                  self.sort_shape if sort_shape is None else sort_shape, _lib.ptr(self.out_keys),
                  ptr("sum"), ptr("count"), ptr("min"), ptr("max"), ptr("mean"),
                  _lib.ptr(self.ngroups), _lib.ptr(self.ws), self.ws.numel(),
                  ctypes.c_void_p(s.cuda_stream))

    def capture(self):
        """Record the pipeline in a HIP graph (replayed by run(graph=True))."""
        if self.parts is not None:
            return [p.capture() for p in self.parts]
        side = torch.cuda.Stream(device=self.dev)
        side.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(side):
            self._launch(stream=side)
        torch.cuda.current_stream(self.dev).wait_stream(side)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch(stream=torch.cuda.current_stream(self.dev))
        self.graph = g
        return g

    def launch(self, graph: bool = False):
        """Enqueue one batch (graph replay or eager launches); no host sync."""
        if self.parts is not None:
            for p in self.parts:
                p.launch(graph)
            return
        if graph:
            if self.graph is None:
                self.capture()
            self.graph.replay()
        else:
            self._launch()

    def result(self) -> BatchResult:
        """Read the group count (one sync), check the device error word (raises
        JfxError when a kernel set it) and split the composite keys."""
        if self.parts is not None:
            rs = [p.result() for p in self.parts]
            offs, base = [torch.zeros(1, dtype=torch.int64, device=self.dev)], 0
            for r in rs:
                offs.append(r.offsets[1:] + base)
                base += int(r.offsets[-1])
            return BatchResult(torch.cat(offs), torch.cat([r.keys for r in rs]),
                               {a: torch.cat([r.values[a] for r in rs]) for a in self.aggs})
        G = int(self.ngroups.item())
        # a key outside the declared range, or a bounded wait that ran out,
        # raised the device error word: the groups are not valid
        code = _lib.device_error()
        if code:
            raise _lib.JfxError(
                f"batched groupby: device error {code:#x}"
                + (" (a key lies outside the declared key_range)" if code & _lib.DEVERR_BOUNDS else "")
                + (" (a cross-workgroup wait timed out)" if code & _lib.DEVERR_LOOKBACK else ""))
        comp = self.out_keys[:G]
        c64 = (comp.long() & 0xFFFFFFFF) if self.comp_dtype == torch.int32 else comp
        kmask = (1 << self.key_bits) - 1
        if self.total_bits == 64:   # logical shift of the top field through the signed view
            frame = (c64 >> self.key_bits) & ((1 << self.frame_bits) - 1)
        else:
            frame = c64 >> self.key_bits
        keys = (c64 & kmask) + self.kmin
        if self.korig != torch.int64:
            keys = keys.to(self.korig)
        counts = torch.bincount(frame, minlength=self.nframes)
        offsets = torch.zeros(self.nframes + 1, dtype=torch.int64, device=self.dev)
        offsets[1:] = torch.cumsum(counts, 0)
        vals = {}
        for a, o in self.outs.items():
            v = o[:G]
            if a in ("sum", "min", "max") and v.dtype != self.vorig:
                v = v.to(self.vorig)
            vals[a] = v
        return BatchResult(offsets, keys, vals)

    def run(self, graph: bool = False) -> BatchResult:
        self.launch(graph)
        return self.result()


def plain_per_frame(keys, values, frame_offsets, agg: str = "sum", frames=None):
    """The plain path the batched plan is measured against: one
    groupby_aggregate call per frame (same kernels, one query at a time).
    Returns [(unique keys, aggregate)] for the frames asked for."""
    from .grouping import groupby_aggregate
    offs = [int(x) for x in torch.as_tensor(frame_offsets).tolist()]
    out = []
    for f in (range(len(offs) - 1) if frames is None else frames):
        a, b = offs[f], offs[f + 1]
        uk, r = groupby_aggregate(keys[a:b], {"v": values[a:b]}, {"v": agg})
        out.append((uk, r["v"]))
    return out
