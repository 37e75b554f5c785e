"""Sort-merge join on the GPU.

Mirrors ParallelSortMergeJoin.join (parallel_algorithms.py:938-1075) and
sort_merge_join (parallel_algorithms.py:1172-1209):
- inner: every matching (left row, right row) pair, ordered by (key asc, left
  row asc, right row asc) -- the row-major order of the reference's L x R
  match matrix over the two stably sorted sides (:987-994);
- left / right: one row per row of the preserved side in sorted order, with
  the FIRST equal key of the other side or no match (:1000-1034);
- outer: one row per distinct key of either side, with first matches
  (:1035-1046).
Payload columns come out as left_<c> / right_<c>; unmatched rows are NaN and
integer payloads become floats, as the reference's jnp.where(valid, col[idx],
nan) makes them (:1053-1073): float32 for <= 32-bit integers, float64 for
64-bit ones.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._tensors import bits_view, key_code, to_device, widen_key
from .sorting import radix_sort

NO_ROW = 0xFFFFFFFF
_NAN32 = 0x7FC00000
_NAN64 = 0x7FF8000000000000


def _sorted_side(keys):
    k, orig = widen_key(to_device(keys))
    sk, _, perm = radix_sort(k, None, True, return_index=True, index_dtype=torch.int32)
    return k, sk, perm, orig


def _stream(dev):
    return ctypes.c_void_p(_lib.stream_handle(dev))


def inner_join_rows(left_keys, right_keys):
    """(keys, left_rows, right_rows) of the inner join; rows are int32 row
    indices into each side (interpreted as uint32)."""
    lk, sl, lperm, lorig = _sorted_side(left_keys)
    rk, sr, rperm, _ = _sorted_side(right_keys)
    if lk.dtype != rk.dtype:
        raise TypeError(f"join keys have different dtypes: {lk.dtype} vs {rk.dtype}")
    dev = lk.device
    L, R = lk.shape[0], rk.shape[0]
    kc = key_code(lk)
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_join_workspace_size", L, ctypes.byref(nb))
    ws = _lib.workspace(nb.value, dev)
    m = ctypes.c_uint64(0)
    _lib.call("jfx_join_count", kc, _lib.ptr(sl), L, _lib.ptr(sr), R, ctypes.byref(m),
              _lib.ptr(ws), ws.numel(), _stream(dev))
    M = m.value
    if M >= (1 << 32):
        raise _lib.JfxError(f"inner join produces {M} pairs; at most 2^32-1 are supported")
    out_l = torch.empty(M, dtype=torch.int32, device=dev)
    out_r = torch.empty(M, dtype=torch.int32, device=dev)
    out_k = torch.empty(M, dtype=lk.dtype, device=dev)
    if M:
        _lib.call("jfx_join_fill", kc, _lib.ptr(sl), _lib.ptr(lperm), L, _lib.ptr(sr),
                  _lib.ptr(rperm), R, _lib.ptr(out_l), _lib.ptr(out_r), _lib.ptr(out_k),
                  _lib.ptr(ws), ws.numel(), _stream(dev))
    if out_k.dtype != lorig:
        out_k = out_k.to(lorig)
    return out_k, out_l, out_r


def _first_match(sorted_probe, sorted_build, build_perm, kc):
    dev = sorted_probe.device
    out = torch.empty(sorted_probe.shape[0], dtype=torch.int32, device=dev)
    _lib.call("jfx_join_first_match", kc, _lib.ptr(sorted_probe), sorted_probe.shape[0],
              _lib.ptr(sorted_build), _lib.ptr(build_perm), sorted_build.shape[0],
              _lib.ptr(out), _stream(dev))
    return out


def payload_source(col, dev) -> torch.Tensor:
    """A payload column as the join emits it: floats as they are (half types
    widened to float32), integers promoted to float64 (8-byte) or float32 —
    the promotion jnp.where(valid, col[idx], nan) applies in the reference."""
    col = to_device(col, dev)
    if col.dtype.is_floating_point:
        return col if col.element_size() in (4, 8) else col.float()
    return col.to(torch.float64 if col.element_size() == 8 else torch.float32)


def fill_values(kc, sl, lperm, L, sr, rperm, R, M, lsrc, l_by_row, rsrc, r_by_row, ws, *,
                keys_dtype, want_rows=(False, False)):
    """The M pairs counted on ws: keys, one payload per side written by the
    fill (jfx_join_fill_values), and the row numbers when asked for."""
    dev = sl.device
    keys = torch.empty(M, dtype=keys_dtype, device=dev)
    lv = torch.empty(M, dtype=lsrc.dtype, device=dev) if lsrc is not None else None
    rv = torch.empty(M, dtype=rsrc.dtype, device=dev) if rsrc is not None else None
    li = torch.empty(M, dtype=torch.int32, device=dev) if want_rows[0] else None
    ri = torch.empty(M, dtype=torch.int32, device=dev) if want_rows[1] else None
    if M:
        _lib.call("jfx_join_fill_values", kc, _lib.ptr(sl), _lib.ptr(lperm), L, _lib.ptr(sr),
                  _lib.ptr(rperm), R, lsrc.element_size() if lsrc is not None else 4,
                  _lib.ptr(lsrc), int(l_by_row), rsrc.element_size() if rsrc is not None else 4,
                  _lib.ptr(rsrc), int(r_by_row), _lib.ptr(li), _lib.ptr(ri), _lib.ptr(keys),
                  _lib.ptr(lv), _lib.ptr(rv), _lib.ptr(ws), ws.numel(), _stream(dev))
    return keys, lv, rv, li, ri


def pack_columns(cols: list) -> torch.Tensor:
    """One record per row holding the columns' bytes in order (4 + 4 or one
    column of up to 8 bytes): int32 or int64 storage."""
    n = cols[0].shape[0]
    width = sum(c.element_size() for c in cols)
    out = torch.empty(n, dtype=torch.int32 if width == 4 else torch.int64, device=cols[0].device)
    ptrs = (ctypes.c_void_p * len(cols))(*[c.data_ptr() for c in cols])
    nbytes = (ctypes.c_int * len(cols))(*[c.element_size() for c in cols])
    _lib.call("jfx_pack_columns", len(cols), ptrs, nbytes, n, _lib.ptr(out), None,
              _stream(cols[0].device))
    return out


# Integer key types the range pre-filter takes, and the smallest side it runs
# on (below it the count's host read costs more than the sort it can save).
_PREFILTER_TYPES = (torch.int32, torch.int64, torch.uint32, torch.uint64)
_PREFILTER_MIN_ROWS = 1 << 16


def _prefilter(k, cols: list, other_sorted, want_rows: bool):
    """Rows of k whose key lies within [other_sorted[0], other_sorted[-1]]
    (jfx_join_prefilter): returns (keys, record or None, rows or None), in row
    order; the record packs cols as pack_columns does."""
    dev = k.device
    n = k.shape[0]
    width = sum(c.element_size() for c in cols)
    out_k = torch.empty_like(k)
    rec = (torch.empty(n, dtype=torch.int32 if width == 4 else torch.int64, device=dev)
           if cols else None)
    rows = torch.empty(n, dtype=torch.int32, device=dev) if want_rows else None
    ptrs = (ctypes.c_void_p * max(1, len(cols)))(*[c.data_ptr() for c in cols])
    nbytes = (ctypes.c_int * max(1, len(cols)))(*[c.element_size() for c in cols])
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_join_prefilter_workspace_size", n, ctypes.byref(nb))
    ws = _lib.workspace(nb.value, dev)
    kept = ctypes.c_uint64(0)
    _lib.call("jfx_join_prefilter", key_code(k), _lib.ptr(k), n, _lib.ptr(other_sorted),
              other_sorted.shape[0], len(cols), ptrs, nbytes, _lib.ptr(out_k), _lib.ptr(rec),
              _lib.ptr(rows), ctypes.byref(kept), _lib.ptr(ws), ws.numel(), _stream(dev))
    m = kept.value
    return out_k[:m], (rec[:m] if rec is not None else None), (rows[:m] if rows is not None else None)


class _JoinSide:
    """One side of the inner join, sorted: its payload either rides through
    the radix sort packed into one record per row (at most 8 bytes of
    columns: read in key order by the fill) or is gathered at the source row
    the sort carried (wider payloads). With other_sorted (the other side's
    sorted keys, integer keys) the rows outside that side's key range are
    dropped first (_prefilter): they have no pair, and the sort then covers
    fewer rows, whose keys often vary in fewer digits."""

    def __init__(self, keys, cols: list, other_sorted=None):
        self.k, self.orig = widen_key(to_device(keys))
        dev = self.k.device
        self.cols = [payload_source(c, dev) for c in cols]
        width = sum(c.element_size() for c in self.cols)
        self.record_bytes = width if self.cols and width <= 8 else 0
        self.perm = self.records = None
        k = self.k
        if other_sorted is not None:
            k, rec, rows = _prefilter(self.k, self.cols if self.record_bytes else [],
                                      other_sorted, bool(self.cols) and not self.record_bytes)
            self.n = k.shape[0]
            if not self.cols:
                self.sorted, _ = radix_sort(k)
            elif self.record_bytes:
                self.sorted, self.records = radix_sort(k, rec)
            else:
                # the source rows ride through the sort as its payload
                self.sorted, self.perm = radix_sort(k, rows)
            return
        self.n = k.shape[0]
        if not self.cols:
            self.sorted, _ = radix_sort(self.k)
        elif self.record_bytes:
            rec = bits_view(self.cols[0]) if len(self.cols) == 1 else pack_columns(self.cols)
            self.sorted, self.records = radix_sort(self.k, rec)
        else:
            self.sorted, _, self.perm = radix_sort(self.k, None, True, return_index=True,
                                                   index_dtype=torch.int32)

    def abi(self, outs: list) -> _lib.JoinSide:
        s = _lib.JoinSide()
        s.rows = _lib.ptr(self.perm).value or None
        s.ncols = len(self.cols)
        s.by_row = 0 if self.record_bytes else 1
        s.record_bytes = self.record_bytes
        s.records = _lib.ptr(self.records).value or None
        off = 0
        for j, (c, o) in enumerate(zip(self.cols, outs)):
            s.cols[j] = c.data_ptr()
            s.col_bytes[j] = c.element_size()
            s.col_offset[j] = off
            s.out_cols[j] = o.data_ptr()
            off += c.element_size()
        return s


def _inner_join_fused(left_keys, left_values: dict, right_keys, right_values: dict):
    """Inner join whose payload columns are all written by the fill
    (jfx_join_fill_columns); payloads of at most 8 bytes per side are carried
    through that side's sort, so the fill reads them in key order."""
    lcols, rcols = list(left_values.items()), list(right_values.items())
    if len(lcols) > _lib.JOIN_MAXCOLS or len(rcols) > _lib.JOIN_MAXCOLS:
        return _inner_join_by_rows(left_keys, left_values, right_keys, right_values)
    lk, _ = widen_key(to_device(left_keys))
    rk, _ = widen_key(to_device(right_keys))
    if lk.dtype != rk.dtype:
        raise TypeError(f"join keys have different dtypes: {lk.dtype} vs {rk.dtype}")
    lvals, rvals = [v for _, v in lcols], [v for _, v in rcols]
    if lk.dtype in _PREFILTER_TYPES and max(lk.shape[0], rk.shape[0]) >= _PREFILTER_MIN_ROWS:
        # the smaller side is sorted first; its key range filters the larger
        if lk.shape[0] >= rk.shape[0]:
            right = _JoinSide(right_keys, rvals)
            left = _JoinSide(left_keys, lvals, other_sorted=right.sorted)
        else:
            left = _JoinSide(left_keys, lvals)
            right = _JoinSide(right_keys, rvals, other_sorted=left.sorted)
    else:
        left = _JoinSide(left_keys, lvals)
        right = _JoinSide(right_keys, rvals)
    dev = left.k.device
    L, R = left.n, right.n
    kc = key_code(left.k)
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_join_workspace_size", L, ctypes.byref(nb))
    ws = _lib.workspace(nb.value, dev)
    m = ctypes.c_uint64(0)
    _lib.call("jfx_join_count", kc, _lib.ptr(left.sorted), L, _lib.ptr(right.sorted), R,
              ctypes.byref(m), _lib.ptr(ws), ws.numel(), _stream(dev))
    M = m.value
    keys = torch.empty(M, dtype=left.k.dtype, device=dev)
    louts = [torch.empty(M, dtype=c.dtype, device=dev) for c in left.cols]
    routs = [torch.empty(M, dtype=c.dtype, device=dev) for c in right.cols]
    if M:
        ls, rs = left.abi(louts), right.abi(routs)
        _lib.call("jfx_join_fill_columns", kc, _lib.ptr(left.sorted), L, ctypes.byref(ls),
                  _lib.ptr(right.sorted), R, ctypes.byref(rs), _lib.ptr(keys), _lib.ptr(ws),
                  ws.numel(), _stream(dev))
    if keys.dtype != left.orig:
        keys = keys.to(left.orig)
    out = {}
    for (c, _), o in zip(lcols, louts):
        out[f"left_{c}"] = o
    for (c, _), o in zip(rcols, routs):
        out[f"right_{c}"] = o
    return keys, out


def _inner_join_by_rows(left_keys, left_values: dict, right_keys, right_values: dict):
    """Inner join through the row pairs and one gather per payload column
    (for more than 8 payload columns on a side)."""
    k, li, ri = inner_join_rows(left_keys, right_keys)
    out = {}
    for c, v in left_values.items():
        out[f"left_{c}"] = gather(v, li)
    for c, v in right_values.items():
        out[f"right_{c}"] = gather(v, ri)
    return k, out


def gather(col: torch.Tensor, rows: torch.Tensor, fill_nan: bool = True) -> torch.Tensor:
    """col[rows] with rows as uint32 (int32 storage); rows == 0xFFFFFFFF give
    NaN. Integer columns come back as floats (the reference's promotion)."""
    src = payload_source(col, rows.device)
    eb = src.element_size()
    out = torch.empty(rows.shape[0], dtype=src.dtype, device=rows.device)
    fill = _NAN32 if eb == 4 else _NAN64
    if rows.shape[0]:
        _lib.call("jfx_gather", eb, _lib.ptr(src), _lib.ptr(rows), rows.shape[0], _lib.ptr(out),
                  fill, _stream(rows.device))
    return out


def sort_merge_join(left_keys, left_values: dict, right_keys, right_values: dict,
                    how: str = "inner"):
    """Single-key join; returns (joined keys, {left_<c>/right_<c>: column})."""
    _lib.require_gpu()
    if how == "inner":
        return _inner_join_fused(left_keys, left_values, right_keys, right_values)
    if how in ("left", "right", "outer"):
        lk, sl, lperm, lorig = _sorted_side(left_keys)
        rk, sr, rperm, rorig = _sorted_side(right_keys)
        if lk.dtype != rk.dtype:
            raise TypeError(f"join keys have different dtypes: {lk.dtype} vs {rk.dtype}")
        kc = key_code(lk)
        if how == "left":
            keys, li = sl, lperm
            ri = _first_match(sl, sr, rperm, kc)
        elif how == "right":
            keys, ri = sr, rperm
            li = _first_match(sr, sl, lperm, kc)
        else:
            from .grouping import groupby_aggregate
            both = torch.cat([lk, rk])
            keys, _ = groupby_aggregate(both, {"c": torch.ones_like(both, dtype=torch.int32)},
                                        {"c": "count"})
            keys = keys.to(lk.dtype)
            li = _first_match(keys, sl, lperm, kc)
            ri = _first_match(keys, sr, rperm, kc)
        if keys.dtype != lorig:
            keys = keys.to(lorig)
    else:
        raise ValueError(f"unsupported join type {how}")
    out = {}
    for c, v in left_values.items():
        out[f"left_{c}"] = gather(v, li)
    for c, v in right_values.items():
        out[f"right_{c}"] = gather(v, ri)
    return keys, out


# ---- multi-column join -------------------------------------------------------

def _dense_codes(lcol: torch.Tensor, rcol: torch.Tensor):
    """int64 codes of one key column over both sides, equal exactly where the
    values compare equal under ==: dense ranks of the union (one radix sort),
    -0.0 and +0.0 sharing a code; NaN rows get distinct negative codes, so they
    match nothing (NaN != NaN in the reference's match matrix)."""
    both = torch.cat([lcol, rcol])
    n = both.shape[0]
    sk, _, perm = radix_sort(both, None, True, return_index=True)
    from .multikey import multikey_heads
    gid, _ = multikey_heads([sk])
    codes = torch.empty(n, dtype=torch.int64, device=both.device)
    codes[perm] = gid
    if both.dtype.is_floating_point:
        nan = torch.isnan(both)
        codes = torch.where(nan, -1 - torch.arange(n, device=both.device), codes)
    return codes[:lcol.shape[0]], codes[lcol.shape[0]:]


def _tuple_codes(lkeys: list, rkeys: list):
    """One int64 code per row for key tuples: equal codes <=> every column
    compares equal (and no NaN). Columns are folded pairwise through dense
    ranks, so the code never needs more than 64 bits."""
    lc, rc = _dense_codes(lkeys[0], rkeys[0])
    for lk, rk in zip(lkeys[1:], rkeys[1:]):
        l2, r2 = _dense_codes(lk, rk)
        nl = lc.shape[0]
        # pair (code_so_far, code_j): both fit 32 bits as dense ranks over L+R
        # rows; NaN rows keep a distinct negative code
        bad = torch.cat([lc, rc]) < 0
        bad |= torch.cat([l2, r2]) < 0
        pair = (torch.cat([lc, rc]) << 32) | torch.cat([l2, r2])
        pair = torch.where(bad, torch.full_like(pair, -1), pair)
        pl, pr = _dense_codes(pair[:nl], pair[nl:])
        m = torch.arange(pair.shape[0], device=pair.device)
        pc = torch.cat([pl, pr])
        pc = torch.where(bad, -1 - m, pc)
        lc, rc = pc[:nl], pc[nl:]
    return lc, rc


def _gather_join_column(arr: torch.Tensor, idx: torch.Tensor, dtype_name: str | None = None):
    """_gather_join_column (parallel_algorithms.py:1254-1278): idx < 0 gives
    -1 for string (dictionary-code) columns, NaN otherwise; integer and bool
    columns become float32 only when some row is unmatched."""
    if idx.shape[0] == 0:
        return arr[:0]
    valid = idx >= 0
    safe = idx.clamp(min=0)
    if dtype_name == "string":
        g = arr[safe]
        return torch.where(valid, g, torch.full_like(g, -1))
    if bool(valid.all()):
        return arr[safe]
    g = arr[safe]
    if not g.dtype.is_floating_point:
        g = g.to(torch.float32)
    return torch.where(valid, g, torch.full_like(g, float("nan")))


def join_multi_column(left_keys: list, left_key_names: list, left_values: dict,
                      right_keys: list, right_key_names: list, right_values: dict,
                      how: str = "inner", left_key_dtypes=None, right_key_dtypes=None,
                      left_value_dtypes=None, right_value_dtypes=None):
    """ParallelSortMergeJoin.join_multi_column (parallel_algorithms.py:826-936).

    Row pairs come in (left row, right row) order -- the row-major order of
    the reference's L x R match matrix over the unsorted rows -- followed by
    the unmatched left rows (left / outer) and then the unmatched right rows
    (right / outer), each ascending, with -1 for the missing side. Keys are
    coalesced (left key where present, else right); payloads follow
    _gather_join_column. Built from the sort-merge join on per-row tuple
    codes instead of the L x R matrix."""
    _lib.require_gpu()
    if how not in ("inner", "left", "right", "outer"):
        raise ValueError(f"unsupported join type {how}")
    lks = [widen_key(to_device(k))[0] for k in left_keys]
    dev = lks[0].device
    rks = [widen_key(to_device(k, dev))[0] for k in right_keys]
    nl, nr = lks[0].shape[0], rks[0].shape[0]
    for i, (a, b) in enumerate(zip(lks, rks)):
        if a.dtype != b.dtype:
            t = torch.promote_types(a.dtype, b.dtype)
            lks[i], rks[i] = a.to(t), b.to(t)
    if nl and nr:
        lc, rc = _tuple_codes(lks, rks)
        _, li, ri = inner_join_rows(lc, rc)
        li = li.long() & 0xFFFFFFFF
        ri = ri.long() & 0xFFFFFFFF
        # (code, left, right) order -> (left, right): a stable sort by left row
        # (the right rows of one left row already ascend)
        _, ri_s, perm = radix_sort(li.to(torch.int32), ri.to(torch.int32), True, return_index=True)
        li = li[perm]
        ri = ri_s.long()
    else:
        li = torch.empty(0, dtype=torch.int64, device=dev)
        ri = torch.empty(0, dtype=torch.int64, device=dev)
    lhas = torch.zeros(nl, dtype=torch.bool, device=dev)
    rhas = torch.zeros(nr, dtype=torch.bool, device=dev)
    if li.shape[0]:
        lhas[li] = True
        rhas[ri] = True
    from .compaction import compact_rows
    lun = compact_rows(~lhas) if nl else li[:0]
    run = compact_rows(~rhas) if nr else ri[:0]
    neg = lambda k: torch.full((k,), -1, dtype=torch.int64, device=dev)
    if how == "inner":
        all_l, all_r = li, ri
    elif how == "left":
        all_l, all_r = torch.cat([li, lun]), torch.cat([ri, neg(lun.shape[0])])
    elif how == "right":
        all_l, all_r = torch.cat([li, neg(run.shape[0])]), torch.cat([ri, run])
    else:
        all_l = torch.cat([li, lun, neg(run.shape[0])])
        all_r = torch.cat([ri, neg(lun.shape[0]), run])
    out_keys = {}
    for name, lk, rk, orig_l in zip(left_key_names, lks, rks, left_keys):
        if all_l.shape[0] == 0:
            out_keys[name] = lk[:0]
            continue
        lv = lk[all_l.clamp(min=0)] if nl else torch.zeros(all_l.shape[0], dtype=lk.dtype, device=dev)
        rv = rk[all_r.clamp(min=0)] if nr else torch.zeros(all_r.shape[0], dtype=rk.dtype, device=dev)
        out_keys[name] = torch.where(all_l >= 0, lv, rv)
    out_vals = {}
    for name, arr in left_values.items():
        dn = (left_value_dtypes or {}).get(name)
        out_vals[f"left_{name}"] = _gather_join_column(to_device(arr, dev), all_l, dn)
    for name, arr in right_values.items():
        dn = (right_value_dtypes or {}).get(name)
        out_vals[f"right_{name}"] = _gather_join_column(to_device(arr, dev), all_r, dn)
    return out_keys, out_vals
