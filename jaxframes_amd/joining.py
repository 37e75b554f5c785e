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
from ._tensors import key_code, to_device, widen_key
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


def gather(col: torch.Tensor, rows: torch.Tensor, fill_nan: bool = True) -> torch.Tensor:
    """col[rows] with rows as uint32 (int32 storage); rows == 0xFFFFFFFF give
    NaN. Integer columns come back as floats (the reference's promotion)."""
    col = to_device(col, rows.device)
    if col.dtype.is_floating_point:
        src = col if col.element_size() in (4, 8) else col.float()
    else:
        src = col.to(torch.float64 if col.element_size() == 8 else torch.float32)
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
        keys, li, ri = inner_join_rows(left_keys, right_keys)
    elif how in ("left", "right", "outer"):
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
