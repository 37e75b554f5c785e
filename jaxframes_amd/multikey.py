"""Multi-column keys on the GPU: lexicographic order, groups and composites.

Mirrors multi_column_lexsort (parallel_algorithms.py:25-66),
ParallelRadixSort.sort_multi_column (:426-461) and the head-flag part of
SortBasedGroupBy.groupby_aggregate_multi_column (:662-731).

When every key column is an integer column and the columns' value ranges fit
in 64 bits together, the columns are packed into one uint64 composite key
(jfx_pack_keys) whose stable sort is the stable lexicographic sort and whose
equal values are equal key tuples, so one radix sort / single-key groupby
replaces the K per-column sorts. Otherwise the reference's scheme runs on the
device radix sort: one stable sort per column, last key first.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._tensors import key_code, to_device, widen_key
from .sorting import radix_sort

_INT_CODES = (_lib.JFX_U32, _lib.JFX_I32, _lib.JFX_U64, _lib.JFX_I64)
_SIGNED = {torch.int32: 32, torch.int64: 64}
_VIEW = {torch.uint32: torch.int32, torch.uint64: torch.int64}


def _as_ascending_list(ascending, k):
    if isinstance(ascending, bool):
        return [ascending] * k
    ascending = list(ascending)
    if len(ascending) != k:
        raise ValueError("ascending must have one entry per key column")
    return ascending


def _ordered_ranges(cols):
    """(ordered min, ordered max) of each integer column as Python ints, with
    one host synchronisation for all columns. ordered() maps the column's
    order onto unsigned integers (signed: + 2^(w-1))."""
    mins, maxs, offs = [], [], []
    for c in cols:
        if c.dtype in _SIGNED:
            t, off = c, 1 << (_SIGNED[c.dtype] - 1)
        else:  # unsigned: flip the top bit so the signed view orders like the value
            w = 32 if c.dtype == torch.uint32 else 64
            t = c.view(_VIEW[c.dtype]) ^ torch.tensor(-(1 << (w - 1)), dtype=_VIEW[c.dtype],
                                                       device=c.device)
            off = 1 << (w - 1)
        mn, mx = torch.aminmax(t)
        mins.append(mn.to(torch.int64))
        maxs.append(mx.to(torch.int64))
        offs.append(off)
    both = torch.stack(mins + maxs).cpu().tolist()
    k = len(cols)
    return [(both[i] + offs[i], both[k + i] + offs[i]) for i in range(k)]


def composite_plan(cols, ascending=True):
    """Column description for jfx_pack_keys, or None when the key columns are
    not all integers, their ranges need more than 64 bits, or a descending
    signed column's max - x would wrap (the reference's order then differs
    from the reversed one, and the per-column path reproduces it)."""
    k = len(cols)
    if k == 0 or k > 8 or cols[0].shape[0] == 0:
        return None
    if any(key_code(c) not in _INT_CODES for c in cols):
        return None
    asc = _as_ascending_list(ascending, k)
    ranges = _ordered_ranges(cols)
    bits = []
    for c, a, (lo, hi) in zip(cols, asc, ranges):
        span = hi - lo
        if not a and c.dtype in _SIGNED and span >= (1 << (_SIGNED[c.dtype] - 1)):
            return None
        bits.append(span.bit_length())
    if sum(bits) > 64:
        return None
    sub = [lo if a else hi for a, (lo, hi) in zip(asc, ranges)]
    return {"types": [key_code(c) for c in cols], "sub": sub, "bits": bits,
            "desc": [0 if a else 1 for a in asc], "dtypes": [c.dtype for c in cols],
            "range_product_bits": sum(bits)}


def _plan_arrays(plan):
    k = len(plan["types"])
    return ((ctypes.c_int * k)(*plan["types"]), (ctypes.c_uint64 * k)(*plan["sub"]),
            (ctypes.c_uint32 * k)(*plan["bits"]), (ctypes.c_int * k)(*plan["desc"]), k)


def pack(cols, plan) -> torch.Tensor:
    """uint64 composite key of each row (jfx_pack_keys)."""
    n = cols[0].shape[0]
    dev = cols[0].device
    out = torch.empty(n, dtype=torch.uint64, device=dev)
    types, sub, bits, desc, k = _plan_arrays(plan)
    ptrs = (ctypes.c_void_p * k)(*[c.data_ptr() for c in cols])
    _lib.call("jfx_pack_keys", ptrs, types, k, sub, bits, desc, n, _lib.ptr(out),
              ctypes.c_void_p(_lib.stream_handle(dev)))
    return out


def unpack(packed: torch.Tensor, plan) -> list:
    """The key columns of each composite key (jfx_unpack_keys)."""
    n = packed.shape[0]
    dev = packed.device
    outs = [torch.empty(n, dtype=dt, device=dev) for dt in plan["dtypes"]]
    types, sub, bits, desc, k = _plan_arrays(plan)
    ptrs = (ctypes.c_void_p * k)(*[o.data_ptr() for o in outs])
    _lib.call("jfx_unpack_keys", _lib.ptr(packed), n, types, k, sub, bits, desc, ptrs,
              ctypes.c_void_p(_lib.stream_handle(dev)))
    return outs


def _descending_transform(col: torch.Tensor) -> torch.Tensor:
    """The reference's descending key (:56-61): -x for floats, max(x) - x in
    the column's own integer type (wrapping like XLA)."""
    if col.dtype.is_floating_point:
        return -col
    if col.dtype in _VIEW:  # unsigned: the same bits through the signed view
        v = col.view(_VIEW[col.dtype])
        w = 32 if col.dtype == torch.uint32 else 64
        flip = torch.tensor(-(1 << (w - 1)), dtype=v.dtype, device=v.device)
        mx = (v ^ flip).max() ^ flip
        return (mx - v).view(col.dtype)
    return col.max() - col


def lexsort_perm(keys, ascending=True) -> torch.Tensor:
    """int64 row order of multi_column_lexsort(keys, ascending).

    One key: the stable ascending argsort (the reference ignores `ascending`
    there, :44-45). Several keys: the composite sort when the columns pack,
    else one stable radix sort per column, last key first."""
    _lib.require_gpu()
    ks = [widen_key(to_device(k))[0] for k in keys]
    if not ks:
        raise ValueError("no key columns")
    n = ks[0].shape[0]
    if any(k.shape[0] != n for k in ks):
        raise ValueError("key columns must have the same length")
    if len(ks) == 1:
        return radix_sort(ks[0], None, True, want_keys=False, return_index=True)[2]
    asc = _as_ascending_list(ascending, len(ks))
    if n == 0:
        return torch.empty(0, dtype=torch.int64, device=ks[0].device)
    plan = composite_plan(ks, asc)
    if plan is not None:
        return radix_sort(pack(ks, plan), None, True, want_keys=False, return_index=True)[2]
    idx = None
    for col, a in zip(reversed(ks), reversed(asc)):
        arr = col if a else _descending_transform(col)
        sk = arr if idx is None else arr[idx]
        _, _, p = radix_sort(sk, None, True, want_keys=False, return_index=True,
                             index_dtype=torch.int32)
        p = p.long()
        idx = p if idx is None else idx[p]
    return idx


def _neq_prev(sk: torch.Tensor) -> torch.Tensor:
    if sk.dtype in _VIEW:
        sk = sk.view(_VIEW[sk.dtype])
    return sk[1:] != sk[:-1]


def multikey_heads(sorted_keys):
    """(group id per row, first row of each group) for rows already in
    lexicographic order: a row starts a group where any key differs from the
    previous row under != (:697-706), so NaN keys are each their own group."""
    n = sorted_keys[0].shape[0]
    dev = sorted_keys[0].device
    flags = torch.zeros(n, dtype=torch.bool, device=dev)
    if n:
        flags[0] = True
    for sk in sorted_keys:
        if n > 1:
            flags[1:] |= _neq_prev(sk)
    starts = torch.nonzero(flags).squeeze(1)
    group_id = torch.cumsum(flags, 0, dtype=torch.int64) - 1
    return group_id, starts


def sort_multi_column(keys, values=None, ascending=True):
    """ParallelRadixSort.sort_multi_column (:426-461): (sorted key columns,
    sorted values or None)."""
    perm = lexsort_perm(keys, ascending)
    ks = [to_device(k) for k in keys]
    sorted_keys = [k[perm] for k in ks]
    sorted_values = None
    if values is not None:
        sorted_values = {c: to_device(v, perm.device)[perm] for c, v in values.items()}
    return sorted_keys, sorted_values
