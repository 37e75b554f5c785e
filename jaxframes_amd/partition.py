"""Row hashing and the local hash partition (one GPU's share of the
repartition exchange).

Mirrors hash_multi_columns (parallel_algorithms.py:69-136) and the per-device
bucketing of collective_hash_repartition (parallel_algorithms.py:185-274,
_bucketize_local_column :163-182). Instead of the reference's padded
[P, local_n] buffer per column, rows come out grouped by bucket (input order
kept inside a bucket) with P+1 offsets: the layout an all-to-all with exact
counts sends.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._tensors import key_code, to_device, widen_key


def _hash_key(t: torch.Tensor) -> torch.Tensor:
    t, _ = widen_key(t)
    return t


def hash_multi_columns(keys) -> torch.Tensor:
    """uint32 hash of every row of one or more key columns."""
    _lib.require_gpu()
    if not isinstance(keys, (list, tuple)):
        keys = [keys]
    if not keys:
        raise ValueError("at least one key column is required")
    cols = [_hash_key(to_device(k)) for k in keys]
    dev = cols[0].device
    n = cols[0].shape[0]
    for c in cols:
        if c.shape[0] != n:
            raise ValueError("key columns must have the same length")
    out = torch.empty(n, dtype=torch.uint32, device=dev)
    ptrs = (ctypes.c_void_p * len(cols))(*[c.data_ptr() for c in cols])
    kts = (ctypes.c_int * len(cols))(*[key_code(c) for c in cols])
    _lib.call("jfx_hash_rows", ptrs, kts, len(cols), n, _lib.ptr(out),
              ctypes.c_void_p(_lib.stream_handle(dev)))
    return out


def hash_partition(keys, nbuckets: int, *, return_keys: bool = True, return_rows: bool = True,
                   return_hashes: bool = False):
    """Partition rows by hash(keys) % nbuckets.

    keys: one key column, or a list of key columns (hashed together as
    hash_multi_columns does). Returns a dict with 'offsets' (int64,
    nbuckets + 1) and, as requested, 'keys' (the single key column
    partitioned), 'rows' (int64 source row of each output row) and 'hashes'
    (uint32, input order).
    """
    _lib.require_gpu()
    if not 1 <= nbuckets <= 2048:
        raise ValueError("nbuckets must be in [1, 2048]")
    multi = isinstance(keys, (list, tuple)) and len(keys) > 1
    if isinstance(keys, (list, tuple)) and len(keys) == 1:
        keys = keys[0]
    if multi:
        k = hash_multi_columns(keys)
        prehashed = 1
        return_keys = False
    else:
        k = _hash_key(to_device(keys))
        prehashed = 0
    dev = k.device
    n = k.shape[0]
    kc = key_code(k)
    out = {"offsets": torch.empty(nbuckets + 1, dtype=torch.int64, device=dev)}
    out_keys = torch.empty_like(k) if return_keys else None
    out_rows = torch.empty(n, dtype=torch.int32, device=dev) if return_rows else None
    out_h = None
    if return_hashes:
        out_h = k if prehashed else torch.empty(n, dtype=torch.uint32, device=dev)
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_hash_partition_workspace_size", kc, n, nbuckets, ctypes.byref(nb))
    ws = _lib.workspace(nb.value, dev)
    _lib.call("jfx_hash_partition", kc, _lib.ptr(k), n, nbuckets, prehashed, _lib.ptr(out_keys),
              _lib.ptr(out_rows), _lib.ptr(out["offsets"]),
              _lib.ptr(out_h if (out_h is not None and not prehashed) else None),
              _lib.ptr(ws), ws.numel(), ctypes.c_void_p(_lib.stream_handle(dev)))
    if out_keys is not None:
        out["keys"] = out_keys
    if out_rows is not None:
        out["rows"] = out_rows.long()
    if out_h is not None:
        out["hashes"] = out_h
    return out


def partition_columns(keys, values: dict, nbuckets: int):
    """Partition a key column and its value columns: returns
    (partitioned keys, {name: partitioned column}, offsets)."""
    r = hash_partition(keys, nbuckets, return_keys=not isinstance(keys, (list, tuple)),
                       return_rows=True)
    rows = r["rows"]
    vals = {c: to_device(v)[rows] for c, v in values.items()}
    if isinstance(keys, (list, tuple)):
        pk = [to_device(k)[rows] for k in keys]
    else:
        pk = r["keys"]
    return pk, vals, r["offsets"]


_INT_PAD = -999999


def logical_pad_value(dtype_name, column: torch.Tensor):
    """The pad value of a column in the padded bucket layout, by the rule of
    _logical_pad_value (parallel_algorithms.py:150-160): dictionary-encoded
    string codes (dtype_name "string") -1, floats NaN, integers -999999, bools
    False, anything else 0 -- as a Python scalar of the column's dtype
    (-999999 wraps modulo 2^bits in unsigned and narrow integer columns, as a
    C cast does)."""
    dt = column.dtype
    if dtype_name == "string":
        return torch.tensor(-1, dtype=torch.int64).to(dt).item()
    if dt.is_floating_point:
        return float("nan")
    if dt == torch.bool:
        return False
    if not dt.is_complex:
        return torch.tensor(_INT_PAD, dtype=torch.int64).to(dt).item()
    return 0


def bucketize_padded(column, offsets: torch.Tensor, rows: torch.Tensor, pad_value=None,
                     dtype_name=None):
    """The reference's padded per-destination layout of one column
    (_bucketize_local_column, parallel_algorithms.py:163-182): [P, n] with row
    r of bucket d at [d, r] and pad_value elsewhere, built from the compact
    layout of hash_partition (offsets, rows). pad_value None takes the
    per-dtype rule of logical_pad_value (dtype_name "string" for string code
    columns). For inputs small enough that a P x n buffer is reasonable; the
    compact layout is what the exchange sends."""
    col = to_device(column)
    if pad_value is None:
        pad_value = logical_pad_value(dtype_name, col)
    n = col.shape[0]
    P = offsets.shape[0] - 1
    counts = offsets[1:] - offsets[:-1]
    bucket = torch.repeat_interleave(torch.arange(P, device=col.device), counts)
    pos = torch.arange(n, device=col.device) - offsets[bucket]
    out = torch.full((P, n), pad_value, dtype=col.dtype, device=col.device)
    out[bucket, pos] = col[rows]
    return out


def collective_hash_repartition(keys: list, values: dict, key_dtypes=None, value_dtypes=None,
                                sharding_spec=None):
    """collective_hash_repartition (parallel_algorithms.py:185-274) on one
    GPU: a single device holds every partition, so the rows stay where they
    are and every row is valid (the reference's mesh-size-1 branch,
    :202-203). The per-device step a mesh would run is hash_partition."""
    if not keys:
        return keys, values, torch.zeros(0, dtype=torch.bool, device=torch.device("cuda"))
    ks = [to_device(k) for k in keys]
    return ks, {c: to_device(v, ks[0].device) for c, v in values.items()}, \
        torch.ones(ks[0].shape[0], dtype=torch.bool, device=ks[0].device)
