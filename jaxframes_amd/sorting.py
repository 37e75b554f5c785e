"""Device radix sort: the HIP onesweep sort behind ParallelRadixSort.sort.

Mirrors ``ParallelRadixSort.sort`` (parallel_algorithms.py:463-581): stable,
ascending by default; descending is the reverse of the ascending order so
ties come out in reverse original order (parallel_algorithms.py:552-555).
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._tensors import bits_view, key_code, to_device, widen_key


def sort_workspace_bytes(kcode: int, pay_bytes: int, n: int) -> int:
    out = ctypes.c_uint64(0)
    _lib.call("jfx_sort_workspace_size", kcode, pay_bytes, n, ctypes.byref(out))
    return out.value


def radix_sort(keys, values=None, ascending: bool = True, *, want_keys: bool = True,
               return_index: bool = False, digit_bits: int = 0, index_dtype=torch.int64):
    """Stable sort of a key column with an optional payload column.

    values: None, or a column whose elements are 4 or 8 bytes (carried
    through the passes as payload), or any other column (permuted by a
    gather of the row index the sort carries).
    return_index: also return the int64 source-row index of each output row.
    index_dtype: torch.int64 (default) or torch.int32 (the kernel's uint32
    row index as stored, no conversion).
    Returns (sorted_keys or None, sorted_values or None[, index]).
    """
    _lib.require_gpu()
    k = to_device(keys)
    dev = k.device
    n = k.shape[0]
    k, orig_dtype = widen_key(k)
    kc = key_code(k)
    v = None if values is None else to_device(values, dev)
    if v is not None and v.shape[0] != n:
        raise ValueError("keys and values must have the same length")
    need_index = return_index or (v is not None and v.element_size() not in (4, 8))
    if n == 0:
        idx = torch.empty(0, dtype=index_dtype, device=dev)
        res_k = k.to(orig_dtype) if want_keys else None
        res_v = None if v is None else v[:0]
        return (res_k, res_v, idx) if return_index else (res_k, res_v)

    out_k = torch.empty_like(k)
    if need_index:
        pay_bytes, pay_in = 4, None
        pay_out = torch.empty(n, dtype=torch.int32, device=dev)
    elif v is not None:
        pay_bytes = v.element_size()
        pay_in = bits_view(v)
        pay_out = torch.empty_like(pay_in)
    else:
        pay_bytes, pay_in, pay_out = 0, None, None
    ws_bytes = sort_workspace_bytes(kc, pay_bytes, n)
    ws = _lib.workspace(ws_bytes, dev)
    _lib.call("jfx_sort", kc, _lib.ptr(k), _lib.ptr(out_k), int(want_keys), pay_bytes,
              _lib.ptr(pay_in), _lib.ptr(pay_out), n, int(not ascending), digit_bits,
              _lib.ptr(ws), ws.numel(), ctypes.c_void_p(_lib.stream_handle(dev)))
    res_k = (out_k.to(orig_dtype) if out_k.dtype != orig_dtype else out_k) if want_keys else None
    idx = None
    if need_index:
        idx = pay_out if index_dtype == torch.int32 else pay_out.long()
        res_v = None if v is None else v[pay_out.long()]
    elif v is not None:
        res_v = pay_out.view(v.dtype)
    else:
        res_v = None
    if return_index:
        return res_k, res_v, idx
    return res_k, res_v


def argsort(keys, ascending: bool = True) -> torch.Tensor:
    """int64 row order of a stable sort (keys not written back)."""
    _, _, idx = radix_sort(keys, None, ascending, want_keys=False, return_index=True)
    return idx
