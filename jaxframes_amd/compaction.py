"""Row filters and stable compaction on the GPU.

- filter_rows: positions of the rows with `column <op> value` (the frame
  layer's boolean filter, core/frame.py:321-338, without a stored mask);
- compact_rows: positions of the nonzero entries of a mask;
- compact_repartitioned_rows: the reference function of that name
  (parallel_algorithms.py:277-293): keep the valid rows of every column, in
  row order, and return the inputs unchanged when every row is valid.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._tensors import to_device

OPS = {">": _lib.CMP_CODES[">"], ">=": _lib.CMP_CODES[">="], "<": _lib.CMP_CODES["<"],
       "<=": _lib.CMP_CODES["<="], "==": _lib.CMP_CODES["=="], "!=": _lib.CMP_CODES["!="]}
_COL_CODE = {torch.uint32: (_lib.JFX_U32, ctypes.c_uint32), torch.int32: (_lib.JFX_I32, ctypes.c_int32),
             torch.float32: (_lib.JFX_F32, ctypes.c_float), torch.uint64: (_lib.JFX_U64, ctypes.c_uint64),
             torch.int64: (_lib.JFX_I64, ctypes.c_int64), torch.float64: (_lib.JFX_F64, ctypes.c_double)}


def _ws(dev):
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_compact_workspace_size", 0, ctypes.byref(nb))
    return _lib.workspace(nb.value, dev)


def _stream(dev):
    return ctypes.c_void_p(_lib.stream_handle(dev))


def filter_rows(column, op: str, value) -> torch.Tensor:
    """int64 positions (ascending) of the rows where `column op value`."""
    _lib.require_gpu()
    if op not in OPS:
        raise ValueError(f"unsupported comparison {op!r}")
    col = to_device(column)
    if col.dtype not in _COL_CODE:
        col = col.to(torch.float32 if col.dtype.is_floating_point else torch.int64)
    code, ctype = _COL_CODE[col.dtype]
    n = col.shape[0]
    dev = col.device
    out = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    cnt = ctypes.c_uint64(0)
    val = ctype(value)
    ws = _ws(dev)
    _lib.call("jfx_filter_rows", code, _lib.ptr(col), OPS[op], ctypes.byref(val), n, _lib.ptr(out),
              ctypes.byref(cnt), _lib.ptr(ws), ws.numel(), _stream(dev))
    return out[:cnt.value].long()


def compact_rows(mask) -> torch.Tensor:
    """int64 positions (ascending) of the nonzero entries of a boolean/byte mask."""
    _lib.require_gpu()
    m = to_device(mask)
    if m.dtype != torch.uint8:
        m = (m != 0).to(torch.uint8)
    n = m.shape[0]
    dev = m.device
    out = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    cnt = ctypes.c_uint64(0)
    ws = _ws(dev)
    _lib.call("jfx_compact_rows", _lib.ptr(m), n, _lib.ptr(out), ctypes.byref(cnt), _lib.ptr(ws),
              ws.numel(), _stream(dev))
    return out[:cnt.value].long()


def take(column: torch.Tensor, rows: torch.Tensor) -> torch.Tensor:
    """column[rows] for device columns (rows int64, all valid)."""
    if column.element_size() in (4, 8) and rows.shape[0]:
        r32 = rows.to(torch.int32)
        out = torch.empty(rows.shape[0], dtype=column.dtype, device=column.device)
        _lib.call("jfx_gather", column.element_size(), _lib.ptr(column), _lib.ptr(r32), rows.shape[0],
                  _lib.ptr(out), 0, _stream(column.device))
        return out
    return column[rows]


def compact_repartitioned_rows(keys: list, values: dict, valid_mask):
    """(keys, values) restricted to the valid rows, row order kept
    (parallel_algorithms.py:277-293)."""
    m = to_device(valid_mask)
    if m.numel() == 0:
        return keys, values
    rows = compact_rows(m)
    if rows.shape[0] == m.shape[0]:
        return keys, values
    ks = [take(to_device(k, m.device), rows) for k in keys]
    vs = {c: take(to_device(v, m.device), rows) for c, v in values.items()}
    return ks, vs


def logical_pad_value(dtype_name: str, dtype) -> float | int | bool:
    """Per-column pad value of the padded repartition layout
    (_logical_pad_value, parallel_algorithms.py:150-160)."""
    if dtype_name == "string":
        return -1
    dt = np.dtype(str(dtype).replace("torch.", ""))
    if dt.kind == "f":
        return float("nan")
    if dt.kind in "iu":
        return -999999
    if dt.kind == "b":
        return False
    return 0
