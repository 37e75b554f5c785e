"""Dictionary-encoded string columns.

Follows core/string_encoding.py and core/string_fastpaths.py of the
reference: a string column is int32 codes into a lexicographically sorted
vocabulary (null -> -1); groupby on a string key aggregates by code over the
non-null rows; a join on a string key aligns both vocabularies and then either
looks every left row up in a code -> right-row table (right keys unique,
pairs in left-row order) or enumerates, code by code, the cartesian product
of the code's left and right rows (many-to-many). Encoding and vocabulary
merges run on the host (they are dictionary work on distinct strings); the
per-row work runs on the GPU kernels of this package (compaction, the
small-key-range groupby, the sort-merge join, the first-match lookup, the
radix sort and gathers).
"""

from __future__ import annotations

from bisect import bisect_left
from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

from . import _lib

SUPPORTED_STRING_GROUPBY_AGGS = {"sum", "mean", "min", "max", "count"}
SUPPORTED_STRING_JOIN_HOW = {"inner", "left", "right", "outer"}


def _is_null(value: Any) -> bool:
    """string_encoding.py:13-19."""
    return value is None or (isinstance(value, float) and np.isnan(value))


def is_string_array(arr: np.ndarray) -> bool:
    """string_encoding.py:22-35: str/bytes arrays, or object arrays of
    strings and nulls."""
    if arr.dtype.kind in ("U", "S"):
        return True
    if arr.dtype != np.object_:
        return False
    return all(_is_null(v) or isinstance(v, str) for v in arr)


def _device():
    return torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() \
        else torch.device("cpu")


def encode_string_array(arr: np.ndarray):
    """(int32 codes on the GPU, vocab): vocab = sorted distinct non-null
    strings, code -1 for null (string_encoding.py:38-51)."""
    obj = np.asarray(arr, dtype=object)
    null = np.fromiter((_is_null(v) for v in obj), dtype=bool, count=obj.shape[0])
    strs = np.array([str(v) for v in obj[~null]], dtype=object)
    vocab = tuple(sorted(set(strs.tolist())))
    codes = np.full(obj.shape[0], -1, dtype=np.int32)
    if vocab:
        index = {s: i for i, s in enumerate(vocab)}
        codes[~null] = np.fromiter((index[s] for s in strs), dtype=np.int32, count=strs.shape[0])
    return torch.from_numpy(codes).to(_device()), vocab


def decode_string_codes(codes, vocab: tuple) -> np.ndarray:
    """Object array of the strings, None for -1 (string_encoding.py:54-62)."""
    c = codes.cpu().numpy() if isinstance(codes, torch.Tensor) else np.asarray(codes)
    table = np.array(list(vocab) + [None], dtype=object)
    return table[np.where(c < 0, len(vocab), c)]


def string_scalar_rank(vocab: tuple, value: Any) -> int:
    """Lexical insertion rank (string_encoding.py:65-69)."""
    if _is_null(value):
        return -1
    return bisect_left(vocab, str(value))


def string_scalar_code(vocab: tuple, value: Any):
    """Exact code, -1 for null, None when absent (string_encoding.py:72-80)."""
    if _is_null(value):
        return -1
    r = string_scalar_rank(vocab, value)
    return r if r < len(vocab) and vocab[r] == str(value) else None


def remap_string_codes(codes: torch.Tensor, source_vocab: tuple, target_vocab: tuple):
    """Codes of source_vocab re-expressed in target_vocab (string_encoding.py:83-94)."""
    if source_vocab == target_vocab:
        return codes
    pos = {v: i for i, v in enumerate(target_vocab)}
    mapping = torch.tensor([pos[v] for v in source_vocab] + [-1], dtype=torch.int32,
                           device=codes.device)
    idx = torch.where(codes < 0, torch.full_like(codes, len(source_vocab)), codes)
    from .compaction import take
    return take(mapping, idx.long())


def align_string_code_arrays(left_codes, left_vocab: tuple, right_codes, right_vocab: tuple):
    """Both code arrays in the merged sorted vocabulary (string_encoding.py:97-112)."""
    if left_vocab == right_vocab:
        return left_codes, right_codes, left_vocab
    merged = tuple(sorted(set(left_vocab).union(right_vocab)))
    return (remap_string_codes(left_codes, left_vocab, merged),
            remap_string_codes(right_codes, right_vocab, merged), merged)


def sortable_string_codes(codes: torch.Tensor, vocab: tuple, descending: bool = False):
    """Sort keys in lexical order with nulls last (string_encoding.py:115-124)."""
    null_rank = len(vocab)
    if not descending:
        return torch.where(codes < 0, torch.full_like(codes, null_rank), codes)
    return torch.where(codes < 0, torch.full_like(codes, null_rank), (len(vocab) - 1) - codes)


# ---- groupby ---------------------------------------------------------------------------

def can_use_string_groupby_fastpath(key_dtype: str, agg_dict: dict, value_dtypes: dict) -> bool:
    """string_fastpaths.py:25-38."""
    if key_dtype != "string" or not agg_dict:
        return False
    return all(agg in SUPPORTED_STRING_GROUPBY_AGGS and value_dtypes[c] not in ("object", "string")
               for c, agg in agg_dict.items())


def groupby_encoded_strings(keys: torch.Tensor, values: dict, agg_funcs: dict, vocab_size: int):
    """Aggregates by code over the rows with a non-null key; observed codes
    ascending (string_fastpaths.py:41-74). count is int32, sum / min / max
    keep the value dtype, mean is at least float32."""
    from .compaction import compact_rows, take
    from .grouping import groupby_aggregate
    valid = compact_rows(keys >= 0)
    if valid.shape[0] == 0:
        out = {}
        for c, agg in agg_funcs.items():
            v = values[c]
            dt = torch.int32 if agg == "count" else (
                torch.promote_types(v.dtype, torch.float32) if agg == "mean" else v.dtype)
            out[c] = torch.empty(0, dtype=dt, device=keys.device)
        return keys[:0], out
    vk = take(keys, valid) if valid.shape[0] != keys.shape[0] else keys
    res = {}
    uk = None
    for c, agg in agg_funcs.items():
        v = values[c]
        vv = take(v, valid) if valid.shape[0] != keys.shape[0] else v
        if agg == "mean":
            # sum / count in the result type (float32 at least), as bincount does
            rt = torch.promote_types(v.dtype, torch.float32)
            uk, r = groupby_aggregate(vk, {"s": vv.to(rt), "n": vv}, {"s": "sum", "n": "count"})
            res[c] = r["s"] / r["n"].to(rt)
        else:
            uk, r = groupby_aggregate(vk, {"v": vv}, {"v": agg})
            res[c] = r["v"].to(torch.int32) if agg == "count" else r["v"]
    return uk.to(keys.dtype), res


# ---- join ----------------------------------------------------------------------------------

@dataclass(frozen=True)
class StringJoinPlan:
    """string_fastpaths.py:17-22: joined codes and row indices (-1: none)."""
    joined_codes: torch.Tensor
    left_indices: torch.Tensor
    right_indices: torch.Tensor


def can_use_string_join_fastpath(join_dtype: str, left_payload_dtypes: dict,
                                 right_payload_dtypes: dict, how: str) -> bool:
    """string_fastpaths.py:77-92."""
    if join_dtype != "string" or how not in SUPPORTED_STRING_JOIN_HOW:
        return False
    return all(d != "object" for d in left_payload_dtypes.values()) and \
        all(d != "object" for d in right_payload_dtypes.values())


def _restore_null(codes: torch.Tensor, null_code: int) -> torch.Tensor:
    return torch.where(codes == null_code, torch.full_like(codes, -1), codes).to(torch.int32)


def _arange(n: int, dev) -> torch.Tensor:
    return torch.arange(n, dtype=torch.int32, device=dev)


def join_encoded_strings(left_keys: torch.Tensor, right_keys: torch.Tensor, num_codes: int,
                         how: str) -> StringJoinPlan:
    """Row indices of a join over aligned codes (string_fastpaths.py:95-119).
    Null keys form one more code and match each other."""
    if how not in SUPPORTED_STRING_JOIN_HOW:
        raise ValueError(f"Unsupported join type: {how}")
    from .grouping import groupby_aggregate
    dev = left_keys.device
    null_code = num_codes
    ld = torch.where(left_keys < 0, torch.full_like(left_keys, null_code), left_keys).to(torch.int32)
    rd = torch.where(right_keys < 0, torch.full_like(right_keys, null_code), right_keys).to(torch.int32)
    unique = True
    if rd.shape[0]:
        _, cnt = groupby_aggregate(rd, {"c": rd}, {"c": "count"})
        unique = int(cnt["c"].max()) <= 1
    if unique:
        return _unique_right_plan(ld, rd, null_code, how, dev)
    return _many_to_many_plan(ld, rd, null_code, how, dev)


def _unique_right_plan(ld, rd, null_code, how, dev) -> StringJoinPlan:
    """string_fastpaths.py:200-243: each left row looks its code up."""
    from .compaction import compact_rows, take
    lookup = torch.full((null_code + 1,), -1, dtype=torch.int32, device=dev)
    if rd.shape[0]:
        lookup[rd.long()] = _arange(rd.shape[0], dev)
    matched = take(lookup, ld.long()) if ld.shape[0] else ld
    pos = compact_rows(matched >= 0).to(torch.int32)
    left_pos = _arange(ld.shape[0], dev)
    if how == "inner":
        return StringJoinPlan(_restore_null(take(ld, pos.long()), null_code), pos,
                              take(matched, pos.long()))
    if how == "left":
        return StringJoinPlan(_restore_null(ld, null_code), left_pos, matched)
    hit = torch.zeros(rd.shape[0], dtype=torch.bool, device=dev)
    if pos.shape[0]:
        hit[take(matched, pos.long()).long()] = True
    unmatched_r = compact_rows(~hit).to(torch.int32)
    none = torch.full((unmatched_r.shape[0],), -1, dtype=torch.int32, device=dev)
    if how == "right":
        codes = torch.cat([take(ld, pos.long()), take(rd, unmatched_r.long())])
        return StringJoinPlan(_restore_null(codes, null_code), torch.cat([pos, none]),
                              torch.cat([take(matched, pos.long()), unmatched_r]))
    codes = torch.cat([ld, take(rd, unmatched_r.long())])
    return StringJoinPlan(_restore_null(codes, null_code), torch.cat([left_pos, none]),
                          torch.cat([matched, unmatched_r]))


def _many_to_many_plan(ld, rd, null_code, how, dev) -> StringJoinPlan:
    """string_fastpaths.py:246-309: per code (ascending), the cartesian
    product of its left rows and right rows (left-major, rows ascending), or
    its unmatched rows for the outer sides. The inner pairs are the
    sort-merge join's (code, left row, right row) order; unmatched rows are
    merged in by a stable sort on the code (a code contributes to one list
    only)."""
    from .compaction import compact_rows, take
    from .joining import _first_match, _sorted_side, inner_join_rows
    from .sorting import radix_sort
    keys, li, ri = inner_join_rows(ld, rd)
    parts_c, parts_l, parts_r = [keys.to(torch.int32)], [li], [ri]
    if how in ("left", "outer"):
        lk, sl, lperm, _ = _sorted_side(ld)
        _, sr, rperm, _ = _sorted_side(rd)
        fm = _first_match(ld, sr, rperm, _lib.JFX_I32) if rd.shape[0] else \
            torch.full((ld.shape[0],), -1, dtype=torch.int32, device=dev)
        lonely = compact_rows(fm == -1).to(torch.int32)
        parts_c.append(take(ld, lonely.long()))
        parts_l.append(lonely)
        parts_r.append(torch.full((lonely.shape[0],), -1, dtype=torch.int32, device=dev))
    if how in ("right", "outer"):
        _, sl, lperm, _ = _sorted_side(ld)
        fm = _first_match(rd, sl, lperm, _lib.JFX_I32) if ld.shape[0] else \
            torch.full((rd.shape[0],), -1, dtype=torch.int32, device=dev)
        lonely = compact_rows(fm == -1).to(torch.int32)
        parts_c.append(take(rd, lonely.long()))
        parts_l.append(torch.full((lonely.shape[0],), -1, dtype=torch.int32, device=dev))
        parts_r.append(lonely)
    codes = torch.cat(parts_c)
    lidx = torch.cat(parts_l)
    ridx = torch.cat(parts_r)
    if len(parts_c) > 1 and codes.shape[0]:
        _, _, order = radix_sort(codes, None, True, want_keys=False, return_index=True)
        codes, lidx, ridx = take(codes, order), take(lidx, order), take(ridx, order)
    return StringJoinPlan(_restore_null(codes, null_code), lidx, ridx)


def gather_join_payload(arr, indices: torch.Tensor, dtype_name: str):
    """string_fastpaths.py:357-378: rows at indices; where an index is -1,
    -1 for a string column, else NaN (integer and bool payloads become
    float32 first). With every index valid the column keeps its dtype."""
    from .compaction import take
    if indices.shape[0] == 0:
        return arr[:0]
    if isinstance(arr, np.ndarray):
        idx = indices.cpu().numpy()
        out = arr[np.maximum(idx, 0)]
        if (idx < 0).any():
            out = out.astype(object)
            out[idx < 0] = None
        return out
    valid = indices >= 0
    g = take(arr, torch.clamp(indices, min=0).long())
    if dtype_name == "string":
        return torch.where(valid, g, torch.full_like(g, -1))
    if bool(valid.all()):
        return g
    if not g.dtype.is_floating_point and not g.dtype.is_complex:
        g = g.to(torch.float32)
    return torch.where(valid, g, torch.full_like(g, float("nan")))
