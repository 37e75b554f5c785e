"""numpy restatement of jaxframes' sort-centric algorithms (the oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker, never by the
product path.

Every function cites the reference code it restates
(``src/jaxframes/distributed/parallel_algorithms.py`` is abbreviated ``pa``).
The reference runs these steps as XLA ops (``jnp.argsort``, ``jnp.bincount``,
``jax.ops.segment_*``); here they are numpy ops with the same integer
semantics. How each function is pinned (golden vectors transcribed from the
reference's tests, closed forms, brute force, pandas, published MurmurHash3
vectors) is listed in ``PINNED_BY`` at the bottom of this file and enforced by
``tests/test_oracle.py``. Float sums are restated in float64 (the FP64 oracle);
the exact float32 bits of XLA's scatter-add are not pinnable without JAX, so
float sums are compared within a stated tolerance.
"""

from __future__ import annotations

import numpy as np

__all__ = [
    "ordered_bits", "radix_sort", "parallel_sort", "multi_column_lexsort",
    "head_flags", "groupby_aggregate", "groupby_aggregate_multi_column",
    "segment_sums_f64", "inner_join_pairs", "inner_join_pairs_bruteforce",
    "sort_merge_join", "fmix32", "hash_multi_columns", "hash_partition",
    "bucketize_padded", "logical_pad_value", "compact_rows", "filter_mask_compact", "join_multi_column",
    "gather_join_column", "encode_string_array", "string_join_plan",
    "groupby_encoded_strings",
    "murmur3_x86_32", "sum_error_metrics", "PINNED_BY",
]

_SIGN32 = np.uint32(0x80000000)
_SIGN64 = np.uint64(0x8000000000000000)


def ordered_bits(keys: np.ndarray) -> np.ndarray:
    """Map keys to unsigned integers whose unsigned order is the key order.

    Signed ints: XOR of the sign bit (``pa:495-501``), inverse at ``pa:569-575``.
    Floats: the reference sorts them with ``jnp.argsort`` (``pa:558-568``), which
    treats -0.0 == +0.0 and puts NaN last; here -0.0 is canonicalised to +0.0
    and every NaN to +NaN, then negatives are bit-inverted and non-negatives get
    the sign bit set (the transform of ``pa:583-603`` on canonical values).
    """
    k = np.asarray(keys)
    dt = k.dtype
    if dt == np.int32:
        return k.view(np.uint32) ^ _SIGN32
    if dt == np.int64:
        return k.view(np.uint64) ^ _SIGN64
    if dt in (np.uint32, np.uint64):
        return k.copy()
    if dt == np.float32 or dt == np.float64:
        ut, sign = (np.uint32, _SIGN32) if dt == np.float32 else (np.uint64, _SIGN64)
        c = np.where(k == 0, dt.type(0), k)
        c = np.where(np.isnan(c), dt.type(np.nan), c).astype(dt)
        bits = c.view(ut)
        neg = (bits & sign) != 0
        return np.where(neg, ~bits, bits | sign).astype(ut)
    raise TypeError(f"unsupported key dtype {dt}")


def radix_sort(keys, values=None, ascending=True, bits_per_pass=8):
    """LSD radix sort as the reference runs it (``pa:463-581``).

    Each pass extracts a ``bits_per_pass`` digit (``pa:336-338``) and applies a
    stable permutation by digit (the stable ``argsort`` of ``pa:400-405``).
    Descending output is the reverse of the ascending one (``pa:552-555``), so
    ties come out in reverse original order. Returns (keys, values, perm).
    """
    k = np.asarray(keys)
    n = k.shape[0]
    ok = ordered_bits(k)
    nbits = ok.dtype.itemsize * 8
    perm = np.arange(n, dtype=np.int64)
    cur = ok
    mask = (1 << bits_per_pass) - 1
    npass = -(-nbits // bits_per_pass)
    for p in range(npass):
        shift = p * bits_per_pass
        digit = ((cur >> ok.dtype.type(shift)) & ok.dtype.type(mask)).astype(
            np.uint16 if bits_per_pass > 8 else np.uint8)
        order = np.argsort(digit, kind="stable")
        cur = cur[order]
        perm = perm[order]
    if not ascending:
        perm = perm[::-1]
    out_v = None if values is None else np.asarray(values)[perm]
    return k[perm], out_v, perm


def parallel_sort(arr, values=None, ascending=True):
    """Module API wrapper (``pa:1079-1114``)."""
    sk, sv, _ = radix_sort(arr, values, ascending)
    return sk if values is None else (sk, sv)


def multi_column_lexsort(keys, ascending=True):
    """Stable LSD over columns, last key first (``pa:25-66``).

    Descending float keys are negated and integer keys become ``max - x``
    (``pa:56-61``), so descending ties keep their original order here.
    """
    if len(keys) == 1:
        return np.argsort(np.asarray(keys[0]), kind="stable")
    if isinstance(ascending, bool):
        ascending = [ascending] * len(keys)
    n = len(keys[0])
    idx = np.arange(n)
    for arr, asc in zip(reversed(keys), reversed(ascending)):
        a = np.asarray(arr)
        if not asc:
            a = -a if a.dtype.kind == "f" else (a.max() - a)
        idx = idx[np.argsort(a[idx], kind="stable")]
    return idx


def head_flags(sorted_keys) -> np.ndarray:
    """Group heads: row 0 and every row whose key differs from its predecessor
    under ``!=`` (``pa:767-770``); NaN keys are therefore each their own group."""
    sk = np.asarray(sorted_keys)
    f = np.ones(sk.shape[0], dtype=bool)
    if sk.shape[0] > 1:
        f[1:] = sk[1:] != sk[:-1]
    return f


def segment_sums_f64(sorted_vals, starts):
    """FP64 per-segment sums (the oracle for float32 Kahan sums)."""
    v = np.asarray(sorted_vals, dtype=np.float64)
    if v.shape[0] == 0:
        return np.zeros(0)
    return np.add.reduceat(v, starts)


def _aggregate(sorted_vals, starts, counts, fn, exact_float=False):
    """One aggregation per segment (``pa:626-660``).

    sum keeps the value dtype (ints wrap like XLA's; floats are returned in
    float64 as the FP64 oracle unless exact_float); mean = sum / max(count, 1);
    count is the row count per group (int64 here; the reference returns it in
    the value dtype, value-equal); min/max are exact and propagate NaN.
    """
    v = np.asarray(sorted_vals)
    if fn == "sum":
        if v.dtype.kind in "iu":
            return np.add.reduceat(v.astype(np.int64), starts).astype(v.dtype)
        if v.dtype == np.bool_:
            return np.add.reduceat(v.astype(np.int64), starts)
        s = segment_sums_f64(v, starts)
        return s.astype(v.dtype) if exact_float else s
    if fn == "mean":
        s = segment_sums_f64(v, starts)
        return s / np.maximum(counts, 1)
    if fn == "count":
        return counts.astype(np.int64)
    if fn == "max":
        return np.maximum.reduceat(v, starts)
    if fn == "min":
        return np.minimum.reduceat(v, starts)
    raise ValueError(f"Unsupported aggregation function: {fn}")


def groupby_aggregate(keys, values, agg_funcs):
    """Sort-based groupby (``pa:733-803``; aggregations ``pa:626-660``).

    Stable sort of the keys with a row-index payload (``pa:762-763``), head
    flags, unique keys at the heads (``pa:773-774``) in ascending key order,
    then one segmented reduction per column. Empty input gives empty outputs
    (``pa:757-758``).
    """
    k = np.asarray(keys)
    vals = {c: np.asarray(v) for c, v in values.items()}
    if k.shape[0] == 0:
        return k[:0], {c: v[:0] for c, v in vals.items()}
    sk, _, perm = radix_sort(k)
    flags = head_flags(sk)
    starts = np.flatnonzero(flags)
    counts = np.diff(np.append(starts, k.shape[0]))
    out = {}
    for col, fn in agg_funcs.items():
        out[col] = _aggregate(vals[col][perm], starts, counts, fn)
    return sk[starts], out


def groupby_aggregate_multi_column(keys, key_names, values, agg_funcs):
    """Multi-key groupby (``pa:662-731``): lexsort, heads OR'ed across keys."""
    ks = [np.asarray(k) for k in keys]
    if not ks or ks[0].shape[0] == 0:
        return ({n: k[:0] for n, k in zip(key_names, ks)},
                {c: np.asarray(v)[:0] for c, v in values.items()})
    idx = multi_column_lexsort(ks)
    n = ks[0].shape[0]
    flags = np.zeros(n, dtype=bool)
    flags[0] = True
    for k in ks:
        sk = k[idx]
        flags[1:] |= sk[1:] != sk[:-1]
    starts = np.flatnonzero(flags)
    counts = np.diff(np.append(starts, n))
    ukeys = {name: k[idx][starts] for name, k in zip(key_names, ks)}
    out = {}
    for col, fn in agg_funcs.items():
        out[col] = _aggregate(np.asarray(values[col])[idx], starts, counts, fn)
    return ukeys, out


def inner_join_pairs(left_keys, right_keys):
    """Inner-join row pairs in the reference's order (``pa:968-994``).

    Both sides are stably sorted with a row-index payload; ``jnp.where`` over the
    L x R match matrix enumerates pairs row-major in (sorted-left position,
    sorted-right position), i.e. (key asc, left row asc, right row asc).
    Closed form here: for each sorted left row, the matching right rows are the
    range [lower_bound, upper_bound) of its key in the sorted right keys.
    Returns (keys, left_rows, right_rows) as int64 row indices.
    """
    lk = np.asarray(left_keys)
    rk = np.asarray(right_keys)
    sl, _, lperm = radix_sort(lk)
    sr, _, rperm = radix_sort(rk)
    osl, osr = ordered_bits(sl), ordered_bits(sr)
    lo = np.searchsorted(osr, osl, side="left")
    hi = np.searchsorted(osr, osl, side="right")
    cnt = (hi - lo).astype(np.int64)
    if sl.dtype.kind == "f":
        cnt[np.isnan(sl)] = 0          # NaN == NaN is False in the match matrix
    m = int(cnt.sum())
    li = np.repeat(np.arange(sl.shape[0], dtype=np.int64), cnt)
    first = np.cumsum(cnt) - cnt
    rj = np.repeat(lo.astype(np.int64) - first, cnt) + np.arange(m, dtype=np.int64)
    return sl[li], lperm[li], rperm[rj]


def inner_join_pairs_bruteforce(left_keys, right_keys):
    """Nested-loop inner join for tiny inputs, pairs in the same order."""
    lk = list(np.asarray(left_keys))
    rk = list(np.asarray(right_keys))
    lord = sorted(range(len(lk)), key=lambda i: (ordered_bits(np.asarray([lk[i]]))[0], i))
    rord = sorted(range(len(rk)), key=lambda j: (ordered_bits(np.asarray([rk[j]]))[0], j))
    keys, li, ri = [], [], []
    for i in lord:
        for j in rord:
            if lk[i] == rk[j]:
                keys.append(lk[i])
                li.append(i)
                ri.append(j)
    return (np.asarray(keys, dtype=np.asarray(left_keys).dtype),
            np.asarray(li, dtype=np.int64), np.asarray(ri, dtype=np.int64))


def _gather_payload(col, idx):
    """Payload gather with NaN for idx < 0 (``pa:1053-1073``): the reference's
    ``jnp.where(valid, col[idx], nan)`` turns integer payloads into floats."""
    c = np.asarray(col)
    idx = np.asarray(idx)
    out_dt = c.dtype
    if c.dtype.kind in "iub":
        out_dt = np.float64 if c.dtype.itemsize == 8 else np.float32
    g = c[np.maximum(idx, 0)].astype(out_dt) if idx.shape[0] else np.zeros(0, out_dt)
    if idx.shape[0]:
        g[idx < 0] = np.nan
    return g


def sort_merge_join(left_keys, left_values, right_keys, right_values, how="inner"):
    """Single-key join (``pa:938-1075``).

    inner: all pairs (see inner_join_pairs). left/right: one row per row of the
    preserved side in sorted order, matched to the FIRST equal key of the other
    side via searchsorted, -1 otherwise (``pa:1000-1034``). outer: one row per
    distinct key of the union with first matches (``pa:1035-1046``). Payload
    columns are gathered with NaN fill and named left_<c> / right_<c>.
    """
    lk, rk = np.asarray(left_keys), np.asarray(right_keys)
    if how == "inner":
        keys, li, ri = inner_join_pairs(lk, rk)
    else:
        sl, _, lperm = radix_sort(lk)
        sr, _, rperm = radix_sort(rk)

        def first_match(sorted_side, perm, probe):
            pos = np.searchsorted(ordered_bits(sorted_side), ordered_bits(probe), side="left")
            ok = pos < sorted_side.shape[0]
            pc = np.minimum(pos, max(sorted_side.shape[0] - 1, 0))
            if sorted_side.shape[0]:
                ok &= sorted_side[pc] == probe
            else:
                ok &= False
            res = np.full(probe.shape[0], -1, dtype=np.int64)
            res[ok] = perm[pc[ok]]
            return res

        if how == "left":
            keys, li = sl, lperm
            ri = first_match(sr, rperm, sl)
        elif how == "right":
            keys, ri = sr, rperm
            li = first_match(sl, lperm, sr)
        elif how == "outer":
            keys = np.unique(np.concatenate([sl, sr]))
            li = first_match(sl, lperm, keys)
            ri = first_match(sr, rperm, keys)
        else:
            raise ValueError(f"unsupported join type {how}")
    out = {}
    for c, v in left_values.items():
        out[f"left_{c}"] = _gather_payload(v, li)
    for c, v in right_values.items():
        out[f"right_{c}"] = _gather_payload(v, ri)
    return keys, out, li, ri


def fmix32(h):
    """MurmurHash3 32-bit finalizer, as applied at ``pa:96-100``."""
    h = np.asarray(h, dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        h ^= h >> np.uint32(16)
        h *= np.uint32(0x85EBCA6B)
        h ^= h >> np.uint32(13)
        h *= np.uint32(0xC2B2AE35)
        h ^= h >> np.uint32(16)
    return h


def _to_u32(key):
    """``key.astype(uint32)`` for ints (modular: the low 32 bits) and the bit
    view for float32 (``pa:88-92``, ``pa:109-118``)."""
    k = np.asarray(key)
    if k.dtype == np.float32:
        return k.view(np.uint32)
    if k.dtype == np.float64:
        return k.astype(np.float32).view(np.uint32)
    return k.astype(np.int64).astype(np.uint64).astype(np.uint32) if k.dtype.kind == "i" \
        else k.astype(np.uint64).astype(np.uint32)


def hash_multi_columns(keys):
    """Row hash (``pa:69-136``). One key: fmix32 of its low 32 bits. K keys:
    seed 0x1337; per key i: h ^= k*(0x9E3779B1+i); h = rotl(h,13);
    h *= 0x85EBCA6B; then fmix32.

    For a single float64 key the reference calls ``key.view(uint32)`` (``pa:91``),
    which reinterprets each float64 as two uint32 words; that length-doubling
    path is not restated (float64 single keys raise here).
    """
    if len(keys) == 1:
        k = np.asarray(keys[0])
        if k.dtype == np.float64:
            raise NotImplementedError("single float64 key hash is not restated")
        return fmix32(_to_u32(k))
    h = np.full(np.asarray(keys[0]).shape[0], 0x1337, dtype=np.uint32)
    with np.errstate(over="ignore"):
        for i, key in enumerate(keys):
            ku = _to_u32(key)
            h ^= ku * np.uint32((0x9E3779B1 + i) & 0xFFFFFFFF)
            h = (h << np.uint32(13)) | (h >> np.uint32(19))
            h *= np.uint32(0x85EBCA6B)
    return fmix32(h)


def hash_partition(keys, nbuckets, hashes=None):
    """Local hash partition (``pa:163-182`` with the target of ``pa:214``).

    bucket = hash % nbuckets; rows are stably grouped by bucket (within a
    bucket they keep their original order). Returns (perm, offsets) where
    offsets has nbuckets+1 entries: the compact form of the reference's padded
    ``[P, local_n]`` buffer.
    """
    h = hash_multi_columns([keys]) if hashes is None else np.asarray(hashes, np.uint32)
    b = (h % np.uint32(nbuckets)).astype(np.int64)
    perm = np.argsort(b, kind="stable")
    counts = np.bincount(b, minlength=nbuckets)
    offsets = np.zeros(nbuckets + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    return perm, offsets


def gather_join_column(arr, idx, dtype_name=None):
    """``_gather_join_column`` (``pa:1254-1278``): -1 for string columns where
    idx < 0; otherwise NaN, integer / bool columns turning float32 only when
    some index is negative."""
    a = np.asarray(arr)
    idx = np.asarray(idx, dtype=np.int64)
    if idx.shape[0] == 0:
        return a[:0]
    valid = idx >= 0
    g = a[np.maximum(idx, 0)]
    if dtype_name == "string":
        return np.where(valid, g, np.asarray(-1, dtype=g.dtype))
    if valid.all():
        return g
    if g.dtype.kind in "iub":
        g = g.astype(np.float32)
    return np.where(valid, g, np.asarray(np.nan, dtype=g.dtype))


def join_multi_column(left_keys, left_key_names, left_values, right_keys, right_key_names,
                      right_values, how="inner", left_value_dtypes=None, right_value_dtypes=None):
    """``ParallelSortMergeJoin.join_multi_column`` (``pa:826-936``) restated with
    the same L x R match matrix (small inputs only): pairs row-major over the
    unsorted rows, then unmatched left rows, then unmatched right rows, -1 for
    the missing side; keys coalesced left-first (``pa:1245-1251``)."""
    lks = [np.asarray(k) for k in left_keys]
    rks = [np.asarray(k) for k in right_keys]
    nl, nr = lks[0].shape[0], rks[0].shape[0]
    m = np.ones((nl, nr), dtype=bool)
    for a, b in zip(lks, rks):
        m &= a[:, None] == b[None, :]
    li, ri = np.nonzero(m)
    lun = np.flatnonzero(~m.any(axis=1))
    run = np.flatnonzero(~m.any(axis=0))
    neg = lambda k: np.full(k, -1, dtype=np.int64)
    if how == "inner":
        al, ar = li, ri
    elif how == "left":
        al, ar = np.concatenate([li, lun]), np.concatenate([ri, neg(lun.size)])
    elif how == "right":
        al, ar = np.concatenate([li, neg(run.size)]), np.concatenate([ri, run])
    else:
        al = np.concatenate([li, lun, neg(run.size)])
        ar = np.concatenate([ri, neg(lun.size), run])
    keys = {}
    for name, a, b in zip(left_key_names, lks, rks):
        if al.shape[0] == 0:
            keys[name] = a[:0]
        else:
            keys[name] = np.where(al >= 0, a[np.maximum(al, 0)] if nl else 0,
                                  b[np.maximum(ar, 0)] if nr else 0)
    vals = {}
    for c, v in left_values.items():
        vals[f"left_{c}"] = gather_join_column(v, al, (left_value_dtypes or {}).get(c))
    for c, v in right_values.items():
        vals[f"right_{c}"] = gather_join_column(v, ar, (right_value_dtypes or {}).get(c))
    return keys, vals


def logical_pad_value(dtype_name, arr):
    """Pad value of the padded bucket layout (``pa:150-160``): string codes
    -1, floats NaN, integers -999999, bools False, else 0, in the column's
    dtype. -999999 is cast to unsigned and narrow integer dtypes with numpy's
    wrapping cast (``jnp.asarray(-999999, dtype=uint32)`` is not pinnable
    without JAX: parity unpinned for those dtypes)."""
    dt = np.asarray(arr).dtype
    if dtype_name == "string":
        return np.array(-1).astype(dt)[()]
    if dt.kind == "f":
        return dt.type(np.nan)
    if dt.kind in "iu":
        return np.array(-999999).astype(dt)[()]
    if dt == np.bool_:
        return np.False_
    return dt.type(0)


def bucketize_padded(arr, targets, num_devices, pad_value):
    """The reference's padded bucket layout (``pa:163-182``): row r of bucket d
    at ``[d, r]``, the rest filled with pad_value."""
    a = np.asarray(arr)
    t = np.asarray(targets).astype(np.int64)
    order = np.argsort(t, kind="stable")
    st = t[order]
    counts = np.bincount(st, minlength=num_devices)
    starts = np.concatenate([[0], np.cumsum(counts[:-1])])
    pos = np.arange(a.shape[0]) - starts[st]
    out = np.full((num_devices, a.shape[0]), pad_value, dtype=a.dtype)
    out[st, pos] = a[order]
    return out


def compact_rows(valid_mask):
    """Row positions kept by ``compact_repartitioned_rows`` (``pa:277-293``)."""
    return np.flatnonzero(np.asarray(valid_mask))


def filter_mask_compact(columns, mask):
    """Boolean-mask row filter (``core/frame.py:321-338``): order preserved."""
    m = np.asarray(mask, dtype=bool)
    return {c: np.asarray(v)[m] for c, v in columns.items()}


def murmur3_x86_32(data: bytes, seed: int = 0) -> int:
    """Full MurmurHash3_x86_32 (Appleby, public domain) in pure Python; its
    finalizer is fmix32. Used only to pin fmix32 against published vectors."""
    c1, c2 = 0xCC9E2D51, 0x1B873593
    h = seed & 0xFFFFFFFF
    n = len(data)
    nblocks = n // 4
    for i in range(nblocks):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = (k * c1) & 0xFFFFFFFF
        k = ((k << 15) | (k >> 17)) & 0xFFFFFFFF
        k = (k * c2) & 0xFFFFFFFF
        h ^= k
        h = ((h << 13) | (h >> 19)) & 0xFFFFFFFF
        h = (h * 5 + 0xE6546B64) & 0xFFFFFFFF
    tail = data[4 * nblocks:]
    k = 0
    if len(tail) >= 3:
        k ^= tail[2] << 16
    if len(tail) >= 2:
        k ^= tail[1] << 8
    if len(tail) >= 1:
        k ^= tail[0]
        k = (k * c1) & 0xFFFFFFFF
        k = ((k << 15) | (k >> 17)) & 0xFFFFFFFF
        k = (k * c2) & 0xFFFFFFFF
        h ^= k
    h ^= n
    return int(fmix32(np.asarray([h], dtype=np.uint32))[0])


def sum_error_metrics(got, sorted_vals, starts):
    """(max relative error, max error / sum|x|) of float sums vs FP64."""
    ref = segment_sums_f64(sorted_vals, starts)
    absx = np.add.reduceat(np.abs(np.asarray(sorted_vals, np.float64)), starts) \
        if len(starts) else np.zeros(0)
    g = np.asarray(got, np.float64)
    err = np.abs(g - ref)
    rel = err / np.maximum(np.abs(ref), np.finfo(np.float64).tiny)
    nrm = err / np.maximum(absx, np.finfo(np.float64).tiny)
    return (float(rel.max()) if rel.size else 0.0, float(nrm.max()) if nrm.size else 0.0)


# ---- dictionary-encoded strings (core/string_encoding.py, core/string_fastpaths.py) ----

def _null_string(v) -> bool:
    return v is None or (isinstance(v, float) and np.isnan(v))


def encode_string_array(arr):
    """(int32 codes, vocab): vocab = sorted distinct non-null strings, -1 for
    null (string_encoding.py:38-51)."""
    obj = np.asarray(arr, dtype=object)
    vocab = tuple(sorted({str(v) for v in obj if not _null_string(v)}))
    pos = {v: i for i, v in enumerate(vocab)}
    codes = np.array([-1 if _null_string(v) else pos[str(v)] for v in obj], dtype=np.int32)
    return codes, vocab


def string_join_plan(left_codes, right_codes, num_codes, how):
    """(joined codes, left rows, right rows), -1 for none, of a join over
    aligned codes; null keys are one more code (string_fastpaths.py:95-119).
    Right codes unique: every left row looks up its right row, pairs in left
    row order (:200-243). Otherwise, code by code ascending, the cartesian
    product of the code's left and right rows, left-major, or its unmatched
    rows (:246-309)."""
    null = num_codes
    ld = np.where(np.asarray(left_codes) < 0, null, left_codes).astype(np.int64)
    rd = np.where(np.asarray(right_codes) < 0, null, right_codes).astype(np.int64)
    restore = lambda c: np.where(np.asarray(c) == null, -1, c).astype(np.int32)
    rcount = np.bincount(rd, minlength=null + 1) if rd.size else np.zeros(null + 1, np.int64)
    if np.all(rcount <= 1):
        lookup = np.full(null + 1, -1, dtype=np.int64)
        lookup[rd] = np.arange(rd.size)
        matched = lookup[ld] if ld.size else np.zeros(0, np.int64)
        m = matched >= 0
        lp = np.arange(ld.size)
        if how == "inner":
            return restore(ld[m]), lp[m].astype(np.int32), matched[m].astype(np.int32)
        if how == "left":
            return restore(ld), lp.astype(np.int32), matched.astype(np.int32)
        hit = np.zeros(rd.size, bool)
        hit[matched[m]] = True
        um = np.flatnonzero(~hit)
        none = np.full(um.size, -1)
        if how == "right":
            return (restore(np.concatenate([ld[m], rd[um]])),
                    np.concatenate([lp[m], none]).astype(np.int32),
                    np.concatenate([matched[m], um]).astype(np.int32))
        return (restore(np.concatenate([ld, rd[um]])), np.concatenate([lp, none]).astype(np.int32),
                np.concatenate([matched, um]).astype(np.int32))
    codes, lo, ro = [], [], []
    top = int(max(ld.max(initial=0), rd.max(initial=0)))
    for c in range(top + 1):
        lrows = np.flatnonzero(ld == c)
        rrows = np.flatnonzero(rd == c)
        if lrows.size and rrows.size:
            codes.append(np.full(lrows.size * rrows.size, c))
            lo.append(np.repeat(lrows, rrows.size))
            ro.append(np.tile(rrows, lrows.size))
        elif lrows.size and how in ("left", "outer"):
            codes.append(np.full(lrows.size, c))
            lo.append(lrows)
            ro.append(np.full(lrows.size, -1))
        elif rrows.size and how in ("right", "outer"):
            codes.append(np.full(rrows.size, c))
            lo.append(np.full(rrows.size, -1))
            ro.append(rrows)
    if not codes:
        e = np.zeros(0, np.int32)
        return e, e, e
    return (restore(np.concatenate(codes)), np.concatenate(lo).astype(np.int32),
            np.concatenate(ro).astype(np.int32))


def groupby_encoded_strings(keys, values, agg_funcs, vocab_size):
    """Observed codes ascending and the aggregates over the non-null rows
    (string_fastpaths.py:41-74, 161-197); sums in float64 (the FP64 oracle),
    count as int32."""
    keys = np.asarray(keys)
    valid = keys >= 0
    vk = keys[valid]
    counts = np.bincount(vk, minlength=vocab_size) if vk.size else np.zeros(vocab_size, np.int64)
    obs = np.flatnonzero(counts > 0)
    out = {}
    for c, agg in agg_funcs.items():
        v = np.asarray(values[c])[valid]
        if agg == "count":
            out[c] = counts[obs].astype(np.int32)
        elif agg in ("sum", "mean"):
            s = np.bincount(vk, weights=v.astype(np.float64), minlength=vocab_size)[obs]
            out[c] = s if agg == "sum" else s / counts[obs]
        elif agg in ("min", "max"):
            f = np.minimum if agg == "min" else np.maximum
            r = np.full(vocab_size, np.inf if agg == "min" else -np.inf)
            f.at(r, vk, v.astype(np.float64))
            out[c] = r[obs]
        else:
            raise ValueError(f"Unsupported aggregation function: {agg}")
    return obs.astype(np.int32), out


# How each function is pinned; tests/test_oracle.py runs every check named here.
PINNED_BY = {
    "ordered_bits": "closed form: unsigned order == numpy order incl. -0.0/NaN",
    "radix_sort": "golden vectors (test_parallel_algorithms.py:29-84, "
                  "test_lazy_operations.py:12-95) + closed form np.argsort(kind='stable')",
    "multi_column_lexsort": "golden vectors (test_lazy_operations.py:39-69) + np.lexsort",
    "groupby_aggregate": "golden vectors (test_parallel_algorithms.py:90-152, 260-296, "
                         "test_lazy_operations.py:101-181) + pandas groupby",
    "groupby_aggregate_multi_column": "pandas groupby on two keys",
    "inner_join_pairs": "golden vectors (test_parallel_algorithms.py:158-208) + "
                        "nested-loop brute force + pandas.merge row count",
    "sort_merge_join": "golden vectors (test_parallel_algorithms.py:210-230, "
                       "test_lazy_operations.py:204-268)",
    "fmix32": "published MurmurHash3_x86_32 vectors (seeded empty inputs reduce to fmix32)",
    "hash_multi_columns": "fmix32 pin + reference test properties "
                          "(test_distributed_multi_column.py:118-179)",
    "hash_partition": "bucketize_padded with padding removed + per-bucket order check",
    "bucketize_padded": "brute-force per-bucket loop",
    "compact_rows": "np.flatnonzero closed form (trivial)",
    "filter_mask_compact": "boolean-index closed form (trivial)",
    "join_multi_column": "golden vectors (test_multi_column_operations.py:206-303) + "
                         "pandas inner merge (same (left row, right row) order) + "
                         "nested-loop brute force for left/right/outer",
    "gather_join_column": "hand-checked cases of the three branches (string / all valid / "
                          "unmatched int -> float32 NaN)",
    "encode_string_array": "golden vectors (test_string_operations.py:12-117) + np.unique",
    "string_join_plan": "golden vectors (test_string_operations.py:55-117) + pandas merge "
                        "(pair multiset; inner many-to-many order = sorted (key, left, right)) "
                        "+ nested-loop brute force of the unique-right lookup",
    "groupby_encoded_strings": "golden vectors (test_string_operations.py:31-53) + pandas groupby "
                               "with dropna",
}
