"""GPU parity of the multi-column paths (composite keys and per-column LSD)
against the oracle: row orders bit-exact, group keys / counts / min / max
exact, sums within the float tolerances of the single-key tests."""

import json
import os

import numpy as np
import pytest
import torch

import oracle
from jaxframes_amd import grouping, multikey
from tests._util import cuda

pytestmark = pytest.mark.gpu

SUM_NORM_TOL = 2e-7   # |sum - f64 sum| / sum |x| for float32 values


def _perm(t):
    return t.cpu().numpy().astype(np.int64)


@pytest.mark.parametrize("dts", [(np.int32, np.int32), (np.int64, np.int32), (np.uint32, np.int64),
                                 (np.int32, np.uint64)])
@pytest.mark.parametrize("asc", [True, [True, False], [False, True], [False, False]])
def test_lexsort_small_ranges_composite(dts, asc):
    rng = np.random.default_rng(5)
    n = 50000
    a = rng.integers(0, 40, n).astype(dts[0])
    b = (rng.integers(0, 300, n) + (10 if np.dtype(dts[1]).kind == "u" else -150)).astype(dts[1])
    exp = oracle.multi_column_lexsort([a, b], asc)
    ka, kb = cuda(a), cuda(b)
    assert multikey.composite_plan([ka, kb], asc) is not None
    got = multikey.lexsort_perm([ka, kb], asc)
    assert np.array_equal(_perm(got), exp)


@pytest.mark.parametrize("asc", [True, [False, True], [True, False]])
def test_lexsort_per_column_path(asc):
    # float and full-range int64 keys do not pack: one radix sort per column
    rng = np.random.default_rng(6)
    n = 30000
    a = rng.choice(np.array([np.nan, -0.0, 0.0, 1.5, -2.0, np.inf], dtype=np.float32), n)
    b = rng.integers(-2**62, 2**62, n).astype(np.int64)
    c = rng.integers(0, 3, n).astype(np.int32)
    keys = [a, b, c]
    exp = oracle.multi_column_lexsort(keys, asc if isinstance(asc, bool) else asc + [True])
    ks = [cuda(k) for k in keys]
    a2 = asc if isinstance(asc, bool) else asc + [True]
    assert multikey.composite_plan(ks, a2) is None
    got = multikey.lexsort_perm(ks, a2)
    assert np.array_equal(_perm(got), exp)


def test_lexsort_descending_wrapping_int_falls_back():
    # max - x wraps in int32 here, as it does in the reference: the per-column
    # path reproduces that order
    k1 = np.array([2**31 - 1, -2**31, 0, 5, -7, 2**31 - 1], dtype=np.int32)
    k2 = np.array([1, 2, 3, 4, 5, 6], dtype=np.int32)
    exp = oracle.multi_column_lexsort([k1, k2], [False, True])
    assert multikey.composite_plan([cuda(k1), cuda(k2)], [False, True]) is None
    assert np.array_equal(_perm(multikey.lexsort_perm([cuda(k1), cuda(k2)], [False, True])), exp)


def test_single_key_lexsort_ignores_ascending():
    k = np.array([3, 1, 2, 1], dtype=np.int32)
    assert np.array_equal(_perm(multikey.lexsort_perm([cuda(k)], False)),
                          oracle.multi_column_lexsort([k], False))


def test_pack_unpack_round_trip():
    rng = np.random.default_rng(7)
    cols = [rng.integers(-1000, 1000, 1000).astype(np.int32),
            rng.integers(2**40, 2**40 + 5000, 1000).astype(np.int64),
            rng.integers(0, 7, 1000).astype(np.uint32)]
    ks = [cuda(c) for c in cols]
    for asc in (True, [False, True, False]):
        plan = multikey.composite_plan(ks, asc)
        back = multikey.unpack(multikey.pack(ks, plan), plan)
        for c, b in zip(cols, back):
            got = b.view(torch.int32 if b.dtype == torch.uint32 else b.dtype).cpu().numpy().view(c.dtype)
            assert np.array_equal(got, c)


def _check_groupby(keys, values, aggs, composite):
    ks = [cuda(k) for k in keys]
    names = [f"k{i}" for i in range(len(keys))]
    assert (multikey.composite_plan([k if k.dtype not in (torch.int8, torch.int16) else k.int()
                                     for k in ks]) is not None) == composite
    eu, er = oracle.groupby_aggregate_multi_column(keys, names, values, aggs)
    gu, gr = grouping.groupby_aggregate_multi_column(ks, names, {c: cuda(v) for c, v in values.items()},
                                                     aggs)
    for nm in names:
        g = gu[nm].cpu().numpy()
        assert g.dtype == eu[nm].dtype
        assert np.array_equal(g.view(np.uint8), eu[nm].view(np.uint8)), nm
    for c, fn in aggs.items():
        g = gr[c].cpu().numpy()
        e = er[c]
        if fn in ("count", "min", "max") or values[c].dtype.kind in "iu" and fn == "sum":
            assert np.array_equal(g, e), (c, fn)
        else:
            # error normalised by the group's sum of |x| (the compensated-sum bound)
            _, ea = oracle.groupby_aggregate_multi_column(
                keys, names, {c: np.abs(values[c].astype(np.float64))}, {c: "sum"})
            _, en = oracle.groupby_aggregate_multi_column(keys, names, {c: values[c]}, {c: "count"})
            absx = ea[c] if fn == "sum" else ea[c] / np.maximum(en[c], 1)
            err = np.abs(g.astype(np.float64) - e)
            assert np.all(err <= SUM_NORM_TOL * np.maximum(absx, 1e-30)), (c, fn, err.max())


def test_groupby_multi_composite_small_range():
    rng = np.random.default_rng(8)
    n = 200000
    keys = [rng.integers(0, 20, n).astype(np.int32), rng.integers(-30, 30, n).astype(np.int64)]
    vals = {"a": rng.standard_normal(n).astype(np.float32), "b": rng.integers(0, 100, n).astype(np.int32)}
    _check_groupby(keys, vals, {"a": "sum", "b": "max"}, composite=True)
    _check_groupby(keys, vals, {"a": "count"}, composite=True)
    _check_groupby(keys, vals, {"b": "min", "a": "mean"}, composite=True)


def test_groupby_multi_composite_large_range():
    rng = np.random.default_rng(9)
    n = 100000
    keys = [rng.integers(0, 5000, n).astype(np.int32), rng.integers(0, 10**6, n).astype(np.int32),
            rng.integers(0, 3, n).astype(np.uint32)]
    vals = {"a": rng.standard_normal(n).astype(np.float32)}
    _check_groupby(keys, vals, {"a": "sum"}, composite=True)


def test_groupby_multi_fallback_float_keys():
    rng = np.random.default_rng(10)
    n = 50000
    keys = [rng.choice(np.array([np.nan, 1.0, 2.5, -0.0, 0.0], dtype=np.float32), n),
            rng.integers(0, 4, n).astype(np.int32)]
    vals = {"a": rng.standard_normal(n).astype(np.float32), "c": rng.integers(0, 9, n).astype(np.int64)}
    _check_groupby(keys, vals, {"a": "sum", "c": "count"}, composite=False)
    _check_groupby(keys, vals, {"c": "max"}, composite=False)


def test_lexsort_reference_vectors():
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")
    with open(path) as f:
        cases = json.load(f)["lexsort"]
    for case in cases:
        keys = [np.asarray(k, dtype=np.int64) for k in case["keys"]]
        perm = _perm(multikey.lexsort_perm([cuda(k) for k in keys], case["ascending"]))
        got = [k[perm].tolist() for k in keys]
        assert got == case["expected"], case["source"]


def test_sort_multi_column_values():
    k1 = np.array([2, 1, 2, 1, 1], dtype=np.int32)
    k2 = np.array([5, 9, 1, 9, 0], dtype=np.int32)
    v = np.array([10., 20., 30., 40., 50.], dtype=np.float32)
    sk, sv = multikey.sort_multi_column([cuda(k1), cuda(k2)], {"v": cuda(v)}, [True, False])
    perm = oracle.multi_column_lexsort([k1, k2], [True, False])
    assert sk[0].cpu().tolist() == k1[perm].tolist()
    assert sk[1].cpu().tolist() == k2[perm].tolist()
    assert sv["v"].cpu().tolist() == v[perm].tolist()
