"""GPU parity of the sort-based groupby against the oracle.

Keys, counts, min and max are bit-exact; integer sums are exact (wrapping in
the value dtype); float sums are TwoSum-compensated and must agree with the
FP64 oracle to a relative error of at most SUM_REL_TOL (float32 values) per
group, or to at most SUM_NORM_TOL of sum(|x|) over the group.
"""

import numpy as np
import pandas as pd
import pytest
import torch

import oracle
from jaxframes_amd import grouping
from jaxframes_amd._tensors import to_numpy
from tests._util import cuda

pytestmark = pytest.mark.gpu

SUM_REL_TOL = 1e-6      # float32 sums: max relative error vs FP64 per group
SUM_NORM_TOL = 2e-7     # float32 sums: max |err| / sum|x| per group
SUM64_NORM_TOL = 1e-15  # float64 sums


def _oracle_groups(keys, vals):
    sk, sv, perm = oracle.radix_sort(keys, vals)
    starts = np.flatnonzero(oracle.head_flags(sk))
    return sk, sv, starts


def _check_float_sums(got, sv, starts, dtype):
    rel, nrm = oracle.sum_error_metrics(got, sv, starts)
    if dtype == np.float32:
        assert nrm <= SUM_NORM_TOL, (rel, nrm)
        ref = oracle.segment_sums_f64(sv, starts)
        # relative error only where the sum is not dominated by cancellation
        absx = np.add.reduceat(np.abs(sv.astype(np.float64)), starts)
        ok = np.abs(ref) > 1e-3 * absx
        r = np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-300)
        assert (r[ok] <= SUM_REL_TOL).all(), r[ok].max()
    else:
        assert nrm <= SUM64_NORM_TOL, (rel, nrm)


@pytest.mark.parametrize("n", [1, 5, 8191, 8192, 8193, 100000, 1000003])
@pytest.mark.parametrize("card", [1, 7, 1000, 100000, 10**9])
def test_groupby_stats_int32_f32(n, card):
    rng = np.random.default_rng(n + card)
    keys = rng.integers(0, card, n).astype(np.int32) if card < 2**31 else \
        rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    vals = rng.standard_normal(n).astype(np.float32)
    uk, res = grouping.groupby_stats(torch.from_numpy(keys).cuda(), torch.from_numpy(vals).cuda())
    sk, sv, starts = _oracle_groups(keys, vals)
    ek = sk[starts]
    assert np.array_equal(uk.cpu().numpy(), ek)
    counts = np.diff(np.append(starts, n))
    assert np.array_equal(res["count"].cpu().numpy(), counts)
    assert np.array_equal(res["min"].cpu().numpy(), np.minimum.reduceat(sv, starts))
    assert np.array_equal(res["max"].cpu().numpy(), np.maximum.reduceat(sv, starts))
    _check_float_sums(res["sum"].cpu().numpy(), sv, starts, np.float32)
    ref_mean = oracle.segment_sums_f64(sv, starts) / counts
    np.testing.assert_allclose(res["mean"].cpu().numpy(), ref_mean, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("kdt", [np.int64, np.uint64, np.float32, np.float64, np.uint32])
@pytest.mark.parametrize("vdt", [np.float64, np.int32, np.int64, np.float32])
def test_groupby_types(kdt, vdt):
    rng = np.random.default_rng(11)
    n = 150001
    base = rng.integers(0, 5000, n)
    keys = (base * 1000003).astype(kdt) if np.dtype(kdt).kind != "f" else (base / 7.0).astype(kdt)
    vals = (rng.standard_normal(n) * 1000).astype(vdt)
    uk, res = grouping.groupby_stats(cuda(keys), cuda(vals))
    sk, sv, starts = _oracle_groups(keys, vals)
    assert np.array_equal(to_numpy(uk), sk[starts])
    assert np.array_equal(res["count"].cpu().numpy(), np.diff(np.append(starts, n)))
    assert np.array_equal(res["min"].cpu().numpy(), np.minimum.reduceat(sv, starts))
    assert np.array_equal(res["max"].cpu().numpy(), np.maximum.reduceat(sv, starts))
    if np.dtype(vdt).kind == "i":
        exp = np.add.reduceat(sv.astype(np.int64), starts).astype(vdt)
        assert np.array_equal(res["sum"].cpu().numpy(), exp)
    else:
        _check_float_sums(res["sum"].cpu().numpy(), sv, starts, vdt)


def test_groupby_float_keys_nan_and_zero():
    # heads use `!=`: every NaN key is its own group, -0.0 and 0.0 share one
    keys = np.array([0.0, np.nan, -0.0, 1.0, np.nan, 1.0, -0.0], dtype=np.float32)
    vals = np.arange(7, dtype=np.float32)
    uk, res = grouping.groupby_stats(torch.from_numpy(keys).cuda(), torch.from_numpy(vals).cuda())
    ek, eres = oracle.groupby_aggregate(keys, {"v": vals}, {"v": "count"})
    got = uk.cpu().numpy()
    assert got.shape == ek.shape
    assert np.array_equal(np.isnan(got), np.isnan(ek))
    assert np.array_equal(got[~np.isnan(got)], ek[~np.isnan(ek)])
    assert np.array_equal(res["count"].cpu().numpy(), eres["v"])


def test_groupby_aggregate_reference_vectors():
    # test_parallel_algorithms.py:90-152
    cuda = lambda x: torch.tensor(x, device="cuda")
    uk, r = grouping.groupby_aggregate(cuda([1, 2, 1, 3, 2, 1, 3]),
                                       {"col1": cuda([10, 20, 30, 40, 50, 60, 70])}, {"col1": "sum"})
    assert uk.tolist() == [1, 2, 3] and r["col1"].tolist() == [100, 70, 110]
    uk, r = grouping.groupby_aggregate(cuda([1, 2, 1, 2, 1]),
                                       {"col1": cuda([10.0, 20.0, 30.0, 40.0, 50.0])}, {"col1": "mean"})
    assert uk.tolist() == [1, 2]
    np.testing.assert_allclose(r["col1"].cpu().numpy(), [30.0, 30.0], rtol=1e-6)
    uk, r = grouping.groupby_aggregate(cuda([1, 2, 1, 2, 1]),
                                       {"col1": cuda([10, 25, 30, 15, 20]), "col2": cuda([5, 8, 2, 9, 7])},
                                       {"col1": "max", "col2": "min"})
    assert uk.tolist() == [1, 2] and r["col1"].tolist() == [30, 25] and r["col2"].tolist() == [2, 8]
    uk, r = grouping.groupby_aggregate(cuda([1, 2, 1, 3, 2, 1]),
                                       {"col1": cuda([10, 20, 30, 40, 50, 60])}, {"col1": "count"})
    assert uk.tolist() == [1, 2, 3] and r["col1"].tolist() == [3, 2, 1]
    with pytest.raises(ValueError):
        grouping.groupby_aggregate(cuda([1]), {"c": cuda([1])}, {"c": "median"})


def test_groupby_multi_column_aggs_vs_pandas():
    rng = np.random.default_rng(2)
    n = 200000
    keys = rng.integers(0, 3000, n).astype(np.int64)
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.integers(-1000, 1000, n).astype(np.int32)
    c = rng.standard_normal(n)
    uk, r = grouping.groupby_aggregate(torch.from_numpy(keys).cuda(),
                                       {"a": torch.from_numpy(a).cuda(), "b": torch.from_numpy(b).cuda(),
                                        "c": torch.from_numpy(c).cuda()},
                                       {"a": "sum", "b": "max", "c": "count"})
    df = pd.DataFrame({"k": keys, "a": a, "b": b, "c": c}).groupby("k").agg({"a": "sum", "b": "max", "c": "count"})
    assert np.array_equal(uk.cpu().numpy(), df.index.values)
    assert np.array_equal(r["b"].cpu().numpy(), df["b"].values)
    assert np.array_equal(r["c"].cpu().numpy(), df["c"].values)
    np.testing.assert_allclose(r["a"].cpu().numpy(), df["a"].values, rtol=2e-4, atol=1e-4)


def test_groupby_config1_vs_pandas():
    # 1M rows, int32 keys uniform in [0,1000), float32 N(0,1)
    rng = np.random.default_rng(0)
    n = 1_000_000
    keys = rng.integers(0, 1000, n, dtype=np.int32)
    vals = rng.standard_normal(n, dtype=np.float32)
    uk, res = grouping.groupby_stats(torch.from_numpy(keys).cuda(), torch.from_numpy(vals).cuda())
    df = pd.DataFrame({"k": keys, "v": vals}).groupby("k")["v"].agg(["sum", "count", "mean"])
    assert np.array_equal(uk.cpu().numpy(), df.index.values)
    assert np.array_equal(res["count"].cpu().numpy(), df["count"].values)
    sk, sv, starts = _oracle_groups(keys, vals)
    _check_float_sums(res["sum"].cpu().numpy(), sv, starts, np.float32)


def test_groupby_empty():
    uk, res = grouping.groupby_aggregate(torch.empty(0, dtype=torch.int32, device="cuda"),
                                         {"v": torch.empty(0, device="cuda")}, {"v": "sum"})
    assert uk.numel() == 0 and res["v"].numel() == 0


def test_no_device_error_after_groupby():
    from jaxframes_amd import _lib
    assert _lib.device_error() == 0


def test_output_capacity_retry_and_small_outputs():
    """jfx_groupby_ex reports the capacity it needs instead of writing past
    small outputs; groupby_aggregate's small-output first try and retry give
    the same groups as full-size outputs."""
    import ctypes
    from jaxframes_amd import _lib
    from jaxframes_amd.grouping import _SMALL_CAP_MIN_ROWS
    rng = np.random.default_rng(21)
    n = _SMALL_CAP_MIN_ROWS + 1000
    wide = torch.from_numpy(rng.integers(-10**9, 10**9, n).astype(np.int64)).cuda()   # general path
    narrow = torch.from_numpy(rng.integers(5, 1500, n).astype(np.int32)).cuda()       # small range
    vals = torch.from_numpy(rng.standard_normal(n).astype(np.float32)).cuda()
    for keys in (wide, narrow):
        uk, res = grouping.groupby_aggregate(keys, {"v": vals}, {"v": "count"})
        ek, ec = np.unique(keys.cpu().numpy(), return_counts=True)
        assert np.array_equal(uk.cpu().numpy(), ek)
        assert np.array_equal(res["v"].cpu().numpy(), ec)
    # the C ABI: too small a capacity writes nothing and reports the need
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_groupby_workspace_size", _lib.JFX_I64, _lib.JFX_F32, n, ctypes.byref(nb))
    ws = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    out_k = torch.full((16,), 7, dtype=torch.int64, device="cuda")
    out_s = torch.full((16,), 7.0, dtype=torch.float32, device="cuda")
    g = ctypes.c_uint64(0)
    st = _lib.lib().jfx_groupby_ex(_lib.JFX_I64, _lib.ptr(wide), _lib.JFX_F32, _lib.ptr(vals), n,
                                   _lib.ptr(out_k), _lib.ptr(out_s), None, None, None, None,
                                   ctypes.byref(g), 16, 0, _lib.ptr(ws), ws.numel(), 0)
    assert st == _lib.JFX_ERR_CAPACITY and g.value == n
    torch.cuda.synchronize()
    assert torch.all(out_k == 7) and torch.all(out_s == 7.0)


def _wrapping_sums(sv, starts):
    """Integer group sums wrapping in the value dtype (the device's rule)."""
    u = np.uint64 if sv.dtype.itemsize == 8 else np.uint32
    wide = sv.view(u).astype(np.uint64)
    return np.add.reduceat(wide, starts).astype(u).view(sv.dtype)


@pytest.mark.parametrize("vdt", [np.float32, np.int32, np.int64, np.uint32, np.uint64, np.float64])
@pytest.mark.parametrize("R", [1, 9, 64, 200, 1000, 1024, 1500, 2048])
def test_small_range_sum_only(vdt, R):
    """Sum-only groupbys over a small key range (R distinct keys, negative
    minimum): float32 and integer values take the LDS-atomic tables, float64
    the tag-round tables; lane-group tables (R <= 256) and 2048-entry tables
    (R > 1024) included. An unaligned view takes the row-by-row loads."""
    rng = np.random.default_rng(R * 7 + np.dtype(vdt).itemsize)
    n = 300_007
    keys = (rng.integers(0, R, n) - 3).astype(np.int32)
    if np.dtype(vdt).kind == "f":
        vals = (rng.standard_normal(n) * 100).astype(vdt)
    else:
        info = np.iinfo(vdt)
        vals = rng.integers(info.min, info.max, n, dtype=vdt, endpoint=True)   # sums wrap
    for off in (0, 1):
        kk, vv = keys[off:], vals[off:]
        uk, res = grouping.groupby_aggregate(cuda(keys)[off:], {"v": cuda(vals)[off:]}, {"v": "sum"})
        sk, sv, starts = _oracle_groups(kk, vv)
        assert np.array_equal(to_numpy(uk), sk[starts])
        got = to_numpy(res["v"])
        if np.dtype(vdt).kind == "f":
            assert got.dtype == vdt
            _check_float_sums(got, sv, starts, vdt)
        else:
            assert np.array_equal(got, _wrapping_sums(sv, starts))


def test_small_range_sum_only_skewed_and_clustered():
    """One key holding most rows (the skew selector may pick the tile-sorted
    kernel) and sorted keys (long register runs) give the same sums."""
    rng = np.random.default_rng(3)
    n = 1_000_003
    skew = np.where(rng.random(n) < 0.9, 17, rng.integers(0, 1000, n)).astype(np.int32)
    clustered = np.sort(rng.integers(0, 1000, n)).astype(np.int32)
    vals = rng.standard_normal(n).astype(np.float32)
    for keys in (skew, clustered):
        uk, res = grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)}, {"v": "sum"})
        sk, sv, starts = _oracle_groups(keys, vals)
        assert np.array_equal(to_numpy(uk), sk[starts])
        _check_float_sums(to_numpy(res["v"]), sv, starts, np.float32)


@pytest.mark.parametrize("vdt", [np.float32, np.float64])
@pytest.mark.parametrize("span", [1, 10**7])
@pytest.mark.parametrize("stats", [False, True])
def test_float_sums_with_infinities(vdt, span, stats):
    """Groups holding +inf, -inf, both, NaN, or (float32) values whose sum
    overflows: +-inf, NaN and inf as the FP64 sum rounded to the value type
    gives them, on the small-range kernels (span 1) and on the sort-based path
    (span 1e7); sum-only and with min / max."""
    rng = np.random.default_rng(5)
    n = 200_003
    g = rng.integers(0, 50, n)
    keys = (g * span).astype(np.int64)
    vals = rng.standard_normal(n).astype(vdt)
    vals[np.flatnonzero(g == 0)[::97]] = np.inf
    vals[np.flatnonzero(g == 1)[::89]] = -np.inf
    both = np.flatnonzero(g == 2)
    vals[both[::50]] = np.inf
    vals[both[25::50]] = -np.inf
    vals[np.flatnonzero(g == 3)[::101]] = np.nan
    if vdt == np.float32:
        vals[np.flatnonzero(g == 4)[:3]] = 3e38
    kt, vt = cuda(keys), cuda(vals)
    if stats:
        uk, res = grouping.groupby_stats(kt, vt)
        got = to_numpy(res["sum"])
    else:
        uk, res = grouping.groupby_aggregate(kt, {"v": vt}, {"v": "sum"})
        got = to_numpy(res["v"])
    sk, sv, starts = _oracle_groups(keys, vals)
    assert np.array_equal(to_numpy(uk), sk[starts])
    with np.errstate(over="ignore", invalid="ignore"):
        ref = np.add.reduceat(sv.astype(np.float64), starts).astype(vdt)
    fin = np.isfinite(ref)
    assert np.array_equal(got[~fin], ref[~fin], equal_nan=True), (got[:6], ref[:6])
    assert (~fin).sum() == (5 if vdt == np.float32 else 4)
    # finite groups against correctly rounded sums (math.fsum)
    import math
    ends = np.append(starts[1:], len(sv))
    exact = np.array([math.fsum(sv[a:b].astype(np.float64)) if f else 0.0
                      for a, b, f in zip(starts, ends, fin)])
    absx = np.add.reduceat(np.abs(sv.astype(np.float64)), starts)
    r = np.abs(got[fin].astype(np.float64) - exact[fin]) / absx[fin]
    assert r.max() <= (SUM_NORM_TOL if vdt == np.float32 else SUM64_NORM_TOL), r.max()


def _dense_keys(pattern, n, rng):
    """int64 keys spread over 2^62 (no small-range or few-groups shortcut) whose
    sorted order has the given group sizes."""
    if pattern == "unique":
        sizes = np.ones(n, dtype=np.int64)
    elif pattern == "pairs":
        sizes = np.full((n + 1) // 2, 2, dtype=np.int64)
    else:  # stretches of single rows and of runs up to 40 rows, mixed
        parts = []
        while sum(int(p.sum()) for p in parts) < n:
            parts.append(np.ones(rng.integers(1, 3000), dtype=np.int64))
            parts.append(rng.integers(1, 41, rng.integers(1, 300)))
        sizes = np.concatenate(parts)
        sizes = sizes[np.cumsum(sizes) <= n]
        sizes = np.append(sizes, n - sizes.sum()) if sizes.sum() < n else sizes
    g = sizes.size
    distinct = np.unique(rng.integers(-2**62, 2**62, g + 64, dtype=np.int64))[:g]
    keys = np.repeat(np.sort(distinct), sizes)[:n]
    return keys[rng.permutation(n)]


@pytest.mark.parametrize("pattern", ["unique", "pairs", "mixed"])
@pytest.mark.parametrize("n", [8193, 100003, 300001])
@pytest.mark.parametrize("vdt", [np.float32, np.int64])
def test_many_groups_staged_writes(pattern, n, vdt):
    """Nearly one group per row: keys, sums, counts and means go out through
    the per-wave staging of the segmented reduction."""
    rng = np.random.default_rng(n + len(pattern))
    keys = _dense_keys(pattern, n, rng)
    if vdt == np.float32:
        vals = rng.standard_normal(n).astype(np.float32)
    else:
        vals = rng.integers(-2**62, 2**62, n, dtype=np.int64)
    sk, sv, starts = _oracle_groups(keys, vals)
    counts = np.diff(np.append(starts, n))
    aggs = {"s": "sum", "c": "count", "m": "mean"}
    uk, res = grouping.groupby_aggregate(cuda(keys), {c: cuda(vals) for c in aggs}, aggs)
    assert np.array_equal(to_numpy(uk), sk[starts])
    assert np.array_equal(to_numpy(res["c"]), counts)
    got = to_numpy(res["s"])
    if vdt == np.float32:
        _check_float_sums(got, sv, starts, np.float32)
        ref_mean = oracle.segment_sums_f64(sv, starts) / counts
        nrm = np.add.reduceat(np.abs(sv.astype(np.float64)), starts) / counts
        assert (np.abs(to_numpy(res["m"]) - ref_mean) <= 4e-7 * nrm + 1e-30).all()
    else:
        ref = _wrapping_sums(sv, starts)
        assert np.array_equal(got, ref)
        assert np.array_equal(to_numpy(res["m"]), ref.astype(np.float64) / counts)
