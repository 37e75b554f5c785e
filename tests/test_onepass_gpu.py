"""GPU parity of the one-pass groupby (integer keys whose range is below 1024;
sum / count / mean of float32 or integer values) against the oracle, and of
its misses (wider ranges), which must give the statistics-first path's
result. Keys and counts are bit-exact, integer sums exact (wrapping in the
value dtype), float32 sums within the FP64-oracle bound of the groupby tests.
"""

import ctypes
import math
import zlib

import numpy as np
import pytest
import torch

import oracle
from jaxframes_amd import _lib, grouping
from jaxframes_amd._tensors import to_numpy
from tests._util import cuda
from tests.test_groupby_gpu import _check_float_sums, _wrapping_sums

pytestmark = pytest.mark.gpu

N = 1 << 18


@pytest.fixture(autouse=True)
def _onepass_at_any_size(monkeypatch):
    # by default the one-pass try starts at 2^16 rows; these tests take it at
    # any size (test_onepass_default_threshold checks the default)
    monkeypatch.setenv("JFX_ONEPASS", "always")


def _groups(keys, vals):
    sk, sv, _ = oracle.radix_sort(keys, vals)
    starts = np.flatnonzero(oracle.head_flags(sk))
    return sk, sv, starts


def _check(keys, vals, uk, res_sum, res_count=None):
    sk, sv, starts = _groups(keys, vals)
    assert np.array_equal(to_numpy(uk), sk[starts])
    if res_count is not None:
        assert np.array_equal(to_numpy(res_count), np.diff(np.append(starts, len(sk))))
    got = to_numpy(res_sum)
    if vals.dtype == np.float64:
        # compensated float64 sums against correctly rounded ones (math.fsum):
        # numpy's own float64 reduceat is the less accurate side here
        ends = np.append(starts[1:], len(sv))
        exact = np.array([math.fsum(sv[a:b]) for a, b in zip(starts, ends)])
        absx = np.add.reduceat(np.abs(sv), starts)
        fin = np.isfinite(exact)
        assert np.array_equal(got[~fin], exact[~fin], equal_nan=True)
        assert (np.abs(got[fin] - exact[fin]) <= 1e-15 * absx[fin]).all()
    elif np.dtype(vals.dtype).kind == "f":
        _check_float_sums(got, sv, starts, vals.dtype)
    else:
        assert np.array_equal(got, _wrapping_sums(sv, starts))


def _kernels_run(fn):
    """fn's result and the names of the library kernels it launched."""
    with _lib.KernelTimer() as kt:
        out = fn()
        torch.cuda.synchronize()
    return out, set(kt.stats().keys())


KEY_CASES = [
    ("int32", np.int32, 0),
    ("int32-neg", np.int32, -500),
    ("int32-top", np.int32, 2**31 - 1000),
    ("int32-bottom", np.int32, -2**31),
    ("int64", np.int64, 10**12),
    ("uint32-top", np.uint32, 2**32 - 1000),
    ("uint64-top", np.uint64, 2**64 - 1000),
]


@pytest.mark.parametrize("name,kdt,base", KEY_CASES, ids=[c[0] for c in KEY_CASES])
@pytest.mark.parametrize("vdt", [np.float32, np.float64, np.int32, np.int64, np.uint64])
def test_onepass_hit(name, kdt, base, vdt):
    rng = np.random.default_rng(zlib.crc32(f"{name}/{np.dtype(vdt).name}".encode()))
    off = rng.integers(0, 1000, N)   # 1000 distinct keys from base on
    if np.dtype(kdt).itemsize == 4:
        keys = (np.int64(base) + off).astype(kdt)
    else:
        keys = (np.uint64(base % 2**64) + off.astype(np.uint64)).view(kdt)
    if np.dtype(vdt).kind == "f":
        vals = rng.standard_normal(N).astype(vdt)
    else:
        info = np.iinfo(vdt)
        vals = rng.integers(info.min, info.max, N, dtype=vdt, endpoint=True)
    (uk, res), ran = _kernels_run(lambda: grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)},
                                                                       {"v": "sum"}))
    assert "onepass_table" in ran and "sort_stats" not in ran, ran
    _check(keys, vals, uk, res["v"])


def test_onepass_count_and_mean():
    rng = np.random.default_rng(1)
    keys = rng.integers(-300, 300, N).astype(np.int32)
    vals = rng.standard_normal(N).astype(np.float32)
    for fn in ("count", "mean"):
        uk, res = grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)}, {"v": fn})
        sk, sv, starts = _groups(keys, vals)
        assert np.array_equal(to_numpy(uk), sk[starts])
        counts = np.diff(np.append(starts, N))
        if fn == "count":
            assert np.array_equal(to_numpy(res["v"]), counts)
        else:
            np.testing.assert_allclose(to_numpy(res["v"]),
                                       oracle.segment_sums_f64(sv, starts) / counts,
                                       rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("case", ["outlier", "sorted_wide", "two_clusters"])
def test_onepass_miss_falls_back(case):
    rng = np.random.default_rng(7)
    if case == "outlier":         # one key far away, late in the input
        keys = rng.integers(0, 1000, N).astype(np.int32)
        keys[N - 12345] = 10**6
    elif case == "sorted_wide":   # every wave sees a narrow slice
        keys = np.sort(rng.integers(0, 10**6, N)).astype(np.int64)
    else:                         # residues of the two clusters collide
        keys = np.where(rng.random(N) < 0.5, rng.integers(0, 100, N),
                        rng.integers(4096, 4196, N)).astype(np.int32)
    vals = rng.standard_normal(N).astype(np.float32)
    (uk, res), ran = _kernels_run(lambda: grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)},
                                                                       {"v": "sum"}))
    # the statistics-first path ran after the miss (with the digit histograms
    # fused into the statistics pass when the span seen was already >= 2048)
    assert "onepass_table" in ran and any(k.startswith("sort_stats") for k in ran), ran
    if case in ("outlier", "sorted_wide"):
        assert "sort_stats_hist" in ran, ran
    _check(keys, vals, uk, res["v"])


def test_onepass_hit_after_miss_and_unaligned():
    """A miss leaves its stop word in the workspace; the next call (same
    workspace size, so the same cached buffer) must still hit. The unaligned
    view takes the row-by-row loads."""
    rng = np.random.default_rng(9)
    wide = rng.integers(0, 10**6, N).astype(np.int32)
    narrow = rng.integers(50, 900, N).astype(np.int32)
    vals = rng.standard_normal(N).astype(np.float32)
    for _ in range(2):
        for keys in (wide, narrow):
            uk, res = grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)}, {"v": "sum"})
            _check(keys, vals, uk, res["v"])
    (uk, res), ran = _kernels_run(lambda: grouping.groupby_aggregate(
        cuda(narrow)[1:], {"v": cuda(vals)[1:]}, {"v": "sum"}))
    assert "onepass_table" in ran and "sort_stats" not in ran
    _check(narrow[1:], vals[1:], uk, res["v"])


def test_onepass_matches_statistics_first_path(monkeypatch):
    rng = np.random.default_rng(11)
    keys = rng.integers(0, 1000, 1_000_003).astype(np.int32)
    vals = rng.standard_normal(keys.size).astype(np.float32)
    uk1, r1 = grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)}, {"v": "sum"})
    monkeypatch.setenv("JFX_ONEPASS", "0")
    (uk2, r2), ran = _kernels_run(lambda: grouping.groupby_aggregate(
        cuda(keys), {"v": cuda(vals)}, {"v": "sum"}))
    assert "onepass_table" not in ran
    assert torch.equal(uk1, uk2)
    a, b = to_numpy(r1["v"]).astype(np.float64), to_numpy(r2["v"]).astype(np.float64)
    sk, sv, starts = _groups(keys, vals)
    nrm = np.add.reduceat(np.abs(sv.astype(np.float64)), starts)
    assert (np.abs(a - b) / nrm).max() <= 4e-7


def test_onepass_capacity_report():
    """Too small an output capacity: the exact group count comes back and
    nothing is written."""
    rng = np.random.default_rng(13)
    keys = torch.from_numpy(rng.integers(0, 700, N).astype(np.int32)).cuda()
    vals = torch.from_numpy(rng.standard_normal(N).astype(np.float32)).cuda()
    G = int(torch.unique(keys).numel())
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_groupby_workspace_size", _lib.JFX_I32, _lib.JFX_F32, N, ctypes.byref(nb))
    ws = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    out_k = torch.full((16,), 7, dtype=torch.int32, device="cuda")
    out_s = torch.full((16,), 7.0, dtype=torch.float32, device="cuda")
    g = ctypes.c_uint64(0)
    st = _lib.lib().jfx_groupby_ex(_lib.JFX_I32, _lib.ptr(keys), _lib.JFX_F32, _lib.ptr(vals), N,
                                   _lib.ptr(out_k), _lib.ptr(out_s), None, None, None, None,
                                   ctypes.byref(g), 16, 0, _lib.ptr(ws), ws.numel(), 0)
    assert st == _lib.JFX_ERR_CAPACITY and g.value == G
    torch.cuda.synchronize()
    assert torch.all(out_k == 7) and torch.all(out_s == 7.0)
    out_k = torch.empty(G, dtype=torch.int32, device="cuda")
    out_s = torch.empty(G, dtype=torch.float32, device="cuda")
    st = _lib.lib().jfx_groupby_ex(_lib.JFX_I32, _lib.ptr(keys), _lib.JFX_F32, _lib.ptr(vals), N,
                                   _lib.ptr(out_k), _lib.ptr(out_s), None, None, None, None,
                                   ctypes.byref(g), G, 0, _lib.ptr(ws), ws.numel(), 0)
    assert st == 0 and g.value == G
    _check(keys.cpu().numpy(), vals.cpu().numpy(), out_k, out_s)


@pytest.mark.parametrize("n", [1, 5, 1000, 65537])
@pytest.mark.parametrize("R", [907, 3])
def test_onepass_small_inputs(n, R):
    """Inputs smaller than one wave step go row by row (no vector loop),
    with 1024-residue and (R = 3) lane-group tables; all five aggregations."""
    rng = np.random.default_rng(n + R)
    keys = rng.integers(-7, R - 7, n).astype(np.int64)
    vals = rng.standard_normal(n).astype(np.float32)
    (uk, res), ran = _kernels_run(lambda: grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)},
                                                                       {"v": "sum"}))
    assert "onepass_table" in ran and "sort_stats" not in ran, ran
    _check(keys, vals, uk, res["v"])
    uk, st = grouping.groupby_stats(cuda(keys), cuda(vals))
    sk, sv, starts = _groups(keys, vals)
    assert np.array_equal(to_numpy(uk), sk[starts])
    assert np.array_equal(to_numpy(st["count"]), np.diff(np.append(starts, n)))
    assert np.array_equal(to_numpy(st["min"]), np.minimum.reduceat(sv, starts))
    assert np.array_equal(to_numpy(st["max"]), np.maximum.reduceat(sv, starts))


@pytest.mark.parametrize("n,onepass", [((1 << 16) - 1, False), ((1 << 16) + 5, True)])
def test_onepass_default_threshold(monkeypatch, n, onepass):
    monkeypatch.delenv("JFX_ONEPASS")
    rng = np.random.default_rng(17)
    keys = rng.integers(0, 1000, n).astype(np.int32)
    vals = rng.standard_normal(n).astype(np.float32)
    (uk, res), ran = _kernels_run(lambda: grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)},
                                                                       {"v": "sum"}))
    assert ("onepass_table" in ran) == onepass and ("sort_stats" in ran) != onepass, ran
    _check(keys, vals, uk, res["v"])


@pytest.mark.parametrize("R", [1, 2, 5, 10, 40, 100, 200])
def test_onepass_lane_group_tables(R):
    """Few distinct keys: the sample picks lane-group tables (16, 64 or 256
    residues each) so that lanes do not collide on one entry."""
    rng = np.random.default_rng(R)
    keys = (rng.integers(0, R, N) + 1_000_000).astype(np.int64)
    vals = rng.standard_normal(N).astype(np.float32)
    (uk, res), ran = _kernels_run(lambda: grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)},
                                                                       {"v": "sum"}))
    assert "onepass_table" in ran and "sort_stats" not in ran, ran
    _check(keys, vals, uk, res["v"])
    (uk, res) = grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)}, {"v": "count"})
    assert np.array_equal(to_numpy(res["v"]), np.unique(keys, return_counts=True)[1])


@pytest.mark.parametrize("rare", [50, 900, 3000])
def test_onepass_narrow_sample_wider_keys(rare):
    """The 64 sampled rows (i * n / 64) hold 4 distinct keys, so the tables
    are 16 residues wide; a few other rows hold a wider key: the range check
    must miss. A span still below 1024 is retried with 1024-residue tables
    (a hit), a wider one goes to the statistics-first path."""
    rng = np.random.default_rng(rare)
    keys = rng.integers(0, 4, N).astype(np.int32)
    sampled = {i * N // 64 for i in range(64)}
    rows = [r for r in (N // 3 + 7, N // 2 + 11, N - 5) if r not in sampled]
    keys[rows] = rare
    vals = rng.standard_normal(N).astype(np.float32)
    with _lib.KernelTimer() as kt:
        uk, res = grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)}, {"v": "sum"})
        torch.cuda.synchronize()
    ran = kt.stats()
    assert ran["onepass_table"][1] == (2 if rare < 1024 else 1), ran
    assert any(k.startswith("sort_stats") for k in ran) == (rare >= 1024), ran
    _check(keys, vals, uk, res["v"])


@pytest.mark.parametrize("lo,hi", [(0, 1025), (-700, 800), (5, 2052)])
def test_onepass_wider_tables(lo, hi):
    """Key spans of 1024..2047 values: the first try misses and the retry
    with 2048-residue tables hits (no statistics pass)."""
    rng = np.random.default_rng(hi)
    keys = rng.integers(lo, hi, N).astype(np.int64)
    keys[:2] = [lo, hi - 1]
    vals = rng.standard_normal(N).astype(np.float32)
    (uk, res), ran = _kernels_run(lambda: grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)},
                                                                       {"v": "sum"}))
    assert "onepass_table_2k" in ran and not any(k.startswith("sort_stats") for k in ran), ran
    _check(keys, vals, uk, res["v"])
    uk, res = grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)}, {"v": "mean"})
    sk, sv, starts = _groups(keys, vals)
    np.testing.assert_allclose(to_numpy(res["v"]),
                               oracle.segment_sums_f64(sv, starts) / np.diff(np.append(starts, N)),
                               rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("hot,R", [(0.5, 1000), (0.8, 3), (0.3, 200), ("zipf", 1000)])
def test_onepass_frequent_key(hot, R):
    """A frequent key makes many lanes of a wave add to one table entry in
    the same instruction (serialised LDS atomics): exact counts, float sums
    to the FP64 bound, exact integer sums."""
    rng = np.random.default_rng(R)
    if hot == "zipf":
        keys = np.minimum(rng.zipf(1.3, N) - 1, R - 1).astype(np.int32)
    else:
        keys = np.where(rng.random(N) < hot, 1, rng.integers(0, R, N)).astype(np.int32)
    for vals in (rng.standard_normal(N).astype(np.float32),
                 rng.integers(-2**31, 2**31 - 1, N, dtype=np.int32)):
        (uk, res), ran = _kernels_run(lambda: grouping.groupby_aggregate(
            cuda(keys), {"v": cuda(vals)}, {"v": "sum"}))
        assert "onepass_table" in ran and "sort_stats" not in ran, ran
        _check(keys, vals, uk, res["v"])
    uk, res = grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)}, {"v": "count"})
    assert np.array_equal(to_numpy(res["v"]), np.unique(keys, return_counts=True)[1])


@pytest.mark.parametrize("vdt", [np.float32, np.float64, np.int32, np.int64, np.uint32, np.uint64])
@pytest.mark.parametrize("R", [3, 1000, 1500])
def test_onepass_min_max(vdt, R):
    """Min / max in the one-pass tables (order-preserving integer codes,
    float NaN winning both): bit-exact against numpy's reductions, with the
    other aggregations of groupby_stats; lane-group tables (R = 3) and the
    2048-residue retry (R = 1500) included."""
    rng = np.random.default_rng(R * 11 + np.dtype(vdt).itemsize)
    keys = (rng.integers(0, R, N) - 40).astype(np.int32)
    if np.dtype(vdt).kind == "f":
        vals = (rng.standard_normal(N) * 100).astype(vdt)
        vals[rng.integers(0, N, 40)] = np.inf
        vals[rng.integers(0, N, 40)] = -np.inf
        vals[rng.integers(0, N, 40)] = -0.0
        vals[np.flatnonzero(keys == keys[7])[:3]] = np.nan    # one group with NaN
    else:
        info = np.iinfo(vdt)
        vals = rng.integers(info.min, info.max, N, dtype=vdt, endpoint=True)
        vals[:2] = [info.min, info.max]
    with _lib.KernelTimer() as kt:
        uk, res = grouping.groupby_stats(cuda(keys), cuda(vals))
        torch.cuda.synchronize()
    ran = kt.stats()
    # float64 with min / max has no 2048-residue tables (LDS): R = 1500 then
    # runs the statistics-first path after the miss
    stats_first = vdt == np.float64 and R == 1500
    assert any(k.startswith("onepass_table") for k in ran), ran
    assert any(k.startswith("sort_stats") for k in ran) == stats_first, ran
    sk, sv, starts = _groups(keys, vals)
    assert np.array_equal(to_numpy(uk), sk[starts])
    assert np.array_equal(to_numpy(res["count"]), np.diff(np.append(starts, N)))
    assert np.array_equal(to_numpy(res["min"]), np.minimum.reduceat(sv, starts), equal_nan=True)
    assert np.array_equal(to_numpy(res["max"]), np.maximum.reduceat(sv, starts), equal_nan=True)
    if np.dtype(vdt).kind != "f":
        assert np.array_equal(to_numpy(res["sum"]), _wrapping_sums(sv, starts))
    for fn, ref in (("min", np.minimum), ("max", np.maximum)):
        uk1, r1 = grouping.groupby_aggregate(cuda(keys), {"v": cuda(vals)}, {"v": fn})
        assert np.array_equal(to_numpy(uk1), sk[starts])
        assert np.array_equal(to_numpy(r1["v"]), ref.reduceat(sv, starts), equal_nan=True)


@pytest.mark.parametrize("R", [8, 1000, 1500])
def test_onepass_sum_only_presence_from_sums(R):
    """float32 sums alone run the count-free tables (group presence from the
    sum entries, which start at -0.0): groups whose rows are all -0.0, a
    single -0.0 or +0.0 row, NaN and infinities are all present, keys with no
    rows are absent, and every sum has the bits of the counted tables'
    (groupby_stats) -- +0.0 for the -0.0 groups, as tables started at +0.0
    give."""
    rng = np.random.default_rng(R + 5)
    keys = rng.integers(0, R, N).astype(np.int32)
    keys[(keys % 7 == 5) & (keys > 2)] = 0             # keys with no rows inside the span
    vals = (rng.standard_normal(N) * 10).astype(np.float32)
    g = np.unique(keys)
    special = {int(g[1]): "all-neg-zero", int(g[-2]): "mixed-zeros"}
    vals[keys == g[1]] = -0.0
    mz = np.flatnonzero(keys == g[-2])
    vals[mz] = np.where(np.arange(mz.size) % 2 == 0, -0.0, 0.0).astype(np.float32)
    # one row each of its own key: -0.0, +0.0 (R >= 1000: keys beyond the rest)
    extra_k, extra_v = np.array([], np.int32), np.array([], np.float32)
    if R >= 1000:
        extra_k = np.array([R, R + 1], np.int32)
        extra_v = np.array([-0.0, 0.0], np.float32)
        special.update({R: "single-neg-zero", R + 1: "single-pos-zero"})
    keys = np.concatenate([keys, extra_k])
    vals = np.concatenate([vals, extra_v])
    for with_specials in (False, True):
        v = vals.copy()
        if with_specials:
            v[np.flatnonzero(keys == g[2])[:2]] = np.nan
            v[np.flatnonzero(keys == g[3])[:1]] = np.inf
            v[np.flatnonzero(keys == g[4])[:1]] = -np.inf
        (uk, res), ran = _kernels_run(lambda: grouping.groupby_aggregate(
            cuda(keys), {"v": cuda(v)}, {"v": "sum"}))
        assert "onepass_table" in ran or "onepass_table_2k" in ran, ran
        sk, sv, starts = _groups(keys, v)
        assert np.array_equal(to_numpy(uk), sk[starts])
        uk2, res2 = grouping.groupby_stats(cuda(keys), cuda(v))
        assert torch.equal(uk, uk2)
        got = to_numpy(res["v"])
        assert np.array_equal(got.view(np.int32), to_numpy(res2["sum"]).view(np.int32))
        for k, what in special.items():
            i = int(np.searchsorted(sk[starts], k))
            assert sk[starts][i] == k, what
            assert got[i] == 0.0 and not np.signbit(got[i]), (what, got[i])
        if not with_specials:
            _check(keys, v, uk, res["v"])


@pytest.mark.parametrize("R", [2, 1000, 1700])
def test_onepass_float64_sums(R):
    """float64 values: sums compensated through ds_add_rtn_f64 (TwoSum of the
    returned entry), against math.fsum; large magnitudes with cancellation."""
    rng = np.random.default_rng(R)
    keys = rng.integers(0, R, N).astype(np.int32)
    vals = rng.standard_normal(N) * np.where(rng.random(N) < 0.1, 1e12, 1.0)
    for agg in ("sum", "mean"):
        (uk, res), ran = _kernels_run(lambda: grouping.groupby_aggregate(
            cuda(keys), {"v": cuda(vals)}, {"v": agg}))
        assert any(k.startswith("onepass_table") for k in ran), ran
        if agg == "sum":
            _check(keys, vals, uk, res["v"])
        else:
            sk, sv, starts = _groups(keys, vals)
            ends = np.append(starts[1:], N)
            exact = np.array([math.fsum(sv[a:b]) for a, b in zip(starts, ends)]) / (ends - starts)
            np.testing.assert_allclose(to_numpy(res["v"]), exact, rtol=1e-12)


def test_kernel_timer_samples_every_kth_launch():
    """KernelTimer(prefix, every=k) times launches 0, k, 2k, ... of the
    kernels passing the prefix filter, and nothing else."""
    rng = np.random.default_rng(31)
    keys = cuda(rng.integers(0, 100, N).astype(np.int32))
    vals = cuda(rng.standard_normal(N).astype(np.float32))
    with _lib.KernelTimer(prefix="onepass_table", every=3) as kt:
        for _ in range(7):
            grouping.groupby_aggregate(keys, {"v": vals}, {"v": "sum"})
        torch.cuda.synchronize()
    st = kt.stats()
    assert set(st) == {"onepass_table"}
    assert st["onepass_table"][1] == 3
    with _lib.KernelTimer(prefix="onepass_table") as kt:
        grouping.groupby_aggregate(keys, {"v": vals}, {"v": "sum"})
        torch.cuda.synchronize()
    assert kt.stats()["onepass_table"][1] == 1
    with pytest.raises(_lib.JfxError):
        _lib.call("jfx_prof_sample", 0)


def test_stream_handle_is_the_current_stream():
    assert _lib.stream_handle() == torch.cuda.current_stream().cuda_stream
    assert _lib.stream_handle(torch.device("cuda")) == torch.cuda.current_stream().cuda_stream
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        assert _lib.stream_handle() == s.cuda_stream
        assert _lib.stream_handle(0) == s.cuda_stream
        assert _lib.stream_handle(torch.device("cuda", 0)) == s.cuda_stream
    assert _lib.stream_handle() == torch.cuda.current_stream().cuda_stream
