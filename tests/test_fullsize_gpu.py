"""GPU parity at the workloads' full sizes, where the kernels take code paths
that small inputs never reach:

- the one-pass groupby at 16M-100M rows, where every wave runs several
  1024-row steps (the step loop, the first-step span check and the stop-word
  polling), including out-of-range keys placed after a wave's first step, in
  its first step and in the last wave's last step;
- the sort-based groupby and the batched groupby beyond 1024 tiles of 8192
  rows, where the segmented reduction's fix-up carries groups from one chunk
  of 1024 tile partials to the next;
- the 100M-key radix sorts (uniform uint64, Zipf(1.1) over 10^6 distinct
  uint64 keys with heavy ties, int32) with 12K-tile decoupled lookbacks;
- the 1e7 x 1e6 join with Zipf(1.2) skew and a hot key on both sides;
- the 2e8-key hash partition into 256 buckets.

Checks are closed forms computed on the host with numpy / pandas, each equal
to the oracle's result by construction:
- groupby: np.bincount over the keys (counts exact; float64 weighted sums as
  the FP64 oracle; |sum - ref| <= 2e-7 * sum|x| per group), pandas min / max;
- sort: the output is the stable sort exactly when (1) the row index is a
  permutation, (2) keys_out == keys_in[index], (3) keys_out is non-decreasing
  and (4) equal neighbours have increasing row indices -- these four fix
  np.argsort(kind="stable"), which tests/test_oracle.py pins to the oracle;
- join: oracle.inner_join_pairs (sort + searchsorted expansion);
- partition: fmix32 of the low 32 key bits (oracle.fmix32), bincount offsets,
  and the stable-within-bucket order.
"""

import numpy as np
import pandas as pd
import pytest
import torch

import oracle
from jaxframes_amd import _lib, batched, datagen, grouping, joining, partition, sorting
from tests._util import cuda
from tests.test_batched_gpu import _check as _check_batched

pytestmark = pytest.mark.gpu

SUM_NORM_TOL = 2e-7


def _host(t):
    if t.dtype == torch.uint64:
        return t.view(torch.int64).cpu().numpy().view(np.uint64)
    if t.dtype == torch.uint32:
        return t.view(torch.int32).cpu().numpy().view(np.uint32)
    return t.cpu().numpy()


def _kernels(fn):
    with _lib.KernelTimer() as kt:
        out = fn()
        torch.cuda.synchronize()
    return out, set(kt.stats().keys())


def _check_dense_groupby(keys, vals, uk, res, nkeys):
    """Groups of non-negative integer keys below nkeys, against bincount."""
    cnt = np.bincount(keys, minlength=nkeys)
    ref = np.bincount(keys, weights=vals.astype(np.float64), minlength=nkeys)
    absx = np.bincount(keys, weights=np.abs(vals.astype(np.float64)), minlength=nkeys)
    present = np.flatnonzero(cnt)
    assert np.array_equal(_host(uk).astype(np.int64), present)
    if "count" in res:
        assert np.array_equal(_host(res["count"]), cnt[present])
    if "sum" in res:
        got = _host(res["sum"]).astype(np.float64)
        err = np.abs(got - ref[present]) / np.maximum(absx[present], 1e-300)
        assert err.max() <= SUM_NORM_TOL, err.max()
    if "mean" in res:
        got = _host(res["mean"]).astype(np.float64)
        c = cnt[present]
        err = np.abs(got - ref[present] / c) / np.maximum(absx[present] / c, 1e-300)
        assert err.max() <= SUM_NORM_TOL, err.max()


def test_headline_groupby_100m():
    """100M int32 keys uniform in [0, 1000), float32 N(0, 1): the headline
    workload, generated as the benchmark generates it."""
    n = 100_000_000
    tk, tv = datagen.groupby_frame(n, 1000, seed=1234)
    (uk, res), ks = _kernels(lambda: grouping.groupby_aggregate(
        tk, {"s": tv, "c": tv, "m": tv}, {"s": "sum", "c": "count", "m": "mean"}))
    assert "onepass_table" in ks and "sort_stats" not in ks, ks
    keys, vals = _host(tk), _host(tv)
    _check_dense_groupby(keys, vals, uk, {"sum": res["s"], "count": res["c"], "mean": res["m"]},
                         1000)
    assert _lib.device_error() == 0


def test_headline_groupby_stats_100m_min_max():
    n = 100_000_000
    tk, tv = datagen.groupby_frame(n, 1000, seed=99)
    uk, res = grouping.groupby_stats(tk, tv)
    keys, vals = _host(tk), _host(tv)
    _check_dense_groupby(keys, vals, uk, res, 1000)
    df = pd.DataFrame({"k": keys, "v": vals}).groupby("k")["v"].agg(["min", "max"])
    assert np.array_equal(_host(res["min"]), df["min"].to_numpy())
    assert np.array_equal(_host(res["max"]), df["max"].to_numpy())


N16 = 16 * 1024 * 1024
# The one-pass kernel deals 1024-row steps to its waves grid-stride (wave w
# takes steps w, w + W, w + 2W, ...); with a sum and a count at 16M rows it
# runs W = 4096 waves, so every wave takes 4 steps and the k-th step of wave w
# holds rows [(k * W + w) * STEP, +STEP).
STEP = 1024
NWAVES = 4096


def _row(wave, k, off):
    return (k * NWAVES + wave) * STEP + off


# name, row of the odd key, odd key, whether the one-pass path still hits
ODD_CASES = [
    ("after-first-step", _row(100, 1, 500), 5000, False),
    ("in-first-step", _row(100, 0, 10), 5000, False),
    ("last-wave-last-step", N16 - 7, 5000, False),
    ("span-1500-retry", _row(2000, 2, 300), 1500, True),
    ("negative-after-first-step", _row(777, 3, 222), -3000, False),
    ("negative-span-1003-hit", _row(777, 3, 222), -3, True),
]


@pytest.mark.parametrize("name,row,odd,hit", ODD_CASES, ids=[c[0] for c in ODD_CASES])
def test_onepass_multistep_odd_key(name, row, odd, hit):
    rng = np.random.default_rng(row)
    keys = rng.integers(0, 1000, N16).astype(np.int32)
    # keep the 64 sampled rows (i * n / 64) inside the common range
    assert row % (N16 // 64) != 0
    keys[row] = odd
    vals = rng.standard_normal(N16).astype(np.float32)
    (uk, res), ks = _kernels(lambda: grouping.groupby_aggregate(
        cuda(keys), {"s": cuda(vals), "c": cuda(vals)}, {"s": "sum", "c": "count"}))
    assert "onepass_table" in ks
    stats = any(k.startswith("sort_stats") for k in ks)
    if hit:
        # spans of 1024..2047 take the 2048-residue retry
        assert ("onepass_table_2k" in ks) == (max(odd, 999) - min(odd, 0) >= 1024), ks
        assert not stats, ks
    else:
        assert stats, ks
    shift = -odd if odd < 0 else 0
    _check_dense_groupby(keys.astype(np.int64) + shift, vals,
                         torch.from_numpy(_host(uk).astype(np.int64) + shift),
                         {"sum": res["s"], "count": res["c"]}, max(odd, 999) + 1 + shift)
    assert _lib.device_error() == 0


@pytest.mark.parametrize("nkeys", [2, 10, 300])
def test_onepass_multistep_lane_group_tables(nkeys):
    """Few distinct keys (lane-group tables of 16 / 64 / 1024 residues) and a
    50% hot key, 4 steps per wave."""
    rng = np.random.default_rng(nkeys)
    keys = rng.integers(0, nkeys, N16).astype(np.int32) + 40
    keys[rng.random(N16) < 0.5] = 41
    vals = rng.standard_normal(N16).astype(np.float32)
    (uk, res), ks = _kernels(lambda: grouping.groupby_aggregate(
        cuda(keys), {"s": cuda(vals), "c": cuda(vals)}, {"s": "sum", "c": "count"}))
    assert "onepass_table" in ks and "sort_stats" not in ks, ks
    _check_dense_groupby(keys, vals, uk, {"sum": res["s"], "count": res["c"]}, nkeys + 41)


def _check_sparse_groupby(keys, vals, uk, res):
    ukeys, inv, cnt = np.unique(keys, return_inverse=True, return_counts=True)
    assert np.array_equal(_host(uk), ukeys)
    v64 = vals.astype(np.float64)
    ref = np.bincount(inv, weights=v64, minlength=ukeys.size)
    absx = np.bincount(inv, weights=np.abs(v64), minlength=ukeys.size)
    if "count" in res:
        assert np.array_equal(_host(res["count"]), cnt)
    got = _host(res["sum"]).astype(np.float64)
    err = np.abs(got - ref) / np.maximum(absx, 1e-300)
    assert err.max() <= SUM_NORM_TOL, err.max()
    if "min" in res:
        df = pd.DataFrame({"g": inv, "v": vals}).groupby("g")["v"].agg(["min", "max"])
        assert np.array_equal(_host(res["min"]), df["min"].to_numpy())
        assert np.array_equal(_host(res["max"]), df["max"].to_numpy())


N20 = 20_000_000   # 2442 tiles of 8192 rows: three chunks of tile partials


@pytest.mark.parametrize("kind", ["5-wide-groups", "1000-wide-groups", "g-near-n", "float-keys"])
def test_general_groupby_beyond_1024_tiles(kind):
    rng = np.random.default_rng(len(kind))
    if kind == "5-wide-groups":
        # groups of ~4M rows each span ~490 tiles, across chunk boundaries
        keys = (rng.integers(0, 5, N20) * 10**12 - 7).astype(np.int64)
    elif kind == "1000-wide-groups":
        keys = (rng.integers(0, 1000, N20) * 3_000_000_007).astype(np.int64)
    elif kind == "g-near-n":
        keys = rng.integers(-2**62, 2**62, N20).astype(np.int64)
        keys[::3] = keys[1::3][: keys[::3].shape[0]]   # some repeats
    else:
        keys = rng.integers(0, 100_000, N20).astype(np.float32) * np.float32(0.25)
    vals = rng.standard_normal(N20).astype(np.float32)
    (uk, res), ks = _kernels(lambda: grouping.groupby_stats(cuda(keys), cuda(vals)))
    assert "segreduce_fix" in ks, ks
    _check_sparse_groupby(keys, vals, uk, res)
    assert _lib.device_error() == 0


def test_batched_groupby_beyond_1024_tiles():
    """64 frames of 300K rows (19.2M rows, 2344 tiles) with cardinalities up
    to 10^6, in one pipeline."""
    rng = np.random.default_rng(64)
    nframes, rows = 64, 300_000
    cards = np.floor(np.exp(rng.uniform(np.log(10), np.log(10**6), nframes))).astype(np.int64)
    cards[::7] = 3          # frames whose few groups span many tiles
    keys = np.concatenate([rng.integers(0, c, rows) for c in cards]).astype(np.int32)
    vals = rng.standard_normal(keys.shape[0]).astype(np.float32)
    offs = np.arange(nframes + 1, dtype=np.int64) * rows
    plan = batched.BatchedGroupBy(cuda(keys), cuda(vals), offs, aggs=("sum", "count", "max"),
                                  digit_bits=8)
    res, ks = _kernels(plan.run)
    assert "segreduce_fix" in ks
    _check_batched(res, keys, vals, offs, ("sum", "count", "max"))
    assert _lib.device_error() == 0


def _check_stable_sort(kin, kout, idx, descending=False):
    n = kin.shape[0]
    assert kout.shape[0] == n and idx.shape[0] == n
    assert idx.min() >= 0 and idx.max() < n
    assert np.bincount(idx, minlength=n).max() == 1          # a permutation
    assert np.array_equal(kout, kin[idx])
    ok = oracle.ordered_bits(kout)
    if descending:
        assert (ok[:-1] >= ok[1:]).all()
        ties = ok[:-1] == ok[1:]
        assert (idx[:-1][ties] > idx[1:][ties]).all()         # ties reversed
    else:
        assert (ok[:-1] <= ok[1:]).all()
        ties = ok[:-1] == ok[1:]
        assert (idx[:-1][ties] < idx[1:][ties]).all()         # stable
    return int(ties.sum())


@pytest.mark.parametrize("kind,descending", [("uniform", False), ("zipf", False), ("zipf", True)])
def test_sort_100m_uint64(kind, descending):
    n = 100_000_000
    tk = datagen.sort_keys(n, kind=kind, seed=1234)
    sk, _, idx = sorting.radix_sort(tk, None, not descending, return_index=True,
                                    index_dtype=torch.int32)
    ties = _check_stable_sort(_host(tk), _host(sk), _host(idx).astype(np.int64), descending)
    if kind == "zipf":
        assert ties > n // 2       # heavy ties: stability is what is tested
    assert _lib.device_error() == 0


def test_sort_100m_int32_with_payload():
    n = 100_000_000
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    tk = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device="cuda", generator=g)
    pay = torch.arange(n, dtype=torch.int32, device="cuda")
    sk, sp = sorting.radix_sort(tk, pay)
    _check_stable_sort(_host(tk), _host(sk), _host(sp).astype(np.int64))


def _check_join(lk, lv, rk, rv, how_many=None):
    k, li, ri = joining.inner_join_rows(lk, rk)
    hk, hl, hr = _host(lk), _host(li).view(np.uint32), _host(ri).view(np.uint32)
    ek, eli, eri = oracle.inner_join_pairs(_host(lk), _host(rk))
    if how_many is not None:
        assert ek.shape[0] == how_many
    assert np.array_equal(_host(k), ek)
    assert np.array_equal(hl.astype(np.int64), eli)
    assert np.array_equal(hr.astype(np.int64), eri)
    keys, out = joining.sort_merge_join(lk, lv, rk, rv, "inner")
    assert np.array_equal(_host(keys), ek)
    for c, v in lv.items():
        assert np.array_equal(_host(out[f"left_{c}"]), _host(v)[eli]), c
    for c, v in rv.items():
        assert np.array_equal(_host(out[f"right_{c}"]), _host(v)[eri]), c
    del hk


def test_join_1e7_x_1e6_zipf():
    """The join workload at a tenth of its size: right keys twice each, 80% of
    the left rows matching by Zipf(1.2) rank, two float32 payloads per side."""
    lk, lv, rk, rv, m = datagen.join_tables(10**7, 10**6, seed=77)
    _check_join(lk, lv, rk, rv, how_many=m)
    assert _lib.device_error() == 0


def test_join_hot_key_both_sides():
    """One key holds ~19% of the left rows and 50 right rows: ~9.5M pairs
    from one key, written by one long run."""
    rng = np.random.default_rng(3)
    nk = 100
    p = 1.0 / np.arange(1, nk + 1) ** 1.2
    p /= p.sum()
    lk = (rng.choice(nk, 1_000_000, p=p) * 7919).astype(np.int64)
    rk = np.concatenate([np.full(50, 0), np.arange(1, nk) * 7919,
                         rng.integers(0, nk, 5000) * 7919]).astype(np.int64)
    rng.shuffle(rk)
    lv = {"a": cuda(rng.standard_normal(lk.shape[0]).astype(np.float32))}
    rv = {"c": cuda(rng.standard_normal(rk.shape[0]).astype(np.float32))}
    _check_join(cuda(lk), lv, cuda(rk), rv)
    assert _lib.device_error() == 0


def test_hash_partition_2e8():
    n = 200_000_000
    tk = datagen.partition_keys(n, seed=1234)
    r = partition.hash_partition(tk, 256, return_keys=True, return_rows=True)
    keys = _host(tk)
    b = (oracle.fmix32((keys & 0xFFFFFFFF).astype(np.uint32)) & np.uint32(255)).astype(np.int64)
    cnt = np.bincount(b, minlength=256)
    offs = np.zeros(257, dtype=np.int64)
    np.cumsum(cnt, out=offs[1:])
    assert np.array_equal(_host(r["offsets"]), offs)
    rows = _host(r["rows"])
    assert np.bincount(rows, minlength=n).max() == 1
    assert np.array_equal(_host(r["keys"]), keys[rows])
    bb = b[rows]
    assert (bb[:-1] <= bb[1:]).all()
    same = bb[:-1] == bb[1:]
    assert (rows[:-1][same] < rows[1:][same]).all()
    assert _lib.device_error() == 0
