"""Run-to-run bit equality of the groupby kernels that add floats with LDS
atomics.

The one-pass tables (k_onepass_table) keep one sum table per wave, so the
adds into an entry come from one wave only: instructions in program order,
and within one ds_add instruction the lanes that hit the same address are
serialised by the LDS in a fixed lane order. The small-range wave tables
(k_wave_table) and the tile-sorted reduction rely on the same two facts; the
sort path and the batched path reduce in a fixed tree. Each case below runs
one call three times (the second after unrelated work on the stream, the
third on another stream) and requires identical bits of every output, on
inputs whose keys collide within single instructions (few distinct keys, one
hot key) as well as spread keys, at sizes where every wave runs many steps.
"""

import zlib

import numpy as np
import pytest
import torch

from jaxframes_amd import _lib, batched, grouping
from tests._util import cuda
from tests.test_onepass_gpu import _kernels_run

pytestmark = pytest.mark.gpu


def _bits(t):
    t = t.contiguous()
    if t.dtype in (torch.float32, torch.int32, torch.uint32):
        return t.view(torch.int32).cpu()
    if t.dtype in (torch.float64, torch.int64, torch.uint64):
        return t.view(torch.int64).cpu()
    return t.cpu()


def _three_runs(fn):
    """fn() three times: as is, after unrelated work queued on the stream,
    and on a second stream (default priority); every output as raw bits, and
    the kernels the first run launched."""
    out1, ran = _kernels_run(fn)
    junk = torch.empty(1 << 26, dtype=torch.float32, device="cuda").normal_()
    junk.mul_(3.0).add_(1.0)
    out2 = fn()
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        out3 = fn()
    s.synchronize()
    del junk
    assert _lib.device_error() == 0
    return [out1, out2, out3], ran


def _flat(out):
    uk, res = out
    return [_bits(uk)] + [_bits(res[c]) for c in sorted(res)]


def _assert_same(outs):
    a = _flat(outs[0])
    for o in outs[1:]:
        b = _flat(o)
        assert len(a) == len(b)
        for x, y in zip(a, b):
            assert torch.equal(x, y)


def _keys(kind, n, rng):
    if kind == "uniform-1000":
        return rng.integers(0, 1000, n).astype(np.int32)
    if kind == "three-keys":
        return rng.integers(7, 10, n).astype(np.int32)
    if kind == "hot-key":
        k = rng.integers(0, 1000, n).astype(np.int32)
        k[rng.random(n) < 0.9] = 123
        return k
    if kind == "wide-range":   # a span of 2000: the 2048-entry small-range tables
        return rng.integers(-900, 1100, n).astype(np.int32)
    if kind == "wide-int64":
        return rng.integers(-2**62, 2**62, n // 64).astype(np.int64)[rng.integers(0, n // 64, n)]
    raise ValueError(kind)


N = 9_000_000  # above 4 * 1024 * 1024 rows: one-pass waves run several steps


@pytest.mark.parametrize("kind", ["uniform-1000", "three-keys", "hot-key"])
@pytest.mark.parametrize("vdt", [np.float32, np.float64])
def test_onepass_sums_bit_identical(kind, vdt):
    rng = np.random.default_rng(zlib.crc32(f"{kind}-{np.dtype(vdt).str}".encode()))
    k = cuda(_keys(kind, N, rng))
    v = cuda(rng.standard_normal(N).astype(vdt) * 100)
    outs, ran = _three_runs(lambda: grouping.groupby_aggregate(k, {"v": v, "w": v},
                                                               {"v": "sum", "w": "mean"}))
    assert "onepass_table" in ran
    _assert_same(outs)


@pytest.mark.parametrize("kind", ["uniform-1000", "hot-key"])
def test_onepass_stats_bit_identical(kind):
    rng = np.random.default_rng(5)
    k = cuda(_keys(kind, N, rng))
    v = cuda(rng.standard_normal(N).astype(np.float32))
    outs, ran = _three_runs(lambda: grouping.groupby_stats(k, v))
    assert "onepass_table" in ran
    _assert_same(outs)


@pytest.mark.parametrize("path,kind", [("wave", "wide-range"), ("wave", "three-keys"),
                                       ("tile", "hot-key"), ("tile", "wide-range")])
def test_small_range_sums_bit_identical(monkeypatch, path, kind):
    monkeypatch.setenv("JFX_ONEPASS", "0")
    monkeypatch.setenv("JFX_SMALL_RANGE", path)
    rng = np.random.default_rng(17)
    k = cuda(_keys(kind, N, rng))
    v = cuda(rng.standard_normal(N).astype(np.float32))
    outs, ran = _three_runs(lambda: grouping.groupby_aggregate(k, {"v": v}, {"v": "sum"}))
    assert "onepass_table" not in ran
    assert any(r.startswith("wave_table" if path == "wave" else "tile_reduce") for r in ran), ran
    _assert_same(outs)


def test_sorted_path_sums_bit_identical():
    rng = np.random.default_rng(23)
    k = cuda(_keys("wide-int64", N, rng))
    v = cuda(rng.standard_normal(N).astype(np.float32))
    outs, ran = _three_runs(lambda: grouping.groupby_stats(k, v))
    # the one-pass try misses on these keys; the sort and the reduction run
    assert "segreduce_tiles" in ran and any(r.startswith("onesweep") for r in ran), ran
    _assert_same(outs)


def test_batched_bit_identical():
    rng = np.random.default_rng(29)
    nframes = 300
    rows = np.floor(np.exp(rng.uniform(0, np.log(200_000), nframes))).astype(np.int64)
    cards = np.floor(np.exp(rng.uniform(np.log(2), np.log(10**5), nframes))).astype(np.int64)
    keys = np.concatenate([rng.integers(0, c, r) for r, c in zip(rows, cards)]).astype(np.int32)
    vals = rng.standard_normal(keys.shape[0]).astype(np.float32)
    offs = np.zeros(nframes + 1, dtype=np.int64)
    offs[1:] = np.cumsum(rows)
    plan = batched.BatchedGroupBy(cuda(keys), cuda(vals), offs, aggs=("sum", "mean", "min"))

    def run():
        r = plan.run()   # the value columns are views of the plan's output buffers
        return r.keys, dict({a: v.clone() for a, v in r.values.items()}, offsets=r.offsets)

    outs, _ = _three_runs(run)
    _assert_same(outs)
