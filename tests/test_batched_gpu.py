"""GPU parity of the batched groupby: every frame equals the oracle's groupby
of that frame (keys, counts, min, max exact; float sums to the single-key
tolerances), eagerly and through HIP-graph replays."""

import numpy as np
import pytest
import torch

import oracle
from jaxframes_amd import _lib, batched
from tests._util import cuda

pytestmark = pytest.mark.gpu

SUM_NORM_TOL = 2e-7


def _frames(seed, nframes, max_rows, max_card, dtype=np.int32, neg=False):
    rng = np.random.default_rng(seed)
    rows = np.floor(np.exp(rng.uniform(0, np.log(max_rows), nframes))).astype(np.int64)
    rows[::17] = 0       # empty frames
    rows[1::19] = 1      # single-row frames
    cards = np.floor(np.exp(rng.uniform(np.log(2), np.log(max_card), nframes))).astype(np.int64)
    keys = np.concatenate([rng.integers(-c if neg else 0, c, r) for r, c in zip(rows, cards)]).astype(dtype)
    vals = rng.standard_normal(keys.shape[0]).astype(np.float32)
    offs = np.zeros(nframes + 1, dtype=np.int64)
    offs[1:] = np.cumsum(rows)
    return keys, vals, offs


def _check(res, keys, vals, offs, aggs):
    got_off = res.offsets.cpu().numpy()
    gk = res.keys.cpu().numpy()
    gv = {a: v.cpu().numpy() for a, v in res.values.items()}
    for f in range(offs.shape[0] - 1):
        a, b = offs[f], offs[f + 1]
        ek, ev = oracle.groupby_aggregate(keys[a:b], {"v": vals[a:b]}, {"v": "count"})
        ga, gb = got_off[f], got_off[f + 1]
        assert np.array_equal(gk[ga:gb], ek), f
        for agg in aggs:
            _, e = oracle.groupby_aggregate(keys[a:b], {"v": vals[a:b]}, {"v": agg})
            g = gv[agg][ga:gb]
            if agg in ("count", "min", "max"):
                assert np.array_equal(g, e["v"]), (f, agg)
            else:
                _, absx = oracle.groupby_aggregate(keys[a:b], {"v": np.abs(vals[a:b].astype(np.float64))},
                                                   {"v": "sum"})
                scale = absx["v"] if agg == "sum" else absx["v"] / np.maximum(ev["v"], 1)
                assert np.all(np.abs(g.astype(np.float64) - e["v"]) <= SUM_NORM_TOL * np.maximum(scale, 1e-30)), (f, agg)


@pytest.mark.parametrize("nframes,max_rows,max_card", [(1, 1000, 10), (64, 5000, 50),
                                                       (300, 20000, 5000), (40, 200000, 10**6)])
def test_batched_matches_per_frame_oracle(nframes, max_rows, max_card):
    keys, vals, offs = _frames(nframes + max_rows, nframes, max_rows, max_card)
    plan = batched.BatchedGroupBy(cuda(keys), cuda(vals), offs, aggs=("sum", "count", "min", "max", "mean"),
                                  digit_bits=8)
    _check(plan.run(), keys, vals, offs, ("sum", "count", "min", "max", "mean"))
    assert _lib.device_error() == 0


def test_batched_signed_int64_keys_and_11_bit_digits():
    keys, vals, offs = _frames(3, 50, 3000, 100000, dtype=np.int64, neg=True)
    plan = batched.BatchedGroupBy(cuda(keys), cuda(vals), offs, aggs=("sum", "count"), digit_bits=11)
    _check(plan.run(), keys, vals, offs, ("sum", "count"))


def test_batched_graph_replays_match_eager():
    keys, vals, offs = _frames(5, 200, 20000, 3000)
    kt, vt = cuda(keys), cuda(vals)
    plan = batched.BatchedGroupBy(kt, vt, offs, aggs=("sum", "count"), digit_bits=8)
    eager = plan.run(graph=False)
    ek, es, ec = eager.keys.clone(), eager.values["sum"].clone(), eager.values["count"].clone()
    plan.capture()
    for i in range(4):
        # eager work of the program between replays
        junk = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        r = plan.run(graph=True)
        assert torch.equal(r.keys, ek), i
        assert torch.equal(r.values["count"], ec), i
        assert torch.equal(r.values["sum"], es), i
    _check(r, keys, vals, offs, ("sum", "count"))


def test_key_outside_declared_range_raises_device_error():
    keys = np.array([0, 1, 2, 50, 3], dtype=np.int32)
    vals = np.ones(5, dtype=np.float32)
    plan = batched.BatchedGroupBy(cuda(keys), cuda(vals), np.array([0, 5]), key_range=(0, 3),
                                  digit_bits=8)
    with pytest.raises(_lib.JfxError, match="outside the declared key_range"):
        plan.run()
    # the word was read and cleared by the check
    assert _lib.device_error() == 0


def test_plain_per_frame_equals_batched():
    keys, vals, offs = _frames(9, 30, 5000, 300)
    kt, vt = cuda(keys), cuda(vals)
    plain = batched.plain_per_frame(kt, vt, offs, "sum")
    res = batched.BatchedGroupBy(kt, vt, offs, aggs=("sum",), digit_bits=8).run()
    for f, (uk, s) in enumerate(plain):
        k, v = res.frame(f)
        assert torch.equal(k, uk)
        assert torch.allclose(v["sum"], s, rtol=1e-5, atol=1e-5)


def test_oversize_batch_is_split_at_frame_boundaries():
    """A table above max_rows runs as several pipelines; the groups equal the
    single-pipeline result frame by frame."""
    from jaxframes_amd import batched, datagen
    keys, vals, offs, cards = datagen.batched_frames(nframes=64, min_rows=1000, max_rows=200_000,
                                                     seed=3)
    n = int(keys.shape[0])
    whole = batched.BatchedGroupBy(keys, vals, offs, aggs=("sum", "count"), digit_bits=8,
                                   tuner=None).run()
    split = batched.BatchedGroupBy(keys, vals, offs, aggs=("sum", "count"), digit_bits=8,
                                   tuner=None, max_rows=n // 3)
    assert split.parts is not None and len(split.parts) >= 3
    for graph in (False, True):
        r = split.run(graph=graph)
        assert torch.equal(r.offsets, whole.offsets)
        assert torch.equal(r.keys, whole.keys)
        assert torch.equal(r.values["count"], whole.values["count"])
        # the same compensated sums; tile boundaries differ between the plans
        d = (r.values["sum"].double() - whole.values["sum"].double()).abs()
        assert torch.all(d <= 1e-6 * (1 + whole.values["sum"].double().abs()))
    with pytest.raises(ValueError):
        batched.BatchedGroupBy(keys, vals, offs, digit_bits=8, tuner=None, max_rows=999)


@pytest.mark.parametrize("digit_bits", [8, 11])
@pytest.mark.parametrize("sort_shape", [0, 1, 2, 3])
def test_batched_every_sort_configuration(digit_bits, sort_shape):
    """Every (digit width, tile shape) the autotuner may choose gives the
    per-frame oracle's groups, for 32-bit and 64-bit composite keys."""
    keys, vals, offs = _frames(100 + sort_shape, 50, 60000, 5000)
    plan = batched.BatchedGroupBy(cuda(keys), cuda(vals), offs, aggs=("sum", "count", "max"),
                                  digit_bits=digit_bits, sort_shape=sort_shape)
    _check(plan.run(), keys, vals, offs, ("sum", "count", "max"))
    wide = keys.astype(np.int64) * 3_000_000_000
    plan = batched.BatchedGroupBy(cuda(wide), cuda(vals), offs, aggs=("sum", "count"),
                                  digit_bits=digit_bits, sort_shape=sort_shape)
    assert plan.total_bits > 32
    _check(plan.run(), wide, vals, offs, ("sum", "count"))
    assert _lib.device_error() == 0


def test_autotuner_picks_and_caches(tmp_path):
    keys, vals, offs = _frames(7, 20, 20000, 300)
    path = tmp_path / "tune.json"
    tuner = batched.Autotuner(str(path))
    plan = batched.BatchedGroupBy(cuda(keys), cuda(vals), offs, aggs=("sum",), tuner=tuner)
    assert (plan.digit_bits, plan.sort_shape) in batched.Autotuner.CANDIDATES
    assert len(tuner.timings[batched.Autotuner.key(plan.n, plan.total_bits)]) == len(
        batched.Autotuner.CANDIDATES)
    again = batched.Autotuner(str(path))
    assert again.lookup(plan.n, plan.total_bits) == (plan.digit_bits, plan.sort_shape)
    _check(plan.run(), keys, vals, offs, ("sum",))
