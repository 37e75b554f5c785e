"""GPU parity of the sort-merge join against the oracle (row pairs bit-exact)."""

import json
import os

import numpy as np
import pytest
import torch

import oracle
from jaxframes_amd import joining
from tests._util import cuda

pytestmark = pytest.mark.gpu


def _rows(t):
    return t.cpu().numpy().view(np.uint32).astype(np.int64)


def _check_inner(lk, rk):
    ek, eli, eri = oracle.inner_join_pairs(lk, rk)
    k, li, ri = joining.inner_join_rows(cuda(lk), cuda(rk))
    got_k = k.cpu().numpy() if k.dtype not in (torch.uint32, torch.uint64) else \
        k.view(torch.int64 if k.dtype == torch.uint64 else torch.int32).cpu().numpy().view(lk.dtype)
    assert got_k.shape == ek.shape, (got_k.shape, ek.shape)
    assert np.array_equal(got_k.view(np.uint8), ek.view(np.uint8))
    assert np.array_equal(_rows(li), eli)
    assert np.array_equal(_rows(ri), eri)


@pytest.mark.parametrize("L,R,card", [(1, 1, 1), (100, 50, 10), (5000, 3000, 100), (100000, 20000, 5000),
                                      (300000, 70000, 10**9), (20000, 200000, 30)])
@pytest.mark.parametrize("dt", [np.int64, np.int32, np.uint64])
def test_inner_join_matches_oracle(L, R, card, dt):
    rng = np.random.default_rng(L + R + card)
    lk = rng.integers(0, card, L).astype(dt)
    rk = rng.integers(0, card, R).astype(dt)
    _check_inner(lk, rk)


def test_inner_join_skewed_and_large_windows():
    rng = np.random.default_rng(9)
    # Zipf-skewed left over right keys that each appear twice (the join config's shape)
    rkeys = rng.choice(np.arange(10**12, 10**12 + 10**6), 20000, replace=False).astype(np.int64)
    rk = np.concatenate([rkeys, rkeys])
    rng.shuffle(rk)
    p = 1.0 / np.arange(1, rkeys.size + 1) ** 1.2
    p /= p.sum()
    lk = np.concatenate([rkeys[rng.choice(rkeys.size, 160000, p=p)],
                         rng.integers(-10**6, 0, 40000).astype(np.int64)])
    rng.shuffle(lk)
    _check_inner(lk, rk)
    # a right run longer than the LDS window: global-search fallback, many pairs per row
    lk = np.full(500, 5, dtype=np.int64)
    rk = np.concatenate([np.full(9000, 5), np.arange(100)]).astype(np.int64)
    _check_inner(lk, rk)


def test_inner_join_float_keys_nan_zero():
    vals = np.array([np.nan, -0.0, 0.0, 1.5, -2.0, np.inf], dtype=np.float32)
    rng = np.random.default_rng(3)
    lk = rng.choice(vals, 3000)
    rk = rng.choice(vals, 2000)
    _check_inner(lk, rk)
    _check_inner(lk.astype(np.float64), rk.astype(np.float64))


def test_inner_join_no_matches_and_empty():
    _check_inner(np.arange(100, dtype=np.int64), np.arange(200, 300, dtype=np.int64))
    k, li, ri = joining.inner_join_rows(torch.empty(0, dtype=torch.int64, device="cuda"),
                                        cuda(np.arange(5, dtype=np.int64)))
    assert k.numel() == 0 and li.numel() == 0


@pytest.mark.parametrize("how", ["inner", "left", "right", "outer"])
def test_join_payloads_vs_oracle(how):
    rng = np.random.default_rng(11)
    lk = rng.integers(0, 300, 4000).astype(np.int64)
    rk = rng.integers(100, 400, 3000).astype(np.int64)
    lv = {"a": rng.standard_normal(4000).astype(np.float32), "b": rng.integers(0, 9, 4000).astype(np.int32)}
    rv = {"c": rng.standard_normal(3000)}
    ek, eout, _, _ = oracle.sort_merge_join(lk, lv, rk, rv, how)
    k, out = joining.sort_merge_join(cuda(lk), {c: cuda(v) for c, v in lv.items()},
                                     cuda(rk), {c: cuda(v) for c, v in rv.items()}, how)
    assert np.array_equal(k.cpu().numpy(), ek)
    for c, e in eout.items():
        g = out[c].cpu().numpy()
        assert g.dtype == e.dtype, (c, g.dtype, e.dtype)
        assert np.array_equal(np.isnan(g), np.isnan(e))
        assert np.array_equal(g[~np.isnan(g)], e[~np.isnan(e)])


def test_join_reference_vectors():
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")
    with open(path) as f:
        cases = json.load(f)["join"]
    for case in cases:
        k, out = joining.sort_merge_join(
            cuda(np.asarray(case["left_keys"])), {c: cuda(np.asarray(v)) for c, v in case["left_values"].items()},
            cuda(np.asarray(case["right_keys"])), {c: cuda(np.asarray(v)) for c, v in case["right_values"].items()},
            case["how"])
        assert k.cpu().numpy().tolist() == case["expected_keys"], case["source"]
        for c, exp in case["expected"].items():
            np.testing.assert_array_equal(out[c].cpu().numpy(), np.asarray(exp, dtype=np.float64), case["source"])


def test_no_device_error_after_join():
    from jaxframes_amd import _lib
    assert _lib.device_error() == 0


PAYLOAD_LAYOUTS = [
    # (left column dtypes, right column dtypes): carried through the sort when
    # a side's columns take at most 8 bytes, else gathered at the source row
    ([], []),
    ([np.float32], [np.float32]),
    ([np.float32, np.float32], [np.float32, np.int32]),       # packed 4 + 4
    ([np.float64], [np.int64]),                               # one 8-byte record
    ([np.float32, np.float32, np.float32], [np.float64, np.float32]),  # by row
    ([np.float32] * 9, [np.int16]),                           # beyond 8 columns
]


@pytest.mark.parametrize("lcols,rcols", PAYLOAD_LAYOUTS,
                         ids=["none", "1+1", "2+2", "8B", "rows", "9cols"])
def test_inner_join_payload_layouts(lcols, rcols):
    rng = np.random.default_rng(len(lcols) * 10 + len(rcols))
    L, R = 30000, 7000
    lk = rng.integers(0, 5000, L).astype(np.int64)
    rk = rng.integers(2000, 9000, R).astype(np.int64)

    def col(dt, n):
        return (rng.standard_normal(n) * 100).astype(dt) if np.dtype(dt).kind == "f" else \
            rng.integers(-1000, 1000, n).astype(dt)
    lv = {f"l{j}": col(dt, L) for j, dt in enumerate(lcols)}
    rv = {f"r{j}": col(dt, R) for j, dt in enumerate(rcols)}
    ek, eout, _, _ = oracle.sort_merge_join(lk, lv, rk, rv, "inner")
    k, out = joining.sort_merge_join(cuda(lk), {c: cuda(v) for c, v in lv.items()},
                                     cuda(rk), {c: cuda(v) for c, v in rv.items()}, "inner")
    assert np.array_equal(k.cpu().numpy(), ek)
    assert sorted(out) == sorted(eout)
    for c, e in eout.items():
        g = out[c].cpu().numpy()
        assert g.dtype == e.dtype, (c, g.dtype, e.dtype)
        assert np.array_equal(g, e), c


def test_inner_join_heavy_tiles_split_into_chunks():
    """Left rows whose key matches 3000 right rows: a 2048-row tile holds
    millions of pairs, written in many fill chunks; pairs and payloads equal
    the oracle's."""
    rng = np.random.default_rng(21)
    lk = np.concatenate([np.full(3000, 7), rng.integers(0, 50, 3000)]).astype(np.int64)
    rng.shuffle(lk)
    rk = np.concatenate([np.full(3000, 7), np.arange(0, 50)]).astype(np.int64)
    rng.shuffle(rk)
    lv = {"a": rng.standard_normal(lk.size).astype(np.float32)}
    rv = {"c": rng.standard_normal(rk.size).astype(np.float32)}
    _check_inner(lk, rk)
    ek, eout, _, _ = oracle.sort_merge_join(lk, lv, rk, rv, "inner")
    k, out = joining.sort_merge_join(cuda(lk), {c: cuda(v) for c, v in lv.items()},
                                     cuda(rk), {c: cuda(v) for c, v in rv.items()}, "inner")
    assert ek.size > 9_000_000
    assert np.array_equal(k.cpu().numpy(), ek)
    for c, e in eout.items():
        assert np.array_equal(out[c].cpu().numpy(), e), c
