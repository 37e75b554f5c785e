"""GPU parity of the multi-column join (join_multi_column) against the oracle's
restatement of the reference's L x R algorithm: row order, coalesced keys and
payloads exact (NaN where unmatched)."""

import json
import os

import numpy as np
import pytest

import oracle
from jaxframes_amd import joining
from tests._util import cuda

pytestmark = pytest.mark.gpu


def _same(g, e):
    g = np.asarray(g)
    assert g.shape == e.shape, (g.shape, e.shape)
    assert g.dtype == e.dtype, (g.dtype, e.dtype)
    if e.dtype.kind == "f":
        assert np.array_equal(np.isnan(g), np.isnan(e))
        assert np.array_equal(g[~np.isnan(g)], e[~np.isnan(e)])
    else:
        assert np.array_equal(g, e)


def _check(lk, lv, rk, rv, how):
    names = [f"k{i}" for i in range(len(lk))]
    ek, ev = oracle.join_multi_column(lk, names, lv, rk, names, rv, how)
    gk, gv = joining.join_multi_column([cuda(k) for k in lk], names, {c: cuda(v) for c, v in lv.items()},
                                       [cuda(k) for k in rk], names, {c: cuda(v) for c, v in rv.items()},
                                       how)
    for nm in names:
        _same(gk[nm].cpu().numpy(), ek[nm])
    for c in ev:
        _same(gv[c].cpu().numpy(), ev[c])


@pytest.mark.parametrize("how", ["inner", "left", "right", "outer"])
@pytest.mark.parametrize("nl,nr,card", [(1, 1, 2), (50, 40, 3), (700, 900, 6), (1500, 300, 40)])
def test_join_multi_int_keys(how, nl, nr, card):
    rng = np.random.default_rng(nl * 7 + nr)
    lk = [rng.integers(0, card, nl).astype(np.int64), rng.integers(0, 3, nl).astype(np.int32).astype(np.int64)]
    rk = [rng.integers(0, card, nr).astype(np.int64), rng.integers(0, 3, nr).astype(np.int64)]
    lv = {"a": rng.standard_normal(nl).astype(np.float32), "b": rng.integers(0, 9, nl).astype(np.int32)}
    rv = {"c": rng.standard_normal(nr)}
    _check(lk, lv, rk, rv, how)


@pytest.mark.parametrize("how", ["inner", "left", "outer"])
def test_join_multi_float_keys_nan_and_signed_zero(how):
    rng = np.random.default_rng(2)
    vals = np.array([np.nan, -0.0, 0.0, 1.5, 2.0], dtype=np.float64)
    lk = [rng.choice(vals, 300), rng.integers(0, 2, 300).astype(np.int64)]
    rk = [rng.choice(vals, 200), rng.integers(0, 2, 200).astype(np.int64)]
    lv = {"a": rng.integers(0, 5, 300).astype(np.int64)}
    rv = {"b": rng.standard_normal(200).astype(np.float32)}
    _check(lk, lv, rk, rv, how)


def test_join_multi_golden_vectors():
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")
    with open(path) as f:
        cases = json.load(f)["join_multi"]
    for case in cases:
        lk = [np.asarray(k, dtype=np.int64) for k in case["left_keys"]]
        rk = [np.asarray(k, dtype=np.int64) for k in case["right_keys"]]
        lv = {c: np.asarray(v) for c, v in case["left_values"].items()}
        rv = {c: np.asarray(v) for c, v in case["right_values"].items()}
        _check(lk, lv, rk, rv, case["how"])
