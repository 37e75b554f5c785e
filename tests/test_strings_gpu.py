"""Dictionary-encoded string columns through JaxFrame on the GPU: the
reference's own string test vectors (tests/test_string_operations.py), and
random keys with nulls and duplicates against the oracle's restatement of the
reference's string join plans and groupby (exact keys, row order, payloads
and counts; float sums within 2e-7 * sum|x|)."""

import numpy as np
import pytest
import torch

import oracle
from jaxframes_amd import strings as S
from jaxframes_amd.frame import JaxFrame

pytestmark = pytest.mark.gpu


def _golden():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")) as f:
        return json.load(f)["strings"]


def _nan_to_none(xs):
    return [None if (isinstance(x, float) and np.isnan(x)) else x for x in xs]


def test_reference_string_vectors():
    for case in _golden():
        kind = case["kind"]
        if kind == "compare":
            df = JaxFrame({"name": np.array(case["values"], dtype=object)})
            assert (df["name"] == case["scalar"]).tolist() == case["eq"]
            assert (df["name"] != case["scalar"]).tolist() == case["ne"]
            assert (df["name"] > case["scalar"]).tolist() == case["gt"]
        elif kind == "sort":
            df = JaxFrame({"name": np.array(case["keys"], dtype=object),
                           "value": np.array(case["values"])})
            r = df.sort_values("name").to_pandas()
            assert r["name"].tolist() == case["sorted_keys"]
            assert r["value"].tolist() == case["sorted_values"]
        elif kind == "groupby":
            df = JaxFrame({"category": np.array(case["keys"], dtype=object),
                           "value": np.array(case["values"])})
            g = df.groupby("category")
            r = (g.sum() if case["agg"] == "sum" else g.count()).to_pandas()
            assert r["category"].tolist() == case["groups"]
            assert r["value"].tolist() == case["result"]
        else:
            strs = lambda xs: np.array(xs, dtype=object) if isinstance(xs[0], str) else np.array(xs)
            left = JaxFrame({"key": np.array(case["left_keys"], dtype=object),
                             "lv": strs(case["left_values"])})
            right = JaxFrame({"key": np.array(case["right_keys"], dtype=object),
                              "rv": strs(case["right_values"])})
            r = left.merge(right, on="key", how=case["how"]).to_pandas()
            if case["sorted_by"] == ["key", "left"]:
                r = r.sort_values(["key", "left_lv"]).reset_index(drop=True)
            elif case["sorted_by"] == ["key"]:
                r = r.sort_values("key").reset_index(drop=True)
            assert r["key"].tolist() == case["keys"]
            assert _nan_to_none(r["left_lv"].tolist()) == case["left"]
            assert _nan_to_none(r["right_rv"].tolist()) == case["right"]


def _words(rng, n, vocab_n, null_frac):
    w = np.array([f"k{x:03d}" for x in rng.integers(0, vocab_n, n)], dtype=object)
    w[rng.random(n) < null_frac] = None
    return w


@pytest.mark.parametrize("unique_right", [True, False])
@pytest.mark.parametrize("how", ["inner", "left", "right", "outer"])
def test_string_merge_matches_oracle(how, unique_right):
    rng = np.random.default_rng(17 + unique_right)
    lk = _words(rng, 3000, 80, 0.05)
    if unique_right:
        rk = np.array([f"k{x:03d}" for x in rng.permutation(120)[:70]] + [None], dtype=object)
    else:
        rk = _words(rng, 400, 120, 0.05)
    lv = rng.integers(-50, 50, lk.size).astype(np.int32)
    rv = rng.standard_normal(rk.size).astype(np.float32)
    r = JaxFrame({"key": lk, "a": lv}).merge(JaxFrame({"key": rk, "b": rv}), on="key", how=how)
    lc, lvoc = oracle.encode_string_array(lk)
    rc, rvoc = oracle.encode_string_array(rk)
    merged = tuple(sorted(set(lvoc) | set(rvoc)))
    pos = {v: i for i, v in enumerate(merged)}
    lc = np.array([-1 if c < 0 else pos[lvoc[c]] for c in lc], dtype=np.int32)
    rc = np.array([-1 if c < 0 else pos[rvoc[c]] for c in rc], dtype=np.int32)
    codes, li, ri = oracle.string_join_plan(lc, rc, len(merged), how)
    assert r.data["key"].cpu().numpy().tolist() == codes.tolist()
    assert r._string_vocabs["key"] == merged
    assert np.array_equal(r.data["left_a"].cpu().numpy(), oracle.gather_join_column(lv, li),
                          equal_nan=True)
    got_b = r.data["right_b"].cpu().numpy()
    exp_b = oracle.gather_join_column(rv, ri)
    assert got_b.dtype == exp_b.dtype
    assert np.array_equal(got_b, exp_b, equal_nan=True)
    pdf = r.to_pandas()
    assert pdf["key"].tolist() == [None if c < 0 else merged[c] for c in codes]


@pytest.mark.parametrize("agg", ["sum", "mean", "count", "min", "max"])
def test_string_groupby_matches_oracle(agg):
    rng = np.random.default_rng(23)
    keys = _words(rng, 200_000, 300, 0.02)
    v = rng.standard_normal(keys.size).astype(np.float32)
    r = JaxFrame({"k": keys, "v": v}).groupby("k").agg({"v": agg})
    codes, vocab = oracle.encode_string_array(keys)
    g, res = oracle.groupby_encoded_strings(codes, {"v": v}, {"v": agg}, len(vocab))
    assert r.data["k"].cpu().numpy().tolist() == g.tolist()
    assert r.to_pandas()["k"].tolist() == [vocab[c] for c in g]
    got = r.data["v"].cpu().numpy()
    if agg == "count":
        assert got.dtype == np.int32 and got.tolist() == res["v"].tolist()
    elif agg in ("min", "max"):
        assert np.array_equal(got.astype(np.float64), res["v"])
    else:
        _, ab = oracle.groupby_encoded_strings(codes, {"v": np.abs(v)}, {"v": "sum"}, len(vocab))
        cnt = np.bincount(codes[codes >= 0])[g] if agg == "mean" else 1
        assert np.all(np.abs(got.astype(np.float64) - res["v"]) <= 2e-7 * ab["v"] / cnt + 1e-30)


def test_string_sort_nulls_last_and_descending():
    rng = np.random.default_rng(29)
    keys = _words(rng, 5000, 40, 0.1)
    v = np.arange(keys.size)
    df = JaxFrame({"k": keys, "v": v})
    codes, vocab = oracle.encode_string_array(keys)
    for asc in (True, False):
        r = df.sort_values("k", ascending=asc)
        sk = np.where(codes < 0, len(vocab), codes if asc else len(vocab) - 1 - codes)
        order = np.argsort(sk, kind="stable")
        assert r.data["v"].cpu().numpy().tolist() == v[order].tolist()
        assert r.to_pandas()["k"].tolist() == [None if c < 0 else vocab[c] for c in codes[order]]
    f = df[df["k"] == "k007"]
    assert set(f.to_pandas()["k"].tolist()) <= {"k007"}
    assert len(f) == int(np.sum(keys == "k007"))
    with pytest.raises(NotImplementedError):
        df.sort_values(["k", "v"])
