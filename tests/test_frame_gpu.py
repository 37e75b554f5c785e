"""JaxFrame on the GPU: the expected vectors of the reference's own frame
tests (tests/test_lazy_operations.py, tests/test_head_tail.py,
tests/test_filter_execution.py), lazy == eager, and both against pandas and
the oracle on random data. Keys, orders, pairs and counts exact; float sums
within 2e-7 * sum|x|."""

import numpy as np
import pandas as pd
import pytest
import torch

import oracle
from jaxframes_amd import lazy as L
from jaxframes_amd.frame import JaxFrame

pytestmark = pytest.mark.gpu


def _np(t):
    return t if isinstance(t, np.ndarray) else t.cpu().numpy()


def _eq(a, b):
    np.testing.assert_array_equal(_np(a), np.asarray(b))


# ---- the reference's vectors ---------------------------------------------------------

@pytest.mark.parametrize("lazy", [True, False])
def test_reference_sort_vectors(lazy):
    d = {"a": np.array([3, 1, 4, 1, 5]), "b": np.array([30.0, 10.0, 40.0, 15.0, 50.0])}
    r = JaxFrame(d, lazy=lazy).sort_values("a").collect()
    _eq(r.data["a"], [1, 1, 3, 4, 5])
    _eq(r.data["b"], [10.0, 15.0, 30.0, 40.0, 50.0])
    _eq(JaxFrame(d, lazy=lazy).sort_values("a", ascending=False).collect().data["a"], [5, 4, 3, 1, 1])
    d2 = {"a": np.array([1, 2, 1, 2, 1]), "b": np.array([3, 1, 2, 3, 1]), "c": np.array([10, 20, 30, 40, 50])}
    r = JaxFrame(d2, lazy=lazy).sort_values(["a", "b"]).collect()
    _eq(r.data["a"], [1, 1, 1, 2, 2])
    _eq(r.data["b"], [1, 2, 3, 1, 3])
    r = JaxFrame(d2, lazy=lazy).sort_values(["a", "b"], ascending=[True, False]).collect()
    _eq(r.data["a"], [1, 1, 1, 2, 2])
    _eq(r.data["b"], [3, 2, 1, 3, 1])
    r = JaxFrame({"a": np.array([-5, 2, -3, 0, 1]), "b": np.array([10, 20, 30, 40, 50])},
                 lazy=lazy).sort_values("a").collect()
    _eq(r.data["a"], [-5, -3, 0, 1, 2])


@pytest.mark.parametrize("lazy", [True, False])
def test_reference_groupby_vectors(lazy):
    d = {"group": np.array([1, 2, 1, 2, 1]), "value": np.array([10, 20, 30, 40, 50])}
    r = JaxFrame(d, lazy=lazy).groupby("group").sum().collect()
    _eq(r.data["group"], [1, 2])
    _eq(r.data["value"], [90, 60])
    r = JaxFrame({"group": d["group"], "value": d["value"].astype(np.float64)},
                 lazy=lazy).groupby("group").mean().collect()
    np.testing.assert_array_almost_equal(_np(r.data["value"]), [30.0, 30.0])
    d3 = {"group": np.array([1, 2, 1, 2, 1]), "value1": np.array([10, 20, 30, 40, 50]),
          "value2": np.array([1.0, 2.0, 3.0, 4.0, 5.0])}
    r = JaxFrame(d3, lazy=lazy).groupby("group").agg({"value1": "sum", "value2": "mean"}).collect()
    _eq(r.data["group"], [1, 2])
    _eq(r.data["value1"], [90, 60])
    np.testing.assert_array_almost_equal(_np(r.data["value2"]), [3.0, 3.0])
    d4 = {"group": np.array([1, 2, 1, 2, 1]), "value": np.array([10, 20, 50, 5, 30])}
    _eq(JaxFrame(d4, lazy=lazy).groupby("group").max().collect().data["value"], [50, 20])
    _eq(JaxFrame(d4, lazy=lazy).groupby("group").min().collect().data["value"], [10, 5])
    d5 = {"group": np.array([1, 2, 1, 2, 1, 1]), "value": np.array([10, 20, 30, 40, 50, 60])}
    _eq(JaxFrame(d5, lazy=lazy).groupby("group").count().collect().data["value"], [4, 2])
    d6 = {"group1": np.array([1, 1, 2, 2, 1]), "group2": np.array([1, 2, 1, 2, 1]),
          "value": np.array([10, 20, 30, 40, 50])}
    r = JaxFrame(d6, lazy=lazy).groupby(["group1", "group2"]).sum().collect()
    _eq(r.data["group1"], [1, 1, 2, 2])
    _eq(r.data["group2"], [1, 2, 1, 2])
    _eq(r.data["value"], [60, 20, 30, 40])


@pytest.mark.parametrize("lazy", [True, False])
def test_reference_merge_vectors(lazy):
    left = {"id": np.array([1, 2, 3]), "value": np.array([10, 20, 30])}
    right = {"id": np.array([2, 3, 4]), "amount": np.array([200, 300, 400])}
    L_, R_ = JaxFrame(left, lazy=lazy), JaxFrame(right, lazy=lazy)
    r = L_.merge(R_, on="id", how="inner").collect()
    _eq(r.data["id"], [2, 3])
    _eq(r.data["left_value"], [20.0, 30.0])       # payloads come back as floats
    _eq(r.data["right_amount"], [200.0, 300.0])
    _eq(L_.merge(R_, on="id", how="left").collect().data["id"], [1, 2, 3])
    _eq(L_.merge(R_, on="id", how="right").collect().data["id"], [2, 3, 4])
    r = L_.merge(R_, on="id", how="outer").collect()
    _eq(r.data["id"], [1, 2, 3, 4])
    assert np.isnan(_np(r.data["right_amount"])[0]) and np.isnan(_np(r.data["left_value"])[3])
    lm = {"key1": np.array([1, 1, 2]), "key2": np.array([1, 2, 1]), "value": np.array([10, 20, 30])}
    rm = {"key1": np.array([1, 2, 2]), "key2": np.array([1, 1, 2]), "amount": np.array([100, 200, 300])}
    r = JaxFrame(lm, lazy=lazy).merge(JaxFrame(rm, lazy=lazy), on=["key1", "key2"]).collect()
    assert len(r.data["key1"]) == 2
    # lazy left, eager right
    r = JaxFrame(left, lazy=True).merge(JaxFrame(right), on="id").collect()
    _eq(r.data["id"], [2, 3])


def test_reference_chains():
    df = JaxFrame({"a": np.array([5, 2, 8, 1, 9]), "b": np.array([10, 20, 30, 40, 50])}, lazy=True)
    _eq(df[df["a"] > 3].sort_values("a").collect().data["a"], [5, 8, 9])
    g = JaxFrame({"group": np.array([2, 1, 2, 1, 2]), "value": np.array([10, 20, 30, 40, 50])},
                 lazy=True)
    _eq(g.sort_values("group").groupby("group").sum().collect().data["group"], [1, 2])
    l2 = JaxFrame({"id": np.array([3, 1, 2]), "value": np.array([30, 10, 20])}, lazy=True)
    r2 = JaxFrame({"id": np.array([2, 1, 3]), "amount": np.array([200, 100, 300])}, lazy=True)
    _eq(l2.merge(r2, on="id").sort_values("id").collect().data["id"], [1, 2, 3])
    c = JaxFrame({"group": np.array([1, 2, 1, 2, 1, 2]), "value": np.array([10, 5, 20, 15, 30, 25]),
                  "score": np.array([1, 2, 3, 4, 5, 6])}, lazy=True)
    r = c[["group", "value"]].sort_values("value").groupby("group").sum().collect()
    _eq(r.data["group"], [1, 2])
    _eq(r.data["value"], [60, 45])
    assert r.columns == ["group", "value"]


@pytest.mark.parametrize("lazy", [True, False])
def test_reference_head_tail(lazy):
    d = {"a": np.arange(10), "b": np.arange(10) * 1.5}
    df = JaxFrame(d, lazy=lazy)
    _eq(df.head().collect().data["a"], [0, 1, 2, 3, 4])
    _eq(df.head(3).collect().data["a"], [0, 1, 2])
    _eq(df.head(100).collect().data["a"], np.arange(10))
    assert len(df.head(0).collect().data["a"]) == 0
    assert len(df.head(-3).collect().data["a"]) == 0
    _eq(df.tail(3).collect().data["a"], [7, 8, 9])
    _eq(df.tail(100).collect().data["a"], np.arange(10))
    assert len(df.tail(0).collect().data["a"]) == 0
    _eq(df.head(7).tail(3).collect().data["a"], [4, 5, 6])
    e = JaxFrame({"a": np.array([], dtype=np.int64)}, lazy=lazy)
    assert len(e.head(3).collect().data["a"]) == 0 and len(e.tail(3).collect().data["a"]) == 0


@pytest.mark.parametrize("lazy", [True, False])
def test_reference_filters(lazy):
    d = {"a": np.array([1, 2, 3, 4, 5]), "b": np.array([10.0, 20.0, 30.0, 40.0, 50.0])}
    df = JaxFrame(d, lazy=lazy)
    cases = [(df["a"] > 3, [4, 5]), (df["a"] < 3, [1, 2]), (df["a"] == 3, [3]),
             (df["a"] != 3, [1, 2, 4, 5]), (df["a"] >= 3, [3, 4, 5]), (df["a"] <= 3, [1, 2, 3]),
             (df["a"] > 10, []), (df["a"] > 0, [1, 2, 3, 4, 5]),
             ((df["a"] > 1) & (df["b"] < 50.0), [2, 3, 4]), ((df["a"] < 2) | (df["a"] > 4), [1, 5]),
             (~(df["a"] == 3), [1, 2, 4, 5]), (df["b"] > 25.5, [3, 4, 5]), (df["a"] > 2.5, [3, 4, 5])]
    for cond, exp in cases:
        r = df[cond].collect()
        _eq(r.data["a"], exp)
        _eq(r.data["b"], np.array(exp, dtype=np.float64) * 10)


# ---- random data: pandas / oracle, lazy == eager ----------------------------------------

def _frame(n, seed, card=50):
    rng = np.random.default_rng(seed)
    return {"k": rng.integers(-card, card, n).astype(np.int64),
            "k2": rng.integers(0, 4, n).astype(np.int32),
            "v": rng.standard_normal(n).astype(np.float32),
            "w": rng.integers(-1000, 1000, n).astype(np.int64)}


def test_pipeline_matches_pandas_and_eager():
    d = _frame(200_000, 1)
    pdf = pd.DataFrame(d)
    lz = JaxFrame(d, lazy=True)
    eg = JaxFrame(d)
    # filter -> groupby (pruned: only k, v gathered)
    q = lz[lz["v"] > 0.25].groupby("k").agg({"v": "sum", "w": "count"})
    r = q.collect()
    e = eg[eg["v"] > 0.25].groupby("k").agg({"v": "sum", "w": "count"})
    m = d["v"] > 0.25
    ek, es = oracle.groupby_aggregate(d["k"][m], {"v": d["v"][m]}, {"v": "sum"})
    _, ea = oracle.groupby_aggregate(d["k"][m], {"v": np.abs(d["v"][m]).astype(np.float64)}, {"v": "sum"})
    _eq(r.data["k"], ek)
    _eq(e.data["k"], ek)
    assert np.all(np.abs(_np(r.data["v"]).astype(np.float64) - es["v"]) <= 2e-7 * ea["v"])
    assert torch.equal(r.data["v"], e.data["v"])
    pc = pdf[pdf.v > 0.25].groupby("k").w.count()
    _eq(r.data["w"], pc.values)
    # sort -> head (fused) equals full sort then slicing, and pandas' stable order
    s = lz.sort_values("w", ascending=True).head(1000).collect()
    full = eg.sort_values("w")
    for c in d:
        assert torch.equal(s.data[c], full.data[c][:1000]), c
    ps = pdf.sort_values("w", kind="stable")
    _eq(full.data["w"], ps.w.values)
    _eq(full.data["v"], ps.v.values)
    t = lz.sort_values("w", ascending=False).tail(500).collect()
    fd = eg.sort_values("w", ascending=False)
    for c in d:
        assert torch.equal(t.data[c], fd.data[c][-500:]), c
    # multi-key sort
    ms = eg.sort_values(["k2", "k"], ascending=[False, True])
    pm = pdf.assign(nk2=-pdf.k2).sort_values(["nk2", "k"], kind="stable")
    _eq(ms.data["k2"], pm.k2.values)
    _eq(ms.data["k"], pm.k.values)
    _eq(ms.data["w"], pm.w.values)


def test_merge_matches_oracle_and_pruning():
    d = _frame(50_000, 2, card=2000)
    rng = np.random.default_rng(3)
    dim = {"k": rng.choice(np.arange(-2500, 2500), 3000, replace=False).astype(np.int64),
           "p": rng.standard_normal(3000).astype(np.float32),
           "q": rng.integers(0, 9, 3000).astype(np.int32)}
    lz = JaxFrame(d, lazy=True).merge(JaxFrame(dim, lazy=True), on="k")
    full = lz.collect()
    ek, eli, eri = oracle.inner_join_pairs(d["k"], dim["k"])
    _eq(full.data["k"], ek)
    _eq(full.data["left_v"], d["v"][eli])
    _eq(full.data["left_w"], d["w"][eli].astype(np.float64))
    _eq(full.data["right_p"], dim["p"][eri])
    _eq(full.data["right_q"], dim["q"][eri].astype(np.float32))
    assert full.columns == ["k", "left_k2", "left_v", "left_w", "right_p", "right_q"]
    # a projection above the join: only the projected payloads are carried
    pr = lz[["k", "right_p"]].collect()
    assert pr.columns == ["k", "right_p"]
    assert torch.equal(pr.data["right_p"], full.data["right_p"])
    eager = JaxFrame(d).merge(JaxFrame(dim), on="k")
    for c in full.columns:
        assert torch.equal(eager.data[c], full.data[c]), c


def test_object_and_string_columns_ride_along():
    names = np.array([f"r{i}" for i in range(6)], dtype=object)          # encoded strings
    objs = np.array([{"i": i} for i in range(6)], dtype=object)          # host objects
    df = JaxFrame({"k": np.array([3, 1, 2, 1, 3, 0]), "name": names, "obj": objs})
    assert df._dtypes["name"] == "string" and df._dtypes["obj"] == "object"
    s = df.sort_values("k")
    pdf = s.to_pandas()
    assert pdf["name"].tolist() == ["r5", "r1", "r3", "r2", "r0", "r4"]
    assert [o["i"] for o in s.data["obj"]] == [5, 1, 3, 2, 0, 4]
    f = df[df["k"] > 1]
    assert f.to_pandas()["name"].tolist() == ["r0", "r2", "r4"]
    assert list(pdf.columns) == ["k", "name", "obj"] and pdf.k.tolist() == [0, 1, 1, 2, 3, 3]
