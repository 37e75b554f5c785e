"""Pins the CPU oracle (numpy restatement) before anything is checked
against it: golden vectors transcribed from the reference's tests, closed
forms, brute force, pandas and published MurmurHash3 vectors. A plausible
mistake in the oracle (dropped tie-break, wrong sign flip, transposed join
order, wrong hash constant) fails at least one check here."""

import json
import os

import numpy as np
import pandas as pd
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)


# ---- sort -----------------------------------------------------------------

def test_sort_golden_vectors():
    g = _golden()["sort"]
    for case in g:
        keys = np.asarray(case["keys"], dtype=case.get("dtype", "int64"))
        vals = np.asarray(case["values"]) if "values" in case else None
        sk, sv, _ = oracle.radix_sort(keys, vals, case.get("ascending", True))
        if np.asarray(case["expected_keys"]).dtype.kind == "f":
            np.testing.assert_allclose(sk, case["expected_keys"], rtol=1e-6)
        else:
            assert sk.tolist() == case["expected_keys"], case["source"]
        if "expected_values" in case:
            assert sv.tolist() == case["expected_values"], case["source"]


def test_sort_seeded_reference_case():
    # test_parallel_algorithms.py:29-41: np.random.seed(42); randint(0, 100, 1000)
    data = np.random.RandomState(42).randint(0, 100, size=1000)
    sk, _, _ = oracle.radix_sort(data)
    assert np.array_equal(sk, np.sort(data))


@pytest.mark.parametrize("dt", [np.int32, np.int64, np.uint32, np.uint64, np.float32, np.float64])
def test_sort_closed_form_stable_argsort(dt):
    rng = np.random.default_rng(1)
    n = 20000
    if np.dtype(dt).kind == "f":
        k = rng.choice(np.array([np.nan, -0.0, 0.0, np.inf, -np.inf, 1.5, -1.5, 2.5], dtype=dt), n)
        k[::7] = rng.standard_normal(len(k[::7])).astype(dt)
    else:
        info = np.iinfo(dt)
        k = rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
        k[::3] = k[0]  # ties
    _, _, perm = oracle.radix_sort(k)
    assert np.array_equal(perm, np.argsort(k, kind="stable"))
    _, _, perm_d = oracle.radix_sort(k, ascending=False)
    assert np.array_equal(perm_d, np.argsort(k, kind="stable")[::-1])


def test_ordered_bits_preserves_order():
    rng = np.random.default_rng(2)
    for dt in (np.int32, np.int64, np.float32, np.float64):
        k = (rng.standard_normal(5000) * 1e6).astype(dt)
        ob = oracle.ordered_bits(k)
        a, b = k[:-1], k[1:]
        oa, obb = ob[:-1], ob[1:]
        assert np.array_equal(a < b, oa < obb)
        assert np.array_equal(a == b, oa == obb)
    z = oracle.ordered_bits(np.array([-0.0, 0.0], dtype=np.float32))
    assert z[0] == z[1]


def test_lexsort_golden_and_closed_form():
    for case in _golden()["lexsort"]:
        keys = [np.asarray(k) for k in case["keys"]]
        idx = oracle.multi_column_lexsort(keys, case["ascending"])
        for col, exp in zip(case["columns"], case["expected"]):
            assert keys[col][idx].tolist() == exp, case["source"]
    rng = np.random.default_rng(3)
    a, b = rng.integers(0, 5, 1000), rng.integers(0, 5, 1000)
    assert np.array_equal(oracle.multi_column_lexsort([a, b]), np.lexsort([b, a]))


# ---- groupby ----------------------------------------------------------------

def test_groupby_golden_vectors():
    for case in _golden()["groupby"]:
        keys = np.asarray(case["keys"])
        vals = {c: np.asarray(v) for c, v in case["values"].items()}
        uk, res = oracle.groupby_aggregate(keys, vals, case["aggs"])
        assert uk.tolist() == case["expected_keys"], case["source"]
        for c, exp in case["expected"].items():
            np.testing.assert_allclose(np.asarray(res[c], dtype=np.float64), exp, rtol=1e-6,
                                       err_msg=case["source"])


def test_groupby_vs_pandas():
    rng = np.random.default_rng(4)
    n = 100000
    k = rng.integers(0, 1000, n).astype(np.int32)
    v = rng.standard_normal(n).astype(np.float32)
    w = rng.integers(-50, 50, n).astype(np.int64)
    uk, res = oracle.groupby_aggregate(k, {"v": v, "w": w, "c": v}, {"v": "sum", "w": "max", "c": "count"})
    df = pd.DataFrame({"k": k, "v": v.astype(np.float64), "w": w}).groupby("k").agg(
        v=("v", "sum"), w=("w", "max"), c=("v", "count"))
    assert np.array_equal(uk, df.index.values)
    np.testing.assert_allclose(res["v"], df["v"].values, rtol=1e-9, atol=1e-9)
    assert np.array_equal(res["w"], df["w"].values)
    assert np.array_equal(res["c"], df["c"].values)


def test_groupby_min_max_mean_vs_pandas():
    rng = np.random.default_rng(5)
    k = rng.integers(-20, 20, 5000)
    v = rng.standard_normal(5000)
    df = pd.DataFrame({"k": k, "v": v}).groupby("k")["v"]
    for fn in ("min", "max", "mean"):
        uk, res = oracle.groupby_aggregate(k, {"v": v}, {"v": fn})
        np.testing.assert_allclose(res["v"], getattr(df, fn)().values, rtol=1e-12)


def test_groupby_multi_column_vs_pandas():
    rng = np.random.default_rng(6)
    a, b = rng.integers(0, 5, 3000), rng.integers(0, 7, 3000)
    v = rng.integers(0, 100, 3000)
    ukd, res = oracle.groupby_aggregate_multi_column([a, b], ["a", "b"], {"v": v}, {"v": "sum"})
    df = pd.DataFrame({"a": a, "b": b, "v": v}).groupby(["a", "b"])["v"].sum().reset_index()
    assert ukd["a"].tolist() == df["a"].tolist() and ukd["b"].tolist() == df["b"].tolist()
    assert res["v"].tolist() == df["v"].tolist()


def test_groupby_empty_and_unknown_agg():
    uk, res = oracle.groupby_aggregate(np.zeros(0, np.int32), {"v": np.zeros(0)}, {"v": "sum"})
    assert uk.size == 0 and res["v"].size == 0
    with pytest.raises(ValueError):
        oracle.groupby_aggregate(np.array([1]), {"v": np.array([1.0])}, {"v": "median"})


# ---- join -------------------------------------------------------------------

def test_join_golden_vectors():
    for case in _golden()["join"]:
        lk, rk = np.asarray(case["left_keys"]), np.asarray(case["right_keys"])
        lv = {c: np.asarray(v) for c, v in case["left_values"].items()}
        rv = {c: np.asarray(v) for c, v in case["right_values"].items()}
        keys, out, _, _ = oracle.sort_merge_join(lk, lv, rk, rv, case["how"])
        assert keys.tolist() == case["expected_keys"], case["source"]
        for c, exp in case["expected"].items():
            got = np.asarray(out[c], dtype=np.float64)
            np.testing.assert_array_equal(np.isnan(got), np.isnan(np.asarray(exp, dtype=np.float64)))
            np.testing.assert_allclose(got[~np.isnan(got)], np.asarray(exp, dtype=np.float64)[~np.isnan(got)])


@pytest.mark.parametrize("seed", range(6))
def test_inner_join_vs_bruteforce(seed):
    rng = np.random.default_rng(seed)
    lk = rng.integers(0, 12, 60).astype(np.int64)
    rk = rng.integers(0, 12, 40).astype(np.int64)
    a = oracle.inner_join_pairs(lk, rk)
    b = oracle.inner_join_pairs_bruteforce(lk, rk)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_inner_join_count_vs_pandas():
    rng = np.random.default_rng(7)
    lk = rng.integers(0, 500, 20000)
    rk = rng.integers(0, 500, 3000)
    keys, li, ri = oracle.inner_join_pairs(lk, rk)
    m = pd.merge(pd.DataFrame({"k": lk, "l": np.arange(lk.size)}),
                 pd.DataFrame({"k": rk, "r": np.arange(rk.size)}), on="k")
    assert keys.size == len(m)
    got = sorted(zip(li.tolist(), ri.tolist()))
    assert got == sorted(zip(m["l"].tolist(), m["r"].tolist()))


# ---- hashing / partition ----------------------------------------------------

def test_fmix32_published_murmur3_vectors():
    # MurmurHash3_x86_32 known answers (published test vectors). The empty
    # input with seed s reduces to fmix32(s), pinning fmix32 directly.
    assert oracle.murmur3_x86_32(b"", 0) == 0
    assert oracle.murmur3_x86_32(b"", 1) == 0x514E28B7
    assert oracle.murmur3_x86_32(b"", 0xFFFFFFFF) == 0x81F16F39
    assert oracle.murmur3_x86_32(b"\x00\x00\x00\x00", 0) == 0x2362F9DE
    assert oracle.murmur3_x86_32(b"aaaa", 0x9747B28C) == 0x5A97808A
    assert oracle.murmur3_x86_32(b"Hello, world!", 1234) == 0xFAF6CDB3
    assert oracle.murmur3_x86_32(b"The quick brown fox jumps over the lazy dog", 0x9747B28C) == 0x2FA826CD
    assert int(oracle.fmix32(np.array([1], np.uint32))[0]) == 0x514E28B7


def test_hash_reference_properties():
    # test_distributed_multi_column.py:118-179
    h = oracle.hash_multi_columns([np.array([1, 1, 2, 2]), np.array([10, 20, 10, 20])])
    assert len(np.unique(h)) == 4
    h = oracle.hash_multi_columns([np.array([1, 2, 1, 2]), np.array([10, 20, 10, 20])])
    assert h[0] == h[2] and h[1] == h[3] and h[0] != h[1]
    i = np.arange(1000)
    h = oracle.hash_multi_columns([i % 100, i % 50])
    c = np.bincount(h % 8, minlength=8)
    assert (c > 1000 / 16).all() and (c < 1000 / 4).all()


def test_hash_single_key_low_bits():
    k = np.array([5, 5 + (1 << 32), -1, (1 << 32) - 1], dtype=np.int64)
    h = oracle.hash_multi_columns([k])
    assert h[0] == h[1] and h[2] == h[3]


@pytest.mark.parametrize("nb", [1, 3, 8, 256])
def test_hash_partition_vs_padded_bucketize(nb):
    rng = np.random.default_rng(nb)
    keys = rng.integers(-10**9, 10**9, 3000).astype(np.int64)
    perm, offs = oracle.hash_partition(keys, nb)
    targets = (oracle.hash_multi_columns([keys]) % np.uint32(nb)).astype(np.int64)
    padded = oracle.bucketize_padded(keys, targets, nb, -999999)
    valid = oracle.bucketize_padded(np.ones(keys.size, bool), targets, nb, False)
    compact = padded[valid]   # row-major: bucket order, then position
    assert np.array_equal(compact, keys[perm])
    assert np.array_equal(np.diff(offs), valid.sum(axis=1))
    # brute-force: each bucket holds its rows in input order
    for b in range(nb):
        rows = perm[offs[b]:offs[b + 1]]
        assert np.array_equal(rows, np.flatnonzero(targets == b))


def test_bucketize_padded_bruteforce():
    arr = np.array([10, 11, 12, 13, 14])
    t = np.array([1, 0, 1, 1, 0])
    out = oracle.bucketize_padded(arr, t, 2, -1)
    assert out.tolist() == [[11, 14, -1, -1, -1], [10, 12, 13, -1, -1]]


def test_compact_and_filter():
    m = np.array([True, False, True, True, False])
    assert oracle.compact_rows(m).tolist() == [0, 2, 3]
    out = oracle.filter_mask_compact({"a": np.arange(5)}, m)
    assert out["a"].tolist() == [0, 2, 3]


def test_pinned_by_covers_public_functions():
    for name in ("radix_sort", "groupby_aggregate", "inner_join_pairs", "fmix32",
                 "hash_partition", "sort_merge_join", "multi_column_lexsort"):
        assert name in oracle.PINNED_BY


def _golden_section(section):
    import json
    path = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")
    with open(path) as f:
        return json.load(f)[section]


def test_join_multi_column_golden_vectors():
    for case in _golden_section("join_multi"):
        lk = [np.asarray(k, dtype=np.int64) for k in case["left_keys"]]
        rk = [np.asarray(k, dtype=np.int64) for k in case["right_keys"]]
        lv = {c: np.asarray(v) for c, v in case["left_values"].items()}
        rv = {c: np.asarray(v) for c, v in case["right_values"].items()}
        names = case["left_key_names"]
        keys, vals = oracle.join_multi_column(lk, names, lv, rk, names, rv, case["how"])
        assert keys[names[0]].shape[0] == case["expected_rows"], case["source"]
        if "expected_sorted_tuples" in case:
            lc, rc = list(lv)[0], list(rv)[0]
            tup = sorted(zip(*(keys[n].tolist() for n in names), vals[f"left_{lc}"].tolist(),
                             vals[f"right_{rc}"].tolist()))
            assert tup == [tuple(t) for t in case["expected_sorted_tuples"]], case["source"]
        if "expected_right_value_isnan" in case:
            assert np.isnan(vals["right_value"]).tolist() == case["expected_right_value_isnan"]


def test_join_multi_column_vs_pandas_and_bruteforce():
    rng = np.random.default_rng(21)
    for _ in range(5):
        nl, nr = rng.integers(1, 60), rng.integers(1, 60)
        lk = [rng.integers(0, 3, nl), rng.integers(0, 4, nl)]
        rk = [rng.integers(0, 3, nr), rng.integers(0, 4, nr)]
        lv = {"x": rng.integers(0, 100, nl).astype(np.int32)}
        rv = {"y": rng.standard_normal(nr)}
        keys, vals = oracle.join_multi_column(lk, ["a", "b"], lv, rk, ["a", "b"], rv, "inner")
        L = pd.DataFrame({"a": lk[0], "b": lk[1], "x": lv["x"], "li": np.arange(nl)})
        R = pd.DataFrame({"a": rk[0], "b": rk[1], "y": rv["y"], "ri": np.arange(nr)})
        m = L.merge(R, on=["a", "b"], how="inner")   # keeps left order, then right order
        assert keys["a"].tolist() == m["a"].tolist() and keys["b"].tolist() == m["b"].tolist()
        assert vals["left_x"].tolist() == m["x"].tolist()
        assert vals["right_y"].tolist() == m["y"].tolist()
        for how in ("left", "right", "outer"):
            keys, vals = oracle.join_multi_column(lk, ["a", "b"], lv, rk, ["a", "b"], rv, how)
            pairs, lun, run = [], [], []
            for i in range(nl):
                for j in range(nr):
                    if lk[0][i] == rk[0][j] and lk[1][i] == rk[1][j]:
                        pairs.append((i, j))
            lm = {i for i, _ in pairs}
            rm = {j for _, j in pairs}
            exp_l = [i for i, _ in pairs] + ([i for i in range(nl) if i not in lm] if how != "right" else []) \
                + ([-1] * len([j for j in range(nr) if j not in rm]) if how in ("right", "outer") else [])
            n_rows = len(exp_l)
            assert keys["a"].shape[0] == n_rows


def test_gather_join_column_branches():
    a = np.array([5, 6, 7], dtype=np.int32)
    assert oracle.gather_join_column(a, np.array([2, 0])).dtype == np.int32
    g = oracle.gather_join_column(a, np.array([2, -1]))
    assert g.dtype == np.float32 and g[0] == 7 and np.isnan(g[1])
    s = oracle.gather_join_column(a, np.array([1, -1]), "string")
    assert s.tolist() == [6, -1]
    f = oracle.gather_join_column(np.array([1.5, 2.5]), np.array([-1, 1]))
    assert np.isnan(f[0]) and f[1] == 2.5


# ---- dictionary-encoded strings ---------------------------------------------------------

def _align(lc, lv, rc, rv):
    """Both code arrays in the merged sorted vocabulary (string_encoding.py:97-112)."""
    if lv == rv:
        return lc, rc, lv
    merged = tuple(sorted(set(lv) | set(rv)))
    pos = {v: i for i, v in enumerate(merged)}
    remap = lambda c, v: np.array([-1 if x < 0 else pos[v[x]] for x in c], dtype=np.int32)
    return remap(lc, lv), remap(rc, rv), merged


def _decode(codes, vocab):
    return [None if c < 0 else vocab[c] for c in codes]


def test_string_golden_vectors():
    for case in _golden_section("strings"):
        kind = case["kind"]
        if kind == "compare":
            codes, vocab = oracle.encode_string_array(np.array(case["values"], dtype=object))
            code = vocab.index(case["scalar"])
            rank = int(np.searchsorted(np.array(vocab, dtype=object), case["scalar"], "left"))
            assert (codes == code).tolist() == case["eq"]
            assert (codes != code).tolist() == case["ne"]
            assert (codes > rank).tolist() == case["gt"]
        elif kind == "sort":
            codes, vocab = oracle.encode_string_array(np.array(case["keys"], dtype=object))
            order = np.argsort(np.where(codes < 0, len(vocab), codes), kind="stable")
            assert _decode(codes[order], vocab) == case["sorted_keys"]
            assert np.asarray(case["values"])[order].tolist() == case["sorted_values"]
        elif kind == "groupby":
            codes, vocab = oracle.encode_string_array(np.array(case["keys"], dtype=object))
            g, r = oracle.groupby_encoded_strings(codes, {"v": np.array(case["values"])},
                                                  {"v": case["agg"]}, len(vocab))
            assert _decode(g, vocab) == case["groups"]
            assert r["v"].tolist() == case["result"]
        else:
            lc, lv = oracle.encode_string_array(np.array(case["left_keys"], dtype=object))
            rc, rv = oracle.encode_string_array(np.array(case["right_keys"], dtype=object))
            lc, rc, vocab = _align(lc, lv, rc, rv)
            codes, li, ri = oracle.string_join_plan(lc, rc, len(vocab), case["how"])
            keys = _decode(codes, vocab)
            lvals = np.asarray(case["left_values"], dtype=object if isinstance(case["left_values"][0], str) else None)
            rvals = np.asarray(case["right_values"], dtype=object if isinstance(case["right_values"][0], str) else None)

            def gather(vals, idx):
                if vals.dtype == object:
                    return [None if i < 0 else vals[i] for i in idx]
                out = oracle.gather_join_column(vals, idx)
                return [None if (isinstance(x, float) and np.isnan(x)) else x for x in out.tolist()]
            rows = list(zip(keys, gather(lvals, li), gather(rvals, ri)))
            if case["sorted_by"] == ["key", "left"]:
                rows.sort(key=lambda t: (t[0], t[1]))
            elif case["sorted_by"] == ["key"]:
                rows.sort(key=lambda t: t[0])
            assert [r[0] for r in rows] == case["keys"]
            assert [r[1] for r in rows] == case["left"]
            assert [r[2] for r in rows] == case["right"]


def test_encode_string_array_matches_np_unique():
    rng = np.random.default_rng(3)
    words = np.array([f"w{i:03d}" for i in rng.integers(0, 60, 500)], dtype=object)
    words[rng.random(500) < 0.1] = None
    codes, vocab = oracle.encode_string_array(words)
    nn = np.array([w is not None for w in words])
    assert list(vocab) == np.unique(words[nn].astype(str)).tolist()
    assert np.all(codes[~nn] == -1)
    assert [vocab[c] for c in codes[nn]] == words[nn].tolist()


def test_string_join_plan_vs_pandas_and_bruteforce():
    rng = np.random.default_rng(5)
    for trial in range(6):
        nv = 12
        lc = rng.integers(-1, nv, 80).astype(np.int32)
        if trial % 2:   # unique right keys (a permutation with a null)
            rc = rng.permutation(np.arange(-1, nv))[: nv - 2].astype(np.int32)
        else:
            rc = rng.integers(-1, nv, 40).astype(np.int32)
        codes, li, ri = oracle.string_join_plan(lc, rc, nv, "inner")
        dom = lambda c: np.where(c < 0, nv, c)
        m = pd.DataFrame({"k": dom(lc), "l": np.arange(lc.size)}).merge(
            pd.DataFrame({"k": dom(rc), "r": np.arange(rc.size)}), on="k")
        got = sorted(zip(dom(codes).tolist(), li.tolist(), ri.tolist()))
        assert got == sorted(zip(m.k.tolist(), m.l.tolist(), m.r.tolist()))
        if trial % 2:   # unique right: left row order, brute force lookup
            exp = [(i, j) for i in range(lc.size) for j in range(rc.size) if dom(lc)[i] == dom(rc)[j]]
            assert list(zip(li.tolist(), ri.tolist())) == exp
        else:           # many-to-many: (code, left, right) order
            assert list(zip(dom(codes).tolist(), li.tolist(), ri.tolist())) == got
        for how in ("left", "right", "outer"):
            c2, l2, r2 = oracle.string_join_plan(lc, rc, nv, how)
            inner = sorted((a, b) for a, b in zip(l2.tolist(), r2.tolist()) if a >= 0 and b >= 0)
            assert inner == sorted(zip(li.tolist(), ri.tolist()))
            if how in ("left", "outer"):
                assert set(range(lc.size)) <= set(l2.tolist())
            if how in ("right", "outer"):
                assert set(range(rc.size)) <= set(r2.tolist())


def test_groupby_encoded_strings_vs_pandas():
    rng = np.random.default_rng(8)
    codes = rng.integers(-1, 20, 1000).astype(np.int32)
    v = rng.standard_normal(1000)
    g, r = oracle.groupby_encoded_strings(codes, {"s": v, "c": v, "m": v},
                                          {"s": "sum", "c": "count", "m": "max"}, 25)
    p = pd.DataFrame({"k": np.where(codes < 0, np.nan, codes), "v": v}).groupby("k", dropna=True).v
    assert g.tolist() == p.sum().index.astype(int).tolist()
    np.testing.assert_allclose(r["s"], p.sum().values, rtol=1e-12)
    assert r["c"].tolist() == p.count().values.tolist() and r["c"].dtype == np.int32
    assert r["m"].tolist() == p.max().values.tolist()


def test_logical_pad_value_known_answers():
    """pa:150-160: string codes -1, floats NaN, integers -999999, bools False
    (a bool column is not an integer column: its pad is False, not a wrapped
    -999999), other dtypes 0."""
    lp = oracle.logical_pad_value
    assert lp("string", np.zeros(3, np.int32)) == -1
    assert lp("string", np.zeros(3, np.int32)).dtype == np.int32
    assert np.isnan(lp("float32", np.zeros(3, np.float32)))
    assert lp("float32", np.zeros(3, np.float32)).dtype == np.float32
    assert np.isnan(lp("float64", np.zeros(3, np.float64)))
    assert lp("int32", np.zeros(3, np.int32)) == -999999
    assert lp("int64", np.zeros(3, np.int64)) == -999999
    b = lp("bool", np.zeros(3, np.bool_))
    assert b is np.False_ or b == False  # noqa: E712
    assert lp("complex", np.zeros(3, np.complex64)) == 0
    # narrow and unsigned integers take -999999 through a wrapping cast
    assert lp("uint32", np.zeros(3, np.uint32)) == 2**32 - 999999
    assert lp("int16", np.zeros(3, np.int16)) == (-999999) % 2**16 - 2**16


def test_partition_pad_rule_matches_oracle_on_host():
    """The product's pad rule (evaluated on CPU tensors, no GPU) against the
    oracle's, for every dtype a frame column can have."""
    import torch
    from jaxframes_amd import partition
    cases = [(np.int8, "int8"), (np.int16, "int16"), (np.int32, "int32"), (np.int64, "int64"),
             (np.uint8, "uint8"), (np.float16, "float16"), (np.float32, "float32"),
             (np.float64, "float64"), (np.bool_, "bool"), (np.int32, "string")]
    for dt, name in cases:
        a = np.zeros(4, dt)
        got = partition.logical_pad_value(name, torch.from_numpy(a))
        exp = oracle.logical_pad_value(name, a)
        if np.dtype(dt).kind == "f":
            assert np.isnan(got) and np.isnan(exp), name
        else:
            assert got == exp, (name, got, exp)
