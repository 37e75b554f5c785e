"""GPU tests of the request loop: every result is the same alone and in the
concurrent mix (keys, counts, join pairs exact; sums exact here as well, the
tolerance stated being 2e-7 * sum|x|), and equals the plain per-query path
and the oracle."""

import numpy as np
import pytest
import torch

import oracle
from jaxframes_amd import serving
from tests._util import cuda

pytestmark = pytest.mark.gpu

SUM_NORM_TOL = 2e-7


@pytest.fixture(scope="module")
def store():
    rng = np.random.default_rng(0)
    st = serving.FrameStore(8 << 30)
    for i, (n, card) in enumerate([(300_000, 50), (1_000_000, 20_000), (2_000_000, 300)]):
        k = rng.integers(-card, card, n).astype(np.int64)
        v = rng.standard_normal(n).astype(np.float32)
        st.load(f"f{i}", cuda(k), cuda(v))
    return st


def _queries(store, nq, seed):
    rng = np.random.default_rng(seed)
    qs = []
    for i in range(nq):
        fr = f"f{rng.integers(0, 3)}"
        if rng.random() < 0.7:
            qs.append(serving.FilterGroupBy(fr, ">", float(rng.uniform(-1, 1))))
        else:
            keys = store[fr].keys
            lo, hi = int(keys.min()), int(keys.max())
            nd = int(rng.integers(1000, 20000))
            dk = rng.choice(np.arange(lo - 100, hi + 100), size=min(nd, hi - lo + 200), replace=False)
            qs.append(serving.Join(fr, cuda(dk.astype(np.int64)), cuda(rng.standard_normal(dk.size).astype(np.float32))))
    return qs


def _same(a, b, exact_sums=False):
    (ka, va), (kb, vb) = a, b
    assert torch.equal(ka, kb)
    for c in va:
        if c in ("count",) or exact_sums:
            assert torch.equal(va[c], vb[c]), c
        else:
            x, y = va[c].double(), vb[c].double()
            assert torch.all((x - y).abs() <= 1e-5 * (1 + y.abs())), c


def test_same_alone_and_in_the_mix(store):
    qs = _queries(store, 40, 1)
    alone = [serving.execute(store, q) for q in qs]
    torch.cuda.synchronize()
    srv = serving.Server(store, slots=4)
    arrivals = serving.poisson_arrivals(len(qs), qps=2000.0, seed=2)
    mixed = serving.replay(srv, qs, arrivals)
    srv.close()
    assert srv.peak_running >= 2          # requests really overlapped
    for q, a, (m, lat) in zip(qs, alone, mixed):
        assert lat > 0
        _same(a, m, exact_sums=True)


def test_served_equals_plain_path(store):
    for q in _queries(store, 12, 3):
        _same(serving.execute(store, q), serving.run_plain(store, q))


def test_filter_groupby_vs_oracle(store):
    fr = store["f1"]
    k = fr.keys.cpu().numpy()
    v = fr.values.cpu().numpy()
    for t in (-0.5, 0.0, 0.7):
        keys, res = serving.execute(store, serving.FilterGroupBy("f1", ">", t))
        m = v > t
        ek, ec = oracle.groupby_aggregate(k[m], {"v": v[m]}, {"v": "count"})
        _, es = oracle.groupby_aggregate(k[m], {"v": v[m]}, {"v": "sum"})
        _, ea = oracle.groupby_aggregate(k[m], {"v": np.abs(v[m]).astype(np.float64)}, {"v": "sum"})
        assert np.array_equal(keys.cpu().numpy(), ek)
        assert np.array_equal(res["count"].cpu().numpy(), ec["v"])
        err = np.abs(res["sum"].cpu().numpy().astype(np.float64) - es["v"])
        assert np.all(err <= SUM_NORM_TOL * np.maximum(ea["v"], 1e-30))


def test_join_vs_oracle(store):
    rng = np.random.default_rng(4)
    fr = store["f0"]
    dk = rng.choice(np.arange(-60, 60), size=80, replace=False).astype(np.int64)
    dv = rng.standard_normal(80).astype(np.float32)
    keys, out = serving.execute(store, serving.Join("f0", cuda(dk), cuda(dv)))
    ek, eli, eri = oracle.inner_join_pairs(fr.keys.cpu().numpy(), dk)
    assert np.array_equal(keys.cpu().numpy(), ek)
    assert np.array_equal(out["left_value"].cpu().numpy(), fr.values.cpu().numpy()[eli])
    assert np.array_equal(out["right_value"].cpu().numpy(), dv[eri])


def test_budget_admission_on_gpu(store):
    q = serving.FilterGroupBy("f2", ">", 0.0)
    fp = serving.footprint(store, q)
    srv = serving.Server(store, slots=4, hbm_budget_bytes=int(fp * 1.5))   # one at a time
    futs = [srv.submit(q) for _ in range(4)]
    res = [f.result()[0] for f in futs]
    srv.close()
    assert srv.peak_running == 1
    for r in res[1:]:
        _same(res[0], r, exact_sums=True)


@pytest.mark.parametrize("op", [">", ">=", "<", "<=", "==", "!="])
def test_indexed_filter_ops_nan_and_signed_zero(op):
    rng = np.random.default_rng(7)
    n = 50000
    k = rng.integers(0, 40, n).astype(np.int64)
    pool = np.array([np.nan, -0.0, 0.0, 1.5, -2.0, np.inf, 0.25], dtype=np.float32)
    v = np.where(rng.random(n) < 0.5, rng.choice(pool, n), rng.standard_normal(n).astype(np.float32))
    v = v.astype(np.float32)
    st = serving.FrameStore(1 << 30)
    st.load("z", cuda(k), cuda(v))
    assert st["z"].indexed
    cmp = {">": np.greater, ">=": np.greater_equal, "<": np.less, "<=": np.less_equal,
           "==": np.equal, "!=": np.not_equal}[op]
    for t in (0.0, -0.0, 1.5, 0.25, float("nan")):
        keys, res = serving.execute(st, serving.FilterGroupBy("z", op, t))
        m = cmp(v, np.float32(t))
        ek, ec = oracle.groupby_aggregate(k[m], {"v": v[m]}, {"v": "count"})
        _, es = oracle.groupby_aggregate(k[m], {"v": v[m]}, {"v": "sum"})
        assert np.array_equal(keys.cpu().numpy(), ek), (op, t)
        assert np.array_equal(res["count"].cpu().numpy(), ec["v"]), (op, t)
        g = res["sum"].cpu().numpy().astype(np.float64)
        e = es["v"]
        assert np.array_equal(np.isnan(g), np.isnan(e)), (op, t)
        assert np.array_equal(g[np.isinf(e)], e[np.isinf(e)]), (op, t)   # +-inf sums exact
        fin = np.isfinite(e)
        assert np.all(np.isfinite(g[fin])), (op, t)
        _, ea = oracle.groupby_aggregate(k[m], {"v": np.abs(v[m]).astype(np.float64)}, {"v": "sum"})
        assert np.all(np.abs(g[fin] - e[fin]) <= SUM_NORM_TOL * np.maximum(ea["v"][fin], 1e-30)), (op, t)


def test_batched_filter_queries_equal_one_at_a_time(store):
    rng = np.random.default_rng(9)
    st = serving.FrameStore(1 << 30)
    st.frames.update(store.frames)
    k = rng.integers(0, 500, 100_000).astype(np.int64)
    st.load("d64", cuda(k), cuda(rng.standard_normal(k.size)))          # float64 values
    ops = [">", ">=", "<", "<=", "==", "!="]
    qs = [serving.FilterGroupBy(f"f{i % 3}" if i % 4 else "d64", ops[i % 6],
                                float(rng.uniform(-2, 2)) if i % 5 else 10.0)   # 10.0: nothing passes >
          for i in range(23)]
    batch = serving.execute_filter_batch(st, qs)
    for q, b in zip(qs, batch):
        _same(serving.execute(st, q), b, exact_sums=True)


def test_server_batches_queued_filter_queries(store):
    qs = [serving.FilterGroupBy(f"f{i % 3}", ">", 0.1 * (i % 7)) for i in range(30)]
    alone = [serving.execute(store, q) for q in qs]
    srv = serving.Server(store, slots=1)
    with srv.cv:                         # queue everything before the worker can admit
        futs = [None] * len(qs)
        for i, q in enumerate(qs):
            srv._rid += 1
            r = serving.Request(q, 0.0, rid=srv._rid, footprint=serving.footprint(store, q))
            srv.pending.append(r)
            futs[i] = r.future
        srv.cv.notify_all()
    res = [f.result()[0] for f in futs]
    srv.close()
    assert max(srv.batches) > 1
    assert srv.admitted_order == sorted(srv.admitted_order)
    for a, r in zip(alone, res):
        _same(a, r, exact_sums=True)


def _served_groups_ref(fr, op, t):
    """jfx_served_groups on every group, then the groups with a passing row."""
    import ctypes
    from jaxframes_amd import _lib
    from jaxframes_amd.compaction import OPS
    G = fr.ngroups
    dev = fr.keys.device
    cnt = torch.empty(G, dtype=torch.int64, device=dev)
    sm = torch.empty(G, dtype=fr.kv_values.dtype, device=dev)
    mean = torch.empty(G, dtype=fr.kv_values.dtype, device=dev)
    mask = torch.empty(G, dtype=torch.uint8, device=dev)
    vt = _lib.JFX_F32 if fr.kv_values.dtype == torch.float32 else _lib.JFX_F64
    val = (ctypes.c_float if vt == _lib.JFX_F32 else ctypes.c_double)(t)
    _lib.call("jfx_served_groups", vt, _lib.ptr(fr.kv_values), _lib.ptr(fr.prefix),
              _lib.ptr(fr.group_starts), _lib.ptr(fr.group_nan_start), _lib.ptr(fr.group_ninf_end),
              _lib.ptr(fr.group_pinf_start), G, OPS[op], ctypes.byref(val), _lib.ptr(cnt),
              _lib.ptr(sm), _lib.ptr(mean), _lib.ptr(mask), 0)
    keep = mask.bool()
    return fr.group_keys[keep], cnt[keep], sm[keep], mean[keep]


def test_served_batch_abi_with_empty_queries(store):
    """jfx_served_batch equals jfx_served_groups + compaction per query, with
    queries of no groups first, in the middle and last."""
    import ctypes
    from jaxframes_amd import _lib
    from jaxframes_amd.compaction import OPS
    spec = [(None, ">", 0.0), ("f0", ">", 0.3), (None, "<", 0.0), ("f1", "<=", -0.2),
            ("f2", "!=", 0.5), ("f1", "==", 0.0), (None, ">", 1.0)]
    Q = len(spec)
    table = (serving.ServedQuery * Q)()
    tot = 0
    for i, (name, op, t) in enumerate(spec):
        if name is None:
            table[i] = serving.ServedQuery(0, 0, 0, 0, 0, 0, 0, 0, OPS[op], t)
        else:
            fr = store[name]
            table[i] = serving.ServedQuery(*fr.index_ptrs, fr.ngroups, OPS[op], t)
            tot += fr.ngroups
    nb = ctypes.c_uint64(0)
    _lib.call("jfx_served_batch_workspace_size", Q, tot, ctypes.byref(nb))
    ws = torch.empty(nb.value, dtype=torch.uint8, device="cuda")
    keys = torch.empty(tot, dtype=torch.int64, device="cuda")
    cnt = torch.empty(tot, dtype=torch.int64, device="cuda")
    sm = torch.empty(tot, dtype=torch.float32, device="cuda")
    mean = torch.empty(tot, dtype=torch.float32, device="cuda")
    starts = (ctypes.c_uint64 * (Q + 1))()
    _lib.call("jfx_served_batch", _lib.JFX_F32, 8, table, Q, _lib.ptr(keys), _lib.ptr(cnt),
              _lib.ptr(sm), _lib.ptr(mean), starts, _lib.ptr(ws), ws.numel(), 0)
    st = list(starts)
    assert st[0] == 0 and all(a <= b for a, b in zip(st, st[1:]))
    for i, (name, op, t) in enumerate(spec):
        a, b = st[i], st[i + 1]
        if name is None:
            assert a == b
            continue
        ek, ec, es, em = _served_groups_ref(store[name], op, t)
        assert torch.equal(keys[a:b], ek)
        assert torch.equal(cnt[a:b], ec)
        assert torch.equal(sm[a:b], es)
        assert torch.equal(mean[a:b], em)
    assert st[Q] == sum(int(_served_groups_ref(store[n], o, t)[0].shape[0])
                        for n, o, t in spec if n is not None)


@pytest.mark.parametrize("kdt,vdt,ddt", [(np.int64, np.float32, np.float32),
                                          (np.int32, np.float64, np.int32),
                                          (np.int64, np.int64, np.float64)])
def test_join_by_groups_duplicates_and_types(kdt, vdt, ddt):
    """Dimension keys repeated (several right rows per group, in row order),
    absent from the frame, and of every integer key width; payloads promoted
    as the reference does (ints to floats)."""
    rng = np.random.default_rng(11)
    n = 120_000
    k = rng.integers(-300, 300, n).astype(kdt)
    v = (rng.standard_normal(n) * 100).astype(vdt)
    st = serving.FrameStore(1 << 30)
    st.load("j", cuda(k), cuda(v))
    dk = rng.integers(-400, 400, 5000).astype(kdt)           # duplicates, misses
    dv = (rng.standard_normal(5000) * 100).astype(ddt)
    keys, out = serving.execute(st, serving.Join("j", cuda(dk), cuda(dv)))
    ek, eli, eri = oracle.inner_join_pairs(k, dk)
    assert np.array_equal(keys.cpu().numpy(), ek)
    lv = out["left_value"].cpu().numpy()
    rv = out["right_value"].cpu().numpy()
    exp_l = v[eli] if np.issubdtype(vdt, np.floating) else v[eli].astype(np.float64)
    exp_r = dv[eri] if np.issubdtype(ddt, np.floating) else dv[eri].astype(np.float32)
    assert lv.dtype == exp_l.dtype and rv.dtype == exp_r.dtype
    assert np.array_equal(lv, exp_l)
    assert np.array_equal(rv, exp_r)
    # equals the plain sort-merge join
    pk, pout = serving.run_plain(st, serving.Join("j", cuda(dk), cuda(dv)))
    assert torch.equal(keys, pk)
    assert torch.equal(out["left_value"], pout["left_value"])
    assert torch.equal(out["right_value"], pout["right_value"])
    # no dimension rows, and no matches
    keys, out = serving.execute(st, serving.Join("j", cuda(dk[:0]), cuda(dv[:0])))
    assert keys.shape[0] == 0 and out["left_value"].shape[0] == 0
    far = (np.arange(50) + 10_000).astype(kdt)
    keys, out = serving.execute(st, serving.Join("j", cuda(far), cuda(far.astype(np.float32))))
    assert keys.shape[0] == 0


def test_served_sums_independent_of_earlier_groups():
    """A group of huge values ahead of small groups in key order: each group's
    sum comes from its own prefix sums, so the small groups are exact (a
    prefix over the whole frame would carry the rounding of 1e23 into them)."""
    keys = np.concatenate([np.zeros(1000, np.int64), np.array([1, 1, 2, 3, 3, 3], np.int64)])
    vals = np.concatenate([np.full(1000, 1e20, np.float32),
                           np.array([1.5, 2.25, -0.75, 1e-3, 2e-3, -5.0], np.float32)])
    perm = np.random.default_rng(1).permutation(keys.size)
    keys, vals = keys[perm], vals[perm]
    st = serving.FrameStore(1 << 30)
    st.load("big", cuda(keys), cuda(vals))
    for op, t in ((">", 0.0), (">=", -1.0), ("!=", 7.0), ("<", 1e19)):
        q = serving.FilterGroupBy("big", op, t)
        uk, res = serving.execute(st, q)
        pk, pres = serving.run_plain(st, q)
        mask = {">": vals > t, ">=": vals >= t, "!=": vals != t, "<": vals < t}[op]
        ek, eres = oracle.groupby_aggregate(keys[mask], {"s": vals[mask]}, {"s": "sum"})
        assert np.array_equal(uk.cpu().numpy(), ek) and torch.equal(uk, pk)
        got = res["sum"].cpu().numpy().astype(np.float64)
        small = ek != 0
        # exact for the small groups (their float32 sums are exact in float64)
        assert np.array_equal(got[small], eres["s"][small].astype(np.float32).astype(np.float64)), (op, got, eres["s"])
        assert torch.equal(res["sum"], pres["sum"]) or np.allclose(
            got, pres["sum"].cpu().numpy(), rtol=1e-6, atol=0)
