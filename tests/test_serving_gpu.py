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
