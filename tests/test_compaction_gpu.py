"""GPU parity of row filters and compaction against the oracle (exact)."""

import numpy as np
import pytest
import torch

import oracle
from jaxframes_amd import compaction, partition
from tests._util import cuda

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 63, 64, 4095, 4096, 4097, 100003, 3_000_017])
@pytest.mark.parametrize("p", [0.0, 0.03, 0.5, 1.0])
def test_compact_rows_matches_flatnonzero(n, p):
    rng = np.random.default_rng(n + int(p * 100))
    m = rng.random(n) < p
    got = compaction.compact_rows(cuda(m.astype(np.uint8)))
    assert np.array_equal(got.cpu().numpy(), oracle.compact_rows(m))


@pytest.mark.parametrize("dt", [np.float32, np.float64, np.int32, np.int64, np.uint32, np.uint64])
@pytest.mark.parametrize("op", [">", ">=", "<", "<=", "==", "!="])
def test_filter_rows_all_ops(dt, op):
    rng = np.random.default_rng(3)
    n = 50000
    if np.dtype(dt).kind == "f":
        col = rng.standard_normal(n).astype(dt)
        col[::97] = np.nan
        thr = dt(0.25)
        col[::101] = thr
    else:
        col = rng.integers(0, 50, n).astype(dt)
        thr = dt(20)
    exp = {">": col > thr, ">=": col >= thr, "<": col < thr, "<=": col <= thr, "==": col == thr,
           "!=": col != thr}[op]
    got = compaction.filter_rows(cuda(col), op, thr.item())
    assert np.array_equal(got.cpu().numpy(), np.flatnonzero(exp))


def test_compact_repartitioned_rows():
    rng = np.random.default_rng(4)
    n = 20000
    k = rng.integers(0, 100, n).astype(np.int64)
    v = rng.standard_normal(n).astype(np.float32)
    m = rng.random(n) < 0.6
    ks, vs = compaction.compact_repartitioned_rows([cuda(k)], {"v": cuda(v)}, cuda(m))
    pos = oracle.compact_rows(m)
    assert np.array_equal(ks[0].cpu().numpy(), k[pos])
    assert np.array_equal(vs["v"].cpu().numpy(), v[pos])
    # every row valid: the inputs come back unchanged (pa:286-288)
    kk = cuda(k)
    ks2, _ = compaction.compact_repartitioned_rows([kk], {}, torch.ones(n, dtype=torch.bool, device="cuda"))
    assert ks2[0] is kk


def test_bucketize_padded_matches_reference_layout():
    rng = np.random.default_rng(5)
    n, P = 3000, 8
    k = rng.integers(-10**9, 10**9, n).astype(np.int64)
    v = rng.standard_normal(n).astype(np.float32)
    r = partition.hash_partition(cuda(k), P)
    got = partition.bucketize_padded(cuda(v), r["offsets"], r["rows"], float("nan")).cpu().numpy()
    targets = oracle.hash_multi_columns([k]) % P
    exp = oracle.bucketize_padded(v, targets, P, np.nan)
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    assert np.array_equal(got[~np.isnan(got)], exp[~np.isnan(exp)])


def test_collective_repartition_single_device_identity():
    k = cuda(np.arange(10, dtype=np.int64))
    ks, vs, mask = partition.collective_hash_repartition([k], {"v": k.float()})
    assert torch.equal(ks[0], k) and bool(mask.all()) and mask.shape[0] == 10
