"""GPU parity of row hashing and the local hash partition (bit-exact)."""

import numpy as np
import pytest
import torch

import oracle
from jaxframes_amd import partition
from tests._util import cuda

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 1000, 8191, 8193, 300001])
@pytest.mark.parametrize("nb", [1, 7, 8, 256, 2048])
@pytest.mark.parametrize("dt", [np.int64, np.int32, np.uint64, np.float32])
def test_partition_matches_oracle(n, nb, dt):
    rng = np.random.default_rng(n * 31 + nb)
    if np.dtype(dt).kind == "f":
        keys = rng.standard_normal(n).astype(dt)
    else:
        info = np.iinfo(dt)
        keys = rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
    r = partition.hash_partition(cuda(keys), nb, return_hashes=True)
    perm, offs = oracle.hash_partition(keys, nb)
    assert np.array_equal(r["offsets"].cpu().numpy(), offs)
    assert np.array_equal(r["rows"].cpu().numpy(), perm)
    got_k = r["keys"].cpu().numpy()
    assert np.array_equal(got_k.view(np.uint8), keys[perm].view(np.uint8))
    h = r["hashes"].view(torch.int32).cpu().numpy().view(np.uint32)
    assert np.array_equal(h, oracle.hash_multi_columns([keys]))


def test_partition_config5_shape_small():
    # the partition config (int64 keys, 256 buckets) at a size the oracle finishes fast
    rng = np.random.default_rng(5)
    keys = rng.integers(-2**63, 2**63 - 1, 2_000_000, dtype=np.int64)
    r = partition.hash_partition(cuda(keys), 256)
    perm, offs = oracle.hash_partition(keys, 256)
    assert np.array_equal(r["offsets"].cpu().numpy(), offs)
    assert np.array_equal(r["rows"].cpu().numpy(), perm)


def test_partition_high_bits_ignored():
    # uint32(key) keeps the low 32 bits: keys equal mod 2^32 share a bucket
    base = np.arange(1000, dtype=np.int64)
    keys = np.concatenate([base, base + (7 << 32), base - (3 << 40)])
    r = partition.hash_partition(cuda(keys), 64, return_hashes=True)
    h = r["hashes"].view(torch.int32).cpu().numpy().view(np.uint32)
    assert np.array_equal(h[:1000], h[1000:2000]) and np.array_equal(h[:1000], h[2000:])


def test_multi_key_hash_matches_oracle_and_reference_properties():
    rng = np.random.default_rng(3)
    a = rng.integers(-10**12, 10**12, 50000).astype(np.int64)
    b = rng.integers(-1000, 1000, 50000).astype(np.int32)
    c = rng.standard_normal(50000).astype(np.float32)
    got = partition.hash_multi_columns([cuda(a), cuda(b), cuda(c)]).view(torch.int32).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, oracle.hash_multi_columns([a, b, c]))
    # test_distributed_multi_column.py:118-179: uniqueness, consistency, 8-way balance
    h = partition.hash_multi_columns([cuda(np.array([1, 1, 2, 2])), cuda(np.array([10, 20, 10, 20]))])
    assert len(torch.unique(h.view(torch.int32))) == 4
    h = partition.hash_multi_columns([cuda(np.array([1, 2, 1, 2])), cuda(np.array([10, 20, 10, 20]))]).view(torch.int32)
    assert h[0] == h[2] and h[1] == h[3] and h[0] != h[1]
    i = np.arange(1000)
    h = partition.hash_multi_columns([cuda(i % 100), cuda(i % 50)]).view(torch.int32).cpu().numpy().view(np.uint32)
    counts = np.bincount(h % 8, minlength=8)
    assert (counts > 1000 / 8 / 2).all() and (counts < 1000 / 8 * 2).all()


def test_multi_key_partition():
    rng = np.random.default_rng(4)
    a = rng.integers(0, 100, 70000).astype(np.int64)
    b = rng.integers(0, 7, 70000).astype(np.int32)
    r = partition.hash_partition([cuda(a), cuda(b)], 16)
    h = oracle.hash_multi_columns([a, b])
    perm, offs = oracle.hash_partition(None, 16, hashes=h)
    assert np.array_equal(r["rows"].cpu().numpy(), perm)
    assert np.array_equal(r["offsets"].cpu().numpy(), offs)


def test_partition_empty():
    r = partition.hash_partition(torch.empty(0, dtype=torch.int64, device="cuda"), 8)
    assert r["offsets"].cpu().numpy().tolist() == [0] * 9


PAD_CASES = [
    # column dtype, dtype name as the frame records it, expected pad (pa:150-160)
    (np.int32, "int32", -999999),
    (np.int64, "int64", -999999),
    (np.float32, "float32", np.nan),
    (np.float64, "float64", np.nan),
    (np.bool_, "bool", False),
    (np.int32, "string", -1),     # dictionary codes of a string column
]


@pytest.mark.parametrize("dt,name,pad", PAD_CASES, ids=[c[1] + "-" + np.dtype(c[0]).name for c in PAD_CASES])
def test_bucketize_padded_default_pad_matches_oracle(dt, name, pad):
    """The padded [P, n] layout with the per-dtype pad rule (pad_value=None)
    against oracle.bucketize_padded with oracle.logical_pad_value."""
    rng = np.random.default_rng(len(name))
    n, nb = 5000, 8
    keys = rng.integers(-10**9, 10**9, n).astype(np.int64)
    if dt == np.bool_:
        col = rng.random(n) < 0.5
    elif np.dtype(dt).kind == "f":
        col = rng.standard_normal(n).astype(dt)
    else:
        col = rng.integers(0, 500, n).astype(dt)
    r = partition.hash_partition(cuda(keys), nb)
    got = partition.bucketize_padded(cuda(col), r["offsets"], r["rows"], dtype_name=name)
    targets = oracle.hash_multi_columns([keys]) % np.uint32(nb)
    epad = oracle.logical_pad_value(name, col)
    exp = oracle.bucketize_padded(col, targets, nb, epad)
    g = got.cpu().numpy()
    assert g.dtype == exp.dtype and g.shape == (nb, n)
    assert np.array_equal(g, exp, equal_nan=np.dtype(dt).kind == "f")
    # the pad fills exactly the slots past each bucket's rows
    counts = np.diff(r["offsets"].cpu().numpy())
    filler = g[nb - 1, counts[nb - 1]:]
    if np.dtype(dt).kind == "f":
        assert np.isnan(pad) and np.isnan(filler).all()
    else:
        assert (filler == pad).all()


@pytest.mark.parametrize("dt", [np.int64, np.int32])
def test_partition_unaligned_view(dt):
    # a view one element into the buffer: the histogram pass takes its
    # scalar-load path (16-byte loads need 16-byte aligned chunks)
    rng = np.random.default_rng(11)
    full = rng.integers(-2**30, 2**30, 200_001).astype(dt)
    keys = full[1:]
    t = cuda(full)[1:]
    assert t.data_ptr() % 16 != 0
    r = partition.hash_partition(t, 256, return_hashes=True)
    perm, offs = oracle.hash_partition(keys, 256)
    assert np.array_equal(r["offsets"].cpu().numpy(), offs)
    assert np.array_equal(r["rows"].cpu().numpy(), perm)
    h = r["hashes"].view(torch.int32).cpu().numpy().view(np.uint32)
    assert np.array_equal(h, oracle.hash_multi_columns([keys]))
