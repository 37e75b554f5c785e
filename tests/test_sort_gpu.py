"""GPU parity of the onesweep radix sort against the oracle (bit-exact)."""

import numpy as np
import pytest
import torch

import oracle
from jaxframes_amd import sorting
from jaxframes_amd._tensors import to_numpy
from tests._util import cuda

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 7, 1000, 8191, 8192, 8193, 65537, 300001]


def _keys(rng, dtype, n, kind):
    if kind == "uniform":
        if dtype == np.float32:
            return rng.standard_normal(n).astype(np.float32) * 1e3
        if dtype == np.float64:
            return rng.standard_normal(n) * 1e6
        info = np.iinfo(dtype)
        return rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
    if kind == "small":   # < 2048 distinct values -> one range-reduced pass
        base = 12345 if dtype != np.uint32 else 7
        return (rng.integers(0, 1000, n) + base).astype(dtype)
    if kind == "tiny":    # < 256 distinct values
        return (rng.integers(-3, 100, n)).astype(dtype) if np.dtype(dtype).kind != "u" \
            else rng.integers(0, 100, n).astype(dtype)
    if kind == "equal":
        return np.full(n, 42, dtype=dtype)
    if kind == "wide_gaps":  # varying high and low bytes only
        hi = rng.integers(0, 4, n).astype(np.uint64) << np.uint64(40)
        return (hi + rng.integers(0, 50, n).astype(np.uint64)).astype(dtype)
    raise ValueError(kind)


def _check(keys_np, ascending=True, payload=None, want_index=True):
    ek, _, eperm = oracle.radix_sort(keys_np, None, ascending)
    t = cuda(keys_np)
    if payload is None:
        sk, _, idx = sorting.radix_sort(t, None, ascending, return_index=True)
        got_k = to_numpy(sk)
        assert np.array_equal(got_k.view(np.uint8), ek.view(np.uint8)), "keys differ"
        assert np.array_equal(idx.cpu().numpy(), eperm), "row order differs"
    else:
        pv = torch.from_numpy(payload).cuda()
        sk, sv = sorting.radix_sort(t, pv, ascending)
        assert np.array_equal(to_numpy(sk).view(np.uint8), ek.view(np.uint8))
        assert np.array_equal(sv.cpu().numpy().view(np.uint8), payload[eperm].view(np.uint8))


@pytest.mark.parametrize("dtype", [np.int32, np.uint32, np.float32, np.int64, np.uint64, np.float64])
@pytest.mark.parametrize("kind", ["uniform", "small", "tiny", "equal"])
@pytest.mark.parametrize("n", [1, 1000, 8193, 300001])
def test_sort_matches_oracle(dtype, kind, n):
    rng = np.random.default_rng(n * 7 + len(kind))
    _check(_keys(rng, dtype, n, kind))


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("ascending", [True, False])
def test_sort_sizes_and_direction(n, ascending):
    rng = np.random.default_rng(n)
    _check(rng.integers(-1000000, 1000000, n).astype(np.int64), ascending)
    _check(rng.integers(0, 50, n).astype(np.int32), ascending)


@pytest.mark.parametrize("pdt", [np.float32, np.int64, np.float64, np.int16])
def test_sort_payload_types(pdt):
    rng = np.random.default_rng(3)
    n = 70001
    keys = rng.integers(0, 1 << 40, n).astype(np.uint64)
    _check(keys, True, payload=(rng.standard_normal(n) * 100).astype(pdt))
    _check(keys, False, payload=(rng.standard_normal(n) * 100).astype(pdt))


def test_sort_digit_skip_and_wide_layout():
    rng = np.random.default_rng(5)
    k = _keys(rng, np.uint64, 200003, "wide_gaps")
    _check(k)
    for db in (8, 11):
        t = cuda(k)
        sk, _, idx = sorting.radix_sort(t, None, True, return_index=True, digit_bits=db)
        ek, _, eperm = oracle.radix_sort(k)
        assert np.array_equal(to_numpy(sk), ek)
        assert np.array_equal(idx.cpu().numpy(), eperm)


def test_sort_float_specials():
    vals = np.array([np.nan, -0.0, 0.0, np.inf, -np.inf, 1.5, -1.5, np.nan, 0.0, -0.0,
                     3.4e38, -3.4e38, 1e-45, -1e-45], dtype=np.float32)
    rng = np.random.default_rng(9)
    k = rng.choice(vals, 50000)
    _check(k)
    _check(k, ascending=False)
    _check(k.astype(np.float64))


def test_sort_empty():
    t = torch.empty(0, dtype=torch.int32, device="cuda")
    sk, sv = sorting.radix_sort(t)
    assert sk.numel() == 0


def test_sort_reference_vectors():
    # test_parallel_algorithms.py:29-84
    rs = np.random.RandomState(42)
    data = rs.randint(0, 100, size=1000)
    sk, _ = sorting.radix_sort(torch.from_numpy(data).cuda())
    assert np.array_equal(sk.cpu().numpy(), np.sort(data))
    k = torch.tensor([5, 2, 8, 1, 9, 3], device="cuda")
    v = torch.tensor([50, 20, 80, 10, 90, 30], device="cuda")
    sk, sv = sorting.radix_sort(k, v)
    assert sk.tolist() == [1, 2, 3, 5, 8, 9] and sv.tolist() == [10, 20, 30, 50, 80, 90]
    sk, _ = sorting.radix_sort(torch.tensor([3, 1, 4, 1, 5, 9, 2, 6], device="cuda"), ascending=False)
    assert sk.tolist() == [9, 6, 5, 4, 3, 2, 1, 1]
    sk, _ = sorting.radix_sort(torch.tensor([3, -1, 4, -5, 0, 2, -3], device="cuda"))
    assert sk.tolist() == [-5, -3, -1, 0, 2, 3, 4]
    sk, _ = sorting.radix_sort(torch.tensor([3.14, 2.71, 1.41, 0.577, 2.71, 3.14], device="cuda"))
    np.testing.assert_allclose(sk.cpu().numpy(), [0.577, 1.41, 2.71, 2.71, 3.14, 3.14], rtol=1e-6)


def test_sort_no_device_error():
    from jaxframes_amd import _lib
    assert _lib.device_error() == 0
