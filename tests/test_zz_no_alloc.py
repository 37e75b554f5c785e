"""GPU test of jfx_groupby_noalloc, the allocation-free one-pass groupby,
against jfx_groupby_ex on the same inputs, both called through the C ABI.

Each case makes two allocation-free calls in a row, the second with a smaller
input, and compares every output with jfx_groupby_ex's bit for bit: at
n >= 2^16 both run the one-pass kernels with the same launch shapes, so keys,
counts, min / max and even the float sums and means must be identical. The
scratch array keeps the previous call's contents, so the second call also
shows that nothing depends on it starting cleared. Error paths: a key span
beyond the one-pass tables, too small outputs, float keys and n >= 2^30.
"""

import ctypes
import zlib

import numpy as np
import pytest
import torch

from jaxframes_amd import _lib
from tests._util import cuda

pytestmark = pytest.mark.gpu

JFX_ERR_UNSUPPORTED = 3
JFX_ERR_CAPACITY = 6

_CODE = {np.dtype(np.uint32): _lib.JFX_U32, np.dtype(np.int32): _lib.JFX_I32,
         np.dtype(np.float32): _lib.JFX_F32, np.dtype(np.uint64): _lib.JFX_U64,
         np.dtype(np.int64): _lib.JFX_I64, np.dtype(np.float64): _lib.JFX_F64}
_OUTS = ("sum", "count", "min", "max", "mean")


def _outputs(keys, vals, cap, aggs):
    kt = keys.dtype
    vt = vals.dtype
    mean_dt = torch.float32 if vals.element_size() == 4 else torch.float64
    dt = {"sum": vt, "count": torch.int64, "min": vt, "max": vt, "mean": mean_dt}
    outs = {a: None for a in _OUTS}
    # cap entries of each type, prefilled with a byte pattern
    for a in aggs:
        outs[a] = torch.full((cap * torch.empty(0, dtype=dt[a]).element_size(),), 0x5A,
                             dtype=torch.uint8, device="cuda").view(dt[a])
    kout = torch.full((cap * keys.element_size(),), 0x5A, dtype=torch.uint8,
                      device="cuda").view(kt)
    return kout, outs


def _call(entry, keys, vals, kcode, vcode, cap, aggs):
    """(status, G, out_keys, {agg: out}) of one call of entry on the current stream."""
    L = _lib.lib()
    n = keys.shape[0]
    kout, outs = _outputs(keys, vals, cap, aggs)
    g = ctypes.c_uint64(12345)
    stream = ctypes.c_void_p(_lib.stream_handle())
    ptrs = [None if outs[a] is None else outs[a].data_ptr() for a in _OUTS]
    if entry == "ex":
        nb = ctypes.c_uint64(0)
        _lib.call("jfx_groupby_workspace_size", kcode, vcode, n, ctypes.byref(nb))
        ws = torch.empty(max(nb.value, 1), dtype=torch.uint8, device="cuda")
        st = L.jfx_groupby_ex(kcode, keys.data_ptr(), vcode, vals.data_ptr(), n, kout.data_ptr(),
                              *ptrs, ctypes.byref(g), cap, 0, ws.data_ptr(), ws.numel(), stream)
    else:
        st = L.jfx_groupby_noalloc(kcode, keys.data_ptr(), vcode, vals.data_ptr(), n,
                                   kout.data_ptr(), *ptrs, ctypes.byref(g), cap, stream)
    torch.cuda.synchronize()
    return st, g.value, kout, outs


def _bits(t):
    """A tensor's bytes, so that NaN payloads and -0.0 compare exactly."""
    return t.contiguous().view(torch.uint8).cpu().numpy()


CASES = [
    # name, key dtype, key base, key span, value dtype, aggregations
    ("i32-1000-f32-all", np.int32, 0, 1000, np.float32, _OUTS),
    ("i32-neg-i32-all", np.int32, -700, 900, np.int32, _OUTS),
    ("i64-1500-f64-sum", np.int64, 10**12, 1500, np.float64, ("sum", "count", "mean")),
    ("u32-3keys-i64-all", np.uint32, 2**32 - 10, 3, np.int64, _OUTS),
    ("u64-top-f32-sum", np.uint64, 2**64 - 2000, 1900, np.float32, ("sum", "mean")),
    ("i64-600-f64-all", np.int64, -(2**40), 600, np.float64, _OUTS),
]


def _inputs(rng, n, kdt, base, span, vdt):
    off = rng.integers(0, span, n).astype(np.uint64)
    if np.dtype(kdt) == np.uint64:
        keys = (np.uint64(base) + off).astype(np.uint64)
    else:
        keys = (np.int64(base) + off.astype(np.int64)).astype(kdt)
    if np.dtype(vdt).kind == "f":
        vals = rng.standard_normal(n).astype(vdt)
        vals[rng.integers(0, n, 5)] = np.nan
    else:
        vals = rng.integers(-10**6, 10**6, n).astype(vdt)
    return keys, vals


@pytest.mark.parametrize("name,kdt,base,span,vdt,aggs", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_groupby_ex_twice(name, kdt, base, span, vdt, aggs):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    kcode, vcode = _CODE[np.dtype(kdt)], _CODE[np.dtype(vdt)]
    for n in (3_000_017, 200_003):   # the second call is smaller
        keys, vals = _inputs(rng, n, kdt, base, span, vdt)
        tk, tv = cuda(keys), cuda(vals)
        st_a, g_a, k_a, o_a = _call("noalloc", tk, tv, kcode, vcode, 4096, aggs)
        st_b, g_b, k_b, o_b = _call("ex", tk, tv, kcode, vcode, 4096, aggs)
        assert st_a == 0, _lib.lib().jfx_last_error()
        assert st_b == 0, _lib.lib().jfx_last_error()
        assert g_a == g_b == len(np.unique(keys)), (g_a, g_b)
        assert np.array_equal(_bits(k_a[:g_a]), _bits(k_b[:g_b]))
        assert np.array_equal(k_a[:g_a].cpu().numpy(), np.unique(keys))
        for a in aggs:
            assert np.array_equal(_bits(o_a[a][:g_a]), _bits(o_b[a][:g_b])), (name, n, a)
        assert _lib.device_error() == 0


def test_noalloc_wide_span_is_rejected_then_recovers():
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 1000, 1_000_003).astype(np.int32)
    keys[777_777] = 5000        # one key far outside the others' range
    vals = rng.standard_normal(keys.shape[0]).astype(np.float32)
    tk, tv = cuda(keys), cuda(vals)
    st, g, _, _ = _call("noalloc", tk, tv, _lib.JFX_I32, _lib.JFX_F32, 8192, ("sum",))
    assert st == JFX_ERR_UNSUPPORTED and g == 0
    assert b"2048" in _lib.lib().jfx_last_error()
    # a smaller in-range input right after the miss (the stop word of the
    # missed call is still in the scratch array)
    small = keys[:300_001].copy()
    small[small == 5000] = 1
    ts, tv2 = cuda(small), cuda(vals[:300_001].copy())
    st_a, g_a, k_a, o_a = _call("noalloc", ts, tv2, _lib.JFX_I32, _lib.JFX_F32, 4096, _OUTS)
    st_b, g_b, k_b, o_b = _call("ex", ts, tv2, _lib.JFX_I32, _lib.JFX_F32, 4096, _OUTS)
    assert st_a == 0 and st_b == 0 and g_a == g_b
    assert np.array_equal(_bits(k_a[:g_a]), _bits(k_b[:g_b]))
    for a in _OUTS:
        assert np.array_equal(_bits(o_a[a][:g_a]), _bits(o_b[a][:g_b])), a


def test_noalloc_capacity_and_argument_errors():
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 1000, 500_000).astype(np.int32)
    vals = rng.standard_normal(keys.shape[0]).astype(np.float32)
    tk, tv = cuda(keys), cuda(vals)
    st, g, _, _ = _call("noalloc", tk, tv, _lib.JFX_I32, _lib.JFX_F32, 100, ("sum",))
    assert st == JFX_ERR_CAPACITY and g == len(np.unique(keys))
    # retry with the capacity returned
    st, g2, k, o = _call("noalloc", tk, tv, _lib.JFX_I32, _lib.JFX_F32, g, ("sum",))
    assert st == 0 and g2 == g
    assert np.array_equal(k[:g].cpu().numpy(), np.unique(keys))
    L = _lib.lib()
    out = ctypes.c_uint64(0)
    s = ctypes.c_void_p(_lib.stream_handle())
    fk = cuda(vals)
    assert L.jfx_groupby_noalloc(_lib.JFX_F32, fk.data_ptr(), _lib.JFX_F32, tv.data_ptr(),
                                 keys.shape[0], None, tv.data_ptr(), None, None, None, None,
                                 ctypes.byref(out), 10, s) == JFX_ERR_UNSUPPORTED
    # rejected on the host before any launch
    assert L.jfx_groupby_noalloc(_lib.JFX_I32, tk.data_ptr(), _lib.JFX_F32, tv.data_ptr(),
                                 1 << 30, None, tv.data_ptr(), None, None, None, None,
                                 ctypes.byref(out), 10, s) == JFX_ERR_UNSUPPORTED
    assert L.jfx_groupby_noalloc(_lib.JFX_I32, tk.data_ptr(), _lib.JFX_F32, tv.data_ptr(),
                                 0, None, tv.data_ptr(), None, None, None, None,
                                 ctypes.byref(out), 10, s) == 0 and out.value == 0
    assert _lib.device_error() == 0
