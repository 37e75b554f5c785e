"""The bandwidth probes' data movement: the copy is exact and the scatter copy
writes every key (and payload) to the run position its header states, so the
ceilings the kernels are read against measure the intended pattern."""

import ctypes

import numpy as np
import pytest
import torch

from jaxframes_amd import _lib

pytestmark = pytest.mark.gpu

TILE = 16384


def _stream():
    return ctypes.c_void_p(_lib.stream_handle(torch.device("cuda", 0)))


def test_copy_probe_is_exact():
    n = 1 << 20
    src = torch.randint(0, 1 << 30, (n,), dtype=torch.int32, device="cuda")
    dst = torch.zeros_like(src)
    for nt in (0, 1):
        dst.zero_()
        _lib.call("jfx_bandwidth_probe", 1, src.data_ptr(), dst.data_ptr(), n * 4, 64, 2, nt,
                  _stream())
        torch.cuda.synchronize()
        assert torch.equal(src, dst)


def test_write_probe_values():
    n16 = 1 << 16
    dst = torch.zeros(n16 * 4, dtype=torch.int32, device="cuda")
    for nt in (0, 1):
        dst.zero_()
        _lib.call("jfx_bandwidth_probe", 4, dst.data_ptr(), dst.data_ptr(), n16 * 16, 64, 4, nt,
                  _stream())
        torch.cuda.synchronize()
        i = np.arange(n16, dtype=np.int64)[:, None] + np.arange(4)[None, :]
        assert np.array_equal(dst.cpu().numpy().reshape(-1, 4).astype(np.int64), i)


@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("runlen", [16, 64, 256])
def test_scatter_probe_positions(mode, runlen):
    tiles = 3
    n = tiles * TILE
    rng = np.random.default_rng(runlen + mode)
    keys = rng.integers(0, 2**63, n, dtype=np.uint64)
    pay = rng.integers(0, 2**32, n, dtype=np.uint32)
    src = np.zeros(n * 3, dtype=np.uint32)
    src[: 2 * n] = keys.view(np.uint32)
    src[2 * n:] = pay
    dsrc = torch.from_numpy(src.view(np.int32)).cuda()
    ddst = torch.zeros_like(dsrc)
    nbytes = n * (12 if mode == 2 else 8)
    _lib.call("jfx_bandwidth_probe", mode, dsrc.data_ptr(), ddst.data_ptr(), nbytes, 0, runlen,
              1, _stream())
    torch.cuda.synchronize()
    out = ddst.cpu().numpy().view(np.uint32)
    j = np.arange(TILE, dtype=np.int64)
    t = np.arange(tiles, dtype=np.int64)[:, None]
    pos = (j // runlen) * (tiles * runlen) + t * runlen + (j % runlen)   # [tiles, TILE]
    exp_k = np.zeros(n, dtype=np.uint64)
    exp_k[pos.ravel()] = keys
    assert np.array_equal(out[: 2 * n].view(np.uint64), exp_k)
    if mode == 2:
        exp_p = np.zeros(n, dtype=np.uint32)
        exp_p[pos.ravel()] = pay
        assert np.array_equal(out[2 * n:], exp_p)
    else:
        assert not out[2 * n:].any()


def test_scatter_probe_rejects_bad_sizes():
    buf = torch.zeros(TILE * 3, dtype=torch.int32, device="cuda")
    with pytest.raises(_lib.JfxError):
        _lib.call("jfx_bandwidth_probe", 2, buf.data_ptr(), buf.data_ptr(), TILE * 12 - 12, 0, 64,
                  0, _stream())
    with pytest.raises(_lib.JfxError):
        _lib.call("jfx_bandwidth_probe", 3, buf.data_ptr(), buf.data_ptr(), TILE * 8, 0, 48, 0,
                  _stream())
