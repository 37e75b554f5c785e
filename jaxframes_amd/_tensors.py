"""Column plumbing: host/DLPack inputs to device tensors, dtype codes."""

from __future__ import annotations

import numpy as np
import torch

from . import _lib

_KEY_CODE = {
    torch.int32: _lib.JFX_I32,
    torch.int64: _lib.JFX_I64,
    torch.float32: _lib.JFX_F32,
    torch.float64: _lib.JFX_F64,
    torch.uint32: _lib.JFX_U32,
    torch.uint64: _lib.JFX_U64,
}

# dtypes widened (order- and value-preserving) to a 32-bit key type
_WIDEN = {
    torch.int8: torch.int32, torch.int16: torch.int32, torch.uint8: torch.int32,
    torch.bool: torch.int32, torch.uint16: torch.int32,
    torch.float16: torch.float32, torch.bfloat16: torch.float32,
}

_BITS_VIEW = {4: torch.int32, 8: torch.int64}


def device() -> torch.device:
    _lib.require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def to_device(x, dev: torch.device | None = None) -> torch.Tensor:
    """numpy / list / DLPack / torch -> 1-D torch tensor on the GPU."""
    if isinstance(x, torch.Tensor) and x.is_cuda:
        # already resident: on the current device, or on dev when given
        if x.get_device() == (torch.cuda.current_device() if dev is None else dev.index):
            return x if x.is_contiguous() else x.contiguous()
    dev = dev or device()
    if isinstance(x, torch.Tensor):
        t = x
    elif isinstance(x, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(x))
    elif hasattr(x, "__dlpack__"):
        t = torch.from_dlpack(x)
    else:
        t = torch.as_tensor(np.asarray(x))
    if t.device != dev:
        t = t.to(dev)
    return t.contiguous()


def key_code(t: torch.Tensor) -> int:
    if t.dtype not in _KEY_CODE:
        raise TypeError(f"unsupported key dtype {t.dtype}")
    return _KEY_CODE[t.dtype]


def widen_key(t: torch.Tensor):
    """Return (key tensor of a supported type, original dtype)."""
    if t.dtype in _KEY_CODE:
        return t, t.dtype
    if t.dtype in _WIDEN:
        return t.to(_WIDEN[t.dtype]), t.dtype
    raise TypeError(f"unsupported key dtype {t.dtype}")


def bits_view(t: torch.Tensor) -> torch.Tensor:
    """Same bytes as a signed integer tensor of the same width (4 or 8)."""
    return t.view(_BITS_VIEW[t.element_size()])


def to_numpy(t: torch.Tensor) -> np.ndarray:
    if t.dtype == torch.uint32:
        return t.view(torch.int32).cpu().numpy().view(np.uint32)
    if t.dtype == torch.uint64:
        return t.view(torch.int64).cpu().numpy().view(np.uint64)
    return t.cpu().numpy()
