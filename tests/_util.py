"""Test helpers: numpy arrays to GPU tensors (including unsigned types)."""
import numpy as np
import torch

_UVIEW = {np.dtype(np.uint32): (np.int32, torch.uint32), np.dtype(np.uint64): (np.int64, torch.uint64)}


def cuda(a):
    a = np.ascontiguousarray(a)
    if a.dtype in _UVIEW:
        sv, tv = _UVIEW[a.dtype]
        return torch.from_numpy(a.view(sv)).cuda().view(tv)
    return torch.from_numpy(a).cuda()
