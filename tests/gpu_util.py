"""Device-buffer plumbing for the GPU parity tests (torch is only the allocator)."""
import numpy as np
import torch

import sealhip

_inited = False


def hip():
    global _inited
    if not _inited:
        if not torch.cuda.is_available():
            raise RuntimeError("GPU tests need a visible gfx950 device")
        torch.cuda.init()
        sealhip.init(0)
        _inited = True
    return sealhip


def to_dev(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return torch.from_numpy(a.view(np.int64).copy()).cuda()


def from_dev(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


def empty_dev(shape):
    return torch.empty(shape, dtype=torch.int64, device="cuda")


def stream():
    return torch.cuda.current_stream().cuda_stream
