"""Device memory and stream plumbing (PyTorch-ROCm tensors as buffers).

Torch is used only to own device memory and to name the current HIP stream;
all arithmetic of the path is in libmi355df.
"""
import ctypes

import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError('electrolites_amd: no ROCm GPU visible; the density-'
                           'fitted path runs only through its HIP kernels')


def device():
    return torch().device('cuda', torch().cuda.current_device())


def stream_ptr():
    """The current HIP stream of torch, as a ``void*`` for the C ABI."""
    return ctypes.c_void_p(torch().cuda.current_stream().cuda_stream)


def ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def is_device_array(x):
    t = torch()
    return isinstance(x, t.Tensor) and x.is_cuda


def to_device(x, dtype=None):
    """numpy / torch -> contiguous device tensor (float64 by default)."""
    t = torch()
    dtype = dtype or t.float64
    if isinstance(x, t.Tensor):
        return x.to(device=device(), dtype=dtype).contiguous()
    return t.from_numpy(np.ascontiguousarray(x)).to(device=device(),
                                                    dtype=dtype).contiguous()


class Workspace:
    """A grow-only scratch buffer per device, reused across calls.  With
    ``grow`` > 1 a buffer that must grow takes at least ``grow`` times its
    old size, so a caller whose needs creep upward (a serving loop whose
    batch gains members) reallocates a few times instead of at every step."""

    def __init__(self, grow=1.0):
        self.grow = float(grow)
        self._buf = {}

    def get(self, nbytes):
        t = torch()
        dev = t.cuda.current_device()
        buf = self._buf.get(dev)
        if buf is None or buf.numel() < nbytes:
            old = 0 if buf is None else buf.numel()
            self._buf.pop(dev, None)
            buf = None
            buf = t.empty(max(int(nbytes), int(old * self.grow), 256), dtype=t.uint8, device=device())
            self._buf[dev] = buf
        return buf

    def release(self):
        self._buf.clear()


WORKSPACE = Workspace()
