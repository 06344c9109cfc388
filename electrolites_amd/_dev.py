"""Device memory and stream plumbing (PyTorch-ROCm tensors as buffers).

Torch is used only to own device memory and to name the current HIP stream;
all arithmetic of the path is in libmi355df.
"""
import ctypes
import threading

import numpy as np

_torch = None

#: Held around every call into the device dense-solver library (the solver
#: thread's batched syevd, torch's SVD of a non-symmetric density).  Several
#: host threads issuing hipSOLVER ``syevd`` at once (a serving loop absorbing
#: requests from a thread pool, one HIP stream each) crashed inside
#: ``hipsolverDsyevd`` with a SIGSEGV; with the lock no two threads are ever
#: inside the solver library together.
SOLVER_LOCK = threading.RLock()


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


def is_foreign_device_array(x):
    """A device array of another library (CuPy, Numba, ...): anything that is
    not a torch tensor and exposes the CUDA array interface."""
    return not isinstance(x, torch().Tensor) and hasattr(x, '__cuda_array_interface__')


def is_device_array(x):
    t = torch()
    return (isinstance(x, t.Tensor) and x.is_cuda) or is_foreign_device_array(x)


def as_torch(x):
    """A torch view of ``x`` without a copy where ``x`` is a device array:
    torch tensors as they are, objects exposing ``__cuda_array_interface__``
    (CuPy's arrays, GPU4PySCF's ``_cderi`` and densities) wrapped in place,
    strides kept.  Anything else is returned unchanged."""
    t = torch()
    if is_foreign_device_array(x):
        cai = x.__cuda_array_interface__
        if cai.get('mask') is not None:
            raise ValueError('masked device arrays are not supported')
        return t.as_tensor(x, device=device())
    return x


def to_host(x):
    """numpy array of a numpy/torch array or of a CUDA-array-interface
    object (copied to the host)."""
    x = as_torch(x)
    if isinstance(x, torch().Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def shape_of(x):
    """Shape of a numpy/torch array or of a CUDA-array-interface object."""
    if is_foreign_device_array(x) and not hasattr(x, 'shape'):
        return tuple(x.__cuda_array_interface__['shape'])
    return tuple(x.shape)


def to_device(x, dtype=None):
    """numpy / torch / CUDA-array-interface object -> contiguous device tensor
    (float64 by default)."""
    t = torch()
    dtype = dtype or t.float64
    x = as_torch(x)
    if isinstance(x, t.Tensor):
        return x.to(device=device(), dtype=dtype).contiguous()
    return t.from_numpy(np.ascontiguousarray(x)).to(device=device(),
                                                    dtype=dtype).contiguous()


def like_input(x, ref):
    """Return the device tensor ``x`` in the family of the input ``ref``:
    a CuPy array for a CuPy input (zero copy), the torch tensor itself for a
    torch input or any other object exposing the CUDA array interface (a
    torch tensor exposes that interface too), and a numpy array for host
    input."""
    if isinstance(ref, torch().Tensor):
        return x
    if is_foreign_device_array(ref):
        mod = type(ref).__module__.split('.')[0]
        if mod == 'cupy':
            import cupy
            return cupy.asarray(x)
        return x
    return x.cpu().numpy()


def is_oom(exc):
    """Whether ``exc`` reports device memory exhaustion (torch's allocator,
    the HIP runtime through this package's library, or CuPy's pool)."""
    t = torch()
    if isinstance(exc, getattr(t.cuda, 'OutOfMemoryError', ())) or isinstance(exc, MemoryError):
        return True
    name = type(exc).__name__
    msg = str(exc)
    return name == 'OutOfMemoryError' or 'out of memory' in msg.lower() or \
        'hipErrorOutOfMemory' in msg or 'cudaErrorMemoryAllocation' in msg


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
