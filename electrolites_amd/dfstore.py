"""Out-of-HBM residency: the cderi in page-locked host memory, a prefix of
its auxiliary rows kept on the device, and the rest streamed per call.

Mirrors the reference's host-resident path for a tensor larger than the
device (electrolites/fastdf.py ``_resident_prefix`` :494-535 and ``_blocks``
:611-671; the build's host store, electrolites/_dfstore.py; the pinned
buffer pool, electrolites/_dfpin.py):

* the build writes each chunk of the cderi into a device staging buffer and
  a 2-D asynchronous copy on a store stream moves it into the ``(naux,
  npairs)`` host tensor, two staging buffers alternating so the next chunk's
  integrals run while the previous chunk crosses PCIe (:class:`HostStore`);
* a get_jk call keeps the first rows of a host tensor on the device, sized
  from the free device memory (``RESIDENT_FRAC``) and cached on the DF
  object, so the per-cycle PCIe traffic is only the remainder
  (:func:`resident_prefix`);
* the remainder is streamed in blocks of aux rows through two device
  buffers: the copy of block ``n + 1`` runs on an I/O stream while block
  ``n`` is contracted (:func:`blocks`).

Page-locked memory comes from ``hipHostMalloc`` (torch's ``pin_memory``),
and released buffers are kept in a pool keyed by size up to ``POOL_BYTES``
so that a second build of the same shape does not pay the allocation again
(the reference pins ordinary memory with ``cudaHostRegister`` instead; that
call is not used here).
"""
import os

import numpy as np

from . import _dev
from . import _lib

#: share of the free device memory a host-resident cderi's device prefix may
#: take; 0 streams every row (electrolites/fastdf.py:255, FASTDF_RESIDENT_FRAC)
RESIDENT_FRAC = float(os.environ.get('FASTDF_RESIDENT_FRAC', 0.5))
#: cap on the prefix rows, 0 for none (FASTDF_RESIDENT_MAX_ROWS, :252)
RESIDENT_MAX_ROWS = int(os.environ.get('FASTDF_RESIDENT_MAX_ROWS', 0))
#: bytes of one streamed block of aux rows (two are in flight)
STREAM_BLOCK_BYTES = int(float(os.environ.get('MDF_STREAM_BLOCK_BYTES', 1 << 30)))


def _mem_total():
    try:
        with open('/proc/meminfo') as f:
            for line in f:
                if line.startswith('MemTotal:'):
                    return int(line.split()[1]) * 1024
    except OSError:
        pass
    return 0


#: bytes of released page-locked buffers kept for reuse (MDF_PIN_POOL_BYTES;
#: default a quarter of the host memory)
POOL_BYTES = int(float(os.environ.get('MDF_PIN_POOL_BYTES', _mem_total() // 4)))

STATS = {'host_stores': 0, 'pinned_allocs': 0, 'pool_hits': 0, 'resident_rows': 0,
         'streamed_rows': 0, 'streamed_bytes': 0, 'blocks': 0}
_POOL = {}          # nbytes -> [flat uint8 pinned tensors]


def stats():
    out = dict(STATS)
    out['pool_bytes'] = sum(k * len(v) for k, v in _POOL.items())
    return out


def _torch_dtype(dtype):
    t = _dev.torch()
    return {np.dtype(np.float64): t.float64, np.dtype(np.float32): t.float32}.get(np.dtype(dtype), dtype) \
        if not isinstance(dtype, t.dtype) else dtype


def pinned_empty(shape, dtype, pin=True):
    """A page-locked host tensor of ``shape``, from the pool when a released
    buffer of the same size is there.  ``pin=False`` gives ordinary memory
    (host-only use and tests without a device)."""
    t = _dev.torch()
    dtype = _torch_dtype(dtype)
    nbytes = int(np.prod(shape)) * t.empty((), dtype=dtype).element_size()
    flat = None
    if pin and _POOL.get(nbytes):
        flat = _POOL[nbytes].pop()
        STATS['pool_hits'] += 1
    if flat is None:
        flat = t.empty(max(nbytes, 1), dtype=t.uint8, pin_memory=pin)
        if pin:
            STATS['pinned_allocs'] += 1
    return flat[:nbytes].view(dtype).view(*shape)


def release(host):
    """Give a host tensor from :func:`pinned_empty` back to the pool (kept
    while the pool stays under ``POOL_BYTES``, else freed).  The caller gives
    the memory up: a later :func:`pinned_empty` of the same size reuses it."""
    t = _dev.torch()
    if not isinstance(host, t.Tensor) or host.is_cuda or not host.is_pinned():
        return
    base = host.untyped_storage()
    nbytes = host.numel() * host.element_size()
    if sum(k * len(v) for k, v in _POOL.items()) + nbytes > POOL_BYTES:
        return
    flat = t.empty(0, dtype=t.uint8).set_(base, 0, (base.nbytes(),))
    _POOL.setdefault(flat.numel(), []).append(flat)


def is_host(cderi):
    t = _dev.torch()
    return isinstance(cderi, np.ndarray) or (isinstance(cderi, t.Tensor) and not cderi.is_cuda)


def as_tensor(cderi):
    """A host cderi as a torch CPU tensor (zero-copy for numpy)."""
    t = _dev.torch()
    return t.from_numpy(cderi) if isinstance(cderi, np.ndarray) else cderi


class HostStore:
    """The build's write-out into a page-locked ``(nrow, npairs)`` host
    tensor (electrolites/_dfstore.py, its 'stage' mode): each chunk is
    computed into one of two device staging buffers and copied out by
    :func:`mdf_copy_2d` on a store stream; the compute stream waits for a
    staging buffer's previous copy before it reuses it."""

    def __init__(self, nrow, npairs, dtype, widest):
        t = _dev.torch()
        self.nrow, self.npairs = nrow, npairs
        self.host = pinned_empty((nrow, npairs), dtype)
        self.stage = [t.empty(max(nrow * widest, 1), dtype=self.host.dtype, device=_dev.device())
                      for _ in range(2)]
        self.elem = self.host.element_size()
        self.stream = t.cuda.Stream()
        self.done = [None, None]
        self.k = 0
        STATS['host_stores'] += 1

    def slot(self, width):
        """A ``(nrow, width)`` staging buffer the current stream may write."""
        t = _dev.torch()
        i = self.k % 2
        if self.done[i] is not None:
            t.cuda.current_stream().wait_event(self.done[i])
        return self.stage[i][:self.nrow * width].view(self.nrow, width)

    def write(self, chunk, p0):
        """Copy the chunk just written on the current stream (from
        :meth:`slot`) into columns ``p0:p0 + width`` of the host tensor."""
        t = _dev.torch()
        width = chunk.shape[1]
        ev = t.cuda.Event()
        ev.record(t.cuda.current_stream())
        self.stream.wait_event(ev)
        e = self.elem
        _lib.call('mdf_copy_2d', self.host.data_ptr() + p0 * e, self.npairs * e, _dev.ptr(chunk), width * e,
                  width * e, self.nrow, self.stream.cuda_stream)
        done = t.cuda.Event()
        done.record(self.stream)
        self.done[self.k % 2] = done
        self.k += 1

    def finish(self):
        self.stream.synchronize()
        self.stage = None
        return self.host


def to_host(cderi):
    """A device cderi copied into a page-locked host tensor (row-contiguous,
    the same values); the caller drops the device tensor afterwards."""
    t = _dev.torch()
    host = pinned_empty(tuple(cderi.shape), cderi.dtype)
    if cderi.numel():
        e = cderi.element_size()
        _lib.call('mdf_copy_2d', host.data_ptr(), cderi.shape[1] * e, _dev.ptr(cderi), cderi.stride(0) * e,
                  cderi.shape[1] * e, cderi.shape[0], _dev.stream_ptr())
        t.cuda.current_stream().synchronize()
    return host


def block_rows(npairs, elem, naux):
    """Aux rows per streamed block."""
    return max(1, min(naux, STREAM_BLOCK_BYTES // max(npairs * elem, 1)))


def resident_rows(naux, npairs, elem, free_bytes, reserve_bytes=0):
    """How many leading rows the device prefix holds: ``RESIDENT_FRAC`` of
    the free device memory less what the call still allocates
    (electrolites/fastdf.py:515-523)."""
    if RESIDENT_FRAC <= 0:
        return 0
    nbytes = int(free_bytes * RESIDENT_FRAC) - int(reserve_bytes)
    rows = min(naux, max(0, nbytes // max(npairs * elem, 1)))
    if RESIDENT_MAX_ROWS:
        rows = min(rows, RESIDENT_MAX_ROWS)
    return int(rows)


def resident_prefix(dfobj, host, reserve_bytes=0):
    """The device copy of the host cderi's leading rows, built once per
    (DF object, host tensor) and reused by later calls; None when no row
    fits."""
    t = _dev.torch()
    cached = getattr(dfobj, '_mdf_resident', None)
    if cached is not None and cached[0] is host:
        buf = cached[1]
    else:
        naux, npairs = host.shape
        elem = host.element_size()
        free = t.cuda.mem_get_info()[0]
        rows = resident_rows(naux, npairs, elem, free, reserve_bytes)
        buf = None
        if rows > 0:
            buf = t.empty((rows, npairs), dtype=host.dtype, device=_dev.device())
            _lib.call('mdf_copy_2d', _dev.ptr(buf), npairs * elem, host.data_ptr(), host.stride(0) * elem,
                      npairs * elem, rows, _dev.stream_ptr())
        try:
            dfobj._mdf_resident = (host, buf)
        except AttributeError:
            pass
    STATS['resident_rows'] = 0 if buf is None else int(buf.shape[0])
    STATS['streamed_rows'] = int(host.shape[0]) - STATS['resident_rows']
    return buf


def blocks(host, resident, blk_rows):
    """Yield ``(block, L0, nL)``: device ``(nL, npairs)`` views of the rows
    ``L0 .. L0 + nL`` of ``host``, first those the resident prefix holds (in
    place), then the rest through two device buffers, the next block's copy
    on an I/O stream overlapping the current block's consumers.  A yielded
    block is valid until the generator is resumed."""
    t = _dev.torch()
    naux, npairs = host.shape
    elem = host.element_size()
    start = 0 if resident is None else int(resident.shape[0])
    for L0 in range(0, start, blk_rows):
        nL = min(blk_rows, start - L0)
        STATS['blocks'] += 1
        yield resident[L0:L0 + nL], L0, nL
    if start >= naux:
        return
    nb = min(blk_rows, naux - start)
    bufs = [t.empty((nb, npairs), dtype=host.dtype, device=_dev.device()) for _ in range(2)]
    comp = t.cuda.current_stream()
    io = t.cuda.Stream()
    ready, free = [None, None], [None, None]
    starts = list(range(start, naux, nb))
    base = host.data_ptr()
    ld = host.stride(0) * elem

    def fetch(i, L0):
        nL = min(nb, naux - L0)
        if free[i] is not None:
            io.wait_event(free[i])
        _lib.call('mdf_copy_2d', _dev.ptr(bufs[i]), npairs * elem, base + L0 * ld, ld, npairs * elem, nL,
                  io.cuda_stream)
        ev = t.cuda.Event()
        ev.record(io)
        ready[i] = ev
        STATS['streamed_bytes'] += nL * npairs * elem

    try:
        fetch(0, starts[0])
        for k, L0 in enumerate(starts):
            i = k % 2
            if k + 1 < len(starts):
                fetch(1 - i, starts[k + 1])
            comp.wait_event(ready[i])
            nL = min(nb, naux - L0)
            STATS['blocks'] += 1
            yield bufs[i][:nL], L0, nL
            ev = t.cuda.Event()
            ev.record(comp)
            free[i] = ev
    finally:
        # a prefetch may still be writing a buffer (also when the consumer
        # stops early): let it land before the buffers go back to the
        # allocator, which hands them out again in the compute stream's order
        io.synchronize()
