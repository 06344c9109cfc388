"""``DF.loop``: the block iterator every other consumer of the cderi uses,
for the tensors this package writes (FP32 storage, host residency).

Mirrors electrolites/fastdffp32.py ``loop`` (:72-141) and the widening of
electrolites/_dffp32.py ``BlockWidener`` (:166-196): GPU4PySCF's ``DF.loop``
yields, per block of auxiliary rows, the unpacked symmetric matrices and the
packed rows, both FP64.  A float32 tensor has no fall-through (a float64
reader would misread it), so the iterator widens each block into one FP64
buffer allocated before the loop (``mdf_widen_f32``) and unpacks from that
(``mdf_unpack_pairs``, the ``fill_symmetric`` of the packed rows: both
triangles written, zeros where a pair is not stored).  A host-resident
tensor is streamed through :func:`electrolites_amd.dfstore.blocks` (the next
block's copy overlapping the current block's consumers).

The patched ``DF.loop`` hands a tensor GPU4PySCF built itself to the
function it replaced, unchanged (the reference delegates its float64 tensors
the same way); this package's own :class:`DF` uses this iterator for every
tensor.

Yielded tensors are views of buffers reused by the next block: a consumer
that keeps one past the next iteration copies it.
"""
import numpy as np

from . import _dev
from . import _lib
from . import dfstore
from . import fastdf

STATS = {'loops': 0, 'blocks': 0, 'widened': 0}
#: GPU4PySCF's DF.loop, kept by patch() for float64 tensors
_orig_loop = None
#: bytes of the unpacked block the default block size aims for
UNPACK_BYTES = 1 << 30


def stats():
    return dict(STATS)


def _nao(dfobj, npairs):
    nao = getattr(dfobj, 'nao', None)
    if nao is None:
        mol = getattr(dfobj, 'mol', None)
        nao = getattr(mol, 'nao', None)
        nao = nao() if callable(nao) else nao
    if nao is None:
        nao = int((np.sqrt(8 * npairs + 1) - 1) // 2)
    return int(nao)


def loop(dfobj, blksize=None, unpack=True):
    """Yield ``(unpacked, packed)`` per block of aux rows: ``packed`` the
    ``(nL, npairs)`` FP64 rows, ``unpacked`` the ``(nL, nao, nao)`` FP64
    symmetric matrices or None when ``unpack`` is false."""
    t = _dev.torch()
    _dev.require_gpu()
    if getattr(dfobj, '_cderi', None) is None:
        dfobj.build()
    cderi = fastdf._cderi_of(dfobj)
    if dfstore.is_host(cderi):
        cderi = dfstore.as_tensor(cderi)
        host = True
    else:
        host = False
    if cderi.dtype not in (t.float64, t.float32) or cderi.dim() != 2 or cderi.stride(1) != 1:
        raise ValueError('DF.loop: expected a pair-contiguous (naux, npairs) float64 or float32 cderi')
    naux, npairs = cderi.shape
    nao = _nao(dfobj, npairs)
    lay = fastdf._layout(dfobj, cderi, nao)
    if blksize is None:
        blksize = UNPACK_BYTES // max(8 * nao * nao, 1) if unpack else dfstore.block_rows(npairs, 8, naux)
    blksize = max(1, min(int(blksize), naux))
    STATS['loops'] += 1
    fp32 = cderi.dtype == t.float32
    wide = t.empty((blksize, npairs), dtype=t.float64, device=_dev.device()) if fp32 else None
    dense = t.empty((blksize, nao, nao), dtype=t.float64, device=_dev.device()) if unpack else None

    def emit(blk, nL):
        STATS['blocks'] += 1
        if fp32:
            STATS['widened'] += 1
            _lib.call('mdf_widen_f32', _dev.ptr(blk), blk.stride(0), nL, npairs, _dev.ptr(wide), npairs,
                      _dev.stream_ptr())
            blk = wide[:nL]
        if not unpack:
            return None, blk
        out = dense[:nL]
        _lib.call('mdf_unpack_pairs', _dev.ptr(blk), nL, blk.stride(0), _dev.ptr(lay.addr), npairs, nao,
                  _dev.ptr(out), nao * nao, _dev.stream_ptr())
        return out, blk

    if not host:
        for L0 in range(0, naux, blksize):
            nL = min(blksize, naux - L0)
            yield emit(cderi[L0:L0 + nL], nL)
        return
    for blk, L0, nL in dfstore.blocks(cderi, None, blksize):
        yield emit(blk, nL)


def patched_loop(self, blksize=None, unpack=True):
    """GPU4PySCF's ``DF.loop`` over the tensors this package's build writes
    (torch tensors: FP64 or FP32, device or page-locked host); a tensor
    GPU4PySCF built itself goes to the function this replaced, unchanged
    (electrolites/fastdffp32.py:84-90 delegates the same way)."""
    t = _dev.torch()
    store = getattr(self, '_cderi', None)
    cderi = None if store is None else fastdf._cderi_of(self, missing_ok=True)
    if _orig_loop is not None and cderi is not None and not isinstance(cderi, t.Tensor):
        yield from _orig_loop(self, blksize, unpack)
        return
    yield from loop(self, blksize, unpack)
