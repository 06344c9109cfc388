"""Coulomb (J) without a stored cderi: the three-centre RI-J build.

Mirrors electrolites/fastdfj3c.py ``get_j`` (:726-784), the J a pure DF
functional (and the Coulomb half of RIJONX) runs on, where GPU4PySCF's
``df_jk.get_j`` never builds a cderi:

    rho_P = sum_ij (ij|P) D_ij           contract_dm      (:388-437)
    c     = j2c^-1 rho                   two triangular solves (:787-798)
    J_ij  = sum_P (ij|P) c_P             contract_auxvec  (:439-530)

Here the three-centre integrals are evaluated chunk by chunk of orbital
shell rows (``mdf_int3c2e``, the cderi build's Rys kernels) into one device
buffer -- only the screened-in pairs, in the compact columns the cderi build
uses, so nothing is zero-filled or streamed for a screened pair -- and
contracted at once by the two halves of the J kernel
(``mdf_j_rho``, ``mdf_j_vj``), twice per call: nothing of size
``naux x npairs`` is stored, so a molecule whose cderi would not fit (or is
not worth building for a J-only run) still gets a GPU J.  The metric solve
is ``c = M^T M rho`` with ``M = L^-1`` (or the eigen filter's ``coef`` for a
non-positive-definite metric), two ``mdf_gemm`` calls; ``M`` and the chunk
plan are cached on the DF object (the reference caches its tables on the
option object for the same reason, :26-55).

Densities follow :func:`fastdf.get_jk`: ``(nao, nao)`` or stacked, numpy or
device, ``hermi`` 1 or 0.
"""
import ctypes

import numpy as np

from . import _dev
from . import _lib
from . import fastdfcderi
from . import ints

STATS = {'get_j': 0, 'chunks': 0}


def stats():
    return dict(STATS)


class _State:
    """Per-molecule device tables: bases, metric factor, shell-pair chunks."""

    def __init__(self, mol, chunk_bytes):
        t = _dev.torch()
        self.mol = mol
        self.ao = fastdfcderi._basis(mol, 'ao')
        self.aux = fastdfcderi._basis(mol, 'aux')
        self.nao, self.naux = self.ao.nao, self.aux.nao
        linv, coef = fastdfcderi.decompose_j2c(self.aux)
        self.M = linv if coef is None else coef          # (nrow, naux)
        bound = getattr(mol, '_mdf_pair_bound', None)
        if bound is None:
            bound = ints.pair_bound(mol.ao)
            mol._mdf_pair_bound = bound
        keep, rows, cols = fastdfcderi.kept_pairs(mol, bound)
        cidx = np.cumsum(keep) - 1
        self.npairs = int(keep.sum())
        self.addr = _dev.to_device(((rows * self.nao + cols)[keep]).astype(np.int32), dtype=t.int32)
        self.tasks = []
        for I0, I1 in fastdfcderi._chunks(mol.ao, self.naux, chunk_bytes):
            tk = ints.PairTasks(mol.ao, I0, I1, bound)
            c0 = int(cidx[tk.p0 - 1]) + 1 if tk.p0 > 0 else 0
            c1 = int(cidx[tk.p1 - 1]) + 1 if tk.p1 > 0 else 0
            if c1 == c0:
                continue
            tk.c0, tk.c1 = c0, c1
            tk.cmap = _dev.to_device((cidx[tk.p0:tk.p1] - c0).astype(np.int32), dtype=t.int32)
            self.tasks.append(tk)
        widest = max([tk.c1 - tk.c0 for tk in self.tasks] + [1])
        self.buf = t.empty(self.naux * widest, dtype=t.float64, device=_dev.device())


def _state(dfobj, chunk_bytes=None):
    st = getattr(dfobj, '_mdf_j3c', None)
    if st is None:
        st = _State(dfobj.mol, chunk_bytes or fastdfcderi.CHUNK_BYTES)
        try:
            dfobj._mdf_j3c = st
        except AttributeError:
            pass
    return st


def _sweep(st, fn):
    """Evaluate each chunk of (P|ij) over its kept pairs into the buffer
    (every column written) and hand it to fn."""
    for tk in st.tasks:
        width = tk.c1 - tk.c0
        blk = st.buf[:st.naux * width]
        ints.int3c2e_chunk(st.ao, st.aux, tk, blk, width, col_map=tk.cmap)
        fn(tk, blk, width)
        STATS['chunks'] += 1


def get_j(dfobj, dm, hermi=1):
    """``vj`` of ``dm`` (shape kept) from three-centre integrals evaluated on
    the fly.  ``dfobj`` needs ``mol`` (with ``ao``/``aux`` shell tables)."""
    from . import fastdf
    _dev.require_gpu()
    t = _dev.torch()
    STATS['get_j'] += 1
    st = _state(dfobj)
    nao, naux, npairs = st.nao, st.naux, st.npairs
    dms3 = _dev.to_device(dm).reshape(-1, nao, nao)
    n = dms3.shape[0]
    d = t.empty((n, npairs), dtype=t.float64, device=_dev.device())
    _lib.call('mdf_pack_dm', _dev.ptr(dms3), n, nao, nao * nao, _dev.ptr(st.addr), npairs,
              1 if hermi == 1 else 0, _dev.ptr(d), npairs, _dev.stream_ptr())
    widest = max(tk.c1 - tk.c0 for tk in st.tasks)
    work = _dev.WORKSPACE.get(_lib.query('mdf_get_j_workspace', naux, widest, n))
    rho = t.zeros((n, naux), dtype=t.float64, device=_dev.device())

    def contract_dm(tk, blk, width):
        _lib.call('mdf_j_rho', _dev.ptr(blk), _lib.DTYPE_F64, naux, width, width,
                  _dev.ptr(d[:, tk.c0:]), n, npairs, _dev.ptr(rho), naux, 1,
                  _dev.ptr(work), work.numel(), _dev.stream_ptr())

    _sweep(st, contract_dm)
    # c = M^T M rho, as rows: s = rho M^T, c = s M
    nrow = st.M.shape[0]
    s = t.empty((n, nrow), dtype=t.float64, device=_dev.device())
    _lib.call('mdf_gemm', 1, 0, 0, n, nrow, naux, 1., _dev.ptr(rho), naux, _dev.ptr(st.M), naux,
              0., _dev.ptr(s), nrow, _dev.stream_ptr())
    c = t.empty((n, naux), dtype=t.float64, device=_dev.device())
    _lib.call('mdf_gemm', 0, 0, 0, n, naux, nrow, 1., _dev.ptr(s), nrow, _dev.ptr(st.M), naux,
              0., _dev.ptr(c), naux, _dev.stream_ptr())
    vjp = t.zeros((n, npairs), dtype=t.float64, device=_dev.device())

    def contract_auxvec(tk, blk, width):
        _lib.call('mdf_j_vj', _dev.ptr(blk), _lib.DTYPE_F64, naux, width, width, _dev.ptr(c), n, naux,
                  _dev.ptr(vjp[:, tk.c0:]), npairs, 1, _dev.ptr(work), work.numel(), _dev.stream_ptr())

    _sweep(st, contract_auxvec)
    vj = t.empty((n, nao, nao), dtype=t.float64, device=_dev.device())
    _lib.call('mdf_unpack_pairs', _dev.ptr(vjp), n, npairs, _dev.ptr(st.addr), npairs, nao, _dev.ptr(vj),
              nao * nao, _dev.stream_ptr())
    return fastdf._finish(vj, _dev.shape_of(dm), dm)
