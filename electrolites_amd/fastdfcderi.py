"""The cderi build on MI355X: the drop-in for ``df._cholesky_eri``.

Mirrors electrolites/fastdfcderi.py ``_cholesky_eri`` (:352-693) and
``_decompose_j2c`` (:173-215) with every arithmetic step a libmi355df kernel:

1. ``j2c = (P|Q)``                          mdf_int2c2e   (Rys, per class)
2. ``j2c = L L^T``                           mdf_potrf     (blocked, MFMA)
   a non-positive pivot switches to the eigen filter of :203-210
   (``v / sqrt(w)`` for ``w > LINEAR_DEP_THR``; the aux count shrinks)
3. ``Linv = L^-1``                           mdf_trtri     (:479-481)
4. per chunk of orbital row shells:
   ``B = (pq|P)`` for the chunk's pairs      mdf_int3c2e   (:604, :667)
   ``cderi[:, chunk] = Linv B``              mdf_trmm      (:237-261)

The result is ``(naux, npairs)`` FP64, pair-contiguous, over the full packed
lower triangle (``p = i(i+1)/2 + j``), with ``cderi_idx = (pair_addr = i*nao + j
(int32), diag_idx)`` as the reference returns (:446-449, :693).  Shell pairs
whose primitive overlap bound is below ``ints.PAIR_SCREEN`` are not computed
and hold zeros.

FP32 storage (the reference's FASTDF_FP32 mode, electrolites/_dffp32.py):
with ``FASTDF_FP32=1`` in the environment (read at import, as there) or
``dtype='float32'``, each chunk is computed in FP64 and stored rounded to
FP32 (``mdf_cast_f32``); J and K widen the values back to FP64 before
arithmetic.  The tensor takes half the HBM and J half the bytes.
"""
import os

import numpy as np

from . import _dev
from . import _lib
from . import eigsolve
from . import ints

STATS = {'builds': 0, 'eigen_fallback': 0, 'chunks': 0, 'pairs_screened': 0,
         'pairs_kept': 0}
LINEAR_DEP_THR = 1e-9
#: device bytes for one chunk of three-centre integrals
CHUNK_BYTES = int(float(os.environ.get('FASTDFCDERI_CHUNK', 4 << 30)))
#: store new cderi tensors in single precision
FP32 = os.environ.get('FASTDF_FP32', '') not in ('', '0')


def stats():
    return dict(STATS)


def _basis(mol, which):
    key = f'_mdf_dev_{which}'
    b = getattr(mol, key, None)
    if b is None:
        b = ints.DeviceBasis(mol.ao if which == 'ao' else mol.aux)
        setattr(mol, key, b)
    return b


class _Timer:
    """Per-stage device time from HIP events on the launch stream."""

    def __init__(self):
        self.events = {}

    def mark(self, name):
        t = _dev.torch()
        ev = t.cuda.Event(enable_timing=True)
        ev.record()
        self.events.setdefault(name, []).append(ev)

    def totals(self):
        _dev.torch().cuda.synchronize()
        out = {}
        for name, evs in self.events.items():
            if name.endswith(':0'):
                base = name[:-2]
                ends = self.events.get(base + ':1', [])
                out[base] = sum(a.elapsed_time(b) for a, b in zip(evs, ends))
        return out


class _NoTimer:
    def mark(self, name):
        pass


def decompose_j2c(aux, timer=None):
    """``(Linv, None)`` from the Cholesky route, or ``(None, coef)`` with
    ``coef = (v / sqrt(w))^T`` from the eigen filter (non-PD metric)."""
    t = _dev.torch()
    tm = timer or _NoTimer()
    n = aux.nao
    tm.mark('int2c2e:0')
    j2c = ints.int2c2e(aux)
    tm.mark('int2c2e:1')
    info = t.zeros(1, dtype=t.int32, device=_dev.device())
    work = t.empty(_lib.query('mdf_potrf_workspace', n), dtype=t.uint8, device=_dev.device())
    low = j2c.clone()
    tm.mark('potrf:0')
    _lib.call('mdf_potrf', _dev.ptr(low), n, n, _dev.ptr(info), _dev.ptr(work),
              work.numel(), _dev.stream_ptr())
    tm.mark('potrf:1')
    if int(info.item()) == 0:
        linv = t.empty((n, n), dtype=t.float64, device=_dev.device())
        w2 = t.empty(_lib.query('mdf_trtri_workspace', n), dtype=t.uint8, device=_dev.device())
        tm.mark('trtri:0')
        _lib.call('mdf_trtri', _dev.ptr(low), n, n, _dev.ptr(work), _dev.ptr(linv), n,
                  _dev.ptr(w2), w2.numel(), _dev.stream_ptr())
        tm.mark('trtri:1')
        return linv, None
    STATS['eigen_fallback'] += 1
    w, v = eigsolve.eigh(j2c)
    keep = w > LINEAR_DEP_THR
    coef = (v[:, keep] / t.sqrt(w[keep])).T.contiguous()
    return None, coef


def _chunks(shells, naux, budget):
    """Row-shell ranges [I0, I1) whose packed pair count fits the budget."""
    ao_loc = shells.ao_loc
    out = []
    I0 = 0
    n = shells.nbas
    while I0 < n:
        I1 = I0 + 1
        while I1 < n:
            i0, i1 = int(ao_loc[I0]), int(ao_loc[I1 + 1])
            if (i1 * (i1 + 1) // 2 - i0 * (i0 + 1) // 2) * naux * 8 > budget:
                break
            I1 += 1
        out.append((I0, I1))
        I0 = I1
    return out


def kept_pairs(mol, bound=None):
    """The function pairs the cderi stores: every (i, j), i >= j, whose shell
    pair passes the screening (``ints.PAIR_SCREEN`` on the shell-pair bound,
    the test :class:`ints.PairTasks` applies), as a mask over the packed lower
    triangle p = i(i+1)/2 + j.  GPU4PySCF's cderi is stored the same way, over
    its screened pair list with the addresses in ``_cderi_idx``."""
    if bound is None:
        bound = getattr(mol, '_mdf_pair_bound', None)
        if bound is None:
            bound = ints.pair_bound(mol.ao)
            mol._mdf_pair_bound = bound
    ao_loc = np.asarray(mol.ao.ao_loc)
    shell_of = np.repeat(np.arange(len(ao_loc) - 1), np.diff(ao_loc))
    rows, cols = np.tril_indices(int(ao_loc[-1]))
    keep = np.asarray(bound)[shell_of[rows], shell_of[cols]] >= ints.PAIR_SCREEN
    return keep, rows, cols


def cholesky_eri(mol, chunk_bytes=None, timer=None, dtype=None, storage='device', layout='pair'):
    """Build the cderi of ``mol`` (orbital basis ``mol.ao``, fitting basis
    ``mol.aux``) on the current device.  Returns ``(cderi, (pair_addr,
    diag_idx))``.  ``timer`` (a :class:`_Timer`) collects per-stage device
    times.  ``dtype`` 'float64' or 'float32' (FP32 storage; default from
    ``FASTDF_FP32``).  ``storage`` 'device' (the tensor in HBM) or 'host'
    (a page-locked host tensor written chunk by chunk through
    :class:`electrolites_amd.dfstore.HostStore`, for a cderi larger than
    the device; electrolites/_dfstore.py).

    Only the screened-in pairs are stored (:func:`kept_pairs`): column c of
    the tensor is the c-th kept pair in packed order, and the returned
    addresses ``i * nao + j`` name them.  Each chunk's integrals are written
    straight into those columns (``mdf_int3c2e`` with a column map), so the
    metric solve and every later pass touch no screened column.

    ``layout`` 'pair' (default) stores the rows pair-contiguous; 'aux' (a
    device tensor only) returns the reference's device layout, the
    ``(naux, npairs)`` transpose of a C-order ``(npairs, naux)`` buffer
    (electrolites/fastdfcderi.py:648, :690): each chunk is computed as for
    'pair' and written out transposed (``mdf_transpose``)."""
    _dev.require_gpu()
    t = _dev.torch()
    tm = timer or _NoTimer()
    fp32 = FP32 if dtype is None else str(dtype).endswith('float32')
    ao, aux = _basis(mol, 'ao'), _basis(mol, 'aux')
    nao, naux = ao.nao, aux.nao
    bound = getattr(mol, '_mdf_pair_bound', None)
    if bound is None:
        bound = ints.pair_bound(mol.ao)
        mol._mdf_pair_bound = bound
    keep, rows, cols = kept_pairs(mol, bound)
    cidx = np.cumsum(keep) - 1                       # compact column of each kept pair
    npairs = int(keep.sum())
    linv, coef = decompose_j2c(aux, timer)
    nrow = naux if coef is None else coef.shape[0]
    if storage not in ('device', 'host'):
        raise ValueError(f'storage {storage!r}: expected device or host')
    on_host = storage == 'host'
    if layout not in ('pair', 'aux') or (layout == 'aux' and on_host):
        raise ValueError(f'layout {layout!r} with storage {storage!r}')
    auxc = layout == 'aux'
    sdt = t.float32 if fp32 else t.float64
    cderi = None
    if auxc:
        cbuf = t.empty((npairs, nrow), dtype=sdt, device=_dev.device())
    elif not on_host:
        cderi = t.empty((nrow, npairs), dtype=sdt, device=_dev.device())
    budget = chunk_bytes or CHUNK_BYTES
    if fp32:
        budget //= 2                       # an FP64 output chunk beside the integrals
    chunks = _chunks(mol.ao, naux, budget)

    def span(I0, I1):                 # full-triangle and compact column ranges of a chunk
        i0, i1 = int(mol.ao.ao_loc[I0]), int(mol.ao.ao_loc[I1])
        p0, p1 = i0 * (i0 + 1) // 2, i1 * (i1 + 1) // 2
        c0 = int(cidx[p0 - 1]) + 1 if p0 > 0 else 0
        c1 = int(cidx[p1 - 1]) + 1 if p1 > 0 else 0
        return p0, p1, c0, c1

    widest = max(max(span(I0, I1)[3] - span(I0, I1)[2] for I0, I1 in chunks), 1)
    buf = t.empty(naux * widest, dtype=t.float64, device=_dev.device())
    out64 = t.empty(nrow * widest, dtype=t.float64, device=_dev.device()) if fp32 or auxc else None
    out32 = t.empty(nrow * widest, dtype=t.float32, device=_dev.device()) if fp32 and auxc else None
    store = None
    if on_host:
        from . import dfstore
        store = dfstore.HostStore(nrow, npairs, sdt, widest)
    for I0, I1 in chunks:
        tasks = ints.PairTasks(mol.ao, I0, I1, bound)
        STATS['pairs_kept'] += tasks.npairs_kept
        STATS['pairs_screened'] += (I1 - I0) * (I0 + I1 + 1) // 2 - tasks.npairs_kept
        p0, p1, c0, c1 = span(I0, I1)
        width = c1 - c0
        if width == 0:
            continue
        blk = buf[:naux * width]
        # the chunk's kept pairs land in its compact columns (every one is
        # written: a kept shell pair writes all its function pairs)
        cmap = _dev.to_device((cidx[p0:p1] - c0).astype(np.int32), dtype=t.int32)
        tm.mark('int3c2e:0')
        ints.int3c2e_chunk(ao, aux, tasks, blk, width, col_map=cmap)
        tm.mark('int3c2e:1')
        if on_host:
            dst, ldd = store.slot(width), width
        elif auxc:
            dst, ldd = (out32 if fp32 else out64), width
        else:
            dst, ldd = cderi[:, c0:], npairs
        # FP64 result straight into the tensor, or into a chunk then rounded
        # and/or transposed
        out, ldo = (out64, width) if fp32 or auxc else (dst, ldd)
        tm.mark('trmm:0')
        if coef is None:
            _lib.call('mdf_trmm', _dev.ptr(linv), naux, naux, _dev.ptr(blk), width, width,
                      _dev.ptr(out), ldo, _dev.stream_ptr())
        else:
            _lib.call('mdf_gemm', 0, 0, 0, nrow, width, naux, 1., _dev.ptr(coef), naux,
                      _dev.ptr(blk), width, 0., _dev.ptr(out), ldo, _dev.stream_ptr())
        tm.mark('trmm:1')
        if fp32:
            _lib.call('mdf_cast_f32', _dev.ptr(out64), width, nrow, width, _dev.ptr(dst), ldd,
                      _dev.stream_ptr())
        if auxc:
            _lib.call('mdf_transpose', _dev.ptr(dst), 1 if fp32 else 0, nrow, width, width,
                      _dev.ptr(cbuf[c0:]), nrow, _dev.stream_ptr())
        if on_host:
            store.write(dst, c0)
        STATS['chunks'] += 1
    if on_host:
        cderi = store.finish()
        STATS['host_builds'] = STATS.get('host_builds', 0) + 1
    if auxc:
        cderi = cbuf.T
    STATS['builds'] += 1
    rk, ck = rows[keep], cols[keep]
    pair_addr = (rk * nao + ck).astype(np.int32)
    diag = np.where(rk == ck)[0]
    return cderi, (pair_addr, diag)


def tile_fill(mol, tile=128):
    """Fraction of the lower-triangle ``tile x tile`` AO blocks that hold at
    least one shell pair kept by the screening (``ints.PAIR_SCREEN``).  The
    reference switches step 1 to its block-sparse tiled form only when this
    fill is at most 0.75 (electrolites/fastdf.py:566-574, ``TILE_MAX_FILL``
    :235); the dense clusters measured here have fill 1.  Host-only."""
    bound = getattr(mol, '_mdf_pair_bound', None)
    if bound is None:
        bound = ints.pair_bound(mol.ao)
        mol._mdf_pair_bound = bound
    kept = np.tril(bound > ints.PAIR_SCREEN)
    blk = np.asarray(mol.ao.ao_loc)[:-1] // tile
    nt = (mol.nao + tile - 1) // tile
    hit = np.zeros((nt, nt), bool)
    I, J = np.nonzero(kept)
    hit[blk[I], blk[J]] = True
    hit |= hit.T
    return float(hit[np.tril_indices(nt)].mean())


def _session_molecules(intopt):
    """``(orbital, auxiliary)`` PySCF molecules of a GPU4PySCF
    ``Int3c2eOpt`` (its ``mol`` and ``auxmol`` are sorted wrappers whose
    ``.mol`` is the original, electrolites/fastdfcderi.py:380-386), or None
    when ``intopt`` is not shaped that way."""
    aux = getattr(intopt, 'auxmol', None)
    mol = getattr(intopt, 'mol', None)
    if aux is None or mol is None or hasattr(mol, 'ao'):
        return None
    mol = getattr(mol, 'mol', mol)
    aux = getattr(aux, 'mol', aux)
    if not all(hasattr(m, a) for m in (mol, aux) for a in ('_atm', '_bas', '_env')):
        return None
    return mol, aux


def _as_session_array(x):
    """A CuPy array over the same device memory when CuPy is importable
    (a GPU4PySCF session), else the torch tensor (which exposes the CUDA
    array interface)."""
    try:
        import cupy
    except ImportError:
        return x
    return cupy.asarray(x)


def _cholesky_eri(intopt, omega=None, use_gpu_memory=None):
    """``df._cholesky_eri``'s signature (electrolites/fastdfcderi.py:352).

    ``intopt`` is either this package's :class:`~electrolites_amd.df.DF`
    (anything with a ``mol`` carrying ``ao``/``aux`` shell tables), which
    gets the pair-contiguous tensor as a torch tensor, or GPU4PySCF's
    ``Int3c2eOpt``, whose original molecules' libcint arrays give the shell
    tables (:mod:`electrolites_amd.gto.libcint`) and which gets the
    reference's return values: ``[cderi]`` with the device tensor in the
    aux-contiguous layout (``out.T``, :690) or, for ``use_gpu_memory=False``,
    a C-order host array (:576-577), and ``(pair_addr, diag_idx)`` as device
    arrays (:446-449) -- CuPy arrays when CuPy is importable.  Range
    separation is not covered."""
    if omega not in (None, 0, 0.):
        raise NotImplementedError('range-separated cderi (omega != 0) is not '
                                  'covered by the MI355X build')
    pair = _session_molecules(intopt)
    if pair is None:
        mol = getattr(intopt, 'mol', intopt)
        # use_gpu_memory False: GPU4PySCF's host-resident tensor (df/df.py),
        # here a page-locked host tensor the J/K calls stream from
        storage = 'host' if use_gpu_memory is False else getattr(intopt, 'storage', 'device')
        cderi, idx = cholesky_eri(mol, storage=storage)
        return [cderi], idx
    from .gto.libcint import LibcintPair
    mol = getattr(intopt, '_mdf_molecules', None)
    if mol is None or mol.pyscf_mol is not pair[0]:
        mol = LibcintPair(*pair)
        try:
            intopt._mdf_molecules = mol
        except AttributeError:
            pass
    t = _dev.torch()
    STATS['session_builds'] = STATS.get('session_builds', 0) + 1
    if use_gpu_memory is False:
        cderi, (pa, diag) = cholesky_eri(mol, storage='host')
        cderi = cderi.numpy()
    else:
        cderi, (pa, diag) = cholesky_eri(mol, layout='aux')
        cderi = _as_session_array(cderi)
    pa = _as_session_array(_dev.to_device(pa, dtype=t.int32))
    diag = _as_session_array(_dev.to_device(diag, dtype=t.int64))
    return [cderi], (pa, diag)
