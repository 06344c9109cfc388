"""Density-fitted J/K on MI355X: the drop-in for ``df_jk.get_jk``.

Mirrors electrolites/fastdf.py (``get_jk`` :754, ``_get_jk`` :805,
``_get_j`` :688, ``_sparse_dm`` :674, declines and fall-through :406-441,
``STATS``/``stats()`` :271-285) with the arithmetic on hand-written gfx950
HIP kernels behind libmi355df's C ABI:

* J: ``mdf_pack_dm`` (pair gather with the symmetry factors), ``mdf_get_j``
  (two streaming passes over the packed cderi), ``mdf_unpack_pairs``;
* K: ``mdf_get_k`` (packed-triangle unpack fused into an FP64-MFMA step-1
  transform, then a symmetric rank-k step 2).

The exchange precision is a per-object knob, ``dfobj.k_precision`` ('fp64',
the default, or 'fp32'), so the signature stays the reference's; the SCF
driver lowers it for early cycles.

``dfobj`` is anything with the reference DF object's fields: ``_cderi``
(per-device mapping to a ``(naux, npairs)`` float64 device array),
``_cderi_idx = (pair_addresses, diag_idx)``, ``build()``.  This package's own
:class:`electrolites_amd.df.DF` is one; a GPU4PySCF ``DF`` object is another
when :func:`electrolites_amd.patch` has installed this function.

Every case this module does not cover (a complex density, a range-separated
``omega``, more than one device, a cderi neither float64 nor float32) is
*declined*: counted in ``stats()`` and handed to the function this one
replaced, or, when there is none (standalone use), raised as
:class:`Declined`.  A float32 cderi (the FP32 storage mode) is covered: its
values are widened to FP64 before arithmetic.
"""
import ctypes
import os
import weakref

import numpy as np

from . import _dev
from . import _lib
from . import dfstore
from . import eigsolve

STATS = {'calls': 0, 'declined': 0, 'j_only': 0, 'mode0': 0, 'mode1': 0,
         'k_calls': 0}
DECLINED = {}

#: the function this one replaced (GPU4PySCF's) once patched; None standalone
_orig_get_jk = None
#: disable switch, read once at import like the reference's FASTDF_OFF
_OFF = os.environ.get('FASTDF_OFF', '0') not in ('', '0')
#: bytes of device scratch the K build may use for its transformed block
K_WORKSPACE_BYTES = int(float(os.environ.get('FASTDF_K_WORKSPACE', 8 << 30)))
#: relative cutoff below which density eigenvalues are dropped (untagged dm)
FACTOR_CUTOFF = 1e-15


class Declined(RuntimeError):
    """This call is outside what the HIP path covers and nothing to fall
    through to exists (standalone use)."""


def stats():
    """What this module did in this process."""
    out = dict(STATS)
    out['declined_reasons'] = dict(DECLINED)
    return out


def _decline(reason):
    STATS['declined'] += 1
    DECLINED[reason] = DECLINED.get(reason, 0) + 1
    return reason


def _fall_through(dfobj, dms, hermi, with_j, with_k, omega, reason):
    _decline(reason)
    cderi = _cderi_of(dfobj, missing_ok=True)
    if cderi is not None and str(getattr(cderi, 'dtype', '')).endswith('float32'):
        # fastdf.py:434-440: a narrow tensor must not reach a float64 reader
        raise RuntimeError(f'the cderi is float32 and this call was declined '
                           f'({reason}); rebuild it in float64')
    if _orig_get_jk is None:
        raise Declined(reason)
    return _orig_get_jk(dfobj, dms, hermi, with_j, with_k, omega=omega)


def _stored_cderi(dfobj, missing_ok=False):
    """The object the DF object stores for the current device: a torch
    tensor, a device array of another library (GPU4PySCF's CuPy tensor) or a
    host array."""
    store = getattr(dfobj, '_cderi', None)
    if store is None:
        if missing_ok:
            return None
        raise ValueError('dfobj has no cderi; call build() first')
    t = _dev.torch()
    dev = t.cuda.current_device() if t.cuda.is_available() else 0
    try:
        return store[dev]
    except (KeyError, IndexError, TypeError):
        if missing_ok:
            return None
        raise


def _cderi_of(dfobj, missing_ok=False):
    """The stored cderi as a torch tensor (a zero-copy view of a CuPy or other
    CUDA-array-interface device array, strides kept); host arrays as they
    are."""
    c = _stored_cderi(dfobj, missing_ok)
    return None if c is None else _dev.as_torch(c)


#: DF objects holding device caches of this module (released on OOM)
_CACHE_OWNERS = weakref.WeakSet()


def _cache_key(dfobj, x, *extra):
    """Key of a cache entry kept on ``dfobj`` for the tensor ``x`` (the
    stored cderi or a view of it): the stored object's identity and the
    memory ``x`` names (address, shape, strides, type), so a rebuilt tensor
    never matches an entry of the one it replaced, even at a reused address
    (the stored object is also held weakly where it allows: a dead
    reference misses)."""
    stored = _stored_cderi(dfobj, missing_ok=True)
    key = (id(stored), x.data_ptr(), tuple(x.shape), tuple(x.stride()), str(x.dtype)) + extra
    try:
        ref = weakref.ref(stored)
    except TypeError:
        ref = None
    return key, ref


def _cache_get(dfobj, name, key):
    cached = getattr(dfobj, name, None)
    if cached is None or cached[0] != key[0]:
        return None
    if cached[1] is not None and cached[1]() is None:
        return None
    return cached[2]


def _cache_put(dfobj, name, key, value):
    try:
        setattr(dfobj, name, (key[0], key[1], value))
    except AttributeError:
        return value
    try:
        _CACHE_OWNERS.add(dfobj)
    except TypeError:
        pass
    return value


def _release_caches():
    """Drop the device memory this module keeps between calls (the device
    prefix of a host cderi, the pair-contiguous copy of an aux-contiguous
    one, the shared workspace); False if nothing was held.  The reference's
    OOM release (electrolites/fastdf.py:779-803): a prefix is replaced by
    the cache's own "no prefix" entry and a dropped copy is not made again
    for that tensor, so the retry streams instead."""
    import gc
    released = 0
    for o in list(_CACHE_OWNERS):
        res = getattr(o, '_mdf_resident', None)
        if res is not None and res[1] is not None:
            o._mdf_resident = (res[0], None)
            STATS['resident_dropped'] = STATS.get('resident_dropped', 0) + 1
            released += 1
        cp = getattr(o, '_mdf_auxc', None)
        if cp is not None and cp[2] is not None:
            o._mdf_auxc = (cp[0], cp[1], None)
            STATS['auxc_copy_dropped'] = STATS.get('auxc_copy_dropped', 0) + 1
            released += 1
    if _dev.WORKSPACE._buf:
        _dev.WORKSPACE.release()
        released += 1
    if not released:
        return False
    gc.collect()
    t = _dev.torch()
    if t.cuda.is_available():
        t.cuda.synchronize()
        t.cuda.empty_cache()
    return True


# --------------------------------------------------------------- pair layout
class _PairLayout:
    """Device-side pair tables of one cderi, cached on the DF object."""

    def __init__(self, pair_addr, nao, npairs):
        t = _dev.torch()
        self.nao, self.npairs = nao, npairs
        full = npairs == nao * (nao + 1) // 2
        addr = None
        if pair_addr is not None:
            addr = _dev.to_device(pair_addr, dtype=t.int32)
            if full:
                rows, cols = np.tril_indices(nao)
                ref = t.from_numpy((rows * nao + cols).astype(np.int32)).to(addr.device)
                full = bool(t.equal(addr, ref))
        self.full = full
        self.addr = None if full else addr
        self.map = None
        if not full:
            if addr is None:
                raise ValueError('a screened cderi needs its pair addresses')
            self.map = t.empty(nao * nao, dtype=t.int32, device=_dev.device())
            _lib.call('mdf_build_pair_map', _dev.ptr(addr), npairs, nao,
                      _dev.ptr(self.map), _dev.stream_ptr())


def _layout(dfobj, cderi, nao):
    """The pair tables of ``dfobj``'s cderi, kept on the DF object under a key
    of the stored tensor and its pair addresses."""
    idx = getattr(dfobj, '_cderi_idx', None)
    pair_addr = None if idx is None else idx[0]
    npairs = cderi.shape[1]
    key = _cache_key(dfobj, cderi, nao, npairs, id(pair_addr))
    lay = _cache_get(dfobj, '_mdf_layout', key)
    if lay is None:
        lay = _cache_put(dfobj, '_mdf_layout', key, _PairLayout(pair_addr, nao, npairs))
    return lay


# ----------------------------------------------------- aux-contiguous storage
#: make a pair-contiguous device copy of an aux-contiguous cderi (GPU4PySCF's
#: device layout) when it fits beside the tensor; otherwise, or with this off,
#: every call transposes blocks of aux rows as it goes (FASTDF_AUXC_COPY)
AUXC_COPY = os.environ.get('FASTDF_AUXC_COPY', '1') not in ('', '0')
#: device bytes kept free beside the copy
AUXC_MARGIN = 4 << 30


def is_aux_contiguous(cderi):
    """A 2-D device tensor whose aux index is the contiguous one: the
    ``(naux, npairs)`` transpose of a C-order ``(npairs, naux)`` buffer, as
    the reference's device build returns it (electrolites/fastdfcderi.py
    :648, :690)."""
    return (cderi.dim() == 2 and cderi.stride(0) == 1 and cderi.shape[0] > 1
            and cderi.stride(1) >= cderi.shape[0])


def transpose_rows(cderi, L0, nL, out):
    """Rows ``L0 .. L0 + nL`` of an aux-contiguous cderi into the
    pair-contiguous ``(nL, npairs)`` tensor ``out`` (``mdf_transpose``)."""
    npairs = cderi.shape[1]
    e = cderi.element_size()
    _lib.call('mdf_transpose', ctypes.c_void_p(cderi.data_ptr() + L0 * e), _dtype_code(cderi),
              npairs, nL, cderi.stride(1), _dev.ptr(out), out.stride(0), _dev.stream_ptr())
    return out


def pair_major(dfobj, cderi):
    """A pair-contiguous ``(naux, npairs)`` device tensor with the values of
    ``cderi``: the tensor itself when it is pair-contiguous; for an
    aux-contiguous one a copy made once and kept on the DF object (None
    when it does not fit, after an OOM release, or with ``AUXC_COPY``
    off)."""
    if cderi.stride(1) == 1:
        return cderi
    if not is_aux_contiguous(cderi):
        return None
    key = _cache_key(dfobj, cderi)
    cached = getattr(dfobj, '_mdf_auxc', None)
    if cached is not None and cached[0] == key[0] and (cached[1] is None or cached[1]() is not None):
        return cached[2]                      # the copy, or None once released
    if not AUXC_COPY:
        return None
    t = _dev.torch()
    nbytes = cderi.numel() * cderi.element_size()
    free = t.cuda.mem_get_info()[0] + t.cuda.memory_reserved() - t.cuda.memory_allocated()
    if free < nbytes + K_WORKSPACE_BYTES + AUXC_MARGIN:
        STATS['auxc_no_room'] = STATS.get('auxc_no_room', 0) + 1
        return None
    out = t.empty(tuple(cderi.shape), dtype=cderi.dtype, device=_dev.device())
    transpose_rows(cderi, 0, cderi.shape[0], out)
    STATS['auxc_copies'] = STATS.get('auxc_copies', 0) + 1
    return _cache_put(dfobj, '_mdf_auxc', key, out)


def auxc_blocks(cderi, blk_rows):
    """Yield ``(block, L0, nL)`` pair-contiguous device blocks of an
    aux-contiguous cderi, each transposed on the current stream into one
    staging buffer (valid until the generator resumes)."""
    t = _dev.torch()
    naux, npairs = cderi.shape
    nb = min(blk_rows, naux)
    buf = t.empty((nb, npairs), dtype=cderi.dtype, device=_dev.device())
    for L0 in range(0, naux, nb):
        nL = min(nb, naux - L0)
        STATS['auxc_blocks'] = STATS.get('auxc_blocks', 0) + 1
        yield transpose_rows(cderi, L0, nL, buf[:nL]), L0, nL


# ------------------------------------------------------------ density factors
def factorize_dm(dms, hermi):
    """``(fl, fr)`` with ``D = fl fr^T`` per density; ``fr is None`` when the
    density is tagged with ``mo_coeff``/``mo_occ`` (mode 0, ``D = F F^T``).

    Restates what GPU4PySCF's ``DFJK.factorize_dm`` provides to the reference
    (call at fastdf.py:829; modes at :837-845): tagged -> ``C_occ sqrt(occ)``;
    untagged -> two factors from ``eigh`` (hermi 1) or an SVD (hermi 0),
    dropping values below ``FACTOR_CUTOFF`` of the largest.
    Returns device tensors shaped ``(n, nao, nocc)``.
    """
    t = _dev.torch()
    nao = dms.shape[-1]
    mo_coeff = getattr(dms, 'mo_coeff', None)
    mo_occ = getattr(dms, 'mo_occ', None)
    mo1 = getattr(dms, 'mo1', None)
    occ_coeff = getattr(dms, 'occ_coeff', None)
    if mo1 is not None and occ_coeff is not None:
        # a response density D = X + X^T with X = mo1 (w occ_coeff)^T, tagged
        # ``symmetrize`` (the CPHF densities of the reference's Hessian
        # tests, tests/test_dfhess.py:100-131): the two factors of X, w = 2
        # for a restricted occ_coeff (nao, nocc), 1 per spin for (2, nao, nocc)
        m1 = _dev.to_device(mo1)
        oc = _dev.to_device(occ_coeff)
        nocc = m1.shape[-1]
        if oc.shape[-1] != nocc or oc.shape[-2] != nao:
            raise ValueError(f'mo1 {tuple(m1.shape)} and occ_coeff {tuple(oc.shape)} do not pair')
        if oc.dim() == 2:
            fl = m1.reshape(-1, nao, nocc)
            fr = (2. * oc).expand(fl.shape[0], nao, nocc)
        else:
            nspin = oc.shape[0]
            fl = m1.reshape(nspin, -1, nao, nocc)
            fr = oc.reshape(nspin, 1, nao, nocc).expand(fl.shape)
            fl = fl.reshape(-1, nao, nocc)
            fr = fr.reshape(-1, nao, nocc)
        return fl.contiguous(), fr.contiguous()
    if mo_coeff is not None and mo_occ is not None:
        c = _dev.to_device(mo_coeff)
        c = c.reshape(-1, nao, c.shape[-1])
        occ = _dev.to_device(mo_occ).reshape(c.shape[0], -1)
        facs = []
        nocc = int(max(int((occ[i] > 0).sum()) for i in range(c.shape[0])))
        out = t.zeros((c.shape[0], nao, nocc), dtype=t.float64, device=c.device)
        for i in range(c.shape[0]):
            sel = occ[i] > 0
            f = c[i][:, sel] * t.sqrt(occ[i][sel])
            out[i, :, :f.shape[1]] = f
        return out.contiguous(), None
    d = _dev.to_device(dms).reshape(-1, nao, nao)
    fls, frs = [], []
    for i in range(d.shape[0]):
        if hermi == 1:
            w, v = eigsolve.eigh(.5 * (d[i] + d[i].T))
            keep = w.abs() > FACTOR_CUTOFF * max(float(w.abs().max()), 1e-300)
            w, v = w[keep], v[:, keep]
            s = w.abs().sqrt()
            fls.append(v * s)
            frs.append(v * (s * t.sign(w)))
        else:
            with _dev.SOLVER_LOCK:
                u, s, vh = t.linalg.svd(d[i])
            keep = s > FACTOR_CUTOFF * max(float(s.max()), 1e-300)
            s = s[keep].sqrt()
            fls.append(u[:, keep] * s)
            frs.append(vh[keep].T * s)
    nocc = max(f.shape[1] for f in fls)
    fl = t.zeros((len(fls), nao, nocc), dtype=t.float64, device=d.device)
    fr = t.zeros_like(fl)
    for i, (a, b) in enumerate(zip(fls, frs)):
        fl[i, :, :a.shape[1]] = a
        fr[i, :, :b.shape[1]] = b
    return fl.contiguous(), fr.contiguous()


# ----------------------------------------------------------------- J and K
def _sparse_dm(dms3, lay, hermi):
    """The density on the stored pairs (fastdf.py:674-686), ``(n, npairs)``."""
    t = _dev.torch()
    n, nao = dms3.shape[0], dms3.shape[-1]
    d = t.empty((n, lay.npairs), dtype=t.float64, device=dms3.device)
    _lib.call('mdf_pack_dm', _dev.ptr(dms3), n, nao, nao * nao,
              _dev.ptr(lay.addr), lay.npairs, 1 if hermi == 1 else 0,
              _dev.ptr(d), lay.npairs, _dev.stream_ptr())
    return d


def _vj(cderi, lay, dms3, hermi):
    t = _dev.torch()
    n, nao = dms3.shape[0], dms3.shape[-1]
    naux, npairs = cderi.shape
    d = _sparse_dm(dms3, lay, hermi)
    rho = t.empty((n, naux), dtype=t.float64, device=dms3.device)
    vjp = t.empty((n, npairs), dtype=t.float64, device=dms3.device)
    nbytes = _lib.query('mdf_get_j_workspace', naux, npairs, n)
    work = _dev.WORKSPACE.get(nbytes)
    _lib.call('mdf_get_j', _dev.ptr(cderi), _dtype_code(cderi), naux, npairs, cderi.stride(0),
              _dev.ptr(d), n, npairs, _dev.ptr(rho), naux, _dev.ptr(vjp),
              npairs, 0, _dev.ptr(work), work.numel(), _dev.stream_ptr())
    vj = t.empty((n, nao, nao), dtype=t.float64, device=dms3.device)
    _lib.call('mdf_unpack_pairs', _dev.ptr(vjp), n, npairs, _dev.ptr(lay.addr),
              npairs, nao, _dev.ptr(vj), nao * nao, _dev.stream_ptr())
    return vj


PRECISIONS = {'fp64': 0, 'fp32': 1, 'bf16x3': 2}


def _dtype_code(cderi):
    """Storage type of a cderi for the library: FP64, or FP32 storage whose
    values are widened to FP64 before arithmetic (the reference's
    FASTDF_FP32 mode, electrolites/_dffp32.py:166-196)."""
    t = _dev.torch()
    if cderi.dtype == t.float64:
        return _lib.DTYPE_F64
    if cderi.dtype == t.float32:
        return _lib.DTYPE_F32
    raise ValueError(f'cderi dtype {cderi.dtype}')


def _vk(cderi, lay, fl, fr, precision='fp64'):
    """``vk[n] = sum_L (A_L fl[n])(A_L fr[n])^T`` on the HIP path.
    ``precision`` 'fp64', 'fp32' (FP32 MFMA for both steps) or 'bf16x3'
    (split-BF16 three-term MFMA for both steps, FP32 accumulation)."""
    t = _dev.torch()
    n, nao, nocc = fl.shape
    naux, npairs = cderi.shape
    prec = PRECISIONS[precision]
    vk = t.empty((n, nao, nao), dtype=t.float64, device=fl.device)
    two = fr is not None
    elem = 8 if prec == 0 else 4
    per_row = nao * nocc * elem * (2 if two else 1)
    rows = max(1, min(naux, K_WORKSPACE_BYTES // max(per_row, 1)))
    # the library cuts naux into passes by shape (the grouped launch cuts the
    # same way); K_WORKSPACE_BYTES can only shorten them
    nbytes = _lib.query('mdf_get_k_workspace', naux, nao, nocc, int(two), prec, rows)
    work = _dev.WORKSPACE.get(nbytes)
    for i in range(n):
        STATS['k_calls'] += 1
        _lib.call('mdf_get_k', _dev.ptr(cderi), _dtype_code(cderi), naux, npairs, cderi.stride(0),
                  _dev.ptr(lay.map), nao, _dev.ptr(fl[i]),
                  _dev.ptr(fr[i]) if two else None, nocc, nocc, prec,
                  _dev.ptr(vk[i]), nao, 0, _dev.ptr(work), nbytes,
                  _dev.stream_ptr())
    return vk


def _jk_blocked(dfobj, cderi, lay, dms3, hermi, with_j, fl, fr, precision, blocks=None,
                blk_rows=None):
    """J and K over a cderi contracted block by block of aux rows
    (electrolites/fastdf.py:805-1026 with ``_blocks`` :611-671), each block
    contracted once for both -- ``mdf_get_j`` adding its rows' share of vj
    (and writing their rho), ``mdf_get_k`` adding their share of vk.  By
    default the blocks come from a host-resident cderi: the device prefix of
    its rows (:func:`dfstore.resident_prefix`, kept between calls) and the
    rest streamed, so the tensor crosses PCIe once per call.  ``blocks`` (a
    generator factory taking ``blk_rows``) supplies other blocks: the
    transposed rows of an aux-contiguous device tensor.  ``fl`` None: J
    only."""
    t = _dev.torch()
    n, nao = dms3.shape[0], dms3.shape[-1]
    naux, npairs = cderi.shape
    elem = cderi.element_size()
    dt = _dtype_code(cderi)
    blk_rows = blk_rows or dfstore.block_rows(npairs, elem, naux)
    with_k = fl is not None
    jbytes = _lib.query('mdf_get_j_workspace', blk_rows, npairs, n) if with_j else 0
    kbytes, prec, two, nocc = 0, PRECISIONS[precision], False, 0
    if with_k:
        nocc, two = fl.shape[-1], fr is not None
        per_row = nao * nocc * (8 if prec == 0 else 4) * (2 if two else 1)
        rows = max(1, min(blk_rows, K_WORKSPACE_BYTES // max(per_row, 1)))
        kbytes = _lib.query('mdf_get_k_workspace', blk_rows, nao, nocc, int(two), prec, rows)
    nbytes = max(jbytes, kbytes)
    if blocks is None:
        # what this call allocates beside the prefix: two streamed blocks,
        # the workspace, the pair vectors and the outputs
        reserve = 2 * blk_rows * npairs * elem + nbytes + 8 * n * (2 * npairs + naux + 2 * nao * nao)
        resident = dfstore.resident_prefix(dfobj, cderi, reserve)
        try:
            _CACHE_OWNERS.add(dfobj)
        except TypeError:
            pass
        it = dfstore.blocks(cderi, resident, blk_rows)
        STATS['host_calls'] = STATS.get('host_calls', 0) + 1
    else:
        it = blocks(blk_rows)
    work = _dev.WORKSPACE.get(nbytes)
    if with_j:
        d = _sparse_dm(dms3, lay, hermi)
        rho = t.empty((n, naux), dtype=t.float64, device=dms3.device)
        vjp = t.zeros((n, npairs), dtype=t.float64, device=dms3.device)
    if with_k:
        vk = t.zeros((n, nao, nao), dtype=t.float64, device=dms3.device)
    for blk, L0, nL in it:
        ld = blk.stride(0)
        if with_j:
            _lib.call('mdf_get_j', _dev.ptr(blk), dt, nL, npairs, ld, _dev.ptr(d), n, npairs,
                      ctypes.c_void_p(rho.data_ptr() + 8 * L0), naux, _dev.ptr(vjp), npairs, 1,
                      _dev.ptr(work), nbytes, _dev.stream_ptr())
        if with_k:
            for i in range(n):
                STATS['k_calls'] += 1
                _lib.call('mdf_get_k', _dev.ptr(blk), dt, nL, npairs, ld, _dev.ptr(lay.map), nao,
                          _dev.ptr(fl[i]), _dev.ptr(fr[i]) if two else None, nocc, nocc, prec,
                          _dev.ptr(vk[i]), nao, 1, _dev.ptr(work), nbytes, _dev.stream_ptr())
    vj = None
    if with_j:
        vj = t.empty((n, nao, nao), dtype=t.float64, device=dms3.device)
        _lib.call('mdf_unpack_pairs', _dev.ptr(vjp), n, npairs, _dev.ptr(lay.addr),
                  npairs, nao, _dev.ptr(vj), nao * nao, _dev.stream_ptr())
    return vj, (vk if with_k else None)


def _jk_host(dfobj, cderi, lay, dms3, hermi, with_j, fl, fr, precision):
    """J/K over a host-resident cderi: :func:`_jk_blocked` on its device
    prefix and streamed remainder."""
    return _jk_blocked(dfobj, cderi, lay, dms3, hermi, with_j, fl, fr, precision)


def _finish(x, shape, ref):
    """``x`` in the input's shape and array family (numpy in, numpy out;
    torch in, torch out; CuPy in, CuPy out: electrolites/fastdf.py:828,
    1017-1024)."""
    return _dev.like_input(x.reshape(tuple(shape)), ref)


#: tags of a density that carry its factors or how to finish K
_TAGS = ('mo_coeff', 'mo_occ', 'mo1', 'occ_coeff', 'symmetrize')


def get_jk(dfobj, dms, hermi=0, with_j=True, with_k=True, omega=None):
    """``(vj, vk)`` of ``dms`` over ``dfobj``'s cderi (df_jk.get_jk's
    signature; electrolites/fastdf.py:754-776).  numpy in gives numpy out; a
    device array in (a torch tensor, a CuPy array or any object exposing the
    CUDA array interface) gives device arrays out.  A call that runs out of
    device memory releases what this module keeps between calls and runs
    once more; nothing else is retried."""
    try:
        return _get_jk(dfobj, dms, hermi, with_j, with_k, omega)
    except Exception as exc:                                 # noqa: BLE001
        if not _dev.is_oom(exc) or not _release_caches():
            raise
    STATS['oom_retries'] = STATS.get('oom_retries', 0) + 1
    return _get_jk(dfobj, dms, hermi, with_j, with_k, omega)


def _as_density(dms):
    """A torch view of a device density from another library (CuPy, ...),
    its factor tags carried over; anything else unchanged."""
    if not _dev.is_foreign_device_array(dms):
        return dms
    out = _dev.as_torch(dms)
    for name in _TAGS:
        v = getattr(dms, name, None)
        if v is not None:
            setattr(out, name, v)
    return out


def _get_jk(dfobj, dms, hermi=0, with_j=True, with_k=True, omega=None):
    if _OFF:
        return _fall_through(dfobj, dms, hermi, with_j, with_k, omega,
                             'FASTDF_OFF')
    t = _dev.torch()
    dms_in = dms
    try:
        dms = _as_density(dms)
    except (TypeError, ValueError, RuntimeError) as exc:
        return _fall_through(dfobj, dms_in, hermi, with_j, with_k, omega,
                             f'unreadable device density ({type(exc).__name__})')
    dt = getattr(dms, 'dtype', None)
    if dt not in (np.float64, t.float64):
        return _fall_through(dfobj, dms_in, hermi, with_j, with_k, omega,
                             'complex density' if dt is not None and
                             'complex' in str(dt) else 'non-float64 density')
    if omega not in (None, 0, 0.):
        return _fall_through(dfobj, dms_in, hermi, with_j, with_k, omega,
                             'range-separated omega')
    if not t.cuda.is_available():
        return _fall_through(dfobj, dms_in, hermi, with_j, with_k, omega,
                             'no GPU')
    if t.cuda.device_count() != 1 and os.environ.get('MDF_ALLOW_MULTI', '0') == '0':
        return _fall_through(dfobj, dms_in, hermi, with_j, with_k, omega,
                             'more than one device')
    symmetrize = int(getattr(dms, 'symmetrize', 0) or 0)
    if symmetrize not in (0, 1):
        return _fall_through(dfobj, dms_in, hermi, with_j, with_k, omega,
                             f'symmetrize={symmetrize}')
    if symmetrize and with_k and (getattr(dms, 'mo1', None) is None or
                                  getattr(dms, 'occ_coeff', None) is None):
        return _fall_through(dfobj, dms_in, hermi, with_j, with_k, omega,
                             'symmetrize without mo1/occ_coeff factors')
    if getattr(dfobj, '_cderi', None) is None:
        dfobj.build()
    cderi = _cderi_of(dfobj)
    host = dfstore.is_host(cderi)
    if host:
        cderi = dfstore.as_tensor(cderi)
    if cderi.dtype not in (t.float64, t.float32):
        return _fall_through(dfobj, dms_in, hermi, with_j, with_k, omega,
                             'cderi neither float64 nor float32')
    if cderi.dtype == t.float32:
        STATS['fp32_store'] = STATS.get('fp32_store', 0) + 1
    if cderi.shape[0] == 0:
        return _fall_through(dfobj, dms_in, hermi, with_j, with_k, omega,
                             'empty cderi slice')
    auxc = None                       # an aux-contiguous device tensor streamed by blocks
    if cderi.dim() != 2:
        return _fall_through(dfobj, dms_in, hermi, with_j, with_k, omega,
                             'cderi not two-dimensional')
    if cderi.stride(1) != 1:
        if host or not is_aux_contiguous(cderi):
            return _fall_through(dfobj, dms_in, hermi, with_j, with_k, omega,
                                 'cderi neither pair- nor aux-contiguous')
        STATS['auxc_calls'] = STATS.get('auxc_calls', 0) + 1
        pm = pair_major(dfobj, cderi)
        if pm is None:
            auxc = cderi
        else:
            cderi = pm
    STATS['calls'] += 1
    shape = _dev.shape_of(dms)
    nao = shape[-1]
    lay = _layout(dfobj, cderi, nao)
    dms3 = _dev.to_device(dms).reshape(-1, nao, nao)
    blocked = host or auxc is not None
    blocks = None
    if auxc is not None:
        def blocks(blk_rows):
            return auxc_blocks(auxc, blk_rows)

    if not with_k:
        STATS['j_only'] += 1
        if blocked:
            vj, _ = _jk_blocked(dfobj, cderi, lay, dms3, hermi, with_j, None, None, 'fp64', blocks)
        else:
            vj = _vj(cderi, lay, dms3, hermi) if with_j else None
        return (None if vj is None else _finish(vj, shape, dms_in)), None

    fl, fr = factorize_dm(dms, hermi)
    if fl.shape[-1] == 0:
        return dms_in, dms_in                   # fastdf.py:850-851
    STATS['mode1' if fr is not None else 'mode0'] += 1
    if fl.shape[0] != dms3.shape[0]:
        raise ValueError(f'{fl.shape[0]} density factors for '
                         f'{dms3.shape[0]} densities')
    vj = _vj(cderi, lay, dms3, hermi) if with_j and not blocked else None
    prec = getattr(dfobj, 'k_precision', 'fp64') or 'fp64'
    if prec not in PRECISIONS:
        raise ValueError(f'k_precision {prec!r}: expected one of {sorted(PRECISIONS)}')
    STATS['k_' + prec] = STATS.get('k_' + prec, 0) + 1
    if blocked:
        vj, vk = _jk_blocked(dfobj, cderi, lay, dms3, hermi, with_j, fl, fr, prec, blocks)
    else:
        vk = _vk(cderi, lay, fl, fr, prec)
    if symmetrize:
        # K of the tagged half X (D = X + X^T) plus its transpose
        # (electrolites/fastdf.py:1020-1021, transpose_sum with hermi 1)
        STATS['symmetrized'] = STATS.get('symmetrized', 0) + 1
        vk = vk + vk.transpose(1, 2)
    return ((None if vj is None else _finish(vj, shape, dms_in)),
            _finish(vk, shape, dms_in))


def tag_array(a, **kwargs):
    """``pyscf.lib.tag_array``: an ndarray view (or the tensor itself) carrying
    ``mo_coeff``/``mo_occ`` attributes, which select the one-factor K."""
    t = _dev.torch()
    if isinstance(a, t.Tensor):
        for k, v in kwargs.items():
            setattr(a, k, v)
        return a
    out = np.asarray(a).view(_Tagged)
    out.__dict__.update(kwargs)
    return out


class _Tagged(np.ndarray):
    """ndarray subclass whose instance dict holds the tags; views and results
    of arithmetic do not inherit them."""
