"""The density-fitted gradient's two-electron term from the cderi the SCF
already built.

Mirrors electrolites/fastdfgrad.py ``jk_energy_per_atom`` (:1229-1325) and
its contraction ``_ejk`` (:870-1182), the drop-in for GPU4PySCF's
``df/grad/rhf.py:_jk_energy_per_atom``: the nuclear derivative, per atom, of
the restricted two-electron energy

    E = 1/2 j_factor tr(D J[D]) - 1/4 k_factor tr(D K[D])

(for an unrestricted density, D = D_a + D_b in J and -1/2 k_factor
sum_s tr(D_s K[D_s]) in K: the sums below run over the two spin factors with
twice the weight)
with J and K density fitted over the metric M = (P|Q).  With the cderi
``B_L = sum_P A[L, P] T_P`` (T_P the three-centre integrals (ij|P); A the
inverse Cholesky factor of M or the eigen filter's coefficients, so that
A^T A = M^-1) and D = F_l F_r^T,

    rho_B[L]  = sum_ij B_L,ij D_ij          c = A^T rho_B          (= M^-1 rho)
    rhok[L]   = F_r^T B_L F_l               dmoo = A^T rhok        (nocc x nocc each)
    W[P, ij]  = j c_P D_ij - 1/2 k (F_l dmoo_P F_r^T)_ij
    Z[P, Q]   = -1/2 j c_P c_Q + 1/4 k tr(dmoo_P dmoo_Q)

    dE/dR = sum_{P,ij} W[P, ij] d(ij|P)/dR + sum_{P,Q} Z[P, Q] d(P|Q)/dR

(the reference's identity A A^T-free form: its ``dm_oo = aux_coef @ rhok``,
:16-45; ``k_half = -0.5 k_factor`` for a restricted density, :900-903).  The
three-centre integrals are not evaluated again: rho_B and rhok come from the
cderi's blocks (:meth:`DF.loop`), and the derivative integrals are
contracted with W and Z inside the Rys kernels (``mdf_int3c2e_ip``,
``mdf_int2c2e_ip``), which add straight into the per-atom gradient.  W is
held in the cderi's own compact column layout (one value per stored pair).

R_L = A_L F_l comes from the K engine's first step (``mdf_k_step1``, the
packed rows read through the pair map) for a device-resident cderi, and from
the unpacked blocks of :meth:`DF.loop` otherwise; the remaining contractions
(F_r^T R_L, the triangular products, W's blocks and Z) run on the package's
FP64 matrix-core GEMM (``mdf_gemm``, ``mdf_gemm_batched``,
``mdf_gemm_pairs_batched``).  With one factor (F_l = F_r) every rhok_L and
dmoo_P is symmetric and is kept as its packed upper triangle
(``mdf_sym_pack`` / ``mdf_sym_unpack``) with the off-diagonal entries times
sqrt 2, so that the dot product of two packed rows is the full trace
product.
"""
import ctypes
import math

import numpy as np

from . import _dev
from . import _lib
from . import fastdf
from . import fastdfcderi
from . import ints

STATS = {'calls': 0, 'declined': 0, 'declined_reasons': {}}
#: the replaced GPU4PySCF methods, by class (kept by patch() for declines)
_orig_methods = {}
#: bytes of the R_L = A_L F_l block buffer
RBLOCK_BYTES = 1 << 31
#: bytes of one block of (nao, nocc) rows F_l dmoo_P of the weight build
WBLOCK_BYTES = 1 << 30
#: aux rows per block of the triangular metric products (host-stored path)
TBLOCK = 1024
#: off-diagonal weight of a packed symmetric block
SQRT2 = math.sqrt(2.)


def stats():
    return dict(STATS)


def _gemm(A, B, C, transb=False, tri_a=0, lower_c=False, alpha=1., beta=0.):
    """``C = alpha A op(B) + beta C`` on the package's FP64 MFMA engine
    (``mdf_gemm``): row-major device matrices with unit column stride."""
    M, K = A.shape
    N = B.shape[0] if transb else B.shape[1]
    _lib.call('mdf_gemm', int(transb), int(tri_a), int(lower_c), M, N, K, float(alpha), _dev.ptr(A),
              A.stride(0), _dev.ptr(B), B.stride(0), float(beta), _dev.ptr(C), C.stride(0), _dev.stream_ptr())
    return C


def _gemm_batched(M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, transb=False, lower_c=False,
                  alpha=1., beta=0.):
    """``mdf_gemm_batched`` over raw device addresses (ints) with strides in
    elements."""
    _lib.call('mdf_gemm_batched', int(transb), 0, int(lower_c), M, N, K, float(alpha), ctypes.c_void_p(A), lda,
              sA, ctypes.c_void_p(B), ldb, sB, float(beta), ctypes.c_void_p(C), ldc, sC, batch,
              _dev.stream_ptr())


def _class_sorted_pairs(mol, addr):
    """The shell pairs (I >= J) whose function pairs the cderi stores, sorted
    by class, and the packed-pair -> column map of the cderi (int32, -1 for a
    pair it does not store).  A shell pair stored in part is an error."""
    sh = mol.ao
    nao, nbas = sh.nao, sh.nbas
    ao_loc = np.asarray(sh.ao_loc)
    addr = np.asarray(addr, dtype=np.int64)
    i, j = np.divmod(addr, nao)
    hi, lo = np.maximum(i, j), np.minimum(i, j)
    cmap = np.full(nao * (nao + 1) // 2, -1, dtype=np.int32)
    cmap[hi * (hi + 1) // 2 + lo] = np.arange(len(addr), dtype=np.int32)
    shell_of = np.repeat(np.arange(nbas), np.diff(ao_loc))
    rows, cols = np.tril_indices(nao)
    sp = shell_of[rows] * nbas + shell_of[cols]
    total = np.bincount(sp, minlength=nbas * nbas)
    have = np.bincount(sp, weights=(cmap >= 0), minlength=nbas * nbas).astype(np.int64)
    if np.any((have > 0) & (have < total)):
        raise ValueError('jk_energy_per_atom: the cderi stores part of a shell pair')
    full = np.nonzero((have == total) & (total > 0))[0]
    I, J = np.divmod(full, nbas)
    l = np.asarray(sh.shl_l)
    cls = l[I] * (ints.LMAX_AO + 1) + l[J]
    order = np.argsort(cls, kind='stable')
    pairs = np.ascontiguousarray(np.stack([I, J], axis=1)[order].astype(np.int32))
    off = np.zeros((ints.LMAX_AO + 1) ** 2 + 1, dtype=np.int64)
    off[1:] = np.cumsum(np.bincount(cls, minlength=(ints.LMAX_AO + 1) ** 2))
    return pairs, off, cmap


def _cache_key(dfobj, mol):
    """What the cached per-cderi tables depend on: the stored tensor (its
    identity and memory, held weakly: ``fastdf._cache_key``), its pair list
    and the basis (positions, exponents, coefficients of both sets)."""
    cderi = fastdf._cderi_of(dfobj)
    idx = getattr(dfobj, '_cderi_idx', None)
    basis = b''.join(np.ascontiguousarray(a).tobytes() for sh in (mol.ao, mol.aux)
                     for a in (sh.shl_xyz, sh.exps, sh.coefs))
    x = _dev.torch().from_numpy(cderi) if isinstance(cderi, np.ndarray) else cderi
    return fastdf._cache_key(dfobj, x, None if idx is None else id(idx[0]), basis)


def _metric_factor(dfobj, mol, aux):
    """``(A, lower, At)``: the matrix the cderi is defined by (cderi = A T;
    the inverse Cholesky factor of the metric, lower triangular, or the
    eigen filter's coefficients) and its transpose, cached on the DF object
    for its current cderi and basis (the reference reuses the build's
    ``aux_coef`` the same way, electrolites/fastdfgrad.py:665-714)."""
    t = _dev.torch()
    key = _cache_key(dfobj, mol)
    cached = fastdf._cache_get(dfobj, '_mdf_grad_metric', key)
    if cached is not None:
        return cached
    linv, coef = fastdfcderi.decompose_j2c(aux)
    A, lower = (t.tril(linv), True) if coef is None else (coef, False)
    return fastdf._cache_put(dfobj, '_mdf_grad_metric', key, (A, lower, A.T.contiguous()))


def _pair_tables(dfobj, mol):
    """The device tables of the derivative launch and the weight build for
    the cderi's pair list (class-sorted shell pairs, packed-pair -> column
    map, column addresses, lower-triangle blocks, shell atoms), cached on the
    DF object like the metric factor."""
    t = _dev.torch()
    key = _cache_key(dfobj, mol)
    cached = fastdf._cache_get(dfobj, '_mdf_grad_pairs', key)
    if cached is not None:
        return cached
    addr = _dev.to_host(dfobj._cderi_idx[0])
    pairs, class_off, cmap = _class_sorted_pairs(mol, addr)
    i32 = lambda a: _dev.to_device(np.ascontiguousarray(a, dtype=np.int32), dtype=t.int32)
    tabs = {
        'ncol': len(addr), 'class_off': class_off,
        'pairs': i32(pairs.reshape(-1)) if len(pairs) else t.zeros(2, dtype=t.int32, device=_dev.device()),
        'cmap': i32(cmap), 'addr': _dev.to_device(addr.astype(np.int64), dtype=t.int64),
        'ao_atom': i32(mol.ao.shl_atom), 'aux_atom': i32(mol.aux.shl_atom),
    }
    fastdf._cache_put(dfobj, '_mdf_grad_pairs', key, tabs)
    return tabs


def _factors(dm, hermi):
    """``[(fl, fr)]`` per density: one for a restricted density ``(nao,
    nao)``, two for the spin densities ``(2, nao, nao)`` of an unrestricted
    one."""
    if hermi != 1:
        raise ValueError('jk_energy_per_atom: a symmetric density (hermi=1) is required')
    nd = dm.dim() if hasattr(dm, 'dim') else np.ndim(dm)
    if nd == 2:
        dms = dm if hasattr(dm, 'mo_coeff') else dm[None]
    elif nd == 3 and dm.shape[0] == 2:
        dms = dm
    else:
        raise ValueError('jk_energy_per_atom: expected one density (nao, nao) or a spin pair (2, nao, nao)')
    fl, fr = fastdf.factorize_dm(dms, 1)
    n = 1 if nd == 2 else 2
    if fl.shape[0] != n:
        raise ValueError('jk_energy_per_atom: density factors do not match the density shape')
    return [(fl[s].contiguous(), (fl[s] if fr is None else fr[s]).contiguous()) for s in range(n)]


class _Stages:
    """Device time per stage from HIP events on the current stream, into
    the caller's dict (ms); a no-op without one."""

    def __init__(self, out):
        self.out, self.ev = out, []

    def mark(self, name):
        if self.out is None:
            return
        t = _dev.torch()
        e = t.cuda.Event(enable_timing=True)
        e.record()
        self.ev.append((name, e))

    def close(self):
        if self.out is None:
            return
        self.ev[-1][1].synchronize()
        for (n, a), (_, b) in zip(self.ev, self.ev[1:]):
            self.out[n] = self.out.get(n, 0.) + a.elapsed_time(b)


def _on_device(cderi):
    t = _dev.torch()
    return isinstance(cderi, t.Tensor) and cderi.is_cuda and cderi.dtype in (t.float64, t.float32) \
        and cderi.dim() == 2 and cderi.stride(1) == 1


def _rho_packed(dfobj, cderi, D, nao, nrow):
    """``rho_B[L] = sum_ij B_L,ij D_ij`` from the packed rows (``mdf_j_rho``
    over the density on the stored pairs, off-diagonal pairs doubled)."""
    t = _dev.torch()
    if cderi.shape[0] != nrow:
        raise ValueError(f'jk_energy_per_atom: the cderi has {cderi.shape[0]} rows, the metric factor {nrow}')
    lay = fastdf._layout(dfobj, cderi, nao)
    npairs = cderi.shape[1]
    d = fastdf._sparse_dm(D[None].contiguous(), lay, 1)
    rho = t.empty(nrow, dtype=t.float64, device=_dev.device())
    nbytes = _lib.query('mdf_get_j_workspace', nrow, npairs, 1)
    work = _dev.WORKSPACE.get(nbytes)
    _lib.call('mdf_j_rho', _dev.ptr(cderi), fastdf._dtype_code(cderi), nrow, npairs, cderi.stride(0),
              _dev.ptr(d), 1, npairs, _dev.ptr(rho), nrow, 0, _dev.ptr(work), nbytes, _dev.stream_ptr())
    return rho

def _same_factor(fl, fr):
    return fl.data_ptr() == fr.data_ptr() and fl.shape == fr.shape and fl.stride() == fr.stride()


def _sym_pack(G, P, n, batch, diag_scale, off_scale):
    """``P[z] = packed upper triangle of s (G[z] + G[z]^T)`` (``mdf_sym_pack``)
    for the rows z < batch of the 2-D P and the contiguous n x n blocks of G."""
    assert P.dim() == 2 and P.shape[0] >= batch
    _lib.call('mdf_sym_pack', n, _dev.ptr(G), n * n, _dev.ptr(P), P.stride(0), float(diag_scale),
              float(off_scale), batch, _dev.stream_ptr())


def _sym_unpack(P, G, n, batch, diag_scale, off_scale):
    """``G[z] = G[z]^T = s P[z]`` unpacked (``mdf_sym_unpack``) from the rows
    z < batch of the 2-D P into contiguous n x n blocks of G."""
    assert P.dim() == 2 and P.shape[0] >= batch
    _lib.call('mdf_sym_unpack', n, _dev.ptr(P), P.stride(0), _dev.ptr(G), n * n, float(diag_scale),
              float(off_scale), batch, _dev.stream_ptr())


def _diag_positions(n):
    """Offsets of the diagonal entries in a packed upper triangle of order n."""
    a = np.arange(n, dtype=np.int64)
    return a * n - a * (a - 1) // 2


def _rhok(dfobj, fl, fr, nao, nrow):
    """``rhok[L] = F_r^T A_L F_l`` for every cderi row L, and whether it is
    packed.  With one factor (F_l = F_r = F) the form is symmetric: A_L =
    T_L + T_L^T with T_L its lower triangle, diagonal halved, so rhok[L] =
    G + G^T with G = F^T T_L F, and T_L F costs about half of A_L F
    (``mdf_k_step1_lower``); rhok then comes back as ``(nrow, nocc (nocc +
    1) / 2)`` packed upper triangles, off-diagonal entries times sqrt 2."""
    t = _dev.torch()
    nocc = fl.shape[1]
    sym = _same_factor(fl, fr)
    shape = (nrow, nocc * (nocc + 1) // 2) if sym else (nrow, nocc, nocc)
    rhok = t.empty(shape, dtype=t.float64, device=_dev.device())
    frT = fr.T.contiguous()
    cderi = fastdf._cderi_of(dfobj)
    if _on_device(cderi):
        # R_L = A_L F_l (or T_L F) by the K engine's first step, straight
        # from the packed rows
        if cderi.shape[0] != nrow:
            raise ValueError(f'jk_energy_per_atom: the cderi has {cderi.shape[0]} rows, the metric factor {nrow}')
        lay = fastdf._layout(dfobj, cderi, nao)
        nb = max(1, min(nrow, RBLOCK_BYTES // max(8 * nao * nocc, 1)))
        R = t.empty((nb, nao, nocc), dtype=t.float64, device=_dev.device())
        G = t.empty((nb, nocc, nocc), dtype=t.float64, device=_dev.device()) if sym else None
        code = fastdf._dtype_code(cderi)
        for L0 in range(0, nrow, nb):
            nL = min(nb, nrow - L0)
            _lib.call('mdf_k_step1_lower' if sym else 'mdf_k_step1', _dev.ptr(cderi[L0]), code, nL,
                      cderi.shape[1], cderi.stride(0), _dev.ptr(lay.map), nao, _dev.ptr(fl), fl.stride(0), nocc,
                      _dev.ptr(R), nocc, nao * nocc, _dev.stream_ptr())
            # F_r^T R_L, one batched product over the block's rows
            out = G if sym else rhok[L0]
            _gemm_batched(nocc, nocc, nao, frT.data_ptr(), nao, 0, R.data_ptr(), nocc, nao * nocc,
                          out.data_ptr(), nocc, nocc * nocc, nL)
            if sym:
                _sym_pack(G, rhok[L0:L0 + nL], nocc, nL, 1., SQRT2)
        return rhok, sym
    L0 = 0
    for dense, _ in dfobj.loop():
        nL = dense.shape[0]
        blk = t.matmul(frT, t.matmul(dense, fl))
        if sym:
            # the full symmetric rhok_L: pack (rhok + rhok^T) / 2
            _sym_pack(blk.contiguous(), rhok[L0:L0 + nL], nocc, nL, .5, .5 * SQRT2)
        else:
            rhok[L0:L0 + nL] = blk
        L0 += nL
    if L0 != nrow:
        raise ValueError(f'jk_energy_per_atom: the cderi has {L0} rows, the metric factor {nrow}')
    return rhok, sym


def _dmoo(A, lower, At, rhok, naux):
    """``dmoo[P] = sum_L A[L, P] rhok[L]`` = A^T rhok as one product; A^T is
    upper triangular on the Cholesky route (only L >= P contributes), and
    the product skips the zero blocks.  Packed rhok rows give packed dmoo
    rows (the product is linear in each column)."""
    t = _dev.torch()
    nrow = rhok.shape[0]
    rk = rhok.reshape(nrow, -1)
    dmoo = t.empty((naux, rk.shape[1]), dtype=t.float64, device=_dev.device())
    _gemm(At, rk, dmoo, tri_a=2 if lower else 0)
    return dmoo


def _trace_products_packed(dmoo):
    """``tr(dmoo_P dmoo_Q)`` for symmetric dmoo from the packed rows (the
    off-diagonal entries already times sqrt 2): the lower 128-tiles of one
    product, mirrored."""
    t = _dev.torch()
    naux = dmoo.shape[0]
    out = t.empty((naux, naux), dtype=t.float64, device=_dev.device())
    _gemm(dmoo, dmoo, out, transb=True, lower_c=True)
    out = t.tril(out)
    return out + t.tril(out, -1).T


def _trace_products(dmoo, naux):
    """``tr(dmoo_P dmoo_Q)`` for all P, Q: the lower 128-tiles of one
    product, mirrored."""
    t = _dev.torch()
    dm_flat = dmoo.reshape(naux, -1)
    dmT_flat = dmoo.transpose(1, 2).contiguous().reshape(naux, -1)
    out = t.empty((naux, naux), dtype=t.float64, device=_dev.device())
    _gemm(dm_flat, dmT_flat, out, transb=True, lower_c=True)
    out = t.tril(out)
    return out + t.tril(out, -1).T


def jk_energy_per_atom(dfobj, dm, j_factor=1., k_factor=1., hermi=1, auxbasis_response=True, timings=None):
    """``(natm, 3)`` nuclear derivative of ``1/2 j tr(D J) - 1/4 k tr(D K)``
    for the restricted density ``dm`` (nao x nao), or of ``1/2 j tr(D J) -
    1/2 k sum_s tr(D_s K[D_s])`` (D = D_a + D_b) for the spin densities ``dm``
    (2, nao, nao) of an unrestricted one (tagged with ``mo_coeff``/``mo_occ``
    or not), from the built cderi of ``dfobj``
    (electrolites/fastdfgrad.py:1229).  ``auxbasis_response`` adds the
    fitting functions' centres (the aux-centre share of the three-centre
    term and the metric term), as GPU4PySCF's flag of that name.
    ``timings``: a dict that receives the device milliseconds of each stage
    ('metric', 'rhok', 'weights', 'ip3c', 'ip2c')."""
    t = _dev.torch()
    tm = _Stages(timings)
    _dev.require_gpu()
    mol = dfobj.mol
    if getattr(dfobj, '_cderi', None) is None:
        dfobj.build()
    nao, natm = mol.nao, mol.natm
    facs = _factors(dm, hermi)
    # K weight per factor set: E_K = -1/4 k tr(D K[D]) restricted,
    # -1/2 k sum_s tr(D_s K[D_s]) over the spin densities
    wk = .25 if len(facs) == 1 else .5
    D = t.empty((nao, nao), dtype=t.float64, device=_dev.device())
    for si, (fl, fr) in enumerate(facs):
        _gemm(fl, fr, D, transb=True, beta=0. if si == 0 else 1.)
    D = .5 * (D + D.T)
    aux = fastdfcderi._basis(mol, 'aux')
    ao = fastdfcderi._basis(mol, 'ao')
    tm.mark('metric')
    A, lower, At = _metric_factor(dfobj, mol, aux)        # cderi = A T
    nrow, naux = A.shape
    tm.mark('rhok')
    cderi = fastdf._cderi_of(dfobj)
    if k_factor == 0 and _on_device(cderi):
        # Coulomb only: rho_B = B d over the packed rows (J's first pass)
        rhoks, rho = [], _rho_packed(dfobj, cderi, D, nao, nrow)
    else:
        # rhok_s from the cderi's blocks; rho_B[L] = tr(A_L D) = sum_s tr(rhok_s[L])
        rhoks = [_rhok(dfobj, fl, fr, nao, nrow) for fl, fr in facs]
        rho = 0
        for (r, packed), (fl, _) in zip(rhoks, facs):
            if packed:
                dpos = t.as_tensor(_diag_positions(fl.shape[1]), device=_dev.device())
                rho = rho + r[:, dpos].sum(dim=1)
            else:
                rho = rho + t.diagonal(r, dim1=1, dim2=2).sum(dim=1)
    tm.mark('weights')
    c = t.empty((1, naux), dtype=t.float64, device=_dev.device())
    _gemm(rho.reshape(1, -1), A, c)                             # c = A^T rho
    c = c.reshape(-1)
    Z = (-.5 * j_factor) * t.outer(c, c)
    # W in the cderi's compact columns: W[P, c] = j c_P D_c, the exchange
    # part added by the pair products (the first of which writes the sum)
    tabs = _pair_tables(dfobj, mol)
    ncol = tabs['ncol']
    W = t.empty((naux, ncol), dtype=t.float64, device=_dev.device())
    Dflat = D.reshape(-1)[tabs['addr']].contiguous()
    jc = (j_factor * c).contiguous()
    if k_factor == 0:
        jb = max(1, min(naux, WBLOCK_BYTES // max(8 * ncol, 1)))
        for P0 in range(0, naux, jb):
            P1 = min(naux, P0 + jb)
            W[P0:P1] = jc[P0:P1, None] * Dflat[None, :]
    else:
        for si, (fl, fr) in enumerate(facs):
            nocc = fl.shape[1]
            rk, sym = rhoks[si]
            rhoks[si] = None
            dmoo = _dmoo(A, lower, At, rk, naux)
            del rk
            if sym:
                # packed rows, off-diagonal entries times sqrt 2: the trace
                # products are plain dot products of the rows
                Z += (wk * k_factor) * _trace_products_packed(dmoo)
            else:
                Z += (wk * k_factor) * _trace_products(dmoo.reshape(naux, nocc, nocc), naux)
            # the exchange part at the stored pairs: G_P = F_l dmoo_P F_r^T
            # as H_P = F_r dmoo_P^T for a block of P (one product), then
            # the lower 128-tiles of F_l H_P^T per P put straight into the
            # cderi's compact columns (one batched product; no G_P is
            # written), over the Coulomb part on the first factor set
            blk = max(1, min(naux, WBLOCK_BYTES // max(8 * nao * nocc, 1)))
            H = t.empty((nao, blk * nocc), dtype=t.float64, device=_dev.device())
            full = t.empty((blk, nocc * nocc), dtype=t.float64, device=_dev.device()) if sym else None
            for P0 in range(0, naux, blk):
                P1 = min(naux, P0 + blk)
                nP = P1 - P0
                Hb = H[:, :nP * nocc]
                if sym:
                    dm_blk = full[:nP]
                    _sym_unpack(dmoo[P0:P1], dm_blk, nocc, nP, 1., 1. / SQRT2)
                else:
                    dm_blk = dmoo[P0:P1]
                _gemm(fr, dm_blk.reshape(nP * nocc, nocc), Hb, transb=True)
                init = (_dev.ptr(jc[P0:P1]), _dev.ptr(Dflat)) if si == 0 else (None, None)
                _lib.call('mdf_gemm_pairs_batched', 1, nao, nocc, float(-2 * wk * k_factor), ctypes.c_void_p(fl.data_ptr()),
                          nocc, 0, ctypes.c_void_p(Hb.data_ptr()), Hb.stride(0), nocc, _dev.ptr(tabs['cmap']),
                          _dev.ptr(W[P0]), W.stride(0), init[0], init[1], nP, _dev.stream_ptr())
            del dmoo, H, full
    del rhoks
    grad = t.zeros((natm, 3), dtype=t.float64, device=_dev.device())
    grad_aux = t.zeros((natm, 3), dtype=t.float64, device=_dev.device())
    pairs_d, cmap_d, class_off = tabs['pairs'], tabs['cmap'], tabs['class_off']
    ao_atom, aux_atom = tabs['ao_atom'], tabs['aux_atom']
    lp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_long))
    if ao.desc.lmax > ints.LMAX_AO or aux.desc.lmax > ints.LMAX_AUX:
        raise ValueError('jk_energy_per_atom: angular momentum beyond the generated classes')
    tm.mark('ip3c')
    _lib.call('mdf_int3c2e_ip', ctypes.byref(ao.desc), ctypes.byref(aux.desc), _dev.ptr(pairs_d),
              lp(class_off), _dev.ptr(aux.sorted), lp(aux.l_off), _dev.ptr(W), W.stride(0),
              _dev.ptr(cmap_d), _dev.ptr(ao_atom), _dev.ptr(aux_atom), natm, _dev.ptr(grad), _dev.ptr(grad_aux),
              _dev.stream_ptr())
    tm.mark('ip2c')
    if auxbasis_response:
        Zc = Z.contiguous()
        _lib.call('mdf_int2c2e_ip', ctypes.byref(aux.desc), _dev.ptr(aux.sorted), lp(aux.l_off),
                  _dev.ptr(Zc), Zc.stride(0), _dev.ptr(aux_atom), natm, _dev.ptr(grad_aux), _dev.stream_ptr())
        grad += grad_aux
    tm.mark('end')
    tm.close()
    STATS['calls'] += 1
    return grad.cpu().numpy()


def _decline_reason(mf, dm, omega):
    t = _dev.torch()
    if omega:
        return 'omega != 0'
    with_df = getattr(mf, 'with_df', None)
    store = getattr(with_df, '_cderi', None)
    if with_df is None or store is None:
        return 'no built cderi'
    cderi = fastdf._cderi_of(with_df, missing_ok=True)
    if not isinstance(cderi, t.Tensor):
        return 'cderi not built by this package'
    nd = dm.dim() if hasattr(dm, 'dim') else np.ndim(dm)
    if not (nd == 2 or (nd == 3 and dm.shape[0] == 2)):
        return 'not one density or a spin pair'
    d = _dev.to_host(dm)
    if np.iscomplexobj(d):
        return 'complex density'
    d = d.reshape(-1, d.shape[-2], d.shape[-1])
    if not all(np.allclose(x, x.T, rtol=0, atol=1e-12 * max(np.abs(x).max(), 1.)) for x in d):
        return 'non-symmetric density'
    return None


def _orig_for(obj):
    for cls in type(obj).__mro__:
        if cls in _orig_methods:
            return _orig_methods[cls]
    return None


def patched_jk_energy_per_atom(self, dm=None, j_factor=1, k_factor=1, omega=0, hermi=0, verbose=None):
    """GPU4PySCF's ``Gradients.jk_energy_per_atom`` for a restricted density
    or an unrestricted spin pair over a cderi this package built
    (electrolites/fastdfgrad.py:1229-1325 replaces the same method on the
    same four classes).  Anything else -- a long-range call, a cderi
    GPU4PySCF built itself, a non-symmetric density -- is counted and handed
    to the method this replaced."""
    mf = self.base
    if dm is None:
        dm = mf.make_rdm1()
    reason = _decline_reason(mf, dm, omega)
    if reason is None:
        return jk_energy_per_atom(mf.with_df, dm, j_factor, k_factor, hermi=1,
                                  auxbasis_response=getattr(self, 'auxbasis_response', True))
    STATS['declined'] += 1
    STATS['declined_reasons'][reason] = STATS['declined_reasons'].get(reason, 0) + 1
    orig = _orig_for(self)
    if orig is None:
        raise fastdf.Declined(f'jk_energy_per_atom: {reason}')
    return orig(self, dm, j_factor, k_factor, omega=omega, hermi=hermi, verbose=verbose)
