"""Device-side basis sets and the Rys integral entry points.

``DeviceBasis`` uploads a :class:`~electrolites_amd.gto.ShellTable` once and
holds the ``mdf_basis`` descriptor libmi355df reads.  ``int3c2e_chunk`` and
``int2c2e`` call the generated Rys kernels; shell-pair screening and the
class-sorted task lists are prepared here on the host.
"""
import ctypes

import numpy as np

from . import _dev
from . import _lib

LMAX_AO = 2
LMAX_AUX = 4
#: a shell pair is skipped when every primitive pair has
#: |c_a c_b| exp(-ab/(a+b) R^2) (pi/(a+b))^(3/2) below this
PAIR_SCREEN = 1e-18


class MdfBasis(ctypes.Structure):
    _fields_ = [('nbas', ctypes.c_int), ('lmax', ctypes.c_int),
                ('nprim', ctypes.c_void_p), ('poff', ctypes.c_void_p),
                ('aooff', ctypes.c_void_p), ('xyz', ctypes.c_void_p),
                ('exps', ctypes.c_void_p), ('coefs', ctypes.c_void_p)]


class DeviceBasis:
    """A shell table resident on the current device."""

    def __init__(self, shells):
        t = _dev.torch()
        self.shells = shells
        self.nbas = shells.nbas
        self.nao = shells.nao
        self.l = np.asarray(shells.shl_l)
        i32 = lambda a: _dev.to_device(np.asarray(a, dtype=np.int32), dtype=t.int32)
        f64 = lambda a: _dev.to_device(np.asarray(a, dtype=np.float64))
        self._nprim = i32(shells.shl_nprim)
        self._poff = i32(shells.shl_prim_off)
        self._aooff = i32(shells.shl_ao_off)
        self._xyz = f64(shells.shl_xyz.reshape(-1))
        self._exps = f64(shells.exps)
        self._coefs = f64(shells.coefs)
        self.desc = MdfBasis(self.nbas, int(self.l.max()) if self.nbas else 0,
                             self._nprim.data_ptr(), self._poff.data_ptr(),
                             self._aooff.data_ptr(), self._xyz.data_ptr(),
                             self._exps.data_ptr(), self._coefs.data_ptr())
        # shells sorted by l (for the aux role)
        order = np.argsort(self.l, kind='stable').astype(np.int32)
        self.sorted = i32(order)
        cnt = np.bincount(self.l, minlength=LMAX_AUX + 1)[:LMAX_AUX + 1]
        self.l_off = np.zeros(LMAX_AUX + 2, dtype=np.int64)
        self.l_off[1:] = np.cumsum(cnt)


def pair_bound(shells):
    """Per shell pair (I, J): max over primitive pairs of
    |c_a c_b| exp(-ab/(a+b) |A-B|^2) (pi/(a+b))^(3/2)  (symmetric)."""
    n = shells.nbas
    npm = int(shells.shl_nprim.max()) if n else 1
    e = np.ones((n, npm))
    c = np.zeros((n, npm))
    for s_ in range(n):
        o, k = shells.shl_prim_off[s_], shells.shl_nprim[s_]
        e[s_, :k] = shells.exps[o:o + k]
        c[s_, :k] = np.abs(shells.coefs[o:o + k])
    xyz = shells.shl_xyz
    out = np.zeros((n, n))
    step = max(1, 4000000 // max(n * npm * npm, 1))
    for I0 in range(0, n, step):
        I1 = min(n, I0 + step)
        r2 = np.sum((xyz[I0:I1, None, :] - xyz[None, :, :]) ** 2, axis=-1)
        ea = e[I0:I1, None, :, None]
        eb = e[None, :, None, :]
        p = ea + eb
        v = (c[I0:I1, None, :, None] * c[None, :, None, :] *
             np.exp(-ea * eb / p * r2[:, :, None, None]) * (np.pi / p) ** 1.5)
        out[I0:I1] = v.reshape(I1 - I0, n, -1).max(axis=-1)
    return out


class PairTasks:
    """Screened shell pairs of the row-shell range [I0, I1), class-sorted."""

    def __init__(self, shells, I0, I1, bound=None, thr=PAIR_SCREEN):
        l = np.asarray(shells.shl_l)
        ii, jj = np.tril_indices(I1)
        sel = ii >= I0
        ii, jj = ii[sel], jj[sel]
        if bound is not None:
            keep = bound[ii, jj] >= thr
            ii, jj = ii[keep], jj[keep]
        pairs = np.stack([ii, jj], axis=1).astype(np.int32)
        cls = l[ii] * (LMAX_AO + 1) + l[jj]
        order = np.argsort(cls, kind='stable')
        self.pairs = np.ascontiguousarray(pairs[order])
        cnt = np.bincount(cls, minlength=(LMAX_AO + 1) ** 2)
        self.class_off = np.zeros((LMAX_AO + 1) ** 2 + 1, dtype=np.int64)
        self.class_off[1:] = np.cumsum(cnt)
        ao_loc = shells.ao_loc
        i0, i1 = int(ao_loc[I0]), int(ao_loc[I1])
        self.p0 = i0 * (i0 + 1) // 2
        self.p1 = i1 * (i1 + 1) // 2
        self.npairs_kept = len(pairs)
        self._dev = None

    @property
    def dev(self):
        """The pair list on the device (uploaded on first use)."""
        if self._dev is None:
            t = _dev.torch()
            self._dev = _dev.to_device(self.pairs.reshape(-1), dtype=t.int32) if \
                self.npairs_kept else t.zeros(2, dtype=t.int32, device=_dev.device())
        return self._dev


def _lp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_long))


def int3c2e_chunk(ao, aux, tasks, out, ldo, col_map=None):
    """Fill ``out[P, p - p0]`` for the chunk's pairs (out pre-zeroed), or
    ``out[P, col_map[p - p0]]`` with a device int32 column map (a cderi that
    stores only the screened-in pairs)."""
    if ao.desc.lmax > LMAX_AO or aux.desc.lmax > LMAX_AUX:
        raise ValueError(f'angular momentum beyond the generated classes '
                         f'(orbital {ao.desc.lmax} > {LMAX_AO} or auxiliary '
                         f'{aux.desc.lmax} > {LMAX_AUX})')
    _lib.call('mdf_int3c2e', ctypes.byref(ao.desc), ctypes.byref(aux.desc),
              _dev.ptr(tasks.dev), _lp(tasks.class_off), _dev.ptr(aux.sorted),
              _lp(aux.l_off), tasks.p0, _dev.ptr(col_map), _dev.ptr(out), ldo, _dev.stream_ptr())


def int2c2e(aux, out=None):
    """The metric ``(P|Q)`` as a ``(naux, naux)`` device tensor."""
    t = _dev.torch()
    n = aux.nao
    if out is None:
        out = t.empty((n, n), dtype=t.float64, device=_dev.device())
    _lib.call('mdf_int2c2e', ctypes.byref(aux.desc), _dev.ptr(aux.sorted),
              _lp(aux.l_off), _dev.ptr(out), out.stride(0), _dev.stream_ptr())
    return out


def int1e(ao, mol, bound=None):
    """``(S, T, V)`` device matrices of the orbital basis of ``mol``."""
    t = _dev.torch()
    n = ao.nao
    tasks = PairTasks(mol.ao, 0, mol.ao.nbas, bound)
    mats = [t.zeros((n, n), dtype=t.float64, device=_dev.device()) for _ in range(3)]
    z = _dev.to_device(mol.atom_charges())
    xyz = _dev.to_device(mol.atom_coords().reshape(-1))
    _lib.call('mdf_int1e', ctypes.byref(ao.desc), n, _dev.ptr(tasks.dev), _lp(tasks.class_off),
              len(mol.symbols), _dev.ptr(z), _dev.ptr(xyz), _dev.ptr(mats[0]), _dev.ptr(mats[1]),
              _dev.ptr(mats[2]), _dev.stream_ptr())
    return tuple(mats)
