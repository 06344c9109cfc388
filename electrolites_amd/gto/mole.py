"""Molecule and shell tables, in PySCF's conventions.

The density-fitted path needs two basis sets on the same atoms (orbital and
auxiliary).  Both are flattened into plain arrays that the HIP library (and
the CPU oracle) read directly:

    shl_l[s], shl_atom[s], shl_nprim[s], shl_prim_off[s], shl_ao_off[s]
    exps[k], coefs[k]                (primitive data, normalised)
    shl_xyz[s, 3]                    (shell centre, bohr)

Conventions follow PySCF (`pyscf/gto/mole.py`, `libcint`), so that matrices
are interchangeable with a PySCF/GPU4PySCF run on the same basis:

* primitives carry ``gto_norm(l, a)`` and each contracted radial function is
  normalised, ``int r^(2l+2) R(r)^2 dr = 1``;
* spherical functions are ``R(r) r^l Y_lm`` with unit-normalised real
  harmonics (no Condon-Shortley phase), ordered ``m = -l..l`` except ``p``,
  which is ``(x, y, z)``;
* Cartesian components are ordered ``lx`` descending, then ``ly``
  descending (``xx, xy, xz, yy, yz, zz`` for ``d``).
"""
import math
from functools import lru_cache

import numpy as np

from . import basis as _basis

BOHR = 0.52917721092          # angstrom per bohr (PySCF's value)

ELEMENTS = {'H': 1, 'He': 2, 'Li': 3, 'Be': 4, 'B': 5, 'C': 6, 'N': 7,
            'O': 8, 'F': 9, 'Ne': 10}


def ncart(l):
    return (l + 1) * (l + 2) // 2


def cart_powers(l):
    """Cartesian exponent triples in libcint order."""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def gaussian_int(n, a):
    """``int_0^inf r^n exp(-a r^2) dr``."""
    return math.gamma((n + 1) * .5) / (2. * a ** ((n + 1) * .5))


def gto_norm(l, a):
    """Radial normalisation of ``r^l exp(-a r^2)``."""
    return 1. / math.sqrt(gaussian_int(2 * l + 2, 2 * a))


def _solid_harmonic(l, m):
    """Real solid harmonic ``r^l Y_lm`` (unit-normalised ``Y``) as a dict
    ``{(lx, ly, lz): coefficient}``.  Racah-normalised form from Helgaker,
    Jorgensen & Olsen, *Molecular Electronic-Structure Theory*, eq. 6.4.47-50,
    rescaled by ``sqrt((2l+1)/4pi)``."""
    am = abs(m)
    vm = 0. if m >= 0 else .5
    nlm = (1. / (2 ** am * math.factorial(l)) *
           math.sqrt(2. * math.factorial(l + am) * math.factorial(l - am) /
                     (2. if m == 0 else 1.)))
    out = {}
    for t in range((l - am) // 2 + 1):
        for u in range(t + 1):
            v = vm
            vmax = math.floor(am / 2. - vm) + vm
            while v <= vmax + 1e-9:
                two_v = int(round(2 * v))
                c = ((-1) ** int(round(t + v - vm)) * .25 ** t *
                     math.comb(l, t) * math.comb(l - t, am + t) *
                     math.comb(t, u) * math.comb(am, two_v))
                lx = 2 * t + am - 2 * u - two_v
                ly = 2 * u + two_v
                lz = l - 2 * t - am
                key = (lx, ly, lz)
                out[key] = out.get(key, 0.) + c * nlm
                v += 1.
    scale = math.sqrt((2 * l + 1) / (4 * math.pi))
    return {k: v * scale for k, v in out.items() if v != 0.}


@lru_cache(maxsize=None)
def cart2sph(l):
    """``(2l+1, ncart(l))`` matrix: spherical = C @ cartesian monomials."""
    ms = [1, -1, 0] if l == 1 else list(range(-l, l + 1))
    powers = cart_powers(l)
    c = np.zeros((2 * l + 1, len(powers)))
    for i, m in enumerate(ms):
        for key, val in _solid_harmonic(l, m).items():
            c[i, powers.index(key)] = val
    c.setflags(write=False)
    return c


def _normalise_shell(l, prims):
    exps = np.array([p[0] for p in prims], dtype=np.float64)
    raw = np.array([p[1] for p in prims], dtype=np.float64)
    c = raw * np.array([gto_norm(l, a) for a in exps])
    s = 0.
    for i, ai in enumerate(exps):
        for j, aj in enumerate(exps):
            s += c[i] * c[j] * gaussian_int(2 * l + 2, ai + aj)
    return exps, c / math.sqrt(s)


class ShellTable:
    """Flattened shells of one basis on one set of atoms."""

    def __init__(self, atom_xyz, atom_symbols, basis_name):
        per_el = _basis.load(basis_name, sorted(set(atom_symbols)))
        l_, atom_, nprim_, off_, aoff_ = [], [], [], [], []
        exps, coefs = [], []
        ao = 0
        for ia, el in enumerate(atom_symbols):
            for l, prims in per_el[el]:
                e, c = _normalise_shell(l, prims)
                l_.append(l)
                atom_.append(ia)
                nprim_.append(len(e))
                off_.append(len(exps))
                aoff_.append(ao)
                exps.extend(e)
                coefs.extend(c)
                ao += 2 * l + 1
        self.shl_l = np.array(l_, dtype=np.int32)
        self.shl_atom = np.array(atom_, dtype=np.int32)
        self.shl_nprim = np.array(nprim_, dtype=np.int32)
        self.shl_prim_off = np.array(off_, dtype=np.int32)
        self.shl_ao_off = np.array(aoff_, dtype=np.int32)
        self.exps = np.array(exps, dtype=np.float64)
        self.coefs = np.array(coefs, dtype=np.float64)
        self.shl_xyz = np.ascontiguousarray(
            np.asarray(atom_xyz, dtype=np.float64)[self.shl_atom])
        self.nbas = len(l_)
        self.nao = ao
        self.lmax = int(self.shl_l.max()) if len(l_) else 0
        self.ao_loc = np.append(self.shl_ao_off, ao).astype(np.int32)

    @classmethod
    def from_shells(cls, atom_xyz, shells):
        """A table from ``(atom, l, exps, coefs)`` tuples whose coefficients
        are already normalised (libcint's stored form)."""
        self = cls.__new__(cls)
        l_, atom_, nprim_, off_, aoff_, exps, coefs = [], [], [], [], [], [], []
        ao = 0
        for ia, l, e, c in shells:
            l_.append(l)
            atom_.append(ia)
            nprim_.append(len(e))
            off_.append(len(exps))
            aoff_.append(ao)
            exps.extend(e)
            coefs.extend(c)
            ao += 2 * l + 1
        self.shl_l = np.array(l_, dtype=np.int32)
        self.shl_atom = np.array(atom_, dtype=np.int32)
        self.shl_nprim = np.array(nprim_, dtype=np.int32)
        self.shl_prim_off = np.array(off_, dtype=np.int32)
        self.shl_ao_off = np.array(aoff_, dtype=np.int32)
        self.exps = np.array(exps, dtype=np.float64)
        self.coefs = np.array(coefs, dtype=np.float64)
        self.shl_xyz = np.ascontiguousarray(
            np.asarray(atom_xyz, dtype=np.float64).reshape(-1, 3)[self.shl_atom])
        self.nbas = len(l_)
        self.nao = ao
        self.lmax = int(self.shl_l.max()) if len(l_) else 0
        self.ao_loc = np.append(self.shl_ao_off, ao).astype(np.int32)
        return self

    def as_dict(self):
        """Plain arrays, for code that must not depend on this class."""
        return dict(l=self.shl_l, atom=self.shl_atom, nprim=self.shl_nprim,
                    prim_off=self.shl_prim_off, ao_off=self.shl_ao_off,
                    exps=self.exps, coefs=self.coefs, xyz=self.shl_xyz,
                    nao=self.nao)


class Mole:
    """A molecule with an orbital basis and (optionally) a fitting basis.

    ``atoms`` is a list of ``(symbol, (x, y, z))``; ``unit`` is 'angstrom' or
    'bohr'.  Mirrors the parts of ``pyscf.gto.Mole`` the DF path reads:
    ``nao``, ``ao_loc``, ``atom_coords()``, ``atom_charges()``,
    ``nelectron``, ``energy_nuc()``.
    """

    def __init__(self, atoms, basis='def2-svp', auxbasis='etjk', charge=0,
                 spin=0, unit='angstrom'):
        self.symbols = [a[0] for a in atoms]
        xyz = np.array([a[1] for a in atoms], dtype=np.float64).reshape(-1, 3)
        if unit.lower().startswith('a'):
            xyz = xyz / BOHR
        self._xyz = xyz
        self.charge = charge
        self.spin = spin
        self.basis = basis
        self.auxbasis = auxbasis
        self.ao = ShellTable(xyz, self.symbols, basis)
        self.aux = ShellTable(xyz, self.symbols, auxbasis) if auxbasis else None

    @property
    def natm(self):
        return len(self.symbols)

    @property
    def nao(self):
        return self.ao.nao

    @property
    def naux(self):
        return self.aux.nao

    @property
    def ao_loc(self):
        return self.ao.ao_loc

    def atom_coords(self):
        return self._xyz.copy()

    def atom_charges(self):
        return np.array([ELEMENTS[s] for s in self.symbols], dtype=np.float64)

    @property
    def nelectron(self):
        return int(round(self.atom_charges().sum())) - self.charge

    @property
    def nocc(self):
        assert self.spin == 0 and self.nelectron % 2 == 0
        return self.nelectron // 2

    def energy_nuc(self):
        z = self.atom_charges()
        r = self._xyz
        e = 0.
        for i in range(self.natm):
            for j in range(i):
                e += z[i] * z[j] / np.linalg.norm(r[i] - r[j])
        return e

    def npair(self):
        n = self.nao
        return n * (n + 1) // 2
