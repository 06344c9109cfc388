"""Shell tables from PySCF's libcint arrays (``mol._atm``, ``mol._bas``,
``mol._env``), so that the build runs on the molecules a GPU4PySCF session
hands it (``Int3c2eOpt.mol.mol`` and ``Int3c2eOpt.auxmol.mol``, the original
orbital and auxiliary molecules the reference's ``_cholesky_eri`` reads,
electrolites/fastdfcderi.py:377-386).

Layout (libcint's public ``cint.h`` slots, which PySCF fills in
``pyscf/gto/mole.py:make_env``):

* ``_atm[ia] = (charge, ptr_coord, nuc_mod, ptr_zeta, ptr_frac, reserved)``,
  coordinates (bohr) at ``_env[ptr_coord : ptr_coord + 3]``;
* ``_bas[ib] = (atom, l, nprim, nctr, kappa, ptr_exp, ptr_coeff, reserved)``,
  exponents at ``_env[ptr_exp : ptr_exp + nprim]`` and coefficients at
  ``_env[ptr_coeff + ictr * nprim + iprim]``, already carrying
  ``gto_norm(l, a)`` and the contracted normalisation -- this package's
  convention for ``ShellTable.coefs``.

A shell with ``nctr`` contractions is ``nctr`` shells here, in order (the
AO order libcint gives them: every m of the first contraction, then the
next).  Spinor shells (``kappa != 0``) and Cartesian molecules are not
covered.
"""
import numpy as np

from .mole import ShellTable

# libcint slot indices
CHARGE_OF, PTR_COORD = 0, 1
ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF, PTR_EXP, PTR_COEFF = range(7)


def shell_table(atm, bas, env, cart=False):
    """A :class:`ShellTable` holding the spherical shells of the arrays."""
    if cart:
        raise NotImplementedError('Cartesian basis functions are not covered')
    atm = np.asarray(atm, dtype=np.int64).reshape(-1, 6)
    bas = np.asarray(bas, dtype=np.int64).reshape(-1, 8)
    env = np.asarray(env, dtype=np.float64)
    xyz = np.array([env[p:p + 3] for p in atm[:, PTR_COORD]], dtype=np.float64).reshape(-1, 3)
    shells = []
    for row in bas:
        if row[KAPPA_OF] != 0:
            raise NotImplementedError('spinor shells (kappa != 0) are not covered')
        l, nprim, nctr = int(row[ANG_OF]), int(row[NPRIM_OF]), int(row[NCTR_OF])
        exps = env[row[PTR_EXP]:row[PTR_EXP] + nprim]
        cs = env[row[PTR_COEFF]:row[PTR_COEFF] + nprim * nctr].reshape(nctr, nprim)
        for ic in range(nctr):
            shells.append((int(row[ATOM_OF]), l, exps.copy(), cs[ic].copy()))
    return ShellTable.from_shells(xyz, shells)


def from_mole(mol):
    """:func:`shell_table` of a PySCF-style molecule (``_atm``, ``_bas``,
    ``_env``, ``cart``)."""
    return shell_table(mol._atm, mol._bas, mol._env, bool(getattr(mol, 'cart', False)))


class LibcintPair:
    """The orbital and auxiliary molecules of a GPU4PySCF session as this
    package's build reads them: ``ao``/``aux`` shell tables, ``nao``,
    ``naux``, ``natm``, ``atom_coords()``; it also holds the build's caches
    (shell-pair bounds, device shell tables), as this package's ``Mole``
    does."""

    def __init__(self, mol, auxmol):
        self.pyscf_mol = mol
        self.ao = from_mole(mol)
        self.aux = from_mole(auxmol)
        coords = getattr(mol, 'atom_coords', None)
        self._xyz = np.asarray(coords() if callable(coords) else
                               self.ao.shl_xyz[:0], dtype=np.float64)

    @property
    def nao(self):
        return self.ao.nao

    @property
    def naux(self):
        return self.aux.nao

    @property
    def natm(self):
        return len(self._xyz)

    def atom_coords(self):
        return self._xyz.copy()
