"""Basis-set data for the density-fitted path.

Two sets are needed by the workloads this package is measured on (water
clusters, with and without a Li+ ion):

* ``def2-svp`` -- the orbital basis, Weigend & Ahlrichs, PCCP 7, 3297 (2005),
  entered by hand for H, Li and O in the published (Turbomole/NWChem) form:
  raw contraction coefficients, normalised at load time the way PySCF does.
  Water gives 24 spherical functions (O [3s2p1d] + 2 x H [2s1p]).

* ``etjk`` -- the auxiliary (fitting) basis.  The published
  def2-universal-jkfit tables are not available offline, so this is an
  even-tempered JK fitting set with the same shell structure per element
  (O [10s7p5d2f1g], H [4s2p1d], Li [10s6p4d2f]) whose exponent ranges span the
  products of the def2-SVP primitives.  A water contributes 109 auxiliary
  functions.  J and K are invariant to any invertible transformation of the
  fitting set, so every parity statement made with it is a statement about the
  algorithm, not about the set; ``parse_nwchem`` accepts the published
  def2-universal-jkfit text for a user who has it.

A basis is ``{element: [(l, [(exponent, coefficient), ...]), ...]}`` with raw
(un-normalised) coefficients; ``mole.py`` normalises.
"""
import re

import numpy as np

ANGULAR = 'spdfghi'

DEF2_SVP = {
    'H': [
        (0, [(13.010701, 0.19682158e-01),
             (1.9622572, 0.13796524),
             (0.44453796, 0.47831935)]),
        (0, [(0.12194962, 1.0)]),
        (1, [(0.8000000, 1.0)]),
    ],
    'Li': [
        (0, [(266.27785516, 0.64920150325e-02),
             (40.069783447, 0.47747863215e-01),
             (9.0559944389, 0.20268796111),
             (2.4503009051, 0.48606574817),
             (0.72209571855, 0.43626977955)]),
        (0, [(0.52810884721e-01, 1.0)]),
        (0, [(0.20960948798e-01, 1.0)]),
        (1, [(1.4500000, 0.25860000000),
             (0.30000000, 1.0)]),
        (1, [(0.82000000e-01, 1.0)]),
    ],
    'O': [
        (0, [(2266.1767785, -0.53431809926e-02),
             (340.87010191, -0.39890039230e-01),
             (77.363135167, -0.17853911985),
             (21.479644940, -0.46427684959),
             (6.6589433124, -0.44309745172)]),
        (0, [(0.80975975668, 1.0)]),
        (0, [(0.25530772234, 1.0)]),
        (1, [(17.721504317, 0.43394573193e-01),
             (3.8635505440, 0.23094120765),
             (1.0480920883, 0.51375311064)]),
        (1, [(0.27641544411, 1.0)]),
        (2, [(1.2000000, 1.0)]),
    ],
}


def even_tempered(lo, hi, n):
    """``n`` exponents in geometric progression from ``hi`` down to ``lo``."""
    if n == 1:
        return [float(np.sqrt(lo * hi))]
    return [float(x) for x in np.geomspace(hi, lo, n)]


def _et_shells(spec):
    out = []
    for l, (lo, hi, n) in enumerate(spec):
        for e in even_tempered(lo, hi, n):
            out.append((l, [(e, 1.0)]))
    return out


# (lowest exponent, highest exponent, count) per angular momentum
_ETJK_SPEC = {
    'H': [(0.22, 15.0, 4), (0.5, 2.0, 2), (1.4, 1.4, 1)],
    'Li': [(0.04, 500.0, 10), (0.06, 8.0, 6), (0.08, 3.0, 4), (0.15, 0.6, 2)],
    'O': [(0.45, 3600.0, 10), (0.4, 60.0, 7), (0.45, 20.0, 5), (0.9, 3.0, 2),
          (2.0, 2.0, 1)],
}

ETJK = {el: _et_shells(spec) for el, spec in _ETJK_SPEC.items()}

_LIBRARY = {
    'def2-svp': DEF2_SVP,
    'etjk': ETJK,
    # the fitting set used in place of the published def2-universal-jkfit
    'def2-universal-jkfit-et': ETJK,
}


def parse_nwchem(text):
    """Parse NWChem-format basis text into ``{element: [(l, prims)]}``.

    Handles ``SP`` shells (split into an s and a p shell) and Fortran ``D``
    exponents; comments start with ``#``.
    """
    basis = {}
    cur = None
    for raw in text.splitlines():
        line = raw.split('#', 1)[0].strip()
        if not line or line.upper().startswith(('BASIS', 'END')):
            continue
        tok = line.split()
        if re.match(r'^[A-Za-z]{1,2}$', tok[0]) and len(tok) == 2 and \
                re.match(r'^[A-Za-z]+$', tok[1]):
            el = tok[0].capitalize()
            kind = tok[1].upper()
            cur = []
            basis.setdefault(el, [])
            if kind == 'SP':
                basis[el].append((0, cur))
                cur_p = []
                basis[el].append((1, cur_p))
                cur = (cur, cur_p)
            else:
                basis[el].append((ANGULAR.index(kind.lower()), cur))
            continue
        nums = [float(t.replace('D', 'E').replace('d', 'e')) for t in tok]
        if isinstance(cur, tuple):
            cur[0].append((nums[0], nums[1]))
            cur[1].append((nums[0], nums[2]))
        else:
            if len(nums) > 2:
                raise ValueError('general contractions are not supported: '
                                 f'{raw!r}')
            cur.append((nums[0], nums[1]))
    return basis


def load(name, elements):
    """The shells of basis ``name`` (a library name, a dict, or NWChem text)
    for each element in ``elements``."""
    if isinstance(name, dict):
        lib = name
    elif name.lower() in _LIBRARY:
        lib = _LIBRARY[name.lower()]
    else:
        lib = parse_nwchem(name)
    out = {}
    for el in elements:
        if el not in lib:
            raise KeyError(f'basis {name!r} has no data for {el}')
        out[el] = lib[el]
    return out
