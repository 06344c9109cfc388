"""Seeded water-cluster geometries (optionally with a Li+ ion).

Recipe, fixed so every run and every test builds the same molecules:

* each water is rigid, O-H 0.9572 A, H-O-H 104.52 degrees, and oriented by a
  uniform random unit quaternion (a normalised 4-vector of N(0,1) draws);
* oxygen positions are uniform in a cube and rejection-sampled so that the
  minimum O-O distance is at least 2.6 A;
* the cube side is given, or set from a density:
  ``side = (n * 18.015 / N_A / rho)^(1/3)``  (12.42 A for 64 waters at
  1.0 g/cm^3, 9.86 A for 32);
* with ``li=True`` a Li+ ion (charge +1, closed shell) sits at the centroid of
  the oxygens; a draw that puts an oxygen within 1.8 A (or a hydrogen within
  1.2 A) of it is discarded and redrawn from the same generator.

All positions come from ``numpy.random.default_rng(seed)``: first the oxygen
positions (in order, with rejections), then one quaternion per water.
"""
import numpy as np

from .mole import Mole

R_OH = 0.9572
THETA_HOH = np.deg2rad(104.52)
N_AVOGADRO = 6.02214076e23
M_WATER = 18.015


def cube_side(n, density=1.0):
    """Cube edge in angstrom holding ``n`` waters at ``density`` g/cm^3."""
    vol_cm3 = n * M_WATER / N_AVOGADRO / density
    return vol_cm3 ** (1. / 3.) * 1e8


def _quat_to_rot(q):
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _water_frame():
    h = THETA_HOH / 2
    return np.array([[R_OH * np.sin(h), 0., R_OH * np.cos(h)],
                     [-R_OH * np.sin(h), 0., R_OH * np.cos(h)]])


def water_cluster_atoms(n, seed=0, side=None, density=1.0, min_oo=2.6,
                        li=False, max_tries=1000000, max_resample=10000):
    """Atoms ``[(symbol, xyz_angstrom)]`` of an ``n``-water cluster."""
    if li and n < 2:
        raise ValueError('a Li+ at the O centroid needs at least 2 waters '
                         '(with one it would sit on the oxygen)')
    if side is None:
        side = cube_side(n, density)
    rng = np.random.default_rng(seed)
    frame = _water_frame()
    for _ in range(max_resample):
        opos = []
        tries = 0
        while len(opos) < n:
            p = rng.uniform(0., side, 3)
            tries += 1
            if tries > max_tries:
                raise RuntimeError(f'could not place {n} waters in a {side} A '
                                   f'cube at min O-O {min_oo} A')
            if opos and np.min(np.linalg.norm(np.array(opos) - p, axis=1)) \
                    < min_oo:
                continue
            opos.append(p)
        opos = np.array(opos)
        atoms = []
        hpos = []
        for i in range(n):
            q = rng.standard_normal(4)
            q /= np.linalg.norm(q)
            h = frame @ _quat_to_rot(q).T + opos[i]
            atoms.append(('O', tuple(opos[i])))
            atoms.append(('H', tuple(h[0])))
            atoms.append(('H', tuple(h[1])))
            hpos.extend(h)
        if not li:
            return atoms
        c = opos.mean(axis=0)
        if (np.min(np.linalg.norm(opos - c, axis=1)) < 1.8 or
                np.min(np.linalg.norm(np.array(hpos) - c, axis=1)) < 1.2):
            continue
        return atoms + [('Li', tuple(c))]
    raise RuntimeError(f'no placement of Li+ among {n} waters in {max_resample} '
                       'resamples (O-Li >= 1.8 A, H-Li >= 1.2 A)')


def water_cluster(n, seed=0, side=None, density=1.0, li=False,
                  basis='def2-svp', auxbasis='etjk'):
    """A :class:`Mole` for an ``n``-water cluster (+ Li+ if ``li``)."""
    atoms = water_cluster_atoms(n, seed=seed, side=side, density=density,
                                li=li)
    return Mole(atoms, basis=basis, auxbasis=auxbasis, charge=1 if li else 0)
