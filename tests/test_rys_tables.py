"""The shipped Rys tables (csrc/rys_tables.inc) reproduce the quadrature:
sum_i w_i u_i^k = F_k(x) for k < 2n, checked with mpmath Boys functions at
points not used to build the tables (no GPU)."""
import os
import re

import mpmath as mp
import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, 'electrolites_amd', 'csrc', 'rys_tables.inc')


def _arr(txt, name):
    m = re.search(name + r'\[[^\]]*\]\s*=\s*\{([^}]*)\}', txt, re.S)
    return np.array([float(v) for v in m.group(1).replace('\n', ' ').split(',') if v.strip()])


@pytest.fixture(scope='module')
def tables():
    txt = open(INC).read()
    deg = int(re.search(r'#define RYS_DEG (\d+)', txt).group(1))
    return dict(deg=deg, xmax=_arr(txt, 'RYS_XMAX').astype(int), off=_arr(txt, 'RYS_OFFSET').astype(int),
                lxu=_arr(txt, 'RYS_LX_U'), lxw=_arr(txt, 'RYS_LX_W'), tab=_arr(txt, 'RYS_TABLE'))


def roots(T, n, x):
    if x >= T['xmax'][n]:
        o = n * (n - 1) // 2
        return T['lxu'][o:o + n] / x, T['lxw'][o:o + n] / np.sqrt(x)
    it = int(x)
    s = 2 * (x - it) - 1
    d1 = T['deg'] + 1
    base = T['off'][n] + it * 2 * n * d1
    vals = []
    for f in range(2 * n):
        c = T['tab'][base + f * d1: base + (f + 1) * d1]
        b1 = b2 = 0.
        for cj in c[:0:-1]:
            b1, b2 = 2 * s * b1 - b2 + cj, b1
        vals.append(s * b1 - b2 + c[0])
    return np.array(vals[:n]), np.array(vals[n:])


def boys(k, x):
    mp.mp.dps = 40
    if x == 0:
        return 1. / (2 * k + 1)
    a = mp.mpf(k) + mp.mpf(1) / 2
    return float(mp.gammainc(a, 0, x) / (2 * mp.mpf(x) ** a))


@pytest.mark.parametrize('n', range(1, 7))
def test_moments(tables, n):
    rng = np.random.default_rng(100 + n)
    xs = list(rng.uniform(0, tables['xmax'][n] + 40, 25)) + [0., 1e-9, float(tables['xmax'][n]) - 1e-9,
                                                             float(tables['xmax'][n])]
    for x in xs:
        u, w = roots(tables, n, x)
        assert np.all(np.diff(u) > 0) and np.all((u > 0) & (u < 1)) and np.all(w > 0)
        for k in range(2 * n):
            ref = boys(k, x)
            got = float(np.sum(w * u ** k))
            assert abs(got - ref) <= 5e-14 * ref, (n, x, k, got, ref)
