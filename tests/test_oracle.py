"""Pins of the CPU oracle (no GPU).

The reference holds no golden vectors for this path (its suites compare
against GPU4PySCF at run time, tests/test_df.py:58), so every oracle function
is pinned by mathematics the reference's algorithm must satisfy: closed forms,
mpmath, finite-difference identities, invariances, library routines, and
PySCF's published constants.
"""
import math

import mpmath as mp
import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O

mp.mp.dps = 40


def boys_mp(m, t):
    if t == 0:
        return mp.mpf(1) / (2 * m + 1)
    a = mp.mpf(m) + mp.mpf(1) / 2
    return mp.gammainc(a, 0, t) / (2 * mp.mpf(t) ** a)


# ------------------------------------------------------------------ Boys
@pytest.mark.parametrize('t', [0., 1e-16, 1e-9, 1e-3, .37, 2., 7.5, 19., 33.,
                               44.99, 45.01, 60., 150., 800.])
def test_boys_vs_mpmath(t):
    F = O.boys(14, t)
    for m in range(15):
        ref = float(boys_mp(m, t))
        assert abs(F[m] - ref) <= 2e-15 * ref, (m, t, F[m], ref)


def test_boys_closed_forms():
    for t in (.5, 3., 30.):
        F = O.boys(3, t)
        assert math.isclose(F[0], .5 * math.sqrt(math.pi / t) * math.erf(math.sqrt(t)), rel_tol=1e-15)
        for m in range(3):   # downward recursion
            assert math.isclose(F[m], (2 * t * F[m + 1] + math.exp(-t)) / (2 * m + 1), rel_tol=1e-14)
    assert np.allclose(O.boys(8, 0.), [1. / (2 * m + 1) for m in range(9)], rtol=0, atol=0)


# ---------------------------------------------------------- cart -> sph
def sphere_gram(l):
    from electrolites_amd.gto import cart_powers
    pw = cart_powers(l)
    g = np.zeros((len(pw), len(pw)))
    for i, a in enumerate(pw):
        for j, b in enumerate(pw):
            e = [a[k] + b[k] for k in range(3)]
            if any(x % 2 for x in e):
                continue
            h = [(x + 1) / 2 for x in e]
            g[i, j] = 2 * math.gamma(h[0]) * math.gamma(h[1]) * math.gamma(h[2]) / math.gamma(sum(h))
    return g


@pytest.mark.parametrize('l', range(6))
def test_cart2sph_orthonormal_and_agrees(l):
    from electrolites_amd.gto import cart2sph
    c = O.cart2sph(l)
    assert np.allclose(c @ sphere_gram(l) @ c.T, np.eye(2 * l + 1), atol=1e-14)
    # the oracle's C table and the package's Python table are two codes of the
    # same formula: they must agree entry for entry
    assert np.allclose(c, cart2sph(l), atol=1e-15, rtol=0)


def test_cart2sph_pyscf_constants():
    """PySCF's real-harmonic coefficients (libcint cart2sph tables) for d/f."""
    d = O.cart2sph(2)     # rows: xy, yz, z2, xz, x2-y2; cols xx xy xz yy yz zz
    assert math.isclose(d[0, 1], 1.092548430592079070, rel_tol=1e-15)
    assert math.isclose(d[2, 5], 0.630783130505040012, rel_tol=1e-15)
    assert math.isclose(d[2, 0], -0.315391565252520002, rel_tol=1e-15)
    assert math.isclose(d[4, 0], 0.546274215296039535, rel_tol=1e-15)
    f = O.cart2sph(3)     # cols xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
    assert math.isclose(f[0, 1], 1.770130769779930531, rel_tol=1e-15)
    assert math.isclose(f[0, 6], -0.590043589926643510, rel_tol=1e-15)
    assert math.isclose(f[1, 4], 2.890611442640554055, rel_tol=1e-15)
    assert math.isclose(f[3, 9], 0.746352665180230782, rel_tol=1e-15)
    assert math.isclose(f[5, 2], 1.445305721320277027, rel_tol=1e-15)


# ------------------------------------------------------- closed forms
def one_shell(ls, centers, exps):
    from electrolites_amd.gto.mole import gto_norm
    d = dict(l=np.array(ls, np.int32), nprim=np.ones(len(ls), np.int32),
             prim_off=np.arange(len(ls), dtype=np.int32), exps=np.array(exps, float),
             xyz=np.array(centers, float),
             coefs=np.array([gto_norm(l, a) for l, a in zip(ls, exps)]))
    ao, o = [], 0
    for l in ls:
        ao.append(o)
        o += 2 * l + 1
    d['ao_off'] = np.array(ao, np.int32)
    d['nao'] = o
    return d


def test_s_closed_forms():
    from electrolites_amd.gto.mole import gto_norm
    a, b, c = 1.3, .7, .9
    A, B, C = np.array([0., 0, 0]), np.array([.3, -.2, .9]), np.array([.4, .1, -.3])
    y00 = 1 / math.sqrt(4 * math.pi)
    na, nb, nc = (gto_norm(0, e) * y00 for e in (a, b, c))
    p = a + b
    mu = a * b / p
    r2 = np.sum((A - B) ** 2)
    K = math.exp(-mu * r2)
    P = (a * A + b * B) / p
    sh = one_shell([0, 0], [A, B], [a, b])
    S = O.int1e('ovlp', sh)
    assert math.isclose(S[0, 1], na * nb * (math.pi / p) ** 1.5 * K, rel_tol=1e-14)
    T = O.int1e('kin', sh)
    assert math.isclose(T[0, 1], mu * (3 - 2 * mu * r2) * S[0, 1], rel_tol=1e-14)
    Zc, Q = 2., np.array([.1, .5, -.4])
    V = O.int1e('nuc', sh, np.array([Zc]), Q[None])
    ref = -Zc * 2 * math.pi / p * K * float(boys_mp(0, p * np.sum((P - Q) ** 2))) * na * nb
    assert math.isclose(V[0, 1], ref, rel_tol=1e-14)
    j3 = O.int3c2e(sh, one_shell([0], [C], [c]))
    al = p * c / (p + c)
    ref = (2 * math.pi ** 2.5 / (p * c * math.sqrt(p + c)) * K *
           float(boys_mp(0, al * np.sum((P - C) ** 2))) * na * nb * nc)
    assert math.isclose(j3[0, 1, 0], ref, rel_tol=1e-14)
    j2 = O.int2c2e(one_shell([0, 0], [C, A], [c, a]))
    al = c * a / (c + a)
    ref = (2 * math.pi ** 2.5 / (c * a * math.sqrt(c + a)) *
           float(boys_mp(0, al * np.sum((A - C) ** 2))) * nc * na)
    assert math.isclose(j2[0, 1], ref, rel_tol=1e-14)


def _fd(fun, x0, h=1e-4):
    return (fun(x0 + h) - fun(x0 - h)) / (2 * h)


def test_p_functions_are_center_derivatives():
    """A primitive p_x Gaussian (x-A_x) e^{-a r_A^2} is (1/2a) d/dA_x of the s
    one: pins the angular machinery against the closed-form-pinned s case
    (a dropped term, wrong sign or index breaks it)."""
    from electrolites_amd.gto.mole import gto_norm
    a, b, c = 1.1, .6, .8
    A, B, C = np.array([.2, -.1, .3]), np.array([-.4, .5, .1]), np.array([.3, .2, -.6])
    y00, y1 = 1 / math.sqrt(4 * math.pi), math.sqrt(3 / (4 * math.pi))
    aux = one_shell([0], [C], [c])

    def s_int(Ax, kind):
        sh = one_shell([0, 0], [[Ax, A[1], A[2]], B], [a, b])
        if kind == '3c':
            return O.int3c2e(sh, aux)[0, 1, 0] / (gto_norm(0, a) * y00)
        if kind == 'nuc':
            return O.int1e('nuc', sh, np.array([1.]), C[None])[0, 1] / (gto_norm(0, a) * y00)
        return O.int1e(kind, sh)[0, 1] / (gto_norm(0, a) * y00)

    shp = one_shell([1, 0], [A, B], [a, b])
    np_ = gto_norm(1, a) * y1
    for kind in ('ovlp', 'kin', 'nuc', '3c'):
        if kind == '3c':
            px = O.int3c2e(shp, aux)[0, 3, 0]
        elif kind == 'nuc':
            px = O.int1e('nuc', shp, np.array([1.]), C[None])[0, 3]
        else:
            px = O.int1e(kind, shp)[0, 3]
        ref = _fd(lambda x: s_int(x, kind), A[0]) / (2 * a)
        assert math.isclose(px / np_, ref, rel_tol=2e-8), kind


def test_kinetic_d_function_second_derivative():
    """The kinetic integral's d_xy component is (1/4a^2) d^2/dA_x dA_y of the
    s one (the operator does not depend on A): pins the kinetic machinery
    for l = 2 beyond the closed-form s case."""
    from electrolites_amd.gto.mole import gto_norm
    a, b = .9, 1.3
    A, B = np.array([.1, .2, -.3]), np.array([.5, -.4, .2])
    y00 = 1 / math.sqrt(4 * math.pi)

    def s_int(ax, ay):
        sh = one_shell([0, 0], [[ax, ay, A[2]], B], [a, b])
        return O.int1e('kin', sh)[0, 1] / (gto_norm(0, a) * y00)

    h = 2e-4
    d2 = (s_int(A[0] + h, A[1] + h) - s_int(A[0] + h, A[1] - h) -
          s_int(A[0] - h, A[1] + h) + s_int(A[0] - h, A[1] - h)) / (4 * h * h)
    shd = one_shell([2, 0], [A, B], [a, b])
    dxy = O.int1e('kin', shd)[0, 5]             # m = -2 is d_xy
    coef = O.cart2sph(2)[0, 1] * gto_norm(2, a)
    assert math.isclose(dxy / coef, d2 / (4 * a * a), rel_tol=1e-6)


def test_d_function_second_derivative():
    """d_xy component: (x-A_x)(y-A_y) e^{-a r^2} = (1/4a^2) d^2/dA_x dA_y of s."""
    from electrolites_amd.gto.mole import gto_norm
    a, c = .9, .7
    A, B, C = np.array([.1, .2, -.3]), np.array([.5, -.4, .2]), np.array([-.2, .3, .4])
    aux = one_shell([0], [C], [c])
    y00 = 1 / math.sqrt(4 * math.pi)

    def s_int(ax, ay):
        sh = one_shell([0, 0], [[ax, ay, A[2]], B], [a, 1.3])
        return O.int3c2e(sh, aux)[0, 1, 0] / (gto_norm(0, a) * y00)

    h = 2e-4
    d2 = (s_int(A[0] + h, A[1] + h) - s_int(A[0] + h, A[1] - h) -
          s_int(A[0] - h, A[1] + h) + s_int(A[0] - h, A[1] - h)) / (4 * h * h)
    shd = one_shell([2, 0], [A, B], [a, 1.3])
    dxy = O.int3c2e(shd, aux)[0, 5, 0]         # m = -2 is d_xy
    coef = O.cart2sph(2)[0, 1] * gto_norm(2, a)
    assert math.isclose(dxy / coef, d2 / (4 * a * a), rel_tol=1e-6)


# ------------------------------------------------- molecule-level pins
@pytest.fixture(scope='module')
def small():
    from electrolites_amd.gto import water_cluster
    mol = water_cluster(2, seed=5, side=4.5, li=True)
    return mol


def test_normalisation(small):
    S = O.int1e('ovlp', small.ao.as_dict())
    assert np.abs(np.diag(S) - 1).max() < 1e-14
    assert np.abs(S - S.T).max() == 0.


def _rotate(mol, R):
    from electrolites_amd.gto import Mole, BOHR
    xyz = mol.atom_coords() @ R.T * BOHR + 1.7
    return Mole([(s, tuple(x)) for s, x in zip(mol.symbols, xyz)], charge=mol.charge)


def test_rotation_translation_invariants(small):
    """Spectra of the metric and of the three-centre matrix are invariant
    under a rigid motion (the spherical functions of a shell mix
    orthogonally)."""
    rng = np.random.default_rng(9)
    R = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    other = _rotate(small, R)
    for m in (small, other):
        m._j2c = O.int2c2e(m.aux.as_dict())
        m._j3 = O.int3c2e(m.ao.as_dict(), m.aux.as_dict()).reshape(m.nao * m.nao, -1)
    e1, e2 = (np.linalg.eigvalsh(m._j2c) for m in (small, other))
    assert np.allclose(e1, e2, rtol=1e-11, atol=1e-13)
    s1, s2 = (np.linalg.svd(m._j3, compute_uv=False) for m in (small, other))
    assert np.allclose(s1, s2, rtol=1e-10, atol=1e-12)


def test_df_identity(small):
    """sum_L cderi[L,p] cderi[L,q] = (p|P) (P|Q)^-1 (Q|q)."""
    j3 = O.int3c2e(small.ao.as_dict(), small.aux.as_dict())
    j2 = O.int2c2e(small.aux.as_dict())
    pa, _ = O.pair_addresses(small.nao)
    cd = O.cholesky_eri(j3, j2, pa)
    rows, cols = np.divmod(pa, small.nao)
    b = j3[rows, cols, :]
    lhs = cd.T @ cd
    rhs = b @ np.linalg.solve(j2, b.T)
    assert np.abs(lhs - rhs).max() < 1e-9 * np.abs(rhs).max()


def test_jk_identities():
    rng = np.random.default_rng(2)
    nao, naux, nocc = 23, 41, 6
    pa, _ = O.pair_addresses(nao)
    cd = rng.standard_normal((naux, len(pa)))
    c = np.linalg.qr(rng.standard_normal((nao, nocc)))[0]
    dm = 2 * c @ c.T
    vj, vk = O.get_jk(cd, pa, dm)
    a = O.unpack(cd, pa, nao)
    # J from the definition sum_P A_P tr(A_P D)
    assert np.allclose(vj, np.einsum('Pij,Pkl,kl->ij', a, a, dm), rtol=1e-13, atol=1e-12)
    # sum_p d_p vj_p = ||rho||^2
    d = O.sparse_dm(dm, pa, 1)[0]
    rho = cd @ d
    assert math.isclose(d @ (cd.T @ rho), rho @ rho, rel_tol=1e-13)
    # factored K equals the direct K; K symmetric; tr(D K) >= 0 for PSD D
    assert np.allclose(vk, O.get_k_factor(cd, pa, c * math.sqrt(2)), rtol=1e-13, atol=1e-12)
    assert np.abs(vk - vk.T).max() < 1e-12 * np.abs(vk).max()
    assert np.sum(dm * vk) >= 0
    # hermi 0 on a symmetric density equals hermi 1
    assert np.allclose(O.get_j(cd, pa, dm, 0), vj, rtol=1e-14, atol=1e-13)


def test_sparse_dm_symmetry_factors():
    dm = np.array([[1., 2.], [3., 4.]])
    pa, _ = O.pair_addresses(2)       # (0,0), (1,0), (1,1)
    assert np.array_equal(O.sparse_dm(dm, pa, 1)[0], [1., 6., 4.])
    assert np.array_equal(O.sparse_dm(dm, pa, 0)[0], [1., 5., 4.])


def test_water_rhf_energy_and_stationarity():
    """DF-RHF of one water: literature-scale energy, converged to 1e-10,
    guess-independent, and a stationary idempotent density."""
    from electrolites_amd.gto import Mole
    m = Mole([('O', (0., 0., 0.)), ('H', (.7571, 0., .5861)), ('H', (-.7571, 0., .5861))])
    s = O.DFSystem(m.ao.as_dict(), m.aux.as_dict(), m.atom_charges(), m.atom_coords(),
                   m.nelectron, m.energy_nuc())
    r = O.rhf(s, conv_tol=1e-11)
    assert r['converged']
    assert abs(r['e_tot'] - (-75.9610)) < 2e-3      # RHF/def2-SVP water ~ -75.961
    dm = r['dm']
    assert np.abs(dm @ s.S @ dm - 2 * dm).max() < 1e-9
    vj, vk = O.get_jk(s.cderi, s.pair_addr, dm)
    F = s.H + vj - .5 * vk
    assert np.abs(F @ dm @ s.S - s.S @ dm @ F).max() < 1e-6
    # a different starting density converges to the same energy
    s.H_guess = s.H + .3 * np.diag(np.linspace(0, 1, s.nao))
    e2 = O.rhf(_with_guess(s), conv_tol=1e-11)['e_tot']
    assert abs(e2 - r['e_tot']) < 1e-9


def _with_guess(s):
    class S2:
        pass
    t = S2()
    t.__dict__.update(s.__dict__)
    return t


def test_nuclear_repulsion_closed_form():
    from electrolites_amd.gto import Mole, BOHR
    m = Mole([('H', (0, 0, 0)), ('H', (0, 0, .74))])
    assert math.isclose(m.energy_nuc(), BOHR / .74, rel_tol=1e-14)
