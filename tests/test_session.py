"""GPU4PySCF-session inputs: PySCF's libcint arrays, the ``Int3c2eOpt``
shape of ``_cholesky_eri``'s argument, and the OOM release-and-retry of
``get_jk`` (electrolites/fastdf.py:754-803).

``libcint_arrays`` builds ``_atm``/``_bas``/``_env`` the way PySCF's
``make_env`` does (pyscf/gto/mole.py): raw coefficients times ``gto_norm``,
then each contraction normalised by the radial overlap, stored contraction
by contraction; coordinates in bohr.  Its normalisation is written here from
the formulas, not taken from the package.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _gaussian_int(n, a):
    return math.gamma((n + 1) / 2.) / (2. * a ** ((n + 1) / 2.))


def _pyscf_coeffs(l, exps, raw):
    """PySCF's stored coefficients (raw: (nprim, nctr))."""
    norm = np.array([1. / math.sqrt(_gaussian_int(2 * l + 2, 2 * a)) for a in exps])
    cs = raw * norm[:, None]
    ee = np.array([[_gaussian_int(2 * l + 2, a + b) for b in exps] for a in exps])
    s = np.einsum('pi,pq,qi->i', cs, ee, cs)
    return cs / np.sqrt(s)


def libcint_arrays(xyz_bohr, charges, shells):
    """``(atm, bas, env)`` for ``shells`` = [(atom, l, exps, raw (nprim, nctr))]."""
    env = [0.] * 20                          # PySCF reserves the first 20 slots
    atm = []
    for z, r in zip(charges, xyz_bohr):
        atm.append([int(z), len(env), 1, len(env) + 3, 0, 0])
        env.extend(list(r) + [0.])
    bas = []
    for ia, l, exps, raw in shells:
        raw = np.asarray(raw, dtype=float).reshape(len(exps), -1)
        pe = len(env)
        env.extend(exps)
        pc = len(env)
        env.extend(_pyscf_coeffs(l, np.asarray(exps), raw).T.reshape(-1))
        bas.append([ia, l, len(exps), raw.shape[1], 0, pe, pc, 0])
    return (np.array(atm, dtype=np.int32), np.array(bas, dtype=np.int32),
            np.array(env, dtype=np.float64))


def _cluster_libcint(mol, which):
    """libcint arrays of this package's molecule, from the basis data."""
    from electrolites_amd.gto import basis as B
    name = mol.basis if which == 'ao' else mol.auxbasis
    per_el = B.load(name, sorted(set(mol.symbols)))
    shells = []
    for ia, el in enumerate(mol.symbols):
        for l, prims in per_el[el]:
            shells.append((ia, l, [p[0] for p in prims], [[p[1]] for p in prims]))
    return libcint_arrays(mol.atom_coords(), mol.atom_charges(), shells)


class PyscfLike:
    """The fields of a PySCF ``Mole`` the build reads."""

    def __init__(self, atm, bas, env, xyz, cart=False):
        self._atm, self._bas, self._env, self.cart = atm, bas, env, cart
        self._xyz = xyz

    def atom_coords(self):
        return self._xyz


class Sorted:
    """GPU4PySCF's sorted-molecule wrapper: ``.mol`` is the original."""

    def __init__(self, mol):
        self.mol = mol


class Int3c2eOptLike:
    def __init__(self, mol, auxmol):
        self.mol, self.auxmol = Sorted(mol), Sorted(auxmol)


def _tables_equal(a, b):
    assert a.nbas == b.nbas and a.nao == b.nao
    for f in ('shl_l', 'shl_atom', 'shl_nprim', 'shl_prim_off', 'shl_ao_off', 'ao_loc'):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.array_equal(a.exps, b.exps)
    np.testing.assert_allclose(a.coefs, b.coefs, rtol=1e-14, atol=0)
    np.testing.assert_allclose(a.shl_xyz, b.shl_xyz, rtol=0, atol=0)


@pytest.mark.parametrize('which', ['ao', 'aux'])
def test_libcint_arrays_give_the_same_shell_tables(which):
    from electrolites_amd.gto import water_cluster, libcint
    mol = water_cluster(3, seed=2, li=True)
    atm, bas, env = _cluster_libcint(mol, which)
    got = libcint.shell_table(atm, bas, env)
    _tables_equal(got, mol.ao if which == 'ao' else mol.aux)


def test_libcint_general_contraction_split():
    """A shell with two contractions is two shells, in libcint's AO order;
    a Cartesian molecule and spinor shells are refused."""
    from electrolites_amd.gto import libcint
    exps = [5., 1.2, .3]
    raw = np.array([[.2, 0.], [.5, .3], [.4, .9]])
    atm, bas, env = libcint_arrays([[0., 0., 0.], [0., 0., 1.4]], [8, 1],
                                   [(0, 2, exps, raw), (1, 0, [.8], [[1.]])])
    tab = libcint.shell_table(atm, bas, env)
    assert tab.nbas == 3 and list(tab.shl_l) == [2, 2, 0] and tab.nao == 11
    assert list(tab.shl_ao_off) == [0, 5, 10]
    for ic in range(2):
        c = tab.coefs[tab.shl_prim_off[ic]:tab.shl_prim_off[ic] + 3]
        np.testing.assert_allclose(c, _pyscf_coeffs(2, np.array(exps), raw)[:, ic], rtol=1e-14)
        # each contracted radial function is normalised
        ee = np.array([[_gaussian_int(6, a + b) for b in exps] for a in exps])
        assert abs(c @ ee @ c - 1.) < 1e-13
    with pytest.raises(NotImplementedError):
        libcint.shell_table(atm, bas, env, cart=True)
    bad = bas.copy()
    bad[0, 4] = -1
    with pytest.raises(NotImplementedError):
        libcint.shell_table(atm, bad, env)


def test_session_shape_detection():
    from electrolites_amd import fastdfcderi
    from electrolites_amd.df import DF
    from electrolites_amd.gto import water_cluster
    mol = water_cluster(1, seed=0)
    assert fastdfcderi._session_molecules(DF(mol)) is None
    atm, bas, env = _cluster_libcint(mol, 'ao')
    a2, b2, e2 = _cluster_libcint(mol, 'aux')
    opt = Int3c2eOptLike(PyscfLike(atm, bas, env, mol.atom_coords()),
                         PyscfLike(a2, b2, e2, mol.atom_coords()))
    m, a = fastdfcderi._session_molecules(opt)
    assert m is opt.mol.mol and a is opt.auxmol.mol


def test_oom_releases_caches_and_retries_once(monkeypatch):
    """An OOM inside a call releases the kept device caches and runs the
    call once more; a second OOM, or one with nothing to release, is
    raised (electrolites/fastdf.py:754-803)."""
    import torch
    from electrolites_amd import _dev, fastdf
    monkeypatch.setattr(_dev, 'WORKSPACE', _dev.Workspace())

    class Obj:
        pass

    o = Obj()
    host = object()
    o._mdf_resident = (host, 'device prefix')
    fastdf._CACHE_OWNERS.add(o)
    calls = []

    def flaky(*a):
        calls.append(a)
        if len(calls) == 1:
            raise torch.cuda.OutOfMemoryError('HIP out of memory')
        return 'vj', 'vk'

    monkeypatch.setattr(fastdf, '_get_jk', flaky)
    assert fastdf.get_jk(o, None) == ('vj', 'vk')
    assert len(calls) == 2
    assert o._mdf_resident == (host, None)           # the prefix given back for good
    # nothing left to release: the OOM propagates
    calls.clear()

    def oom(*a):
        calls.append(a)
        raise torch.cuda.OutOfMemoryError('HIP out of memory')

    monkeypatch.setattr(fastdf, '_get_jk', oom)
    with pytest.raises(torch.cuda.OutOfMemoryError):
        fastdf.get_jk(o, None)
    assert len(calls) == 1
    # other errors are not retried
    calls.clear()

    def boom(*a):
        calls.append(a)
        raise ValueError('not memory')

    o._mdf_resident = (host, 'device prefix')
    fastdf._CACHE_OWNERS.add(o)
    monkeypatch.setattr(fastdf, '_get_jk', boom)
    with pytest.raises(ValueError):
        fastdf.get_jk(o, None)
    assert len(calls) == 1 and o._mdf_resident[1] == 'device prefix'


@pytest.mark.gpu
@pytest.mark.parametrize('fp32', [False, True])
def test_session_cholesky_eri_layout_and_values(fp32, monkeypatch):
    """``_cholesky_eri`` on an ``Int3c2eOpt``-shaped argument: the
    aux-contiguous device tensor of the reference (out.T of a C-order
    (npairs, naux) buffer) with the pair-contiguous build's values, device
    pair tables, and J/K through ``get_jk`` on it equal to the oracle's."""
    import torch
    from electrolites_amd import fastdf, fastdfcderi
    from electrolites_amd.gto import water_cluster
    monkeypatch.setattr(fastdfcderi, 'FP32', fp32)
    mol = water_cluster(2, seed=3, side=5.0, li=True)
    atm, bas, env = _cluster_libcint(mol, 'ao')
    a2, b2, e2 = _cluster_libcint(mol, 'aux')
    opt = Int3c2eOptLike(PyscfLike(atm, bas, env, mol.atom_coords()),
                         PyscfLike(a2, b2, e2, mol.atom_coords()))
    (cderi,), (pa, diag) = fastdfcderi._cholesky_eri(opt)
    assert cderi.is_cuda and cderi.stride(0) == 1 and cderi.stride(1) == cderi.shape[0]
    assert cderi.dtype == (torch.float32 if fp32 else torch.float64)
    dt = 'float32' if fp32 else 'float64'
    # the same tensor as the pair-contiguous build from the same shell tables
    same, _ = fastdfcderi.cholesky_eri(opt._mdf_molecules, dtype=dt)
    assert torch.equal(cderi, same)
    # and, to rounding of the coefficients, as the build from the package's
    # own molecule (its normalisation of the basis data was done separately)
    ref_c, (ref_pa, ref_diag) = fastdfcderi.cholesky_eri(mol, dtype=dt)
    assert float((cderi.double() - ref_c.double()).abs().max() / ref_c.double().abs().max()) < \
        (1e-6 if fp32 else 1e-12)
    assert np.array_equal(pa.cpu().numpy(), ref_pa) and np.array_equal(diag.cpu().numpy(), ref_diag)

    class SessionDF:
        _cderi = [cderi]
        _cderi_idx = (pa, diag)

    rng = np.random.default_rng(9)
    nocc = mol.nelectron // 2
    c = np.linalg.qr(rng.standard_normal((mol.nao, nocc)))[0]
    dm = fastdf.tag_array(2 * c @ c.T, mo_coeff=c, mo_occ=np.full(nocc, 2.))
    vj, vk = fastdf.get_jk(SessionDF(), dm, hermi=1)
    store = cderi.double().cpu().numpy()
    rj, rk = O.get_jk(store, ref_pa.astype(np.int64), np.asarray(dm))
    assert np.abs(vj - rj).max() / np.abs(rj).max() < 1e-12
    assert np.abs(vk - rk).max() / np.abs(rk).max() < 1e-12
