"""The drop-in boundary without a GPU: patch registry over a stand-in
GPU4PySCF module tree (the reference's own boundary test walks GPU4PySCF for
stale aliases, tests/test_patch_aliases.py:1-33), decline bookkeeping, and
the library's error reporting."""
import sys
import types

import numpy as np
import pytest


@pytest.fixture
def fake_gpu4pyscf(monkeypatch):
    calls = []

    def orig_get_jk(dfobj, dms, hermi=0, with_j=True, with_k=True, omega=None):
        calls.append(('get_jk', omega))
        return 'orig-vj', 'orig-vk'

    def orig_cholesky_eri(intopt, omega=None, use_gpu_memory=None):
        calls.append(('cholesky_eri', omega))
        return 'orig'

    root = types.ModuleType('gpu4pyscf')
    df_pkg = types.ModuleType('gpu4pyscf.df')
    df_jk = types.ModuleType('gpu4pyscf.df.df_jk')
    df_mod = types.ModuleType('gpu4pyscf.df.df')
    def orig_get_j(dfobj, dm, hermi=1):
        calls.append(('get_j', hermi))
        return 'orig-vj'

    def orig_loop(self, blksize=None, unpack=True):
        calls.append(('loop', blksize))
        yield 'orig-block'

    def orig_jk_energy_per_atom(self, dm=None, j_factor=1, k_factor=1, omega=0, hermi=0, verbose=None):
        calls.append(('jk_energy_per_atom', omega))
        return 'orig-ejk'

    grad_pkg = types.ModuleType('gpu4pyscf.df.grad')
    grad_mods = {}
    for kind in ('rhf', 'rks', 'uhf', 'uks'):
        m = types.ModuleType(f'gpu4pyscf.df.grad.{kind}')
        m.Gradients = type('Gradients', (), {'jk_energy_per_atom': orig_jk_energy_per_atom})
        setattr(grad_pkg, kind, m)
        grad_mods[kind] = m
    df_pkg.grad = grad_pkg
    df_jk.get_jk = orig_get_jk
    df_jk.get_j = orig_get_j
    df_mod._cholesky_eri = orig_cholesky_eri
    df_mod.DF = type('DF', (), {'loop': orig_loop})
    root.df = df_pkg
    df_pkg.df_jk = df_jk
    df_pkg.df = df_mod
    for name, mod in [('gpu4pyscf', root), ('gpu4pyscf.df', df_pkg),
                      ('gpu4pyscf.df.df_jk', df_jk), ('gpu4pyscf.df.df', df_mod),
                      ('gpu4pyscf.df.grad', grad_pkg)] + \
            [(f'gpu4pyscf.df.grad.{k}', m) for k, m in grad_mods.items()]:
        monkeypatch.setitem(sys.modules, name, mod)
    import electrolites_amd
    from electrolites_amd import fastdf, fastdfcderi, fastdfj3c, fastdffp32, fastdfgrad
    monkeypatch.setattr(fastdfgrad, '_orig_methods', {})
    monkeypatch.setattr(fastdf, '_orig_get_jk', None)
    monkeypatch.setattr(fastdffp32, '_orig_loop', None)
    monkeypatch.setattr(fastdfcderi, '_orig_cholesky_eri', None, raising=False)
    monkeypatch.setattr(fastdfj3c, '_orig_get_j', None, raising=False)
    monkeypatch.setattr(electrolites_amd, '_patched', [])
    return types.SimpleNamespace(df_jk=df_jk, df=df_mod, calls=calls, orig_get_jk=orig_get_jk,
                                 orig_cholesky_eri=orig_cholesky_eri, orig_get_j=orig_get_j,
                                 orig_loop=orig_loop, grad=grad_mods,
                                 orig_jk_energy_per_atom=orig_jk_energy_per_atom)


def test_patch_installs_and_keeps_originals(fake_gpu4pyscf):
    import electrolites_amd
    from electrolites_amd import fastdf, fastdfcderi
    from electrolites_amd import fastdfj3c
    from electrolites_amd import fastdffp32
    assert electrolites_amd.check_targets() == {'fastdf': None, 'fastdfcderi': None, 'fastdfj3c': None,
                                                'fastdffp32': None, 'fastdfgrad': None}
    assert electrolites_amd.patch() == ('fastdf', 'fastdfcderi', 'fastdfj3c', 'fastdffp32', 'fastdfgrad')
    assert fake_gpu4pyscf.df.DF.loop is fastdffp32.patched_loop
    assert fastdffp32._orig_loop is fake_gpu4pyscf.orig_loop
    # a tensor GPU4PySCF built itself (not a torch tensor) goes to its own loop
    obj = fake_gpu4pyscf.df.DF()
    obj._cderi = {0: np.zeros((3, 6))}
    assert list(obj.loop(2)) == ['orig-block'] and fake_gpu4pyscf.calls[-1] == ('loop', 2)
    assert fake_gpu4pyscf.df_jk.get_jk is fastdf.get_jk
    assert fake_gpu4pyscf.df_jk.get_j is fastdfj3c.get_j
    assert fastdfj3c._orig_get_j is fake_gpu4pyscf.orig_get_j
    assert fake_gpu4pyscf.df._cholesky_eri is fastdfcderi._cholesky_eri
    assert fastdf._orig_get_jk is fake_gpu4pyscf.orig_get_jk
    assert fastdfcderi._orig_cholesky_eri is fake_gpu4pyscf.orig_cholesky_eri
    # patching twice keeps the true originals (no alias of ourselves)
    electrolites_amd.patch()
    assert fastdf._orig_get_jk is fake_gpu4pyscf.orig_get_jk


def test_declined_call_falls_through(fake_gpu4pyscf):
    import electrolites_amd
    from electrolites_amd import fastdf
    electrolites_amd.patch('fastdf')
    before = fastdf.stats()['declined']
    dm = np.zeros((4, 4), dtype=np.complex128)
    out = fake_gpu4pyscf.df_jk.get_jk(object(), dm, hermi=1)
    assert out == ('orig-vj', 'orig-vk')
    st = fastdf.stats()
    assert st['declined'] == before + 1 and st['declined_reasons'].get('complex density')
    out = fake_gpu4pyscf.df_jk.get_jk(object(), np.zeros((4, 4)), hermi=1, omega=.3)
    assert out == ('orig-vj', 'orig-vk')
    assert fake_gpu4pyscf.calls[-1] == ('get_jk', .3)


def test_missing_target_is_reported(fake_gpu4pyscf, monkeypatch):
    import electrolites_amd
    monkeypatch.delattr(fake_gpu4pyscf.df, '_cholesky_eri')
    rep = electrolites_amd.check_targets()
    assert rep['fastdf'] is None and 'gpu4pyscf.df.df._cholesky_eri' in rep['fastdfcderi']
    with pytest.raises(RuntimeError):
        electrolites_amd.patch()
    assert electrolites_amd.patch(strict=False) == ('fastdf', 'fastdfj3c', 'fastdffp32', 'fastdfgrad')


def test_standalone_decline_raises(monkeypatch):
    from electrolites_amd import fastdf
    monkeypatch.setattr(fastdf, '_orig_get_jk', None)
    with pytest.raises(fastdf.Declined):
        fastdf.get_jk(object(), np.zeros((3, 3), dtype=np.complex128))


def test_narrow_tensor_never_falls_through(monkeypatch):
    """fastdf.py:434-440: a declined call on a float32 cderi must raise
    instead of reaching a float64 reader (covered calls run the FP32-storage
    kernels)."""
    import torch
    from electrolites_amd import fastdf

    class DF:
        _cderi = {0: torch.zeros((2, 3), dtype=torch.float32)}

    monkeypatch.setattr(fastdf, '_orig_get_jk', lambda *a, **k: ('x', 'y'))
    with pytest.raises(RuntimeError):
        fastdf.get_jk(DF(), np.zeros((2, 2), dtype=np.complex128))


def test_library_error_message():
    from electrolites_amd import _lib
    with pytest.raises(_lib.MDFError, match='bad sizes'):
        _lib.call('mdf_get_k', None, 0, -1, 0, 0, None, 0, None, None, 0, 0, 0, None, 0, 0, None, 0, None)


def test_tile_fill_dense_configs():
    """The block-sparse tiled step 1 of the reference (taken only at AO tile
    fill <= 0.75) does not apply to the dense 1 g/cm^3 clusters: every
    128 x 128 AO block holds kept shell pairs, and more than 90 % of the
    16 x 16 blocks the reference's gate counts."""
    from electrolites_amd import fastdfcderi
    from electrolites_amd.gto import water_cluster
    for n in (32, 64):
        mol = water_cluster(n, seed=0)
        assert fastdfcderi.tile_fill(mol) == 1.0
        # at the reference's own 16-AO tile (electrolites/fastdf.py:216) the
        # fill is still far above its 0.75 gate (0.977 at (H2O)64)
        assert fastdfcderi.tile_fill(mol, tile=16) > .9
    # a sparse arrangement (two waters 60 A apart) is not dense
    from electrolites_amd.gto.mole import Mole
    far = Mole([('O', (0., 0., 0.)), ('H', (0.76, 0.59, 0.)), ('H', (-0.76, 0.59, 0.)),
                ('O', (60., 0., 0.)), ('H', (60.76, 0.59, 0.)), ('H', (59.24, 0.59, 0.))])
    assert fastdfcderi.tile_fill(far, tile=8) < 0.75


def test_kept_pairs_cover_the_chunk_tasks():
    """The cderi's stored columns (fastdfcderi.kept_pairs) are exactly the
    function pairs of the shell pairs the integral chunks compute, chunk by
    chunk, so every stored column is written and no computed pair is lost
    (host-only: the screening bound and the task lists)."""
    from electrolites_amd import fastdfcderi, ints
    from electrolites_amd.gto import water_cluster
    mol = water_cluster(6, seed=2, side=7.0, li=True)
    bound = ints.pair_bound(mol.ao)
    keep, rows, cols = fastdfcderi.kept_pairs(mol, bound)
    ao_loc = np.asarray(mol.ao.ao_loc)
    covered = np.zeros_like(keep)
    tri = lambda i, j: i * (i + 1) // 2 + j
    for I0, I1 in fastdfcderi._chunks(mol.ao, mol.naux, 400000):
        tk = ints.PairTasks(mol.ao, I0, I1, bound)
        for ish, jsh in tk.pairs:
            for i in range(ao_loc[ish], ao_loc[ish + 1]):
                for j in range(ao_loc[jsh], ao_loc[jsh + 1]):
                    if j <= i:
                        p = tri(i, j)
                        assert tk.p0 <= p < tk.p1
                        covered[p] = True
    assert np.array_equal(covered, keep)
    assert 0 < keep.sum() < keep.size                 # something is screened at this size


def test_gradient_method_patched_and_declines(fake_gpu4pyscf):
    """All four gradient classes get the method (each binds its own copy,
    electrolites/__init__.py:133-143); a long-range call and a cderi
    GPU4PySCF built itself go to the replaced method, counted."""
    import electrolites_amd
    from electrolites_amd import fastdfgrad
    electrolites_amd.patch('fastdfgrad')
    for kind in ('rhf', 'rks', 'uhf', 'uks'):
        cls = fake_gpu4pyscf.grad[kind].Gradients
        assert cls.jk_energy_per_atom is fastdfgrad.patched_jk_energy_per_atom
        assert fastdfgrad._orig_methods[cls] is fake_gpu4pyscf.orig_jk_energy_per_atom
    electrolites_amd.patch('fastdfgrad')               # twice: originals kept
    cls = fake_gpu4pyscf.grad['rhf'].Gradients
    assert fastdfgrad._orig_methods[cls] is fake_gpu4pyscf.orig_jk_energy_per_atom

    class WithDF:
        _cderi = {0: np.zeros((3, 6))}                # a GPU4PySCF-built (non-torch) cderi

    g = cls()
    g.base = types.SimpleNamespace(with_df=WithDF(), make_rdm1=lambda: np.eye(3))
    before = fastdfgrad.stats()['declined']
    assert g.jk_energy_per_atom(np.eye(3)) == 'orig-ejk'
    assert g.jk_energy_per_atom(np.eye(3), omega=.2) == 'orig-ejk'
    assert fake_gpu4pyscf.calls[-1] == ('jk_energy_per_atom', .2)
    st = fastdfgrad.stats()
    assert st['declined'] == before + 2
    assert st['declined_reasons'].get('omega != 0') and st['declined_reasons'].get('cderi not built by this package')
