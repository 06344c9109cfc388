"""The DF gradient's two-electron term (electrolites/fastdfgrad.py
``jk_energy_per_atom`` :1229; the reference's own test compares it with
GPU4PySCF's function and checks the whole gradient by finite differences,
tests/test_dfgrad.py:1-40).

GPU4PySCF is not here, so the check is the finite-difference one: the
derivative of the oracle's DF energy

    E_J = 1/2 rho^T M^-1 rho,   E_K = -1/4 sum_PQ M^-1_PQ tr(T_P D T_Q D)

(rho_P = sum_ij (ij|P) D_ij, T_P = (ij|P), M = (P|Q), all from the oracle's
integrals) under a displacement of one atom's shells in both bases, by a
five-point stencil.  The energy restatement itself is pinned on the CPU
against the oracle's J/K from its cderi (1/2 tr(D J) - 1/4 tr(D K))."""
import numpy as np
import pytest

from oracle import oracle as O


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _copy(d):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in d.items()}


def energy_parts(ao, aux, D):
    """(E_J, E_K) of the oracle's density-fitted energy for density D."""
    T = O.int3c2e(ao, aux)
    M = O.int2c2e(aux)
    rho = np.einsum('ijP,ij->P', T, D)
    ej = .5 * rho @ np.linalg.solve(M, rho)
    naux = T.shape[2]
    G = np.einsum('ijP,jk->Pik', T, D)                 # T_P D
    X = G.reshape(naux, -1) @ G.transpose(0, 2, 1).reshape(naux, -1).T
    ek = -.25 * np.trace(np.linalg.solve(M, X))
    return ej, ek


def energy_parts_spin(ao, aux, Da, Db):
    """(E_J of Da + Db, -1/2 sum_s tr(D_s K[D_s])) of the oracle's DF energy
    for an unrestricted density (each spin's exchange twice the restricted
    form's -1/4 weight)."""
    ej, _ = energy_parts(ao, aux, Da + Db)
    return ej, 2 * (energy_parts(ao, aux, Da)[1] + energy_parts(ao, aux, Db)[1])


def fd_gradient(mol, D, h=2e-3, atoms=None, energy=None):
    """(natm, 3, 2) five-point derivative of (E_J, E_K) per atom coordinate
    (``energy(ao, aux)``, default the restricted parts for D)."""
    energy = energy or (lambda ao, aux: energy_parts(ao, aux, D))
    ao0, aux0 = mol.ao.as_dict(), mol.aux.as_dict()
    atoms = range(mol.natm) if atoms is None else atoms
    out = np.zeros((mol.natm, 3, 2))
    for a in atoms:
        for x in range(3):
            vals = []
            for s in (-2, -1, 1, 2):
                ao, aux = _copy(ao0), _copy(aux0)
                ao['xyz'][ao['atom'] == a, x] += s * h
                aux['xyz'][aux['atom'] == a, x] += s * h
                vals.append(np.array(energy(ao, aux)))
            out[a, x] = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
    return out


def _density(nao, nocc, seed):
    rng = np.random.default_rng(seed)
    c = np.linalg.qr(rng.standard_normal((nao, nocc)))[0]
    return 2 * c @ c.T, c


@pytest.fixture(scope='module')
def mol():
    from electrolites_amd.gto import water_cluster
    return water_cluster(2, seed=3, side=4.0, li=True)


def test_energy_restatement_matches_oracle_jk(mol):
    """1/2 rho M^-1 rho - 1/4 tr(M^-1 X) equals 1/2 tr(D J) - 1/4 tr(D K)
    of the oracle's cderi path (the energy the finite differences use)."""
    ao, aux = mol.ao.as_dict(), mol.aux.as_dict()
    D, _ = _density(mol.nao, 6, 1)
    ej, ek = energy_parts(ao, aux, D)
    cderi = O.cholesky_eri(O.int3c2e(ao, aux), O.int2c2e(aux))
    pa, _ = O.pair_addresses(mol.nao)
    vj, vk = O.get_jk(cderi, pa, D[None], hermi=1)
    assert abs(ej - .5 * np.sum(D * vj[0])) < 1e-10 * abs(ej)
    assert abs(ek + .25 * np.sum(D * vk[0])) < 1e-10 * abs(ek)


def test_fd_harness_translation_invariance():
    """The finite-difference gradient of a free molecule sums to zero over
    atoms (a displaced-shell bookkeeping mistake breaks it)."""
    from electrolites_amd.gto import water_cluster
    m = water_cluster(1, seed=0)
    D, _ = _density(m.nao, 5, 2)
    g = fd_gradient(m, D)
    assert np.abs(g.sum(axis=0)).max() < 1e-8
    assert np.abs(g).max() > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize('tagged', [True, False])
def test_gradient_matches_finite_differences(mol, tagged):
    """J-only, K-only and J - 1/2 K gradients against the five-point
    derivative of the oracle's energy, to 1e-7 Eh/bohr."""
    from electrolites_amd import fastdf, fastdfgrad
    from electrolites_amd.df import DF
    D, c = _density(mol.nao, 8, 4)
    fd = fd_gradient(mol, D)
    dm = fastdf.tag_array(D, mo_coeff=c, mo_occ=np.full(8, 2.)) if tagged else D
    df = DF(mol).build()
    for jf, kf in [(1., 0.), (0., 1.), (1., .5)]:
        g = fastdfgrad.jk_energy_per_atom(df, dm, j_factor=jf, k_factor=kf)
        ref = jf * fd[..., 0] + kf * fd[..., 1]
        err = np.abs(g - ref).max()
        print(f'j={jf} k={kf} max|g - fd| = {err:.2e}, max|g| = {np.abs(g).max():.3f}', flush=True)
        assert err < 1e-7
        assert np.abs(g.sum(axis=0)).max() < 1e-9          # translation invariance


@pytest.mark.gpu
def test_gradient_without_aux_response(mol):
    """Without the fitting functions' response only the orbital centres'
    share of the three-centre term remains; it equals the full gradient
    less the aux centres' shares (the metric term and the aux centre of the
    three-centre term), which a free molecule's sum rule fixes: the orbital
    share alone no longer sums to zero."""
    from electrolites_amd import fastdfgrad
    from electrolites_amd.df import DF
    D, _ = _density(mol.nao, 8, 5)
    df = DF(mol).build()
    full = fastdfgrad.jk_energy_per_atom(df, D, 1., .5)
    orb = fastdfgrad.jk_energy_per_atom(df, D, 1., .5, auxbasis_response=False)
    assert np.abs(full.sum(axis=0)).max() < 1e-9
    assert np.abs(orb.sum(axis=0)).max() > 1e-4
    assert np.abs(full - orb).max() > 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize('storage,dtype,tol', [('host', 'float64', 1e-11), ('device', 'float32', 1e-5)])
def test_gradient_storage_modes(mol, monkeypatch, storage, dtype, tol):
    """The gradient from a host-resident cderi equals the device one (the
    same rows, streamed); from FP32 storage it stays within the narrow
    tensor's rounding."""
    import electrolites_amd.fastdfcderi as FC
    from electrolites_amd import fastdfgrad
    from electrolites_amd.df import DF
    D, _ = _density(mol.nao, 8, 6)
    ref = fastdfgrad.jk_energy_per_atom(DF(mol).build(), D, 1., .5)
    monkeypatch.setattr(FC, 'FP32', dtype == 'float32')
    df = DF(mol, storage=storage).build()
    g = fastdfgrad.jk_energy_per_atom(df, D, 1., .5)
    assert np.abs(g - ref).max() < tol * max(np.abs(ref).max(), 1.)
    if storage == 'host':
        df.reset()


@pytest.mark.gpu
def test_patched_method_runs_the_hip_path(mol):
    """The method patch() installs on GPU4PySCF's gradient classes runs this
    package's gradient for a restricted density over its own cderi."""
    import types
    from electrolites_amd import fastdfgrad
    from electrolites_amd.df import DF
    D, _ = _density(mol.nao, 8, 7)
    df = DF(mol).build()
    g = types.SimpleNamespace(base=types.SimpleNamespace(with_df=df, make_rdm1=lambda: D),
                              auxbasis_response=True)
    before = fastdfgrad.stats()['calls']
    out = fastdfgrad.patched_jk_energy_per_atom(g, None, 1, .5)
    ref = fastdfgrad.jk_energy_per_atom(df, D, 1., .5)
    assert fastdfgrad.stats()['calls'] == before + 2
    assert np.abs(out - ref).max() < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize('nao,naux,nocc,nL,screened,dtype', [(37, 19, 5, 19, False, 'float64'),
                                                              (64, 130, 17, 41, True, 'float64'),
                                                              (200, 77, 90, 77, True, 'float32')])
def test_k_step1_entry(nao, naux, nocc, nL, screened, dtype):
    """mdf_k_step1 (K's first step alone, the gradient's R_L = A_L F) equals
    the unpacked rows times F (numpy) on ragged shapes, full-triangle and
    screened pair lists, FP64 and FP32 storage (widened exactly); and
    mdf_k_step1_lower equals T_L F with T_L the lower triangle of A_L,
    diagonal halved (A_L = T_L + T_L^T)."""
    import torch
    from electrolites_amd import _dev, _lib, fastdf
    rng = np.random.default_rng(nao + naux)
    pa, _ = O.pair_addresses(nao)
    if screened:
        pa = np.sort(rng.choice(pa, size=len(pa) * 2 // 3, replace=False))
    cderi = rng.standard_normal((naux, len(pa))).astype(dtype)
    f = rng.standard_normal((nao, nocc))
    ref = np.matmul(O.unpack(cderi.astype(np.float64), pa, nao)[:nL], f)

    class DFo:
        _cderi = {0: torch.from_numpy(cderi).cuda()}
        _cderi_idx = (pa.astype(np.int32), None)

    c = DFo._cderi[0]
    lay = fastdf._layout(DFo(), c, nao)
    fd = torch.from_numpy(f).cuda()
    A = O.unpack(cderi.astype(np.float64), pa, nao)[:nL]
    T = np.tril(A, -1) + .5 * A * np.eye(nao)
    assert np.allclose(T + T.transpose(0, 2, 1), A)
    for entry, want in (('mdf_k_step1', ref), ('mdf_k_step1_lower', np.matmul(T, f))):
        R = torch.full((nL, nao, nocc + 3), np.nan, dtype=torch.float64, device='cuda')
        _lib.call(entry, _dev.ptr(c), fastdf._dtype_code(c), nL, c.shape[1], c.stride(0),
                  _dev.ptr(lay.map), nao, _dev.ptr(fd), fd.stride(0), nocc, _dev.ptr(R), nocc + 3,
                  nao * (nocc + 3), _dev.stream_ptr())
        got = R.cpu().numpy()
        assert np.all(np.isnan(got[..., nocc:]))              # nothing written past nocc
        assert rel(got[..., :nocc], want) < 1e-13, entry


@pytest.mark.gpu
def test_gradient_eigen_route(mol, monkeypatch):
    """The metric's eigen route (cderi = coef T with coef^T coef = M^-1, the
    reference's branch for a metric that is not positive definite,
    electrolites/fastdfcderi.py:203-210) gives the same gradient as the
    Cholesky route: the build and the gradient take the same coefficients."""
    import torch
    import electrolites_amd.fastdfcderi as FC
    from electrolites_amd import fastdfgrad, ints
    from electrolites_amd.df import DF
    D, _ = _density(mol.nao, 8, 8)
    ref = fastdfgrad.jk_energy_per_atom(DF(mol).build(), D, 1., .5)

    def eigen_route(aux, timer=None):
        w, v = torch.linalg.eigh(ints.int2c2e(aux))
        keep = w > FC.LINEAR_DEP_THR
        return None, (v[:, keep] / torch.sqrt(w[keep])).T.contiguous()

    monkeypatch.setattr(FC, 'decompose_j2c', eigen_route)
    df = DF(mol).build()
    g = fastdfgrad.jk_energy_per_atom(df, D, 1., .5)
    assert df._mdf_grad_metric[2][1] is False                # the general (non-triangular) branch ran
    assert np.abs(g - ref).max() < 1e-7



@pytest.mark.gpu
@pytest.mark.parametrize('tagged', [True, False])
def test_unrestricted_gradient_matches_finite_differences(mol, tagged):
    """Spin densities (2, nao, nao) with different occupations: J of the
    total density, K per spin (-1/2 k sum_s tr(D_s K[D_s])), against the
    five-point derivative of the oracle's energy; a closed-shell pair
    (D_a = D_b = D/2) equals the restricted gradient of D."""
    from electrolites_amd import fastdf, fastdfgrad
    from electrolites_amd.df import DF
    rng = np.random.default_rng(9)
    q = np.linalg.qr(rng.standard_normal((mol.nao, 9)))[0]
    ca, cb = q[:, :9], q[:, :7]
    Da, Db = ca @ ca.T, cb @ cb.T
    fd = fd_gradient(mol, None, energy=lambda ao, aux: energy_parts_spin(ao, aux, Da, Db))
    dm = np.stack([Da, Db])
    if tagged:
        mo = np.stack([q[:, :9], q[:, :9]])
        occ = np.array([[1.] * 9, [1.] * 7 + [0.] * 2])
        dm = fastdf.tag_array(dm, mo_coeff=mo, mo_occ=occ)
    df = DF(mol).build()
    for jf, kf in [(1., 0.), (0., 1.), (1., .5)]:
        g = fastdfgrad.jk_energy_per_atom(df, dm, j_factor=jf, k_factor=kf)
        ref = jf * fd[..., 0] + kf * fd[..., 1]
        err = np.abs(g - ref).max()
        print(f'UHF j={jf} k={kf} max|g - fd| = {err:.2e}, max|g| = {np.abs(g).max():.3f}', flush=True)
        assert err < 1e-7
        assert np.abs(g.sum(axis=0)).max() < 1e-9
    D, _ = _density(mol.nao, 8, 10)
    gu = fastdfgrad.jk_energy_per_atom(df, np.stack([D / 2, D / 2]), 1., 1.)
    gr = fastdfgrad.jk_energy_per_atom(df, D, 1., 1.)
    assert np.abs(gu - gr).max() < 1e-10


@pytest.mark.parametrize('n', [1, 2, 7, 320])
def test_packed_diagonal_positions(n):
    """The diagonal offsets the gradient reads rho_B from in the packed
    upper triangles (row-major over a <= b, the layout of mdf_sym_pack) are
    where numpy's triu_indices order puts a == b, and a packed symmetric
    matrix with its off-diagonal entries times sqrt 2 has the Frobenius dot
    product of the full ones (the trace products tr(dmoo_P dmoo_Q))."""
    from electrolites_amd.fastdfgrad import SQRT2, _diag_positions
    a, b = np.triu_indices(n)
    assert np.array_equal(_diag_positions(n), np.nonzero(a == b)[0])
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal((2, n, n))
    x, y = x + x.T, y + y.T
    w = np.where(a == b, 1., SQRT2)
    assert abs((w * x[a, b]) @ (w * y[a, b]) - (x * y).sum()) < 1e-12 * n * n
