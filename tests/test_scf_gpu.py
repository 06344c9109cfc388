"""GPU one-electron integrals and the device DF-RHF energy against the
oracle (energy bar 1e-9 Eh, the reference suite's SCF bar,
tests/test_df.py:304-348)."""
import numpy as np
import pytest

from oracle import oracle as O


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope='module')
def mol():
    from electrolites_amd.gto import water_cluster
    return water_cluster(2, seed=3, side=5.0, li=True)


@pytest.mark.gpu
def test_int1e(mol):
    from electrolites_amd import ints
    ao = ints.DeviceBasis(mol.ao)
    S, T, V = (m.cpu().numpy() for m in ints.int1e(ao, mol))
    d = mol.ao.as_dict()
    assert rel(S, O.int1e('ovlp', d)) < 1e-13
    assert rel(T, O.int1e('kin', d)) < 1e-13
    assert rel(V, O.int1e('nuc', d, mol.atom_charges(), mol.atom_coords())) < 1e-13


@pytest.mark.gpu
def test_rhf_energy_matches_oracle(mol):
    from electrolites_amd.scf import RHF
    mf = RHF(mol, conv_tol=1e-11)
    e = mf.kernel()
    assert mf.converged
    s = O.DFSystem(mol.ao.as_dict(), mol.aux.as_dict(), mol.atom_charges(), mol.atom_coords(),
                   mol.nelectron, mol.energy_nuc())
    ref = O.rhf(s, conv_tol=1e-11)
    assert ref['converged']
    assert abs(e - ref['e_tot']) < 1e-9, (e, ref['e_tot'])


@pytest.mark.gpu
@pytest.mark.parametrize('low', ['fp32', 'bf16x3'])
def test_mixed_precision_rhf(mol, low):
    """FP32 or split-BF16 K until the DIIS error drops below 1e-5, FP64
    finishing cycles: the energy matches the all-FP64 run to 1e-8 Eh."""
    from electrolites_amd.scf import RHF
    e64 = RHF(mol, conv_tol=1e-10).kernel()
    mf = RHF(mol, conv_tol=1e-10, k_precision='mixed', low_precision=low)
    e = mf.kernel()
    assert mf.converged
    precs = [h['k_precision'] for h in mf.history]
    assert precs[0] == low and precs[-1] == 'fp64'
    assert abs(e - e64) < 1e-8, (e, e64)


@pytest.mark.gpu
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_purified_scf_matches_diagonalised(mol, precision):
    """Cycles whose occupied space comes from SP2 purification (no
    diagonalisation) reach the energy of the diagonalised SCF (1e-9 Eh) and
    the same converged density; the purified basis is orthonormal."""
    import torch
    from electrolites_amd.scf import RHF
    ref = RHF(mol, conv_tol=1e-10, k_precision=precision)
    e_ref = ref.kernel()
    mf = RHF(mol, conv_tol=1e-10, k_precision=precision, purify_min=0)
    e = mf.kernel()
    assert mf.converged and mf.purified >= mf.cycles - 2
    assert abs(e - e_ref) < 1e-9, (e, e_ref)
    assert float((mf.dm - ref.dm).abs().max()) < 1e-6
    q = mf._q_occ
    assert float((q.T @ q - torch.eye(q.shape[1], dtype=q.dtype, device=q.device)).abs().max()) < 1e-12


@pytest.mark.gpu
def test_conv_check(mol):
    """Converged on an FP64 J/K, the cycle's energy is the result; with
    conv_check one more FP64 J/K of the new density gives it.  The two agree
    to 1e-9 Eh (the energy is stationary at convergence), and conv_check
    costs exactly one more J/K."""
    from electrolites_amd.scf import RHF
    out = []
    for check in (False, True):
        mf = RHF(mol, conv_tol=1e-10, conv_check=check)
        calls = [0]
        get_jk = mf.with_df.get_jk

        def counted(*a, **k):
            calls[0] += 1
            return get_jk(*a, **k)
        mf.with_df.get_jk = counted
        e = mf.kernel()
        assert mf.converged
        out.append((e, mf.cycles, calls[0], mf.history[-1]['e_tot']))
    (ea, ca, na, ha), (eb, cb, nb, hb) = out
    assert ca == cb and na == ca and nb == cb + 1
    assert ea == ha
    assert abs(ea - eb) < 1e-9, (ea, eb)


@pytest.mark.gpu
def test_projector_basis():
    """The purification's basis from the projector alone (pivoted Cholesky,
    then one Cholesky-QR pass): for a random rank-nocc projector P on the
    device, Q^T Q = I and Q Q^T = P to 1e-12; a matrix of another rank is
    refused."""
    import torch
    from electrolites_amd.scf import RHF
    rng = np.random.default_rng(5)
    n, nocc = 230, 47
    q, _ = np.linalg.qr(rng.standard_normal((n, nocc)))
    P = torch.from_numpy(q @ q.T).cuda()
    mf = RHF.__new__(RHF)
    mf.nocc, mf.concurrent_diag = nocc, False
    Q = mf._projector_basis(P)
    Qh = Q.cpu().numpy()
    assert Qh.shape == (n, nocc)
    assert np.abs(Qh.T @ Qh - np.eye(nocc)).max() < 1e-12
    assert np.abs(Qh @ Qh.T - q @ q.T).max() < 1e-12
    mf.nocc = nocc + 1
    assert mf._projector_basis(P) is None
