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
