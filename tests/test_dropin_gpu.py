"""The drop-in on GPU4PySCF-shaped inputs, against the FP64 oracle.

GPU4PySCF hands ``df_jk.get_jk`` CuPy densities and a DF object whose
device cderi is the ``(naux, npairs)`` transpose of a C-order
``(npairs, naux)`` buffer (electrolites/fastdfcderi.py:648, :690), and the
reference returns device arrays for device input (electrolites/fastdf.py:828,
1017-1024).  CuPy is not installed here; ``CAIArray`` stands in for it: an
object that exposes only ``__cuda_array_interface__`` (and carries tags as
attributes, as ``pyscf.lib.tag_array`` does).

Covered: the aux-contiguous tensor through the kept pair-contiguous copy and
through per-call block transposes; densities and tags as interface objects;
the ``symmetrize`` tag of a CPHF response density (X = mo1 (w occ)^T,
D = X + X^T: electrolites/fastdf.py:830, 1020-1021; the densities of the
reference's tests/test_dfhess.py:100-131) for restricted and unrestricted
shapes, compared with K of the full density from the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O

TOL = 1e-12


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


class CAIArray:
    """A device array known only through the CUDA array interface."""

    def __init__(self, tensor):
        self._t = tensor

    @property
    def __cuda_array_interface__(self):
        return self._t.__cuda_array_interface__


def cai_host(x):
    """numpy view of a CUDA-array-interface result."""
    import torch
    return torch.as_tensor(x, device='cuda').cpu().numpy()


class AuxContigDF:
    """A DF object holding the reference's device layout: ``_cderi[0]`` is
    ``buf.T`` for a C-order ``(npairs, naux)`` buffer, as an interface
    object."""

    def __init__(self, cderi, pair_addr, nao):
        import torch
        buf = torch.from_numpy(np.ascontiguousarray(cderi.T)).cuda()       # (npairs, naux)
        self._buf = buf
        self._cderi = {0: CAIArray(buf.T)}
        rows, cols = np.divmod(pair_addr, nao)
        self._cderi_idx = (CAIArray(torch.from_numpy(pair_addr.astype(np.int32)).cuda()),
                           np.where(rows == cols)[0])

    def build(self):
        raise AssertionError('cderi given')


@pytest.fixture(scope='module')
def water2():
    from electrolites_amd.gto import water_cluster
    mol = water_cluster(2, seed=1, side=4.0)
    ao, aux = mol.ao.as_dict(), mol.aux.as_dict()
    pa, _ = O.pair_addresses(mol.nao)
    cderi = O.cholesky_eri(O.int3c2e(ao, aux), O.int2c2e(aux), pa)
    # a screened-looking pair list: every third off-diagonal pair dropped,
    # addresses given in the upper triangle for half the kept pairs (the
    # reference's scatter writes both triangles either way)
    rows, cols = np.divmod(pa, mol.nao)
    keep = (rows == cols) | (np.arange(len(pa)) % 3 != 0)
    sub = cderi[:, keep]
    r, c = rows[keep], cols[keep]
    flip = np.arange(len(r)) % 2 == 1
    pa_s = np.where(flip, c * mol.nao + r, r * mol.nao + c).astype(np.int64)
    return mol.nao, cderi, pa, sub, pa_s


def _dm(nao, nocc, seed, tagged=True):
    import torch
    rng = np.random.default_rng(seed)
    c = np.linalg.qr(rng.standard_normal((nao, nocc)))[0]
    dm = CAIArray(torch.from_numpy(2 * c @ c.T).cuda())
    if tagged:
        dm.mo_coeff = CAIArray(torch.from_numpy(c).cuda())
        dm.mo_occ = CAIArray(torch.full((nocc,), 2., dtype=torch.float64).cuda())
    return dm, 2 * c @ c.T


@pytest.mark.gpu
@pytest.mark.parametrize('copy', [True, False])
@pytest.mark.parametrize('screened', [False, True])
def test_aux_contiguous_interface_arrays(water2, monkeypatch, copy, screened):
    """CUDA-array-interface density and aux-contiguous cderi through the
    patched-style ``get_jk``: the kept pair-contiguous copy, and (copy off)
    per-call block transposes over several blocks."""
    from electrolites_amd import dfstore, fastdf
    nao, cderi, pa, sub, pa_s = water2
    c_use, pa_use = (sub, pa_s) if screened else (cderi, pa)
    monkeypatch.setattr(fastdf, 'AUXC_COPY', copy)
    # four streamed blocks of aux rows when copying is off
    monkeypatch.setattr(dfstore, 'STREAM_BLOCK_BYTES', c_use.shape[1] * 8 * (c_use.shape[0] // 4 + 1))
    df = AuxContigDF(c_use, pa_use, nao)
    before = dict(fastdf.STATS)
    for tagged in (True, False):
        dm, dm_np = _dm(nao, 9, 5 + tagged, tagged)
        vj, vk = fastdf.get_jk(df, dm, hermi=1)
        assert hasattr(vj, '__cuda_array_interface__') and hasattr(vk, '__cuda_array_interface__')
        rj, rk = O.get_jk(c_use, pa_use, dm_np)
        assert rel(cai_host(vj), rj) < TOL
        assert rel(cai_host(vk), rk) < (TOL if tagged else 1e-11)
    vj, none = fastdf.get_jk(df, dm, hermi=1, with_k=False)
    assert none is None and rel(cai_host(vj), rj) < TOL
    assert fastdf.STATS['auxc_calls'] - before.get('auxc_calls', 0) == 3
    if copy:
        assert fastdf.STATS.get('auxc_copies', 0) - before.get('auxc_copies', 0) == 1
    else:
        assert fastdf.STATS.get('auxc_blocks', 0) - before.get('auxc_blocks', 0) >= 12


@pytest.mark.gpu
def test_transpose_entry():
    """``mdf_transpose`` on ragged shapes, FP64 and FP32, into a padded
    destination (bit-exact: a copy)."""
    import ctypes
    import torch
    from electrolites_amd import _dev, _lib
    rng = np.random.default_rng(3)
    for rows, cols in ((1, 1), (63, 65), (130, 7), (1000, 129)):
        for dt, code in ((torch.float64, 0), (torch.float32, 1)):
            src = torch.from_numpy(rng.standard_normal((rows, cols + 3))).to(dt).cuda()
            dst = torch.full((cols, rows + 5), -7., dtype=dt, device='cuda')
            _lib.call('mdf_transpose', _dev.ptr(src), code, rows, cols, cols + 3, _dev.ptr(dst), rows + 5,
                      _dev.stream_ptr())
            torch.cuda.synchronize()
            assert torch.equal(dst[:, :rows], src[:, :cols].T)
            assert bool((dst[:, rows:] == -7.).all())


def _cphf_dm(nao, nocc, n_dm, seed, uhf):
    """A CPHF-style response density and its factors, built as the
    reference's tests build them (tests/test_dfhess.py:110-131)."""
    rng = np.random.default_rng(seed)
    if uhf:
        orbo = np.stack([np.linalg.qr(rng.standard_normal((nao, nocc)))[0] for _ in range(2)])
        mo1 = rng.standard_normal((2, n_dm, nao, nocc)) * 1e-2
        x = np.einsum('snpi,sqi->snpq', mo1, orbo)
    else:
        orbo = np.linalg.qr(rng.standard_normal((nao, nocc)))[0]
        mo1 = rng.standard_normal((n_dm, nao, nocc)) * 1e-2
        x = np.einsum('npi,qi->npq', mo1, orbo * 2)
    return x + np.swapaxes(x, -1, -2), mo1, orbo


@pytest.mark.gpu
@pytest.mark.parametrize('uhf', [False, True])
@pytest.mark.parametrize('device_in', [False, True])
def test_symmetrize_response_density(water2, uhf, device_in):
    """K of D = X + X^T from the half X's factors plus the transpose, against
    the oracle's K of the full density; J of D itself."""
    import torch
    from electrolites_amd import fastdf
    nao, cderi, pa, _, _ = water2
    dm, mo1, orbo = _cphf_dm(nao, 7, 3, 17 + uhf, uhf)
    if device_in:
        d = CAIArray(torch.from_numpy(np.ascontiguousarray(dm)).cuda())
        d.mo1 = CAIArray(torch.from_numpy(mo1).cuda())
        d.occ_coeff = CAIArray(torch.from_numpy(orbo).cuda())
        d.symmetrize = 1
    else:
        d = fastdf.tag_array(dm, mo1=mo1, occ_coeff=orbo, symmetrize=1)
    from tests.test_jk_gpu import FakeDF
    df = FakeDF(cderi, pa, nao)
    n0 = fastdf.STATS.get('symmetrized', 0)
    vj, vk = fastdf.get_jk(df, d, hermi=1)
    assert fastdf.STATS['symmetrized'] == n0 + 1
    if device_in:
        vj, vk = cai_host(vj), cai_host(vk)
    assert vk.shape == dm.shape and vj.shape == dm.shape
    flat = dm.reshape(-1, nao, nao)
    rj, rk = O.get_jk(cderi, pa, flat)
    assert rel(vj.reshape(-1, nao, nao), rj) < TOL
    assert rel(vk.reshape(-1, nao, nao), rk) < TOL


@pytest.mark.gpu
def test_symmetrize_without_factors_declines(water2):
    from electrolites_amd import fastdf
    from tests.test_jk_gpu import FakeDF
    nao, cderi, pa, _, _ = water2
    dm = fastdf.tag_array(np.eye(nao), symmetrize=1)
    with pytest.raises(fastdf.Declined):
        fastdf.get_jk(FakeDF(cderi, pa, nao), dm, hermi=1)
