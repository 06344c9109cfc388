"""Out-of-HBM residency and the block iterator: the page-locked host store,
the device prefix plus streamed remainder, and ``DF.loop`` (the reference's
``_resident_prefix`` / ``_blocks``, electrolites/fastdf.py:494-535,611-671;
``DF.loop`` for narrow tensors, electrolites/fastdffp32.py:72-141; its tests
compare the streamed J/K with the device-resident ones, tests/test_df.py)."""
import numpy as np
import pytest

from oracle import oracle as O


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ------------------------------------------------------------------ host-only
def test_resident_rows_rule(monkeypatch):
    from electrolites_amd import dfstore
    monkeypatch.setattr(dfstore, 'RESIDENT_FRAC', .5)
    monkeypatch.setattr(dfstore, 'RESIDENT_MAX_ROWS', 0)
    # 0.5 x 10000 bytes less 1000 reserved -> 4000 bytes of 80-byte rows
    assert dfstore.resident_rows(1000, 10, 8, 10000, 1000) == 50
    assert dfstore.resident_rows(30, 10, 8, 10000, 1000) == 30          # capped by naux
    assert dfstore.resident_rows(1000, 10, 8, 10000, 6000) == 0         # reserve exceeds the share
    monkeypatch.setattr(dfstore, 'RESIDENT_MAX_ROWS', 7)
    assert dfstore.resident_rows(1000, 10, 8, 10000, 1000) == 7
    monkeypatch.setattr(dfstore, 'RESIDENT_FRAC', 0.)
    assert dfstore.resident_rows(1000, 10, 8, 10000, 0) == 0


def test_block_rows_rule(monkeypatch):
    from electrolites_amd import dfstore
    monkeypatch.setattr(dfstore, 'STREAM_BLOCK_BYTES', 1000)
    assert dfstore.block_rows(10, 8, 100) == 12
    assert dfstore.block_rows(10, 8, 5) == 5
    assert dfstore.block_rows(1000, 8, 5) == 1                          # one row at least


def test_unpinned_host_buffer():
    import torch
    from electrolites_amd import dfstore
    x = dfstore.pinned_empty((3, 5), np.float32, pin=False)
    assert x.shape == (3, 5) and x.dtype == torch.float32 and not x.is_cuda
    assert dfstore.is_host(x) and dfstore.is_host(np.zeros(2))


# ----------------------------------------------------------------------- GPU
@pytest.fixture(scope='module')
def mol():
    from electrolites_amd.gto import water_cluster
    return water_cluster(3, seed=21, side=5.5, li=True)


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_host_store_build_equals_device(mol, dtype):
    """The host store holds exactly the device build's values (several
    chunks through the two staging buffers)."""
    import torch
    from electrolites_amd import fastdfcderi
    dev, idx = fastdfcderi.cholesky_eri(mol, chunk_bytes=300000, dtype=dtype)
    host, idx2 = fastdfcderi.cholesky_eri(mol, chunk_bytes=300000, dtype=dtype, storage='host')
    assert not host.is_cuda and host.is_pinned() and host.dtype == dev.dtype
    assert torch.equal(host, dev.cpu())
    assert np.array_equal(idx[0], idx2[0])


def _dms(nao, nocc, n, rng, tagged):
    from electrolites_amd import fastdf
    cs = [np.linalg.qr(rng.standard_normal((nao, nocc)))[0] for _ in range(n)]
    d = np.array([2 * c @ c.T for c in cs])
    if tagged:
        d = fastdf.tag_array(d, mo_coeff=np.array(cs), mo_occ=np.full((n, nocc), 2.))
    return d


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
@pytest.mark.parametrize('frac,max_rows', [(0., 0), (.5, 37), (.5, 0)])
@pytest.mark.parametrize('tagged,n_dm', [(True, 1), (False, 2)])
def test_host_resident_jk(mol, monkeypatch, dtype, frac, max_rows, tagged, n_dm):
    """J/K streamed from the host store (no prefix, a 37-row prefix, the
    whole tensor resident) in blocks of a few rows equal the device-resident
    J/K of the same tensor to 1e-12 (the same products, summed block by
    block)."""
    import torch
    from electrolites_amd import dfstore
    from electrolites_amd.df import DF
    monkeypatch.setattr(dfstore, 'RESIDENT_FRAC', frac)
    monkeypatch.setattr(dfstore, 'RESIDENT_MAX_ROWS', max_rows)
    import electrolites_amd.fastdfcderi as FC
    monkeypatch.setattr(FC, 'FP32', dtype == 'float32')
    dev = DF(mol).build()
    host = DF(mol, storage='host').build()
    assert dfstore.is_host(host.cderi())
    npairs = host.cderi().shape[1]             # the stored (screened-in) pairs
    monkeypatch.setattr(dfstore, 'STREAM_BLOCK_BYTES', 29 * npairs * (8 if dtype == 'float64' else 4))
    rng = np.random.default_rng(5 + n_dm)
    dm = _dms(mol.nao, 9, n_dm, rng, tagged)
    before = dfstore.stats()['blocks']
    for _ in range(2):                     # second call reuses the cached prefix
        vj, vk = host.get_jk(dm, hermi=1)
        rj, rk = dev.get_jk(dm, hermi=1)
        assert rel(vj, rj) < 1e-12 and rel(vk, rk) < 1e-12
    # and against the oracle on the stored values (widened for FP32)
    stored = host.cderi().double().numpy()
    pa = np.asarray(host._cderi_idx[0]).astype(np.int64)
    oj, ok = O.get_jk(stored, pa, np.asarray(dm))
    assert rel(vj, oj) < 1e-12 and rel(vk, ok) < 1e-12
    st = dfstore.stats()
    naux = host.cderi().shape[0]
    res = 0 if frac == 0 else (min(max_rows, naux) if max_rows else naux)
    assert st['resident_rows'] == res and st['streamed_rows'] == naux - res
    assert st['blocks'] - before >= 2 * -(-naux // 29)
    vj2, _ = host.get_jk(dm, hermi=1, with_k=False)
    assert rel(vj2, rj) < 1e-12
    host.reset()


@pytest.mark.gpu
def test_host_resident_low_precision_k(mol, monkeypatch):
    """Split-BF16 K streamed from the host store stays within the mode's
    bar against the FP64 oracle."""
    from electrolites_amd import dfstore
    from electrolites_amd.df import DF
    monkeypatch.setattr(dfstore, 'RESIDENT_FRAC', 0.)
    host = DF(mol, storage='host').build()
    monkeypatch.setattr(dfstore, 'STREAM_BLOCK_BYTES', 50 * host.cderi().shape[1] * 8)
    cderi = host.cderi().numpy()
    pa = host._cderi_idx[0]
    dm = _dms(mol.nao, 9, 1, np.random.default_rng(3), True)
    host.k_precision = 'bf16x3'
    _, vk = host.get_jk(dm, hermi=1)
    rk = O.get_k(cderi, pa, np.asarray(dm)[0])
    assert rel(vk[0], rk) < 1e-5


def _unpack_rows(cderi, pa, nao):
    out = np.zeros((cderi.shape[0], nao, nao))
    i, j = np.divmod(pa, nao)
    out[:, i, j] = cderi
    out[:, j, i] = cderi
    return out


@pytest.mark.gpu
@pytest.mark.parametrize('storage', ['device', 'host'])
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_df_loop_blocks(mol, monkeypatch, storage, dtype):
    """``DF.loop`` visits every aux row once, in order, with the packed rows
    (widened exactly for FP32 storage) and their symmetric unpacking equal
    to the tensor's; ragged last block."""
    import torch
    import electrolites_amd.fastdfcderi as FC
    from electrolites_amd.df import DF
    monkeypatch.setattr(FC, 'FP32', dtype == 'float32')
    df = DF(mol, storage=storage).build()
    cderi = df.cderi()
    ref = cderi.cpu().double().numpy()
    pa = df._cderi_idx[0]
    naux = ref.shape[0]
    got, L0 = [], 0
    for dense, packed in df.loop(blksize=41):
        nL = packed.shape[0]
        assert packed.dtype == torch.float64 and dense.shape == (nL, mol.nao, mol.nao)
        p = packed.cpu().numpy()
        assert np.array_equal(p, ref[L0:L0 + nL])
        assert np.array_equal(dense.cpu().numpy(), _unpack_rows(ref[L0:L0 + nL], pa, mol.nao))
        got.append(nL)
        L0 += nL
    assert L0 == naux and got[:-1] == [41] * (len(got) - 1)
    n = sum(p.shape[0] for _, p in df.loop(blksize=1000, unpack=False))
    assert n == naux
    if storage == 'host':
        df.reset()
