"""The benchmark configurations on the HIP path, each against the oracle or
what the mathematics fixes.

* config 1 -- (H2O)4, 6 A cube, seed 0, FP64 DF-RHF to 1e-10 Eh: energy
  against the oracle's DF-RHF (1e-9 Eh) and J/K of the converged density
  against the oracle's J/K on the same cderi (1e-12);
* config 2 -- (H2O)64 at 1 g/cm^3 (nao 1536), rank-320 PSD density, the
  full get_jk: J/K over a block of aux rows against the oracle on those rows
  (J and K are sums over aux rows, so a block checks the kernels at the full
  nao and pair count; 1e-12), and over the whole tensor the identities
  ``sum d.vj = |rho|^2`` (rho from a library GEMV), ``K = K^T`` and
  ``tr(D K) = sum_L |F^T A_L F|^2 >= 0``;
* config 3 -- (H2O)32 mixed precision (FP32, then split-BF16, K while the
  DIIS error is above 1e-5, FP64 finish): |dE| against the all-FP64 run
  <= 1e-8 Eh, and the per-cycle J/K error against FP64 J/K of the same
  density recorded and bounded;
* config 4 -- the 48-molecule mix, 1-8 densities each (UHF pairs and
  full-rank response densities), one grouped launch: every item
  bit-identical to its single get_jk, the smallest items against the
  oracle;
* config 5 -- a Poisson stream over the mix with requests above nao 400
  (device diagonalisations from the serving loop's worker threads): each
  energy the same alone and in the mix, and within 1e-8 Eh of the plain
  FP64 solve.
"""
import numpy as np
import pytest

from oracle import oracle as O


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


class ResidentDF:
    """A DF object over a given device cderi and its pair addresses."""

    def __init__(self, cderi, idx):
        self._cderi = {0: cderi}
        self._cderi_idx = idx

    def build(self):
        raise AssertionError('cderi given')


@pytest.mark.gpu
def test_config1_energy_and_jk():
    import torch
    from electrolites_amd import fastdf
    from electrolites_amd.gto import water_cluster
    from electrolites_amd.scf import RHF
    mol = water_cluster(4, seed=0, side=6.0)
    print(f'config 1: (H2O)4 nao {mol.nao} naux {mol.naux}', flush=True)
    mf = RHF(mol, conv_tol=1e-10)
    e = mf.kernel()
    assert mf.converged
    s = O.DFSystem(mol.ao.as_dict(), mol.aux.as_dict(), mol.atom_charges(), mol.atom_coords(),
                   mol.nelectron, mol.energy_nuc())
    ref = O.rhf(s, conv_tol=1e-10)
    assert ref['converged'] and abs(e - ref['e_tot']) < 1e-9, (e, ref['e_tot'])
    # J/K of the converged density against the oracle on the GPU-built cderi
    df = mf.with_df
    dm = mf.dm
    vj, vk = df.get_jk(dm, hermi=1)
    store = df.cderi().cpu().numpy()
    pa = np.asarray(df._cderi_idx[0]).astype(np.int64)
    rj, rk = O.get_jk(store, pa, dm.cpu().numpy())
    assert rel(vj.cpu(), rj) < 1e-12 and rel(vk.cpu(), rk) < 1e-12
    # and the oracle's cderi is the same tensor to the build's bar
    ao, aux = mol.ao.as_dict(), mol.aux.as_dict()
    full_pa, _ = O.pair_addresses(mol.nao)
    oc = O.cholesky_eri(O.int3c2e(ao, aux), O.int2c2e(aux), full_pa)
    pos = np.searchsorted(full_pa, pa)
    assert rel(store, oc[:, pos]) < 1e-9


@pytest.fixture(scope='module')
def h2o64():
    import torch
    from electrolites_amd import fastdfcderi
    from electrolites_amd.gto import water_cluster
    mol = water_cluster(64, seed=0)
    print(f'config 2: (H2O)64 nao {mol.nao} naux {mol.naux}; building the cderi', flush=True)
    cderi, idx = fastdfcderi.cholesky_eri(mol)
    torch.cuda.synchronize()
    rng = np.random.default_rng(0)
    nocc = 320
    c = np.linalg.qr(rng.standard_normal((mol.nao, nocc)))[0]
    yield mol, cderi, idx, c
    del cderi
    torch.cuda.empty_cache()


@pytest.mark.gpu
def test_config2_row_block_against_oracle(h2o64):
    """J and K over 40 aux rows in the middle of the (H2O)64 tensor, through
    the same entry points, against the oracle on those rows."""
    import torch
    from electrolites_amd import fastdf
    mol, cderi, idx, c = h2o64
    nao, nocc = mol.nao, c.shape[1]
    L0, nb = 3001, 40
    blk = cderi[L0:L0 + nb]                            # a row view, stride npairs
    dm = torch.from_numpy(2 * c @ c.T).cuda()
    fastdf.tag_array(dm, mo_coeff=torch.from_numpy(c).cuda(),
                     mo_occ=torch.full((nocc,), 2., dtype=torch.float64, device='cuda'))
    vj, vk = fastdf.get_jk(ResidentDF(blk, idx), dm, hermi=1)
    rows = blk.cpu().numpy()
    pa = np.asarray(idx[0]).astype(np.int64)
    rj = O.get_j(rows, pa, 2 * c @ c.T)
    rk = O.get_k_factor(rows, pa, c * np.sqrt(2.))
    assert rel(vj.cpu(), rj) < 1e-12
    assert rel(vk.cpu(), rk) < 1e-12


@pytest.mark.gpu
def test_config2_full_identities(h2o64):
    import torch
    from electrolites_amd import fastdf
    mol, cderi, idx, c = h2o64
    nao, nocc = mol.nao, c.shape[1]
    C = torch.from_numpy(c).cuda()
    dm = 2 * C @ C.T
    fastdf.tag_array(dm, mo_coeff=C, mo_occ=torch.full((nocc,), 2., dtype=torch.float64, device='cuda'))
    df = ResidentDF(cderi, idx)
    vj, vk = fastdf.get_jk(df, dm, hermi=1)
    # sum_pq D_pq J_pq = |rho|^2, rho = L d with the pair vector of the
    # symmetry factors (library GEMV, independent of the J kernels)
    pa = torch.from_numpy(np.asarray(idx[0]).astype(np.int64)).cuda()
    rows, cols = pa // nao, pa % nao
    d = 2. * dm[rows, cols]
    d[rows == cols] *= .5
    rho = cderi @ d
    e_j = float((dm * vj).sum())
    assert abs(e_j - float(rho @ rho)) < 1e-11 * abs(e_j)
    # K symmetric and tr(D K) = sum_L |F^T A_L F|_F^2 (F = C sqrt 2), which
    # the K kernels never form; over the aux rows in slices
    assert float((vk - vk.T).abs().max()) <= 1e-13 * float(vk.abs().max())
    F = C * 2 ** .5
    s = 0.
    for L0 in range(0, cderi.shape[0], 256):
        a = torch.zeros((min(256, cderi.shape[0] - L0), nao, nao), dtype=torch.float64, device='cuda')
        a[:, rows, cols] = cderi[L0:L0 + a.shape[0]]
        a[:, cols, rows] = cderi[L0:L0 + a.shape[0]]
        m = F.T @ a @ F
        s += float((m * m).sum())
        del a, m
    tr = float((dm * vk).sum())
    assert tr > 0 and abs(tr - s) < 1e-11 * s, (tr, s)


@pytest.mark.gpu
@pytest.mark.parametrize('low,bound', [('fp32', 5e-6), ('bf16x3', 5e-5)])
def test_config3_mixed_precision(low, bound):
    from electrolites_amd.gto import water_cluster
    from electrolites_amd.scf import RHF
    mol = water_cluster(32, seed=0)
    print(f'config 3: (H2O)32 nao {mol.nao} naux {mol.naux}, low precision {low}', flush=True)
    mf64 = RHF(mol, conv_tol=1e-10)
    e64 = mf64.kernel()
    assert mf64.converged
    mf = RHF(mol, conv_tol=1e-10, k_precision='mixed', low_precision=low, df=mf64.with_df,
             precision_report=True)
    e = mf.kernel()
    assert mf.converged
    assert abs(e - e64) <= 1e-8, (e, e64)
    low_cycles = [h for h in mf.history if h['k_precision'] == low]
    assert low_cycles and mf.history[-1]['k_precision'] == 'fp64'
    for h in low_cycles:
        assert h['j_rel_err'] == 0.                 # J is FP64 arithmetic on an FP64 tensor
        assert 0. < h['k_rel_err'] < bound, h
    print(f'config 3 {low}: {len(low_cycles)} low cycles, max K rel err '
          f'{max(h["k_rel_err"] for h in low_cycles):.2e}, |dE| {abs(e - e64):.2e}', flush=True)


@pytest.mark.gpu
def test_config4_batch_mix():
    import torch
    from electrolites_amd import batch, fastdf, fastdfcderi, serve
    mols = serve.molecule_mix(48, seed=0)
    print(f'config 4: 48 molecules, nao {min(m.nao for m in mols)}-{max(m.nao for m in mols)}', flush=True)
    rng = np.random.default_rng(1)
    dfs, dms, host = [], [], []
    for mol in mols:
        cderi, idx = fastdfcderi.cholesky_eri(mol)
        df = ResidentDF(cderi, idx)
        nao, nocc = mol.nao, mol.nocc
        n_dm = int(rng.integers(1, 9))
        nuhf = min(2, n_dm)
        cs = np.array([np.linalg.qr(rng.standard_normal((nao, nocc)))[0] for _ in range(nuhf)])
        uhf = torch.from_numpy(np.einsum('nik,njk->nij', cs, cs)).cuda()
        fastdf.tag_array(uhf, mo_coeff=torch.from_numpy(cs).cuda(),
                         mo_occ=torch.ones((nuhf, nocc), dtype=torch.float64, device='cuda'))
        dfs.append(df)
        dms.append(uhf)
        host.append((mol, cs, None))
        if n_dm > 2:
            a = rng.standard_normal((n_dm - 2, nao, nao)) * 1e-3
            resp = .5 * (a + a.transpose(0, 2, 1))
            dfs.append(df)
            dms.append(torch.from_numpy(resp).cuda())
            host.append((mol, None, resp))
    got = batch.batched_get_jk(dfs, dms, hermi=1)
    torch.cuda.synchronize()
    print(f'config 4: {len(dms)} items, {sum(d.shape[0] for d in dms)} densities', flush=True)
    for df, dm, (vj, vk) in zip(dfs, dms, got):
        rj, rk = fastdf.get_jk(df, dm, hermi=1)
        assert torch.equal(vj.reshape(rj.shape), rj), 'J differs alone vs in the batch'
        assert torch.equal(vk.reshape(rk.shape), rk), 'K differs alone vs in the batch'
    # the smallest molecules' items against the oracle
    order = np.argsort([m.nao for m, _, _ in host], kind='stable')
    for i in order[:4]:
        mol, cs, resp = host[i]
        df, (vj, vk) = dfs[i], got[i]
        store = df._cderi[0].cpu().numpy()
        pa = np.asarray(df._cderi_idx[0]).astype(np.int64)
        d = np.einsum('nik,njk->nij', cs, cs) if resp is None else resp
        rj, rk = O.get_jk(store, pa, d)
        assert rel(vj.cpu().reshape(rj.shape), rj) < 1e-12
        assert rel(vk.cpu().reshape(rk.shape), rk) < (1e-12 if resp is None else 1e-11)


@pytest.mark.gpu
def test_config5_serving_stream():
    """Poisson arrivals (8 req/s) over eight molecules of the mix, three of
    them above nao 400 whose diagonalisations go to the device eigensolver
    thread from the absorb threads at once: each result is the same alone
    and in the mix, and within 1e-8 Eh of the plain FP64 driver."""
    from electrolites_amd import serve
    from electrolites_amd.scf import RHF
    mols = serve.molecule_mix(8, seed=6)
    assert sum(m.nao > 400 for m in mols) >= 3
    arr = serve.poisson_arrivals(len(mols), 8., seed=2)
    reqs = [serve.Request(m, arr[i], rid=i, k_precision='mixed') for i, m in enumerate(mols)]
    srv = serve.Server(conv_tol=1e-10, host_diag_max=400, purify_min=None)
    print(f'config 5: {len(mols)} requests nao {sorted(m.nao for m in mols)}, '
          f'device diagonalisation above {srv.host_diag_max}', flush=True)
    stats = srv.serve(reqs)
    assert stats['n'] == len(reqs) and stats['converged'] == len(reqs)
    assert max(tr[1] for tr in srv.trace) >= 3, 'requests did not overlap'
    assert stats['p99_s'] >= stats['p50_s'] > 0
    for r in sorted(reqs, key=lambda r: r.rid):
        alone = serve.Server(conv_tol=1e-10, host_diag_max=400, purify_min=None)
        a = serve.Request(r.mol, k_precision='mixed')
        alone.serve([a])
        assert abs(r.e_tot - a.e_tot) <= 1e-12, (r.rid, r.e_tot, a.e_tot)
        assert r.cycles == a.cycles
        e64 = RHF(r.mol, conv_tol=1e-10).kernel()
        assert abs(r.e_tot - e64) <= 1e-8, (r.rid, r.e_tot, e64)
        print(f'config 5: request {r.rid} nao {r.mol.nao}: |dE| vs FP64 {abs(r.e_tot - e64):.1e}', flush=True)
