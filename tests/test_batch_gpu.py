"""Grouped J/K: every item equals its single-molecule get_jk exactly (the
alone-vs-in-the-mix requirement), and a HIP-graph replay equals the eager run
on every replay."""
import numpy as np
import pytest

from oracle import oracle as O


class FakeDF:
    def __init__(self, cderi, pa, nao):
        import torch
        self._cderi = {0: torch.from_numpy(np.ascontiguousarray(cderi)).cuda()}
        rows, cols = np.divmod(pa, nao)
        self._cderi_idx = (pa.astype(np.int32), np.where(rows == cols)[0])


def make_batch(seed=0):
    from electrolites_amd import fastdf
    rng = np.random.default_rng(seed)
    dfs, dms = [], []
    # (nao, naux, nocc, n_dm, tagged, screened)
    spec = [(37, 60, 5, 1, True, False), (72, 330, 15, 2, True, False), (129, 500, 40, 3, False, False),
            (90, 150, 12, 1, True, True), (161, 610, 45, 1, True, False), (210, 260, 170, 1, True, False)]
    for nao, naux, nocc, n_dm, tagged, screened in spec:
        pa, _ = O.pair_addresses(nao)
        if screened:
            rows, cols = np.divmod(pa, nao)
            keep = (rng.random(len(pa)) > .3) | (rows == cols)
            pa = pa[keep]
        cd = rng.standard_normal((naux, len(pa))) * .1
        dfs.append(FakeDF(cd, pa, nao))
        cs = [np.linalg.qr(rng.standard_normal((nao, nocc)))[0] for _ in range(n_dm)]
        d = np.array([2 * c @ c.T for c in cs])
        if tagged:
            d = fastdf.tag_array(d, mo_coeff=np.array(cs), mo_occ=np.full((n_dm, nocc), 2.))
        dms.append(d)
    return dfs, dms


@pytest.mark.gpu
@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
def test_grouped_equals_single(precision):
    import torch
    from electrolites_amd import batch, fastdf
    dfs, dms = make_batch(1)
    got = batch.batched_get_jk(dfs, dms, hermi=1, precision=precision)
    for df, dm, (vj, vk) in zip(dfs, dms, got):
        df.k_precision = precision
        rj, rk = fastdf.get_jk(df, dm, hermi=1)
        assert np.array_equal(vj.cpu().numpy(), rj), 'J differs alone vs in the batch'
        assert np.array_equal(vk.cpu().numpy(), rk), 'K differs alone vs in the batch'


@pytest.mark.gpu
def test_grouped_matches_oracle():
    from electrolites_amd import batch
    dfs, dms = make_batch(2)
    got = batch.batched_get_jk(dfs, dms, hermi=1)
    for df, dm, (vj, vk) in zip(dfs, dms, got):
        cd = df._cderi[0].cpu().numpy()
        pa = df._cderi_idx[0]
        rj, rk = O.get_jk(cd, pa, np.asarray(dm))
        assert np.abs(vj.cpu().numpy() - rj).max() <= 1e-12 * np.abs(rj).max()
        assert np.abs(vk.cpu().numpy() - rk).max() <= 1e-11 * np.abs(rk).max()


@pytest.mark.gpu
def test_graph_replay():
    """Capture once, replay three times with eager work in between; every
    replay matches the eager result."""
    import torch
    from electrolites_amd import batch, fastdf
    dfs, dms = make_batch(3)
    entries = []
    for df, dm in zip(dfs, dms):
        cd = fastdf._cderi_of(df)
        nao = dm.shape[-1]
        lay = fastdf._layout(df, cd, nao)
        dms3 = torch.from_numpy(np.asarray(dm).reshape(-1, nao, nao)).cuda()
        fl, fr = fastdf.factorize_dm(dm, 1)
        d = fastdf._sparse_dm(dms3, lay, 1)
        entries.append(batch.Entry(cd, nao, dms3.shape[0], fl.shape[-1], dm_pairs=d, fl=fl, fr=fr,
                                   pair_map=lay.map))
    g = batch.GroupedJK(entries, bn=128)
    g.run()
    torch.cuda.synchronize()
    ref = [(e.vj_pairs.clone(), e.vk.clone()) for e in entries]
    g.capture_graph()
    for rep in range(3):
        for e in entries:
            e.vj_pairs.zero_()
            e.vk.zero_()
        junk = torch.randn(1000, 1000, device='cuda', dtype=torch.float64) @ \
            torch.randn(1000, 1000, device='cuda', dtype=torch.float64)
        torch.cuda.synchronize()
        g.run()
        torch.cuda.synchronize()
        for e, (rj, rk) in zip(entries, ref):
            assert torch.equal(e.vj_pairs, rj), rep
            assert torch.equal(e.vk, rk), rep
    g.close()


def _entries(dfs, dms):
    import torch
    from electrolites_amd import batch, fastdf
    entries = []
    for df, dm in zip(dfs, dms):
        cd = fastdf._cderi_of(df)
        nao = dm.shape[-1]
        lay = fastdf._layout(df, cd, nao)
        dms3 = torch.from_numpy(np.asarray(dm).reshape(-1, nao, nao)).cuda()
        fl, fr = fastdf.factorize_dm(dm, 1)
        d = fastdf._sparse_dm(dms3, lay, 1)
        entries.append(batch.Entry(cd, nao, dms3.shape[0], fl.shape[-1], dm_pairs=d, fl=fl, fr=fr,
                                   pair_map=lay.map))
    return entries


@pytest.mark.gpu
@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
def test_passes_and_waves(precision, monkeypatch):
    """Items cut into several aux passes (a small pass bound) and run in one
    wave or one wave per item: K equals the single entry's bit for bit."""
    import torch
    from electrolites_amd import batch, fastdf
    monkeypatch.setenv('MDF_K_BLOCK_BYTES', str(300_000))
    dfs, dms = make_batch(4)
    ref = []
    for df, dm in zip(dfs, dms):
        df.k_precision = precision
        ref.append(fastdf.get_jk(df, dm, hermi=1))
    nk = sum(np.asarray(dm).reshape(-1, dm.shape[-1], dm.shape[-1]).shape[0] for dm in dms)
    for limit, waves in ((0, 1), (1, nk)):
        entries = _entries(dfs, dms)
        g = batch.GroupedJK(entries, precision=precision, tune=False, work_limit=limit)
        info = g.tiles()
        assert info['k_waves'] == waves
        assert info['k_launches'] > 3 * waves, 'expected several aux passes'
        g.run()
        torch.cuda.synchronize()
        for e, (rj, rk) in zip(entries, ref):
            assert np.array_equal(e.vk.cpu().numpy().reshape(rk.shape), rk), (limit, precision)
        g.close()
    if precision == 'fp64':
        for df, dm, (rj, rk) in zip(dfs, dms, ref):
            _, ok = O.get_jk(df._cderi[0].cpu().numpy(), df._cderi_idx[0], np.asarray(dm))
            assert np.abs(rk - ok).max() <= 1e-11 * np.abs(ok).max()


@pytest.mark.gpu
@pytest.mark.parametrize('precision', ['fp64', 'bf16x3'])
def test_grouped_fp32_storage(precision):
    """A batch of FP32-stored cderis: every item equals its single-entry
    get_jk bit for bit."""
    import torch
    from electrolites_amd import batch, fastdf
    dfs, dms = make_batch(5)
    for df in dfs:
        df._cderi = {0: df._cderi[0].float()}
    got = batch.batched_get_jk(dfs, dms, hermi=1, precision=precision)
    for df, dm, (vj, vk) in zip(dfs, dms, got):
        df.k_precision = precision
        rj, rk = fastdf.get_jk(df, dm, hermi=1)
        assert np.array_equal(vj.cpu().numpy(), rj)
        assert np.array_equal(vk.cpu().numpy(), rk)
