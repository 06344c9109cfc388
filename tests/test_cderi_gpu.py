"""GPU parity of the metric linear algebra and the whole cderi build.

Linear algebra is checked against LAPACK (scipy) on random SPD matrices of
ragged sizes; the cderi against the oracle's (MD integrals + scipy Cholesky
and triangular solve), the reference suite's bar being 1e-9 on the tensor
(tests/test_dfcderi.py:51).  J/K built on the GPU cderi are checked against
the oracle's J/K on its own cderi.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def spd(n, rng, cond=1e4):
    q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    w = np.geomspace(1., 1. / cond, n)
    return (q * w) @ q.T


def _call(name, *args):
    from electrolites_amd import _lib
    return _lib.call(name, *args)


@pytest.mark.gpu
@pytest.mark.parametrize('n', [1, 37, 128, 129, 300, 700])
def test_potrf_trtri(n):
    import torch
    from electrolites_amd import _dev, _lib
    rng = np.random.default_rng(n)
    a = spd(n, rng)
    A = torch.from_numpy(a.copy()).cuda()
    info = torch.zeros(1, dtype=torch.int32, device='cuda')
    work = torch.empty(_lib.query('mdf_potrf_workspace', n), dtype=torch.uint8, device='cuda')
    _call('mdf_potrf', _dev.ptr(A), n, n, _dev.ptr(info), _dev.ptr(work), work.numel(), None)
    assert int(info.item()) == 0
    low = np.tril(A.cpu().numpy())
    ref = scipy.linalg.cholesky(a, lower=True)
    assert rel(low, ref) < 1e-12
    X = torch.empty((n, n), dtype=torch.float64, device='cuda')
    w2 = torch.empty(_lib.query('mdf_trtri_workspace', n), dtype=torch.uint8, device='cuda')
    _call('mdf_trtri', _dev.ptr(A), n, n, _dev.ptr(work), _dev.ptr(X), n, _dev.ptr(w2), w2.numel(), None)
    x = X.cpu().numpy()
    assert np.abs(np.triu(x, 1)).max() == 0.
    assert rel(x, scipy.linalg.solve_triangular(ref, np.eye(n), lower=True)) < 1e-10
    # C = Linv B (triangular multiply)
    b = rng.standard_normal((n, 333))
    B = torch.from_numpy(b).cuda()
    C = torch.empty_like(B)
    _call('mdf_trmm', _dev.ptr(X), n, n, _dev.ptr(B), 333, 333, _dev.ptr(C), 333, None)
    assert rel(C.cpu().numpy(), x @ b) < 1e-13


@pytest.mark.gpu
def test_potrf_reports_non_pd():
    import torch
    from electrolites_amd import _dev, _lib
    rng = np.random.default_rng(1)
    a = spd(200, rng)
    a[150, 150] = -1.
    A = torch.from_numpy(a).cuda()
    info = torch.zeros(1, dtype=torch.int32, device='cuda')
    work = torch.empty(_lib.query('mdf_potrf_workspace', 200), dtype=torch.uint8, device='cuda')
    _call('mdf_potrf', _dev.ptr(A), 200, 200, _dev.ptr(info), _dev.ptr(work), work.numel(), None)
    assert int(info.item()) > 0


@pytest.mark.gpu
@pytest.mark.parametrize('transb,tri_a,lower_c', [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 0, 1)])
def test_gemm_variants(transb, tri_a, lower_c):
    import torch
    from electrolites_amd import _dev
    rng = np.random.default_rng(transb * 4 + tri_a * 2 + lower_c)
    M, K = 300, 211
    N = M if lower_c else 257
    a = rng.standard_normal((M, K))
    if tri_a:
        K = M
        a = np.tril(rng.standard_normal((M, M)))
    b = rng.standard_normal((N, K) if transb else (K, N))
    c = rng.standard_normal((M, N))
    A, B, C = (torch.from_numpy(x.copy()).cuda() for x in (a, b, c))
    _call('mdf_gemm', transb, tri_a, lower_c, M, N, K, ctypes.c_double(-.5), _dev.ptr(A), K, _dev.ptr(B),
          b.shape[1], ctypes.c_double(2.), _dev.ptr(C), N, None)
    ref = -.5 * a @ (b.T if transb else b) + 2. * c
    got = C.cpu().numpy()
    if lower_c:
        ti = np.arange(M)[:, None] // 128
        tj = np.arange(N)[None, :] // 128
        mask = ti >= tj
        assert rel(got[mask], ref[mask]) < 1e-13
    else:
        assert rel(got, ref) < 1e-13


@pytest.fixture(scope='module')
def built():
    from electrolites_amd.gto import water_cluster
    from electrolites_amd.df import DF
    mol = water_cluster(2, seed=3, side=5.0, li=True)
    df = DF(mol).build()
    ao, aux = mol.ao.as_dict(), mol.aux.as_dict()
    pa, _ = O.pair_addresses(mol.nao)
    ref = O.cholesky_eri(O.int3c2e(ao, aux), O.int2c2e(aux), pa)
    return mol, df, ref, pa


@pytest.mark.gpu
@pytest.mark.parametrize('chunk_bytes', [None, 200000])
def test_cderi_matches_oracle(built, chunk_bytes):
    from electrolites_amd import fastdfcderi
    mol, df, ref, pa = built
    cderi, idx = fastdfcderi.cholesky_eri(mol, chunk_bytes=chunk_bytes)
    got = cderi.cpu().numpy()
    # the tensor stores the screened-in pairs, in packed order, named by idx
    pos = np.searchsorted(pa, idx[0])
    assert np.array_equal(pa[pos], idx[0]) and np.all(np.diff(idx[0]) > 0)
    assert got.shape == (ref.shape[0], len(pos))
    assert rel(got, ref[:, pos]) < 1e-9
    # the dropped pairs carry nothing (the oracle computes them unscreened)
    dropped = np.setdiff1d(np.arange(ref.shape[1]), pos)
    assert np.abs(ref[:, dropped]).max(initial=0.) < 1e-12 * np.abs(ref).max()
    rows, cols = np.divmod(idx[0], mol.nao)
    assert np.array_equal(idx[1], np.nonzero(rows == cols)[0])


@pytest.mark.gpu
def test_jk_on_built_cderi(built):
    from electrolites_amd import fastdf
    mol, df, ref, pa = built
    rng = np.random.default_rng(4)
    nocc = mol.nelectron // 2
    c = np.linalg.qr(rng.standard_normal((mol.nao, nocc)))[0]
    dm = fastdf.tag_array(2 * c @ c.T, mo_coeff=c, mo_occ=np.full(nocc, 2.))
    vj, vk = df.get_jk(dm)
    rj, rk = O.get_jk(ref, pa, np.asarray(dm))
    assert rel(vj, rj) < 1e-11
    assert rel(vk, rk) < 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize('chunk_bytes', [None, 200000])
def test_fp32_storage_build(built, chunk_bytes):
    """The FP32-storage build (FASTDF_FP32) stores exactly the FP64 build's
    values rounded to nearest FP32, chunk by chunk."""
    import torch
    from electrolites_amd import fastdfcderi
    mol, df, ref, pa = built
    c64, _ = fastdfcderi.cholesky_eri(mol, chunk_bytes=chunk_bytes)
    c32, idx = fastdfcderi.cholesky_eri(mol, chunk_bytes=chunk_bytes, dtype='float32')
    assert c32.dtype == torch.float32 and c32.shape == c64.shape
    assert torch.equal(c32, c64.float())
    assert np.array_equal(idx[0], fastdfcderi.cholesky_eri(mol, chunk_bytes=chunk_bytes)[1][0])


@pytest.mark.gpu
def test_eigen_fallback_on_singular_metric(monkeypatch):
    """A fitting set with one shell duplicated has an exactly singular
    metric; shifted by -1e-11 (planted on the host, so the pivot is negative
    whatever the rounding) the Cholesky factorisation fails and the build
    takes the eigen filter (electrolites/fastdfcderi.py:203-210): its cderi
    spans the same fit as the oracle's eigen route (the Gram matrix
    cderi^T cderi and J/K agree; the rows differ by an orthogonal rotation)."""
    import torch
    from electrolites_amd import fastdfcderi, fastdf, ints
    from electrolites_amd.gto import Mole, basis as B
    from electrolites_amd.gto.cluster import water_cluster_atoms
    atoms = water_cluster_atoms(2, seed=3, side=5.0)
    aux = B.load('etjk', ['O', 'H'])
    aux['H'] = aux['H'] + [aux['H'][1]]               # one shell twice on every H
    mol = Mole(atoms, basis='def2-svp', auxbasis=aux)
    shift = 1e-11
    orig = ints.int2c2e
    monkeypatch.setattr(ints, 'int2c2e', lambda b: orig(b) - shift * torch.eye(b.nao, dtype=torch.float64,
                                                                              device='cuda'))
    n0 = fastdfcderi.STATS['eigen_fallback']
    cderi, idx = fastdfcderi.cholesky_eri(mol)
    assert fastdfcderi.STATS['eigen_fallback'] == n0 + 1
    got = cderi.cpu().numpy()
    ao_d, aux_d = mol.ao.as_dict(), mol.aux.as_dict()
    j2c = O.int2c2e(aux_d) - shift * np.eye(mol.naux)
    with pytest.raises(np.linalg.LinAlgError):
        np.linalg.cholesky(j2c)
    pa = np.asarray(idx[0]).astype(np.int64)
    ref = O.cholesky_eri(O.int3c2e(ao_d, aux_d), j2c, pa)
    assert got.shape == ref.shape and got.shape[0] < mol.naux      # the null directions dropped
    assert rel(got.T @ got, ref.T @ ref) < 1e-9
    rng = np.random.default_rng(2)
    nocc = mol.nelectron // 2
    c = np.linalg.qr(rng.standard_normal((mol.nao, nocc)))[0]
    dm = fastdf.tag_array(2 * c @ c.T, mo_coeff=c, mo_occ=np.full(nocc, 2.))

    class D:
        _cderi = {0: cderi}
        _cderi_idx = idx

    vj, vk = fastdf.get_jk(D(), dm, hermi=1)
    rj, rk = O.get_jk(ref, pa, np.asarray(dm))
    assert rel(vj, rj) < 1e-9 and rel(vk, rk) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize('transb,tri_a,lower_c', [(0, 2, 0), (1, 2, 0), (1, 0, 1), (0, 0, 0)])
def test_gemm_batched(transb, tri_a, lower_c):
    """mdf_gemm_batched against numpy: upper-triangular A (the K loop of a
    row tile starts at its first row), shared operands (stride 0), ragged
    sizes, padded leading dimensions, beta accumulation."""
    import torch
    from electrolites_amd import _dev, _lib
    rng = np.random.default_rng(10 * transb + 3 * tri_a + lower_c)
    nb, M = 5, 261
    N = M if lower_c else 300
    K = M if tri_a else 147
    a = rng.standard_normal((nb, M, K + 3))
    if tri_a == 2:
        a[:, :, :K] = np.triu(a[:, :, :K])
    b = rng.standard_normal((N, K + 2) if transb else (K, N + 2))          # shared: stride 0
    c = rng.standard_normal((nb, M, N + 1))
    A, B, C = (torch.from_numpy(x.copy()).cuda() for x in (a, b, c))
    _lib.call('mdf_gemm_batched', transb, tri_a, lower_c, M, N, K, ctypes.c_double(.7),
              _dev.ptr(A), K + 3, M * (K + 3), _dev.ptr(B), b.shape[1], 0, ctypes.c_double(-1.5),
              _dev.ptr(C), N + 1, M * (N + 1), nb, None)
    got = C.cpu().numpy()
    bb = b[:, :K].T if transb else b[:K, :N]
    for i in range(nb):
        ref = .7 * a[i, :, :K] @ bb + -1.5 * c[i, :, :N]
        g = got[i, :, :N]
        if lower_c:
            mask = (np.arange(M)[:, None] // 128) >= (np.arange(N)[None, :] // 128)
            assert rel(g[mask], ref[mask]) < 1e-13
        else:
            assert rel(g, ref) < 1e-13
        assert np.array_equal(got[i, :, N:], c[i, :, N:])          # the padding column untouched


@pytest.mark.gpu
@pytest.mark.parametrize('transb', [0, 1])
@pytest.mark.parametrize('init', [False, True])
def test_gemm_pairs_batched(transb, init):
    """mdf_gemm_pairs_batched against numpy: each P_z = A_z op(B) is put at
    the stored pairs of row z of W only (a ragged n over three 128-tiles, a
    column map with 60 % of the pairs screened out and the stored ones in a
    shuffled column order, a shared B, padded W rows): added to W's value,
    or over init_scale[z] init_vals[c] with W's old value not read (W
    starts as NaN there); every other entry of W keeps its value."""
    import torch
    from electrolites_amd import _dev, _lib
    rng = np.random.default_rng(20 + transb + 2 * init)
    nb, n, K = 4, 261, 37
    a = rng.standard_normal((nb, n, K + 1))
    b = rng.standard_normal((n, K + 2) if transb else (K, n + 2))
    npk = n * (n + 1) // 2
    keep = rng.random(npk) < .4
    ncol = int(keep.sum())
    col = np.full(npk, -1, dtype=np.int32)
    col[keep] = rng.permutation(ncol)
    w = rng.standard_normal((nb, ncol + 5))
    if init:
        w[:, :ncol] = np.nan
    sc, vals = rng.standard_normal(nb), rng.standard_normal(ncol)
    A, B, W, S, V = (torch.from_numpy(x.copy()).cuda() for x in (a, b, w, sc, vals))
    C = torch.from_numpy(col).cuda()
    _lib.call('mdf_gemm_pairs_batched', transb, n, K, ctypes.c_double(-.6), _dev.ptr(A), K + 1, n * (K + 1),
              _dev.ptr(B), b.shape[1], 0, _dev.ptr(C), _dev.ptr(W), ncol + 5, _dev.ptr(S) if init else None,
              _dev.ptr(V) if init else None, nb, None)
    bb = b[:, :K].T if transb else b[:K, :n]
    i, j = np.tril_indices(n)
    pk = i * (i + 1) // 2 + j
    sel = col[pk] >= 0
    ref = w.copy()
    if init:
        ref[:, :ncol] = sc[:, None] * vals[None, :]
    for z in range(nb):
        p = a[z, :, :K] @ bb
        ref[z, col[pk[sel]]] += -.6 * p[i[sel], j[sel]]
    assert rel(W.cpu().numpy(), ref) < 1e-13
    with pytest.raises(RuntimeError, match='go together'):
        _lib.call('mdf_gemm_pairs_batched', transb, n, K, ctypes.c_double(-.6), _dev.ptr(A), K + 1, n * (K + 1),
                  _dev.ptr(B), b.shape[1], 0, _dev.ptr(C), _dev.ptr(W), ncol + 5, _dev.ptr(S), None, nb, None)


@pytest.mark.gpu
@pytest.mark.parametrize('n,batch', [(1, 1), (31, 3), (32, 2), (33, 5), (320, 4), (5, 70000)])
def test_sym_pack_unpack(n, batch):
    """mdf_sym_pack / mdf_sym_unpack against numpy's upper-triangle order
    (row-major over a <= b) with diagonal and off-diagonal scales, at sizes
    on, below and above the 32-tiles and a batch beyond the grid's y limit;
    padded batch strides leave the padding untouched."""
    import torch
    from electrolites_amd import _dev, _lib
    rng = np.random.default_rng(n + batch)
    if batch > 1000:
        batch = 66000      # beyond 65535 blocks in y: the kernel strides the batch
    npk = n * (n + 1) // 2
    g = rng.standard_normal((batch, n * n + 3))
    p = rng.standard_normal((batch, npk + 2))
    G, P = torch.from_numpy(g.copy()).cuda(), torch.from_numpy(p.copy()).cuda()
    _lib.call('mdf_sym_pack', n, _dev.ptr(G), n * n + 3, _dev.ptr(P), npk + 2, ctypes.c_double(.5),
              ctypes.c_double(2 ** .5), batch, None)
    gm = g[:, :n * n].reshape(batch, n, n)
    a, b2 = np.triu_indices(n)
    s = np.where(a == b2, .5, 2 ** .5)
    ref = p.copy()
    ref[:, :npk] = s * (gm[:, a, b2] + gm[:, b2, a])
    assert np.array_equal(P.cpu().numpy(), ref)
    G2 = torch.from_numpy(g.copy()).cuda()
    _lib.call('mdf_sym_unpack', n, _dev.ptr(P), npk + 2, _dev.ptr(G2), n * n + 3, ctypes.c_double(2.),
              ctypes.c_double(2 ** -.5), batch, None)
    full = np.empty((batch, n, n))
    u = np.where(a == b2, 2., 2 ** -.5) * ref[:, :npk]
    full[:, a, b2] = u
    full[:, b2, a] = u
    out = g.copy()
    out[:, :n * n] = full.reshape(batch, -1)
    assert np.array_equal(G2.cpu().numpy(), out)
    with pytest.raises(RuntimeError, match='batch stride'):
        _lib.call('mdf_sym_pack', n, _dev.ptr(G), n * n - 1, _dev.ptr(P), npk + 2, ctypes.c_double(1.),
                  ctypes.c_double(1.), 2, None)


@pytest.mark.gpu
@pytest.mark.parametrize('n', [1, 63, 64, 150, 257])
def test_sp2_step(n):
    """mdf_sp2_begin / mdf_sp2_step against numpy: from a symmetric X with
    its spectrum in [0, 1], each step takes X^2 or 2X - X^2 by the trace rule
    (|tr X^2 - nocc| >= |2 tr X - tr X^2 - nocc| -> 2X - X^2), Y is
    symmetric to the bit (padded row strides; the padding untouched), the
    per-tile sums add up to tr(Y) and ||Y||_F^2, and 60 steps from a
    Gershgorin-scaled Fock-like matrix give the spectral projector onto its
    nocc lowest eigenvectors (1e-10)."""
    import torch
    from electrolites_amd import _dev, _lib
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    a = a + a.T + np.diag(np.arange(n) * 3.)
    nocc = max(1, n // 3)
    d = np.diag(a)
    off = np.abs(a).sum(1) - np.abs(d)
    emin, emax = (d - off).min(), (d + off).max()
    # (one function: the bounds meet, and any value in (0, 1) is the input)
    x = (emax * np.eye(n) - a) / (emax - emin) if n > 1 else np.full((1, 1), .6)
    nt = _lib.query('mdf_sp2_tiles', n)
    ld = n + 3
    X = torch.full((2, n, ld), 7., dtype=torch.float64).cuda()
    X[0, :, :n] = torch.from_numpy(x)
    part = torch.zeros((2, nt, 2), dtype=torch.float64).cuda()
    _lib.call('mdf_sp2_begin', n, _dev.ptr(X[0]), ld, _dev.ptr(part[0]), None)
    s = part[0].sum(0).cpu().numpy()
    assert abs(s[0] - np.trace(x)) < 1e-10 * n and abs(s[1] - (x * x).sum()) < 1e-10 * n
    ref, cur = x.copy(), 0
    for step in range(60):
        _lib.call('mdf_sp2_step', n, ctypes.c_double(nocc), _dev.ptr(X[cur]), ld, _dev.ptr(part[cur]),
                  _dev.ptr(X[1 - cur]), ld, _dev.ptr(part[1 - cur]), None)
        cur = 1 - cur
        x2 = ref @ ref
        up = abs(np.trace(x2) - nocc) >= abs(2 * np.trace(ref) - np.trace(x2) - nocc)
        ref = 2 * ref - x2 if up else x2
        if step < 3:
            got = X[cur].cpu().numpy()
            assert np.array_equal(got[:, :n], got[:, :n].T)
            assert np.all(got[:, n:] == 7.)
            assert rel(got[:, :n], ref) < 1e-12
            s = part[cur].sum(0).cpu().numpy()
            assert abs(s[0] - np.trace(ref)) < 1e-10 * n and abs(s[1] - (ref * ref).sum()) < 1e-10 * n
    w, v = np.linalg.eigh(a)
    proj = v[:, :nocc] @ v[:, :nocc].T
    assert np.abs(X[cur].cpu().numpy()[:, :n] - proj).max() < 1e-10
    with pytest.raises(RuntimeError, match='alias'):
        _lib.call('mdf_sp2_step', n, ctypes.c_double(nocc), _dev.ptr(X[0]), ld, _dev.ptr(part[0]),
                  _dev.ptr(X[0]), ld, _dev.ptr(part[1]), None)
