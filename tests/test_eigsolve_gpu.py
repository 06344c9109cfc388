"""The device eigensolver service (rocSOLVER strided-batched syevd behind
one thread) against LAPACK, and its determinism: a matrix gives the same
bits alone and batched with others, from one thread or many at once (the
serving loop's pattern, which crashed hipSOLVER when its calls ran
concurrently)."""
import concurrent.futures

import numpy as np
import pytest


def _sym(rng, n, scale=1.):
    a = rng.standard_normal((n, n)) * scale
    return .5 * (a + a.T)


@pytest.mark.gpu
@pytest.mark.parametrize('n', [1, 5, 63, 64, 65, 130, 500])
def test_eigh_matches_lapack(n):
    import torch
    from electrolites_amd import eigsolve
    rng = np.random.default_rng(n)
    a = _sym(rng, n)
    w, v = eigsolve.eigh(torch.from_numpy(a).cuda())
    w, v = w.cpu().numpy(), v.cpu().numpy()
    ref = np.linalg.eigvalsh(a)
    norm = np.abs(ref).max()
    assert np.abs(w - ref).max() <= 1e-13 * norm * max(1, n ** .5)
    assert np.abs(a @ v - v * w).max() <= 1e-12 * norm
    assert np.abs(v.T @ v - np.eye(n)).max() <= 1e-12


@pytest.mark.gpu
def test_eigh_alone_equals_batched_and_threaded():
    import torch
    from electrolites_amd import eigsolve
    rng = np.random.default_rng(3)
    mats = [torch.from_numpy(_sym(rng, n)).cuda() for n in (200, 220, 250, 256, 700, 650, 90)]
    alone = [eigsolve.eigh(m) for m in mats]
    b0 = eigsolve.STATS['batches']

    def run(m):
        return eigsolve.eigh(m)

    with concurrent.futures.ThreadPoolExecutor(16) as ex:
        for _ in range(3):
            got = list(ex.map(run, mats * 3))
            for (w, v), (wa, va) in zip(got, alone * 3):
                assert torch.equal(w, wa) and torch.equal(v, va)
    # the threads' requests shared batches
    assert eigsolve.STATS['batches'] - b0 < 3 * len(mats) * 3


@pytest.mark.gpu
def test_eigh_rejects_bad_input():
    import torch
    from electrolites_amd import eigsolve
    with pytest.raises(ValueError):
        eigsolve.eigh(torch.zeros((3, 4), dtype=torch.float64, device='cuda'))
    with pytest.raises(ValueError):
        eigsolve.eigh(torch.zeros((3, 3), dtype=torch.float32, device='cuda'))
    w, v = eigsolve.eigh(torch.zeros((0, 0), dtype=torch.float64, device='cuda'))
    assert w.numel() == 0 and v.numel() == 0
