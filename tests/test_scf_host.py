"""Host-side pieces of the SCF driver that need no GPU."""
import numpy as np
import pytest


@pytest.mark.parametrize('n,rank', [(1, 1), (40, 7), (90, 45), (300, 61)])
def test_projector_factor(n, rank):
    """projector_factor: for a random symmetric projector of the given rank,
    W W^T reproduces it, W^T W = I (W is an orthonormal basis of its range)
    and P W = W; a wrong rank and a matrix of another rank are refused."""
    from electrolites_amd.scf import projector_factor
    rng = np.random.default_rng(n + rank)
    q, _ = np.linalg.qr(rng.standard_normal((n, rank)))
    P = q @ q.T
    W = projector_factor(P.copy(), rank)
    assert W.shape == (n, rank)
    assert np.abs(W @ W.T - P).max() < 1e-13
    assert np.abs(W.T @ W - np.eye(rank)).max() < 1e-13
    assert np.abs(P @ W - W).max() < 1e-13
    assert projector_factor(P.copy(), rank + 1) is None
    if rank > 1:
        assert projector_factor(P.copy(), rank - 1) is None


def test_projector_factor_concurrent_needs_single_thread_blas():
    """The concurrent form refuses to run outside single_thread_blas (whose
    one-thread LAPACK keeps the result independent of the caller), and runs
    inside it."""
    from electrolites_amd.scf import projector_factor, single_thread_blas
    P = np.eye(3)
    P[2, 2] = 0.
    with pytest.raises(RuntimeError, match='single_thread_blas'):
        projector_factor(P.copy(), 2, concurrent=True)
    with single_thread_blas():
        W = projector_factor(P.copy(), 2, concurrent=True)
    assert np.allclose(W @ W.T, P)
