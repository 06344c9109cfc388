"""The no-allocation J entry point (``mdf_get_j_noalloc``: partial sums in a
file-scope device array instead of a workspace) against ``mdf_get_j`` on the
same inputs, bit for bit (the same kernels and plan): two calls in a row, the
second with a smaller input, FP64 and FP32 storage, several densities; and a
size whose partials exceed the array rejected with a clear error."""
import ctypes

import numpy as np
import pytest


def _case(gen, naux, npairs, n_dm, dtype):
    import torch
    cd = torch.randn((naux, npairs + 3), generator=gen, dtype=torch.float64, device='cuda').to(dtype)
    d = torch.randn((n_dm, npairs), generator=gen, dtype=torch.float64, device='cuda')
    return cd, d


def _run(entry, cd, d, naux, npairs, code, acc_seed=None):
    import torch
    from electrolites_amd import _dev, _lib
    n_dm = d.shape[0]
    rho = torch.full((n_dm, naux + 1), 7., dtype=torch.float64, device='cuda')
    vj = torch.full((n_dm, npairs + 2), 3., dtype=torch.float64, device='cuda')
    args = (_dev.ptr(cd), code, naux, npairs, cd.stride(0), _dev.ptr(d), n_dm, npairs,
            _dev.ptr(rho), naux + 1, _dev.ptr(vj), npairs + 2, 1)
    if entry == 'mdf_get_j':
        work = torch.empty(_lib.query('mdf_get_j_workspace', naux, npairs, n_dm), dtype=torch.uint8,
                           device='cuda')
        _lib.call(entry, *args, _dev.ptr(work), work.numel(), _dev.stream_ptr())
    else:
        _lib.call(entry, *args, _dev.stream_ptr())
    torch.cuda.synchronize()
    return rho.cpu().numpy(), vj.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_equals_workspace_entry():
    import torch
    gen = torch.Generator(device='cuda')
    gen.manual_seed(7)
    # (H2O)64's shape first, then a smaller one: the array keeps the first
    # call's partials, which the second must not see
    for naux, npairs, n_dm, dt, code in ((6976, 378596, 4, torch.float64, 0), (300, 5003, 2, torch.float64, 0),
                                         (1000, 77777, 3, torch.float32, 1), (37, 129, 1, torch.float64, 0)):
        cd, d = _case(gen, naux, npairs, n_dm, dt)
        ra, va = _run('mdf_get_j', cd, d, naux, npairs, code)
        rb, vb = _run('mdf_get_j_noalloc', cd, d, naux, npairs, code)
        assert np.array_equal(ra, rb) and np.array_equal(va, vb), (naux, npairs, n_dm)
        # accumulate=1 adds vj onto the 3.0 fill; the padding is untouched
        assert np.all(vb[:, npairs:] == 3.) and np.all(rb[:, naux:] == 7.)
        # rho (written, not accumulated) against a library product
        ref = (d @ cd[:, :npairs].double().T).cpu().numpy()
        assert np.abs(rb[:, :naux] - ref).max() <= 1e-12 * np.abs(ref).max()
        del cd, d
        torch.cuda.empty_cache()


@pytest.mark.gpu
def test_noalloc_rejects_oversize():
    import torch
    from electrolites_amd import _dev, _lib
    # (naux 20000, 4e6 pairs, 4 densities) needs ~1.4 GiB of partials: refused
    # on the host before any launch, so tiny stand-in buffers are never read
    buf = torch.zeros(16, dtype=torch.float64, device='cuda')
    with pytest.raises(_lib.MDFError, match='scratch array'):
        _lib.call('mdf_get_j_noalloc', _dev.ptr(buf), 0, 20000, 4000000, 4000000, _dev.ptr(buf), 4, 4000000,
                  _dev.ptr(buf), 20000, _dev.ptr(buf), 4000000, 0, _dev.stream_ptr())
