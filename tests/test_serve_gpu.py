"""Continuous-batching serving: each request's energy is the same alone and
in the mix (requests arriving while others are mid-solve), the resident bytes
never exceed the admission budget, and the mixed-precision schedule lands
within 1e-8 Eh of the all-FP64 solve."""
import numpy as np
import pytest


def _mols():
    from electrolites_amd.gto import water_cluster
    return [water_cluster(2, seed=11, side=5.0), water_cluster(3, seed=12, side=5.5, li=True),
            water_cluster(1, seed=13, side=4.0), water_cluster(4, seed=14, side=6.0),
            water_cluster(2, seed=15, side=5.0, li=True)]


def test_footprint_and_streams():
    from electrolites_amd import serve
    mols = serve.molecule_mix(6, seed=3, lo=3, hi=40)
    assert [m.natm for m in mols] == [m.natm for m in serve.molecule_mix(6, seed=3, lo=3, hi=40)]
    for i, m in enumerate(mols):
        nw = (m.natm - (i % 2)) // 3
        assert 3 <= nw <= 40 and m.charge == i % 2
        nao, naux = m.nao, m.naux
        # the stored (screened-in) pairs, not the full triangle
        from electrolites_amd.fastdfcderi import kept_pairs
        kept = int(kept_pairs(m)[0].sum())
        assert kept <= nao * (nao + 1) // 2
        assert 8 * naux * kept < serve.footprint(m) < 8 * naux * kept + 8 * 40 * nao * nao
    t = serve.poisson_arrivals(2000, rate=4., seed=1)
    assert np.all(np.diff(t) >= 0) and abs(t[-1] / 2000 - .25) < .03
    s = serve.Server(hbm_budget=1e6)
    with pytest.raises(ValueError, match='does not fit'):
        s.submit(serve.Request(mols[0]))
    from electrolites_amd.gto import water_cluster
    with pytest.raises(ValueError, match='at least 2 waters'):
        water_cluster(1, li=True)


@pytest.mark.gpu
@pytest.mark.parametrize('schedule,purify_min', [('async', None), ('lockstep', None), ('async', 60)])
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_alone_equals_mix(precision, schedule, purify_min):
    from electrolites_amd import serve
    from electrolites_amd.scf import RHF
    mols = _mols()
    # in the mix: a new request joins every other J/K launch while earlier
    # ones run (built at admission, so it is active at once)
    srv = serve.Server(conv_tol=1e-10, schedule=schedule, async_build=False, purify_min=purify_min)
    reqs = [serve.Request(m, rid=i, k_precision=precision) for i, m in enumerate(mols)]
    pending = list(reqs)
    peak = 0
    while pending or srv.busy:
        # a new request every other tick, or at once when the server drained
        if pending and (srv.ticks % 2 == 0 or not srv.busy):
            srv.submit(pending.pop(0))
        srv.step()
        peak = max(peak, len(srv.active))
        assert srv.ticks < 2000
    srv.close()
    assert peak >= 3, 'requests did not overlap'
    for r in reqs:
        assert r.converged
        alone = serve.Server(conv_tol=1e-10, schedule=schedule, purify_min=purify_min)
        a = serve.Request(r.mol, k_precision=precision)
        alone.serve([a])
        # the stand-alone driver with the server's diagonalisation choice
        mf = RHF(r.mol, conv_tol=1e-10, k_precision=precision, host_diag_max=srv.host_diag_max,
                 low_precision=srv.low_precision, purify_min=srv.purify_min)
        e = mf.kernel()
        assert abs(r.e_tot - a.e_tot) <= 1e-12, (r.rid, r.e_tot, a.e_tot)
        assert abs(r.e_tot - e) <= 1e-12, (r.rid, r.e_tot, e)
        assert r.cycles == a.cycles == mf.cycles
        # the plain driver (device diagonalisation): the same SCF solution
        assert abs(r.e_tot - RHF(r.mol, conv_tol=1e-10, k_precision=precision).kernel()) <= 1e-9
        if precision == 'mixed':
            assert srv.low_precision in [h['k_precision'] for h in r.history]
            e64 = RHF(r.mol, conv_tol=1e-10).kernel()
            assert abs(r.e_tot - e64) <= 1e-8


@pytest.mark.gpu
def test_admission_budget():
    from electrolites_amd import serve
    mols = _mols()
    reqs = [serve.Request(m, rid=i, k_precision='mixed') for i, m in enumerate(mols)]
    big = sorted(r.bytes for r in reqs)
    budget = big[-1] + big[-2]                  # at most two of the largest resident
    srv = serve.Server(hbm_budget=budget, conv_tol=1e-9)
    stats = srv.serve(reqs)
    assert stats['n'] == len(reqs) and stats['converged'] == len(reqs)
    assert max(tr[3] for tr in srv.trace) <= budget
    assert srv.resident == 0
    for r in reqs:
        assert r.t_admit >= r.arrival and r.t_done > r.t_admit


@pytest.mark.gpu
def test_plans_never_outlive_their_requests():
    """Grouped plans are keyed by never-reused slot serials and dropped when
    a member retires: while requests retire and others join (the pattern
    that, keyed by object ids, once reran a retired request's plan on a new
    request), every cached plan holds active requests only, and every
    energy equals the request's energy alone."""
    from electrolites_amd import serve
    mols = _mols() * 2
    srv = serve.Server(conv_tol=1e-9, async_build=False)
    reqs = [serve.Request(m, rid=i) for i, m in enumerate(mols)]
    pending = list(reqs)
    while pending or srv.busy:
        if pending and len(srv.active) < 3:
            srv.submit(pending.pop(0))
        srv.step()
        live = {s.serial for s in srv.active}
        for key in srv._plans:
            assert set(key[1:]) <= live, (key, live)
        assert srv.ticks < 5000
    srv.close()
    alone = {}
    for r in reqs:
        if id(r.mol) not in alone:
            a = serve.Request(r.mol)
            serve.Server(conv_tol=1e-9).serve([a])
            alone[id(r.mol)] = a.e_tot
        assert abs(r.e_tot - alone[id(r.mol)]) <= 1e-12, (r.rid, r.e_tot, alone[id(r.mol)])
