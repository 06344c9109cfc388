"""Benchmark of the density-fitted J/K path on one MI355X.

Workload (one "step" = one FP64 ``get_jk`` call, J and K, over a resident
cderi): a 64-water cluster at 1.0 g/cm^3 (seed 0), def2-SVP orbital basis
(nao 1536) with the 109-function-per-water fitting set (naux 6976), the
screened-in pairs of the packed triangle (378,596 of 1,180,416; 21.1 GB in
FP64), and a mo-tagged random PSD density of rank nocc = 320.  The cderi is built on the GPU first (Rys
int2c2e/int3c2e, Cholesky, L^-1, triangular multiply) and its build is timed
per stage; ``--synthetic-cderi`` replaces it with seeded N(0,1) values over
the full triangle (65.9 GB) for quick kernel timing.

``df_gradient_2e`` times the DF gradient's two-electron term (j_factor 1,
k_factor 0.5 on the same density) from the same cderi, per stage (metric factor, rhok,
weights, derivative integrals).

Prints ONE JSON line.  ``roofline`` is the K build (the dominant kernel
group) against FP64 MFMA peak; ``roofline_j`` is the J build against HBM
bandwidth.  Per-call durations come from HIP events recorded on the stream
the library launches on (torch's current stream).

``serving`` (skipped with ``--no-serve``) is the north-star workload at one
rate: 48 DF-RHF energy requests over the molecule mix ((H2O)3-40, Li+ on
half), Poisson arrivals at 8 req/s, served by the continuous-batching loop in
mixed precision (split-BF16 K until the DIIS error is below 1e-5, FP64
finish), beside the same stream solved one request at a time by the plain
FP64 driver: solves/hr, p50/p99 latency, max |dE| against FP64.

``cpu_baseline`` times the CPU restatement of the same algorithm (numpy/
OpenBLAS FP64: rho = L d, J = L^T rho; per aux block unpack + Lij,jk->Lik +
rank-k update), min of 3 after a correctness gate against the GPU result,
with the CPU model, BLAS and thread count recorded: J and K over a sample
of the (H2O)64 tensor's own rows scaled to the full naux (``value``), the
whole (H2O)32 get_jk, and the whole (H2O)4 DF-RHF solve.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK = 8.0e12          # B/s
FP64_MFMA_PEAK = 78.6e12   # flop/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--nwater', type=int, default=64)
    ap.add_argument('--nocc', type=int, default=320)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--synthetic-cderi', action='store_true')
    ap.add_argument('--cpu-sample-j', type=int, default=768)
    ap.add_argument('--cpu-sample-k', type=int, default=192)
    ap.add_argument('--no-serve', action='store_true', help='getjk: skip the serving leg')
    ap.add_argument('--serve-requests', type=int, default=48, help='getjk: serving-leg requests')
    ap.add_argument('--serve-rate', type=float, default=8., help='getjk: serving-leg Poisson rate (req/s)')
    ap.add_argument('--workload', choices=('getjk', 'batch', 'serve', 'grad'), default='getjk',
                    help='getjk: one (H2O)64 get_jk (default); batch: one grouped J/K launch '
                         'over a molecule mix; serve: continuous-batching DF-SCF serving; '
                         'grad: the DF gradient two-electron term from the built cderi')
    ap.add_argument('--nmol', type=int, default=48, help='batch: molecules')
    ap.add_argument('--requests', type=int, default=24, help='serve: requests per rate')
    ap.add_argument('--rates', default='0.5,2,8', help='serve: Poisson rates (req/s)')
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--no-host', action='store_true', help='getjk: skip the host-resident (streamed) line')
    ap.add_argument('--no-grad', action='store_true',
                    help='getjk: skip the DF gradient two-electron term line')
    ap.add_argument('--host-frac', type=float, default=.25,
                    help='getjk: share of the aux rows moved to page-locked host memory for the streamed line')
    ap.add_argument('--graphs', action='store_true', help='serve: replay grouped runs as HIP graphs')
    ap.add_argument('--copy-probe', type=int, default=0,
                    help='getjk: first run a plain device copy of this many MiB (the reference '
                         'kernel that calibrates HBM byte counters)')
    return ap.parse_args()


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def measured_traffic(nwater):
    """HBM bytes per call of the K and J builds from the committed counter
    summary (two rocprofv3 --pmc passes, FETCH_SIZE then WRITE_SIZE, scaled by
    a plain copy of known size); None when no summary covers this workload."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                        f'hbm_traffic_h2o{nwater}.md')
    out = {}
    try:
        with open(path) as f:
            for line in f:
                if line.startswith('bench-traffic-'):
                    key, val = line.split(':', 1)
                    out[key.strip()[len('bench-traffic-'):]] = float(val)
    except OSError:
        pass
    return out


def _cpu_info():
    """CPU model and the BLAS numpy calls (threadpoolctl), for the record."""
    model = None
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    blas = None
    try:
        from threadpoolctl import threadpool_info
        for lib in threadpool_info():
            if lib.get('user_api') == 'blas':
                blas = {'library': lib.get('internal_api'), 'version': lib.get('version'),
                        'threads': lib.get('num_threads')}
                break
    except Exception:                                         # noqa: BLE001
        pass
    return model, blas


def _min_of(n, fn, gate=None):
    """Best wall time of ``n`` runs of ``fn`` (the CPU protocol's min of 3)
    and the first result; ``gate`` checks that result before the other runs
    (it raises on a wrong answer, so nothing wrong is ever reported)."""
    best, first = None, None
    for i in range(n):
        t0 = time.perf_counter()
        out = fn()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
        if i == 0:
            first = out
            if gate is not None:
                gate(out)
    return best, first


def _cpu_jk(O, rows, pa, nao, dm, f, blk=128):
    """The CPU restatement of PySCF's DF J/K on host rows (numpy/OpenBLAS
    FP64): J = L^T (L d); K per block of aux rows: unpack, ``Lij,jk->Lik``,
    then the rank-k update ``K += R R^T``."""
    d = O.sparse_dm(dm, pa, 1)[0]
    rho = rows @ d
    vjp = rows.T @ rho
    vk = np.zeros((nao, nao))
    for L0 in range(0, rows.shape[0], blk):
        a = O.unpack(rows[L0:L0 + blk], pa, nao)
        r = np.matmul(a, f)
        r2 = r.transpose(1, 0, 2).reshape(nao, -1)
        vk += r2 @ r2.T
    vj = np.zeros((nao, nao))
    rr, cc = np.divmod(pa, nao)
    vj[rr, cc] = vjp
    vj[cc, rr] = vjp
    return vj, vk


def cpu_baseline(sample, args):
    """The CPU restatement (the oracle's numpy/OpenBLAS FP64 J/K, PySCF's CPU
    DF algorithm) timed on the host cores, following the reference's CPU
    protocol (docs/CPU_BASELINE.md: min of 3, a correctness gate first,
    cores, CPU model and BLAS recorded):

    * config 2: J over ``--cpu-sample-j`` and K over ``--cpu-sample-k`` aux
      rows of the GPU-built (H2O)64 tensor at the full shape, gated against
      the GPU's J/K over the same rows, scaled by naux/sample;
    * config 3: the whole (H2O)32 get_jk (no scaling), gated against the GPU;
    * config 1: the oracle's DF-RHF solve of (H2O)4 (integrals, cderi, SCF),
      gated against the GPU energy, as CPU solves/hr."""
    from oracle import oracle as O
    nthreads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    model, blas = _cpu_info()
    out = {'unit': 'get_jk/hr', 'cores': nthreads, 'kind': 'port', 'cpu_model': model, 'blas': blas,
           'timing': 'min of 3 after a correctness gate'}
    nao, naux, nocc = sample['nao'], sample['naux'], sample['nocc']
    pa = sample['pa']
    c = sample['c']
    dm = 2 * c @ c.T
    f = c * np.sqrt(2.)
    # config 2, sampled rows
    jrows, krows = sample['jrows'], sample['krows']
    d = O.sparse_dm(dm, pa, 1)[0]
    gates = {}

    def gate_of(name, gj, gk, limit=1e-11):
        def check(res):
            cj, ck = res
            g = max(float(np.abs(cj - gj).max() / np.abs(gj).max()),
                    float(np.abs(ck - gk).max() / np.abs(gk).max()))
            gates[name] = g
            if not g < limit:
                raise RuntimeError(f'CPU baseline gate {name} failed: {g:.2e}')
        return check

    def jpass():
        rho = jrows @ d
        return jrows.T @ rho

    tj, _ = _min_of(3, jpass)
    tk, _ = _min_of(3, lambda: _cpu_jk(O, krows, pa, nao, dm, f),
                    gate_of('config2', sample['gpu_j'], sample['gpu_k']))
    gate = gates['config2']
    tj *= naux / jrows.shape[0]
    tk *= naux / krows.shape[0]
    per_call = tj + tk
    out.update({'value': 3600. / per_call,
                'sample': f'J over {jrows.shape[0]} and K over {krows.shape[0]} of the {naux} aux rows of the '
                          f'GPU-built (H2O)64 tensor at the full shape (nao {nao}, {len(pa)} stored pairs, '
                          f'nocc {nocc}), scaled by naux/sample',
                'j_ms': tj * 1e3, 'k_ms': tk * 1e3, 'gate_rel_err_vs_gpu': gate})
    log(f'cpu: config 2 J {tj:.2f} s K {tk:.1f} s per call (gate {gate:.1e})')
    # config 3, whole
    c3 = sample.get('config3')
    if c3 is not None:
        rows3, pa3, c3m = c3['rows'], c3['pa'], c3['c']
        dm3 = 2 * c3m @ c3m.T
        f3 = c3m * np.sqrt(2.)
        t3, _ = _min_of(3, lambda: _cpu_jk(O, rows3, pa3, c3['nao'], dm3, f3),
                        gate_of('config3', c3['gpu_j'], c3['gpu_k']))
        g3 = gates['config3']
        out['config3_get_jk'] = {'cpu_ms': t3 * 1e3, 'gpu_ms': c3['gpu_ms'], 'nao': c3['nao'],
                                 'naux': rows3.shape[0], 'npairs': len(pa3), 'nocc': c3m.shape[1],
                                 'gate_rel_err_vs_gpu': g3, 'gpu_over_cpu': t3 * 1e3 / c3['gpu_ms']}
        log(f'cpu: config 3 get_jk {t3 * 1e3:.0f} ms (GPU {c3["gpu_ms"]:.1f} ms; gate {g3:.1e})')
    # config 1, a whole DF-RHF solve
    c1 = sample.get('config1')
    if c1 is not None:
        mol = c1['mol']

        def solve():
            s = O.DFSystem(mol.ao.as_dict(), mol.aux.as_dict(), mol.atom_charges(), mol.atom_coords(),
                           mol.nelectron, mol.energy_nuc())
            return O.rhf(s, conv_tol=1e-10)

        def check(r):
            if not (r['converged'] and abs(r['e_tot'] - c1['gpu_e']) < 1e-9):
                raise RuntimeError(f'CPU baseline config 1 gate failed: {r["e_tot"]} vs {c1["gpu_e"]}')

        t1, r1 = _min_of(3, solve, check)
        de = abs(r1['e_tot'] - c1['gpu_e'])
        out['config1_rhf'] = {'cpu_s': t1, 'cpu_solves_per_hr': 3600. / t1, 'gpu_s': c1['gpu_s'],
                              'gpu_solves_per_hr': 3600. / c1['gpu_s'], 'abs_dE_vs_gpu': de,
                              'e_tot': r1['e_tot']}
        log(f'cpu: config 1 DF-RHF {t1:.2f} s (GPU {c1["gpu_s"]:.2f} s; |dE| {de:.1e})')
    return out


def serving_leg(args):
    """Continuous-batching serving at one Poisson rate over the 48-molecule
    mix (mixed precision: split-BF16 K until the DIIS error is below 1e-5,
    FP64 finish), then the same stream solved one request at a time with the
    plain FP64 driver; solves/hr, p50/p99 latency and max |dE| of each."""
    import torch
    from electrolites_amd import serve
    mols = serve.molecule_mix(args.serve_requests, seed=args.seed)
    warm = serve.molecule_mix(1, seed=99, lo=3, hi=3)
    serve.Server().serve([serve.Request(warm[0])])
    serve.serve_one_at_a_time([serve.Request(warm[0])])
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    arr = serve.poisson_arrivals(len(mols), args.serve_rate, seed=args.seed + 7)
    reqs = [serve.Request(m, arr[i], rid=i, k_precision='mixed') for i, m in enumerate(mols)]
    srv = serve.Server()
    s = srv.serve(reqs)
    s['max_active'] = max(tr[1] for tr in srv.trace) if srv.trace else 0
    s['ticks'] = srv.ticks
    log(f'serving: {s["solves_per_hr"]:.0f} solves/hr p50 {s["p50_s"]:.2f} s p99 {s["p99_s"]:.2f} s')
    torch.cuda.empty_cache()
    base = [serve.Request(m, arr[i], rid=i, k_precision='fp64') for i, m in enumerate(mols)]
    b = serve.serve_one_at_a_time(base)
    de = max(abs(r.e_tot - q.e_tot) for r, q in zip(sorted(reqs, key=lambda r: r.rid), base))
    log(f'plain FP64 one at a time: {b["solves_per_hr"]:.0f} solves/hr p50 {b["p50_s"]:.2f} s '
        f'p99 {b["p99_s"]:.2f} s; max |dE| {de:.1e}')
    # the same requests all at once: the server's capacity (the plain path
    # never waits for an arrival in either stream, its span is its work)
    torch.cuda.empty_cache()
    burst = [serve.Request(m, 0., rid=i, k_precision='mixed') for i, m in enumerate(mols)]
    sb = serve.Server().serve(burst)
    deb = max(abs(r.e_tot - q.e_tot) for r, q in zip(sorted(burst, key=lambda r: r.rid), base))
    log(f'serving, all {len(mols)} at once: {sb["solves_per_hr"]:.0f} solves/hr p50 {sb["p50_s"]:.2f} s '
        f'p99 {sb["p99_s"]:.2f} s; max |dE| {deb:.1e}')
    return {'workload': f'{len(mols)} requests over the molecule mix ((H2O)3-40, Li+ on half), '
                        f'Poisson {args.serve_rate} req/s',
            'solves_per_hr': s['solves_per_hr'], 'p50_s': s['p50_s'], 'p99_s': s['p99_s'],
            'plain_fp64_solves_per_hr': b['solves_per_hr'], 'plain_fp64_p50_s': b['p50_s'],
            'plain_fp64_p99_s': b['p99_s'], 'speedup_vs_plain_fp64': s['solves_per_hr'] / b['solves_per_hr'],
            'max_abs_dE_vs_fp64': de, 'converged': s['converged'], 'max_active': s['max_active'],
            'ticks': s['ticks'], 'cycles': s['cycles'], 'plain_fp64_cycles': b['cycles'],
            'burst': {'solves_per_hr': sb['solves_per_hr'], 'p50_s': sb['p50_s'], 'p99_s': sb['p99_s'],
                      'speedup_vs_plain_fp64': sb['solves_per_hr'] / b['solves_per_hr'],
                      'max_abs_dE_vs_fp64': deb}}


class _ResidentDF:
    """The two fields the DF path reads, around a cderi already in HBM."""

    def __init__(self, cderi, idx=None):
        import torch
        self._cderi = {torch.cuda.current_device(): cderi}
        self._cderi_idx = idx

    def build(self):
        pass


def bench_grad(args):
    """The DF gradient's two-electron term (``fastdfgrad.jk_energy_per_atom``,
    j_factor 1, k_factor 0.5 on the getjk workload's density) from the (H2O)64 cderi built on
    the GPU: ``--warmup`` untimed calls, then ``--steps`` timed calls."""
    import torch
    from electrolites_amd import _dev, fastdf, fastdfgrad
    from electrolites_amd.df import DF
    from electrolites_amd.gto import water_cluster
    _dev.require_gpu()
    mol = water_cluster(args.nwater, seed=0)
    nao, nocc = mol.nao, args.nocc
    log(f'(H2O){args.nwater}: nao {nao} naux {mol.naux} natm {mol.natm} nocc {nocc}; building the cderi')
    df = DF(mol).build()
    rng = np.random.default_rng(0)
    c = np.linalg.qr(rng.standard_normal((nao, nocc)))[0]
    dm = torch.from_numpy(2 * c @ c.T).cuda()
    fastdf.tag_array(dm, mo_coeff=torch.from_numpy(c).cuda(),
                     mo_occ=torch.full((nocc,), 2., dtype=torch.float64, device='cuda'))
    for _ in range(args.warmup):
        fastdfgrad.jk_energy_per_atom(df, dm, 1., .5)
    torch.cuda.synchronize()
    stages = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g = fastdfgrad.jk_energy_per_atom(df, dm, 1., .5, timings=stages)
        log(f'gradient call done: sum over atoms {np.abs(g.sum(axis=0)).max():.1e}')
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {
        'metric': 'DF gradient two-electron term calls/hr, (H2O)64 def2-SVP FP64',
        'value': 3600. / (wall / args.steps), 'unit': 'gradient_2e/hr',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': wall / args.steps * 1e3, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic geometry: seeded (H2O)64 cluster, cderi built on the GPU; random rank-320 PSD density',
        'config': {'workload': f'(H2O){args.nwater} DF gradient 2e term FP64', 'nao': nao, 'naux': mol.naux,
                   'natm': mol.natm, 'nocc': nocc},
        'stages_ms_per_call': {k: v / args.steps for k, v in stages.items()},
        'translation_residual': float(np.abs(g.sum(axis=0)).max()),
    }
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """One grouped J/K launch over a molecule mix: ``--nmol`` molecules with
    a log-uniform number of waters in [3, 40], every second one with a Li+ at
    the centroid; each carries 1-8 densities: UHF alpha/beta (mo-tagged,
    random orthonormal occupied orbitals) and random symmetric response
    densities (N(0,1) entries scaled by 1e-3, full rank: two-factor K).
    Timed: the grouped run (autotuned step-1 width) against the same items
    through the single-molecule entry one after another."""
    import torch
    from electrolites_amd import _dev, batch, fastdf, fastdfcderi, serve
    _dev.require_gpu()
    mols = serve.molecule_mix(args.nmol, seed=args.seed)
    rng = np.random.default_rng(args.seed + 1)
    sets = []
    t0 = time.perf_counter()
    for mol in mols:
        cderi, idx = fastdfcderi.cholesky_eri(mol)
        df = _ResidentDF(cderi, idx)
        nao, nocc = mol.nao, mol.nocc
        n_dm = int(rng.integers(1, 9))
        nuhf = min(2, n_dm)
        cs = np.array([np.linalg.qr(rng.standard_normal((nao, nocc)))[0] for _ in range(nuhf)])
        uhf = torch.from_numpy(np.einsum('nik,njk->nij', cs, cs)).cuda()
        fastdf.tag_array(uhf, mo_coeff=torch.from_numpy(cs).cuda(),
                         mo_occ=torch.ones((nuhf, nocc), dtype=torch.float64, device='cuda'))
        sets.append((df, uhf))
        if n_dm > 2:
            a = rng.standard_normal((n_dm - 2, nao, nao)) * 1e-3
            sets.append((df, torch.from_numpy(.5 * (a + a.transpose(0, 2, 1))).cuda()))
    torch.cuda.synchronize()
    log(f'{len(mols)} cderi built in {time.perf_counter() - t0:.1f} s; '
        f'{len(sets)} items, {sum(d.shape[0] for _, d in sets)} densities')

    entries = []
    for df, dm in sets:
        cd = fastdf._cderi_of(df)
        nao = dm.shape[-1]
        lay = fastdf._layout(df, cd, nao)
        fl, fr = fastdf.factorize_dm(dm, 1)
        d = fastdf._sparse_dm(dm, lay, 1)
        e = batch.Entry(cd, nao, dm.shape[0], fl.shape[-1], dm_pairs=d, fl=fl, fr=fr, pair_map=lay.map)
        e.layout = lay
        entries.append(e)
    t0 = time.perf_counter()
    g = batch.GroupedJK(entries, tune=True)
    torch.cuda.synchronize()
    log(f'plan + autotune {time.perf_counter() - t0:.1f} s: bn {g.bn}, tiles {g.tiles()}')
    for _ in range(args.warmup):
        g.run()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(args.steps):
        g.run()
    ev[1].record()
    torch.cuda.synchronize()
    t_grouped = ev[0].elapsed_time(ev[1]) * 1e-3 / args.steps

    # the same items one molecule at a time through the single entry
    ev2 = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev2[0].record()
    singles = [fastdf.get_jk(df, dm, hermi=1) for df, dm in sets]
    ev2[1].record()
    torch.cuda.synchronize()
    t_single = ev2[0].elapsed_time(ev2[1]) * 1e-3
    same = all(torch.equal(e.vk.reshape(vk.shape), vk) for e, (vj, vk) in zip(entries, singles))
    vj_same = True
    for e, (vj, vk) in zip(entries, singles):
        full = torch.empty((e.n_dm, e.nao, e.nao), dtype=torch.float64, device='cuda')
        from electrolites_amd import _lib
        _lib.call('mdf_unpack_pairs', _dev.ptr(e.vj_pairs), e.n_dm, e.npairs, _dev.ptr(e.layout.addr),
                  e.npairs, e.nao, _dev.ptr(full), e.nao * e.nao, _dev.stream_ptr())
        vj_same &= bool(torch.equal(full.reshape(vj.shape), vj))
    torch.cuda.synchronize()
    fk = 0.
    for e in entries:
        two = 2 if e.fr is not None else 1
        fk += e.n_dm * (2. * two * e.naux * e.nao * e.nao * e.nocc +
                        (1. if two == 2 else .5) * 2 * e.naux * e.nocc * e.nao * e.nao)
    ndens = sum(e.n_dm for e in entries)
    out = {
        'metric': 'grouped J/K densities per second (48-molecule mix, FP64)',
        'value': ndens / t_grouped, 'unit': 'densities/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': t_grouped * 1e3,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic geometries and densities; cderi built on the GPU from the integrals',
        'config': {'workload': f'{len(mols)} molecules, (H2O)3-40 (+Li+ on half), 1-8 densities each',
                   'items': len(entries), 'densities': ndens, 'bn': g.bn, 'tiles': g.tiles(),
                   'nao': [m.nao for m in mols]},
        'single_entry_s': t_single, 'grouped_s': t_grouped, 'speedup_vs_single': t_single / t_grouped,
        'bit_identical_to_single': {'k': bool(same), 'j': bool(vj_same)},
        'k_tflops': fk / t_grouped / 1e12, 'k_frac_fp64_peak': fk / t_grouped / FP64_MFMA_PEAK,
    }
    g.close()
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """Continuous-batching DF-SCF serving over the molecule mix (as
    ``bench_batch``): for each Poisson rate, ``--requests`` energy requests
    served with mixed precision (split-BF16 K until the DIIS error is below
    the switch threshold, FP64 finish), then the same stream solved one
    request at a time with the plain FP64 driver.  Reports solves/hr and
    p50/p99 latency per rate."""
    import torch
    from electrolites_amd import _dev, serve
    _dev.require_gpu()
    mols = serve.molecule_mix(args.requests, seed=args.seed)
    warm = serve.molecule_mix(1, seed=99, lo=3, hi=3)
    serve.Server().serve([serve.Request(warm[0])])
    serve.serve_one_at_a_time([serve.Request(warm[0])])
    torch.cuda.synchronize()
    rates = [float(x) for x in args.rates.split(',')]
    per_rate = {}
    for rate in rates:
        torch.cuda.empty_cache()
        arr = serve.poisson_arrivals(len(mols), rate, seed=args.seed + 7)
        reqs = [serve.Request(m, arr[i], rid=i, k_precision='mixed') for i, m in enumerate(mols)]
        srv = serve.Server(graphs=args.graphs)
        s = srv.serve(reqs)
        s['max_active'] = max(tr[1] for tr in srv.trace) if srv.trace else 0
        s['ticks'] = srv.ticks
        torch.cuda.empty_cache()
        base = [serve.Request(m, arr[i], rid=i) for i, m in enumerate(mols)]
        b = serve.serve_one_at_a_time(base)
        de = max(abs(r.e_tot - q.e_tot) for r, q in zip(sorted(reqs, key=lambda r: r.rid), base))
        per_rate[str(rate)] = {'serving': s, 'one_at_a_time_fp64': b, 'max_abs_dE_vs_fp64': de,
                               'speedup': s['solves_per_hr'] / b['solves_per_hr']}
        log(f'rate {rate}: serving {s["solves_per_hr"]:.0f}/hr p50 {s["p50_s"]:.2f} p99 {s["p99_s"]:.2f}; '
            f'fp64 one-at-a-time {b["solves_per_hr"]:.0f}/hr p50 {b["p50_s"]:.2f} p99 {b["p99_s"]:.2f}; '
            f'max |dE| {de:.2e}')
    top = per_rate[str(max(rates))]
    out = {
        'metric': 'DF-SCF solves/hr, continuous batching, mixed precision (molecule mix)',
        'value': top['serving']['solves_per_hr'], 'unit': 'solves/hr',
        'n_gpus': 1, 'steps': len(mols), 'warmup': 1,
        'ms_per_step': 3.6e6 / top['serving']['solves_per_hr'],
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64+bf16x3',
        'data': 'synthetic geometries (seeded water clusters, Li+ on half); Poisson arrivals',
        'config': {'workload': f'{len(mols)} requests, (H2O)3-40 (+Li+ on half), rates {rates} req/s',
                   'nao': [m.nao for m in mols]},
        'speedup_vs_fp64_one_at_a_time': top['speedup'],
        'per_rate': per_rate,
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f'bench.py measures one GPU only (asked for --gpus {args.gpus}); '
              'the path shards along the aux index across GPUs but is built '
              'and measured on one', file=sys.stderr)
        sys.exit(2)
    if args.workload == 'batch':
        return bench_batch(args)
    if args.workload == 'serve':
        return bench_serve(args)
    if args.workload == 'grad':
        return bench_grad(args)
    import torch
    from electrolites_amd import _dev, _lib, fastdf
    from electrolites_amd.gto import water_cluster
    _dev.require_gpu()
    mol = water_cluster(args.nwater, seed=0)
    nao, naux = mol.nao, mol.naux
    npairs = nao * (nao + 1) // 2
    nocc = args.nocc
    log(f'(H2O){args.nwater}: nao {nao} naux {naux} npairs {npairs} nocc {nocc}')

    build = None
    idx = None                                    # pair addresses: None = the full triangle
    if args.synthetic_cderi:
        g = torch.Generator(device='cuda')
        g.manual_seed(0)
        cderi = torch.empty((naux, npairs), dtype=torch.float64, device='cuda')
        for r0 in range(0, naux, 512):
            cderi[r0:r0 + 512].normal_(generator=g)
    else:
        from electrolites_amd import fastdfcderi
        log('building the cderi on the GPU')
        timer = fastdfcderi._Timer()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cderi, idx = fastdfcderi.cholesky_eri(mol, timer=timer)
        npairs = cderi.shape[1]                   # the screened-in pairs the tensor stores
        torch.cuda.synchronize()
        wall_b = time.perf_counter() - t0
        st = timer.totals()
        fm = float(npairs) * naux * naux
        build = {'wall_s': wall_b, 'stages_ms': st,
                 'trmm_tflops': fm / (st['trmm'] * 1e-3) / 1e12 if st.get('trmm') else None,
                 'metric_flop': fm, 'stats': fastdfcderi.stats()}
        log(f'cderi built in {wall_b:.2f} s: {st}')
    rng = np.random.default_rng(0)
    c = np.linalg.qr(rng.standard_normal((nao, nocc)))[0]
    dm = torch.from_numpy(2 * c @ c.T).cuda()
    fastdf.tag_array(dm, mo_coeff=torch.from_numpy(c).cuda(),
                     mo_occ=torch.full((nocc,), 2., dtype=torch.float64, device='cuda'))

    class DF:
        _cderi = {torch.cuda.current_device(): cderi}
        _cderi_idx = idx

        def build(self):
            pass

    dfobj = DF()
    torch.cuda.synchronize()
    if args.copy_probe:
        src = torch.ones(args.copy_probe << 17, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f'copy probe: {args.copy_probe} MiB read + {args.copy_probe} MiB written')
        del src, dst
    log('cderi resident; warmup')

    def one(ev):
        ev[0].record()
        vj = fastdf.get_jk(dfobj, dm, hermi=1, with_k=False)[0]
        ev[1].record()
        fl, fr = fastdf.factorize_dm(dm, 1)
        lay = fastdf._layout(dfobj, cderi, nao)
        ev[2].record()
        vk = fastdf._vk(cderi, lay, fl, fr)
        ev[3].record()
        return vj, vk

    for _ in range(args.warmup):
        one([torch.cuda.Event(enable_timing=True) for _ in range(4)])
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)]
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        one(evs[k])
        log(f'step {k} issued')
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tj = np.mean([e[0].elapsed_time(e[1]) for e in evs]) * 1e-3
    tk = np.mean([e[2].elapsed_time(e[3]) for e in evs]) * 1e-3
    ms_step = wall / args.steps * 1e3

    # the same K in the mixed-precision modes, outside the timed steps
    fl, fr = fastdf.factorize_dm(dm, 1)
    lay = fastdf._layout(dfobj, cderi, nao)
    vk64 = fastdf._vk(cderi, lay, fl, fr, 'fp64')
    low = {}
    for prec in ('fp32', 'bf16x3'):
        fastdf._vk(cderi, lay, fl, fr, prec)
        e32 = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e32[0].record()
        vkl = fastdf._vk(cderi, lay, fl, fr, prec)
        e32[1].record()
        torch.cuda.synchronize()
        low[prec] = (e32[0].elapsed_time(e32[1]) * 1e-3,
                     float((vkl - vk64).abs().max() / vk64.abs().max()))
        del vkl
        log(f'K {prec}: {low[prec][0] * 1e3:.1f} ms, rel err {low[prec][1]:.2e}')
    del vk64
    tk32, k32_err = low['fp32']

    # FP32 storage (the FASTDF_FP32 mode): the same tensor rounded to FP32,
    # widened to FP64 in the kernels
    c32 = cderi.float()

    class DF32:
        _cderi = {torch.cuda.current_device(): c32}
        _cderi_idx = idx

        def build(self):
            pass

    df32 = DF32()
    vj64 = fastdf.get_jk(dfobj, dm, hermi=1, with_k=False)[0]
    fastdf.get_jk(df32, dm, hermi=1, with_k=False)
    e3 = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    e3[0].record()
    vj32 = fastdf.get_jk(df32, dm, hermi=1, with_k=False)[0]
    e3[1].record()
    vk32s = fastdf._vk(c32, fastdf._layout(df32, c32, nao), fl, fr, 'fp64')
    e3[2].record()
    torch.cuda.synchronize()
    fp32_store = {'j_ms': e3[0].elapsed_time(e3[1]), 'k_ms': e3[1].elapsed_time(e3[2]),
                  'cderi_GB': naux * npairs * 4 / 1e9,
                  'j_rel_err_vs_fp64_store': float((vj32 - vj64).abs().max() / vj64.abs().max())}
    bj32 = naux * npairs * 8 + 16 * (npairs + naux)
    fp32_store['j_hbm_frac'] = bj32 / (fp32_store['j_ms'] * 1e-3) / HBM_PEAK
    log(f'FP32 storage: J {fp32_store["j_ms"]:.2f} ms ({100 * fp32_store["j_hbm_frac"]:.1f} % HBM), '
        f'K {fp32_store["k_ms"]:.1f} ms')
    del c32, vj32, vk32s, df32

    # J without the cderi (three-centre integrals contracted on the fly)
    ri_j = None
    if not args.synthetic_cderi:
        from electrolites_amd import fastdfj3c

        class MolDF:
            pass

        mdf = MolDF()
        mdf.mol = mol
        fastdfj3c.get_j(mdf, dm, hermi=1)                  # tables, metric factor
        e4 = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e4[0].record()
        vjr = fastdfj3c.get_j(mdf, dm, hermi=1)
        e4[1].record()
        torch.cuda.synchronize()
        ri_j = {'ms': e4[0].elapsed_time(e4[1]), 'chunks': len(mdf._mdf_j3c.tasks),
                'rel_err_vs_cderi_j': float((vjr - vj64).abs().max() / vj64.abs().max())}
        log(f'RI-J without a cderi: {ri_j["ms"]:.1f} ms ({ri_j["chunks"]} chunks per sweep)')
        del vjr, mdf
    del vj64

    # out-of-HBM residency: a leading share of the aux rows moved to
    # page-locked host memory and streamed through get_jk with no device
    # prefix, against the same rows resident in HBM
    host_line = None
    if not args.no_host:
        from electrolites_amd import dfstore
        rows = max(1, int(naux * args.host_frac))
        dev_part = cderi[:rows]
        t0 = time.perf_counter()
        hpart = dfstore.to_host(dev_part)
        t_pin = time.perf_counter() - t0

        class HostDF:
            _cderi = {torch.cuda.current_device(): hpart}
            _cderi_idx = idx

            def build(self):
                pass

        class DevDF:
            _cderi = {torch.cuda.current_device(): dev_part}
            _cderi_idx = idx

            def build(self):
                pass

        hdf, ddf = HostDF(), DevDF()
        old = dfstore.RESIDENT_FRAC
        dfstore.RESIDENT_FRAC = 0.
        try:
            fastdf.get_jk(hdf, dm, hermi=1)
            torch.cuda.synchronize()
            b0 = dfstore.stats()['streamed_bytes']
            t0 = time.perf_counter()
            hj, hk = fastdf.get_jk(hdf, dm, hermi=1)
            torch.cuda.synchronize()
            t_host = time.perf_counter() - t0
            streamed = dfstore.stats()['streamed_bytes'] - b0
        finally:
            dfstore.RESIDENT_FRAC = old
        fastdf.get_jk(ddf, dm, hermi=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dj, dk = fastdf.get_jk(ddf, dm, hermi=1)
        torch.cuda.synchronize()
        t_dev = time.perf_counter() - t0
        host_line = {'rows': rows, 'GB': rows * npairs * 8 / 1e9, 'streamed_GB': streamed / 1e9,
                     'host_ms': t_host * 1e3, 'device_ms': t_dev * 1e3,
                     'stream_GBps': streamed / t_host / 1e9, 'copy_to_host_s': t_pin,
                     'block_MB': dfstore.STREAM_BLOCK_BYTES / 1e6,
                     'rel_diff_vs_device': max(float((hj - dj).abs().max() / dj.abs().max()),
                                               float((hk - dk).abs().max() / dk.abs().max()))}
        log(f'host-resident get_jk over {rows} aux rows ({host_line["GB"]:.1f} GB streamed): '
            f'{t_host * 1e3:.0f} ms = {host_line["stream_GBps"]:.1f} GB/s over PCIe; the same rows in HBM '
            f'{t_dev * 1e3:.1f} ms; rel diff {host_line["rel_diff_vs_device"]:.1e}')
        del hj, hk, dj, dk, hdf, ddf, hpart, dev_part

    # the DF gradient's two-electron term (j_factor 1, k_factor 0.5) from
    # the same cderi: one warm call, then one timed call with its stages
    grad_line = None
    if not args.no_grad and not args.synthetic_cderi:
        from electrolites_amd import fastdfgrad
        from electrolites_amd.df import DF as _DF
        gdf = _DF(mol)
        gdf._cderi = {torch.cuda.current_device(): cderi}
        gdf._cderi_idx = idx
        fastdfgrad.jk_energy_per_atom(gdf, dm, 1., .5)
        torch.cuda.synchronize()
        stages = {}
        t0 = time.perf_counter()
        g = fastdfgrad.jk_energy_per_atom(gdf, dm, 1., .5, timings=stages)
        t_grad = time.perf_counter() - t0
        grad_line = {'wall_ms': t_grad * 1e3, 'stages_ms': stages, 'natm': mol.natm,
                     'max_abs_grad': float(np.abs(g).max()),
                     'translation_residual': float(np.abs(g.sum(axis=0)).max())}
        log(f'DF gradient 2e term: {t_grad * 1e3:.0f} ms ({stages}); sum over atoms '
            f'{grad_line["translation_residual"]:.1e}')
        del gdf, g
        torch.cuda.empty_cache()

    S = naux * npairs * 8
    bj = 2 * S + 16 * (npairs + naux)
    traffic = measured_traffic(args.nwater) if args.nocc == 320 else {}
    fk = 2. * naux * nao * nao * nocc + 1. * naux * nocc * nao * (nao + 1)
    out = {
        'metric': 'get_jk (J+K) calls/hr, (H2O)64 def2-SVP FP64, single density',
        'value': 3600. / (wall / args.steps),
        'unit': 'get_jk/hr',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': ms_step,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64',
        'data': ('synthetic (seeded N(0,1) cderi of the (H2O)64 shape; random '
                 'rank-320 PSD density)') if args.synthetic_cderi else
                ('synthetic geometry: seeded (H2O)64 cluster, cderi built on the '
                 'GPU from its integrals; random rank-320 PSD density'),
        'config': {'workload': f'(H2O){args.nwater} get_jk FP64', 'nao': nao,
                   'naux': naux, 'npairs': npairs, 'nocc': nocc, 'n_dm': 1,
                   'cderi_GB': S / 1e9},
        'get_jk_ms': {'j': tj * 1e3, 'k': tk * 1e3, 'k_fp32': tk32 * 1e3,
                      'k_fp32_rel_err_vs_fp64': k32_err, 'k_bf16x3': low['bf16x3'][0] * 1e3,
                      'k_bf16x3_rel_err_vs_fp64': low['bf16x3'][1]},
        'cderi_build': build,
        'fp32_storage': fp32_store,
        'ri_j_no_cderi': ri_j,
        'host_resident': host_line,
        'df_gradient_2e': grad_line,
        'roofline': {'bound': 'mfma', 'achieved': fk / tk / 1e12,
                     'peak': FP64_MFMA_PEAK / 1e12, 'unit': 'TFLOP/s',
                     'frac': fk / tk / FP64_MFMA_PEAK, 'traffic': traffic.get('k'),
                     'kernel': 'mdf_get_k (k_step1 + k_step2 + k_finalize)',
                     'algorithmic': fk},
        'roofline_j': {'bound': 'hbm', 'achieved': bj / tj / 1e9,
                       'peak': HBM_PEAK / 1e9, 'unit': 'GB/s',
                       'frac': bj / tj / HBM_PEAK, 'traffic': traffic.get('j'),
                       'kernel': 'mdf_get_j (j_rho_partial + j_vj_pass)',
                       'algorithmic': bj},
    }
    sample = None
    if not args.no_cpu and not args.synthetic_cderi:
        # the CPU baseline's rows of this tensor and the GPU's J/K over them
        # (its correctness gate), before the tensor is freed
        mid = naux // 2
        jrows = cderi[mid:mid + args.cpu_sample_j].cpu().numpy()
        kblk = cderi[mid:mid + args.cpu_sample_k]
        gj, gk = fastdf.get_jk(_ResidentDF(kblk, idx), dm, hermi=1)
        sample = {'nao': nao, 'naux': naux, 'nocc': nocc, 'pa': np.asarray(idx[0]).astype(np.int64), 'c': c,
                  'jrows': jrows, 'krows': kblk.cpu().numpy(), 'gpu_j': gj.cpu().numpy(),
                  'gpu_k': gk.cpu().numpy()}
        del kblk, gj, gk
    del cderi
    DF._cderi = None
    torch.cuda.empty_cache()
    if not args.no_serve and not args.synthetic_cderi:
        log('serving leg')
        out['serving'] = serving_leg(args)
        torch.cuda.empty_cache()
    if sample is not None:
        log('cpu baseline: config 3 and config 1 on the GPU for the gates')
        sample['config3'] = _config3_sample()
        sample['config1'] = _config1_sample()
        torch.cuda.empty_cache()
        out['cpu_baseline'] = cpu_baseline(sample, args)
    print(json.dumps(out), flush=True)


def _config3_sample():
    """(H2O)32: its GPU-built cderi on the host, a rank-160 density and the
    GPU get_jk of it (timed), for the CPU baseline's whole-call leg."""
    import torch
    from electrolites_amd import fastdf, fastdfcderi
    from electrolites_amd.gto import water_cluster
    mol = water_cluster(32, seed=0)
    cderi, idx = fastdfcderi.cholesky_eri(mol)
    nocc = mol.nocc
    c = np.linalg.qr(np.random.default_rng(3).standard_normal((mol.nao, nocc)))[0]
    dm = torch.from_numpy(2 * c @ c.T).cuda()
    fastdf.tag_array(dm, mo_coeff=torch.from_numpy(c).cuda(),
                     mo_occ=torch.full((nocc,), 2., dtype=torch.float64, device='cuda'))
    df = _ResidentDF(cderi, idx)
    fastdf.get_jk(df, dm, hermi=1)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    vj, vk = fastdf.get_jk(df, dm, hermi=1)
    ev[1].record()
    torch.cuda.synchronize()
    out = {'nao': mol.nao, 'rows': cderi.cpu().numpy(), 'pa': np.asarray(idx[0]).astype(np.int64), 'c': c,
           'gpu_j': vj.cpu().numpy(), 'gpu_k': vk.cpu().numpy(), 'gpu_ms': ev[0].elapsed_time(ev[1])}
    del cderi, df, vj, vk
    return out


def _config1_sample():
    """(H2O)4 in a 6 A cube: the GPU DF-RHF energy and its wall time
    (cderi build, integrals and SCF; a warm solve first)."""
    import torch
    from electrolites_amd.gto import water_cluster
    from electrolites_amd.scf import RHF
    RHF(water_cluster(4, seed=0, side=6.0), conv_tol=1e-10).kernel()
    mol = water_cluster(4, seed=0, side=6.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e = RHF(mol, conv_tol=1e-10).kernel()
    torch.cuda.synchronize()
    return {'mol': mol, 'gpu_e': e, 'gpu_s': time.perf_counter() - t0}


if __name__ == '__main__':
    main()
