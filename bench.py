"""Benchmark of the density-fitted J/K path on one MI355X.

Workload (one "step" = one FP64 ``get_jk`` call, J and K, over a resident
cderi): a 64-water cluster at 1.0 g/cm^3 (seed 0), def2-SVP orbital basis
(nao 1536) with the 109-function-per-water fitting set (naux 6976), the full
packed triangle (npairs 1,180,416, 65.9 GB in FP64), and a mo-tagged random
PSD density of rank nocc = 320.  The cderi is built on the GPU first (Rys
int2c2e/int3c2e, Cholesky, L^-1, triangular multiply) and its build is timed
per stage; ``--synthetic-cderi`` replaces it with seeded N(0,1) values of the
same shape for quick kernel timing.

Prints ONE JSON line.  ``roofline`` is the K build (the dominant kernel
group) against FP64 MFMA peak; ``roofline_j`` is the J build against HBM
bandwidth.  Per-call durations come from HIP events recorded on the stream
the library launches on (torch's current stream).

``cpu_baseline`` times the CPU restatement of the same algorithm (numpy/
OpenBLAS FP64: rho = L d, J = L^T rho; per aux block unpack + Lij,jk->Lik +
rank-k update) on a bounded sample of aux rows, scaled to the full naux.
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
    ap.add_argument('--cpu-sample-j', type=int, default=96)
    ap.add_argument('--cpu-sample-k', type=int, default=24)
    return ap.parse_args()


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def cpu_baseline(mol_nao, naux, npairs, nocc, args):
    """CPU restatement (oracle leg) on a sample of aux rows, scaled."""
    from oracle import oracle as O
    nthreads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    rng = np.random.default_rng(1)
    nao = mol_nao
    pa, _ = O.pair_addresses(nao)
    c = np.linalg.qr(rng.standard_normal((nao, nocc)))[0]
    dm = 2 * c @ c.T
    f = c * np.sqrt(2.)
    d = O.sparse_dm(dm, pa, 1)[0]
    sj = args.cpu_sample_j
    blk = rng.standard_normal((sj, npairs))
    t0 = time.perf_counter()
    rho = blk @ d
    vjp = blk.T @ rho
    tj = (time.perf_counter() - t0) * naux / sj
    del blk, vjp
    sk = args.cpu_sample_k
    blk = rng.standard_normal((sk, npairs))
    t0 = time.perf_counter()
    a = O.unpack(blk, pa, nao)                        # (sk, nao, nao)
    r = np.matmul(a, f)                               # (sk, nao, nocc)
    r2 = r.transpose(1, 0, 2).reshape(nao, -1)
    vk = r2 @ r2.T
    tk = (time.perf_counter() - t0) * naux / sk
    del a, r, r2, blk
    per_call = tj + tk
    return {'value': 3600. / per_call, 'unit': 'get_jk/hr', 'cores': nthreads,
            'kind': 'port',
            'sample': f'J over {sj} and K over {sk} of {naux} aux rows at the '
                      f'full (nao {nao}, npairs {npairs}, nocc {nocc}) shape, '
                      f'numpy/OpenBLAS FP64, scaled by naux/sample',
            'j_ms': tj * 1e3, 'k_ms': tk * 1e3}


def main():
    args = parse()
    if args.gpus != 1:
        print(f'bench.py measures one GPU only (asked for --gpus {args.gpus}); '
              'the path shards along the aux index across GPUs but is built '
              'and measured on one', file=sys.stderr)
        sys.exit(2)
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
        cderi, _ = fastdfcderi.cholesky_eri(mol, timer=timer)
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
        _cderi_idx = None

        def build(self):
            pass

    dfobj = DF()
    torch.cuda.synchronize()
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

    # the same K in FP32 (the mixed-precision cycles), outside the timed steps
    fl, fr = fastdf.factorize_dm(dm, 1)
    lay = fastdf._layout(dfobj, cderi, nao)
    fastdf._vk(cderi, lay, fl, fr, 'fp32')
    e32 = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    e32[0].record()
    vk32 = fastdf._vk(cderi, lay, fl, fr, 'fp32')
    e32[1].record()
    vk64 = fastdf._vk(cderi, lay, fl, fr, 'fp64')
    torch.cuda.synchronize()
    tk32 = e32[0].elapsed_time(e32[1]) * 1e-3
    k32_err = float((vk32 - vk64).abs().max() / vk64.abs().max())
    del vk32, vk64

    S = naux * npairs * 8
    bj = 2 * S + 16 * (npairs + naux)
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
                      'k_fp32_rel_err_vs_fp64': k32_err},
        'cderi_build': build,
        'roofline': {'bound': 'mfma', 'achieved': fk / tk / 1e12,
                     'peak': FP64_MFMA_PEAK / 1e12, 'unit': 'TFLOP/s',
                     'frac': fk / tk / FP64_MFMA_PEAK, 'traffic': None,
                     'kernel': 'mdf_get_k (k_step1 + k_step2 + k_finalize)',
                     'algorithmic': fk},
        'roofline_j': {'bound': 'hbm', 'achieved': bj / tj / 1e9,
                       'peak': HBM_PEAK / 1e9, 'unit': 'GB/s',
                       'frac': bj / tj / HBM_PEAK, 'traffic': None,
                       'kernel': 'mdf_get_j (j_rho_partial + j_vj_pass)',
                       'algorithmic': bj},
    }
    if not args.no_cpu:
        del cderi
        DF._cderi = None
        torch.cuda.empty_cache()
        log('cpu baseline')
        out['cpu_baseline'] = cpu_baseline(nao, naux, npairs, nocc, args)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
