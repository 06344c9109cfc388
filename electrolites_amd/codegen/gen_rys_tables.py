"""Generate the Rys quadrature tables used by the HIP integral kernels.

For n roots and argument x, the Rys quadrature ``{(u_i, w_i)}`` (``u = t^2``
in [0, 1)) is exact for ``int_0^1 P(t^2) exp(-x t^2) dt`` with ``P`` of degree
< 2n, i.e. ``sum_i w_i u_i^k = F_k(x)`` (Boys functions) for ``k < 2n``.

Reference values are computed in 60-digit arithmetic (mpmath): moments
``F_k(x) = gamma_lower(k + 1/2, x) / (2 x^(k+1/2))``, the Hankel moment matrix
is Cholesky-factored, the Jacobi matrix of the recurrence coefficients is
diagonalised (Golub-Welsch), and the weights are ``F_0 (first eigenvector
component)^2``.

Table layout (per n): x in [0, XMAX_n) is split into unit intervals; on each,
every root and weight is a Chebyshev series of degree DEG in
``s = 2 (x - x0) - 1``.  For x >= XMAX_n the large-x limit is used:
``u_i = h_i^2 / x``, ``w_i = H_i / sqrt(x)`` with (h_i, H_i) the positive
nodes and weights of the 2n-point Gauss-Hermite rule (the integral extends to
infinity with a relative error below 1e-16 there).  XMAX_n is the smallest
multiple of 5 for which that limit matches the exact quadrature to 1e-15
relative on a dense check grid.

The branch structure (tabulated middle, asymptotic large-x) follows the Rys
root code the reference carries (electrolites/kernels/cuda/gvhf-rys/
rys_roots.cu:24-60); the tables themselves are this script's own.

Run: ``python -m electrolites_amd.codegen.gen_rys_tables`` -> writes
``csrc/rys_tables.inc`` and ``codegen/rys_tables_meta.json``.
"""
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import mpmath as mp

NMAX = 6
DEG = 13
DPS = 60
HERE = os.path.dirname(os.path.abspath(__file__))
OUT_INC = os.path.join(HERE, '..', 'csrc', 'rys_tables.inc')
OUT_META = os.path.join(HERE, 'rys_tables_meta.json')


def boys_mp(k, x):
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1) / (2 * k + 1)
    a = mp.mpf(k) + mp.mpf(1) / 2
    return mp.gammainc(a, 0, x) / (2 * x ** a)


def rys_exact(n, x):
    """(u[0..n-1], w[0..n-1]) ascending in u, as mpf."""
    with mp.workdps(DPS):
        m = [boys_mp(k, x) for k in range(2 * n + 1)]
        H = mp.matrix(n + 1, n + 1)
        for i in range(n + 1):
            for j in range(n + 1):
                H[i, j] = m[i + j]
        R = mp.cholesky(H).T          # upper: H = R^T R
        J = mp.matrix(n, n)
        for j in range(n):
            a = R[j, j + 1] / R[j, j]
            if j > 0:
                a -= R[j - 1, j] / R[j - 1, j - 1]
            J[j, j] = a
            if j < n - 1:
                b = R[j + 1, j + 1] / R[j, j]
                J[j, j + 1] = b
                J[j + 1, j] = b
        E, Q = mp.eigsy(J)
        pairs = sorted((E[i], m[0] * Q[0, i] ** 2) for i in range(n))
        return [p[0] for p in pairs], [p[1] for p in pairs]


def hermite_limit(n):
    """Positive nodes/weights of the 2n-point Gauss-Hermite rule (mpf)."""
    with mp.workdps(DPS):
        nodes, weights = [], []
        N = 2 * n
        # Jacobi matrix of physicists' Hermite: offdiag sqrt(k/2)
        J = mp.matrix(N, N)
        for k in range(1, N):
            J[k - 1, k] = J[k, k - 1] = mp.sqrt(mp.mpf(k) / 2)
        E, Q = mp.eigsy(J)
        for i in range(N):
            if E[i] > 0:
                nodes.append(E[i])
                weights.append(mp.sqrt(mp.pi) * Q[0, i] ** 2)
        pairs = sorted(zip(nodes, weights))
        return [p[0] for p in pairs], [p[1] for p in pairs]


def asym(n, x, hl):
    h, H = hl
    x = mp.mpf(x)
    return [hi ** 2 / x for hi in h], [Hi / mp.sqrt(x) for Hi in H]


def cheb_nodes(deg):
    return [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / (deg + 1)) for k in range(deg + 1)]


def fit_interval(args):
    n, x0 = args
    with mp.workdps(DPS):
        nodes = cheb_nodes(DEG)
        vals = []
        for s in nodes:
            x = mp.mpf(x0) + (s + 1) / 2
            u, w = rys_exact(n, x)
            vals.append(u + w)
        coefs = []
        for f in range(2 * n):
            cf = []
            for j in range(DEG + 1):
                acc = mp.mpf(0)
                for k, s in enumerate(nodes):
                    acc += vals[k][f] * mp.cos(j * mp.acos(s))
                acc *= mp.mpf(2) / (DEG + 1)
                if j == 0:
                    acc /= 2
                cf.append(acc)
            coefs.append(cf)
        return n, x0, coefs


def cheb_eval(cf, s):
    # Clenshaw
    b1 = b2 = 0.
    for c in reversed(cf[1:]):
        b1, b2 = 2 * s * b1 - b2 + c, b1
    return s * b1 - b2 + cf[0]


def find_xmax(n, hl):
    for X in range(15, 200, 5):
        ok = True
        for k in range(0, 41):
            x = X + k * 2.5
            u, w = rys_exact(n, x)
            ua, wa = asym(n, x, hl)
            for a, b in zip(u + w, ua + wa):
                if abs(a - b) > 1e-15 * abs(a):
                    ok = False
                    break
            if not ok:
                break
        if ok:
            return X
    raise RuntimeError(f'no large-x threshold for n={n}')


def main():
    jobs = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    meta = {'nmax': NMAX, 'deg': DEG, 'xmax': {}, 'offset': {}, 'hermite': {}}
    hls = {n: hermite_limit(n) for n in range(1, NMAX + 1)}
    for n in range(1, NMAX + 1):
        meta['xmax'][n] = find_xmax(n, hls[n])
        h, H = hls[n]
        meta['hermite'][n] = {'u_times_x': [float(hi ** 2) for hi in h],
                              'w_times_sqrtx': [float(Hi) for Hi in H]}
        print(f'n={n}: XMAX={meta["xmax"][n]}', flush=True)
    tasks = [(n, x0) for n in range(1, NMAX + 1) for x0 in range(meta['xmax'][n])]
    results = {}
    with ProcessPoolExecutor(jobs) as ex:
        for n, x0, cf in ex.map(fit_interval, tasks, chunksize=4):
            results[(n, x0)] = cf
    # layout: table[n][interval][function][DEG+1], functions u_0..u_{n-1}, w_0..w_{n-1}
    flat = []
    off = 0
    for n in range(1, NMAX + 1):
        meta['offset'][n] = off
        for x0 in range(meta['xmax'][n]):
            for f in range(2 * n):
                flat.extend(float(c) for c in results[(n, x0)][f])
                off += DEG + 1
    # verification at random points against fresh exact values
    import random
    rnd = random.Random(7)
    worst = 0.
    for n in range(1, NMAX + 1):
        for _ in range(60):
            x = rnd.uniform(0, meta['xmax'][n] + 30)
            u, w = rys_exact(n, x)
            if x < meta['xmax'][n]:
                x0 = int(x)
                s = 2 * (x - x0) - 1
                base = meta['offset'][n] + x0 * 2 * n * (DEG + 1)
                got = [cheb_eval(flat[base + f * (DEG + 1): base + (f + 1) * (DEG + 1)], s)
                       for f in range(2 * n)]
            else:
                ua, wa = asym(n, x, hls[n])
                got = [float(v) for v in ua + wa]
            for g, e in zip(got, u + w):
                worst = max(worst, abs(g - float(e)) / abs(float(e)))
    meta['verify_worst_rel'] = worst
    print(f'verification: worst relative error {worst:.3e}', flush=True)
    with open(OUT_INC, 'w') as f:
        f.write('// Generated by electrolites_amd/codegen/gen_rys_tables.py; do not edit.\n')
        f.write(f'// Rys roots/weights: Chebyshev degree {DEG} on unit intervals of x,\n')
        f.write(f'// verified worst relative error {worst:.2e} against 60-digit values.\n')
        f.write(f'#define RYS_NMAX {NMAX}\n#define RYS_DEG {DEG}\n')
        f.write('constexpr int RYS_XMAX[RYS_NMAX + 1] = {0, ' +
                ', '.join(str(meta['xmax'][n]) for n in range(1, NMAX + 1)) + '};\n')
        f.write('constexpr int RYS_OFFSET[RYS_NMAX + 1] = {0, ' +
                ', '.join(str(meta['offset'][n]) for n in range(1, NMAX + 1)) + '};\n')
        ux = [v for n in range(1, NMAX + 1) for v in meta['hermite'][n]['u_times_x']]
        wx = [v for n in range(1, NMAX + 1) for v in meta['hermite'][n]['w_times_sqrtx']]
        f.write('// large-x limit, n(n-1)/2 + i: u = RYS_LX_U / x, w = RYS_LX_W / sqrt(x)\n')
        f.write('__device__ const double RYS_LX_U[] = {' + ', '.join(repr(v) for v in ux) + '};\n')
        f.write('__device__ const double RYS_LX_W[] = {' + ', '.join(repr(v) for v in wx) + '};\n')
        f.write(f'__device__ const double RYS_TABLE[{len(flat)}] = {{\n')
        for i in range(0, len(flat), 4):
            f.write('    ' + ', '.join(repr(v) for v in flat[i:i + 4]) + ',\n')
        f.write('};\n')
    with open(OUT_META, 'w') as f:
        json.dump(meta, f, indent=1)
    print('wrote', OUT_INC, len(flat), 'doubles')


if __name__ == '__main__':
    main()
