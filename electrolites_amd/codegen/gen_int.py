"""Generate the Rys-quadrature integral kernels (csrc/mdf_int_gen.hip).

One kernel per angular class, fully unrolled for that class:

* ``int3c_<la><lb><lc>``: three-centre ``(ij|P)`` over the shell pairs of class
  ``(la, lb)`` and the auxiliary shells of angular momentum ``lc``.  One thread
  per (shell pair, aux shell, aux spherical component m); ``m`` is the slowest
  task index, so a wavefront normally shares it and the ``switch (m)`` below
  does not diverge.  Each thread accumulates the Cartesian bra block of its m
  over the primitive triples, transforms the bra to spherical, and writes its
  values into the (naux, pair-chunk) output at pair ``i(i+1)/2 + j`` (i >= j).
* ``int2c_<la><lc>``: the metric ``(P|Q)`` with the same machinery (bra = one
  aux shell and a unit s function of exponent 0).

Rys recurrences (u = t^2 a root, p/q bra/ket exponents, P/Q their centres):
    B00 = u / (2(p+q)),  B10 = (1 - q u / (p+q)) / (2p),  B01 = (1 - p u/(p+q)) / (2q)
    C00 = (P - A) - q u (P - Q) / (p+q),   C00' = (Q - C) + p u (P - Q) / (p+q)
    G(e+1, f) = C00 G(e, f) + e B10 G(e-1, f) + f B00 G(e, f-1)
    G(e, f+1) = C00' G(e, f) + f B01 G(e, f-1) + e B00 G(e-1, f)
and the horizontal transfer I(a, b+1) = I(a+1, b) + (A - B) I(a, b).  The
prefactor 2 pi^(5/2) / (p q sqrt(p+q)) exp(-ab/p |A-B|^2) times the weight is
folded into G_z(0, 0).  (Standard Rys scheme; the class-per-kernel layout and
the generator-driven unrolling follow electrolites/codegen/gen_j3c.py:158-372,
which does the same for McMurchie-Davidson.)

Run: ``python -m electrolites_amd.codegen.gen_int``.
"""
import os

import numpy as np

from electrolites_amd.gto.mole import cart2sph, cart_powers, ncart

LMAX_AO = 2
LMAX_AUX = 4
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, '..', 'csrc', 'mdf_int_gen.hip')
THREADS = 128
#: a three-centre class whose Cartesian bra block times the aux components
#: holds at most this many accumulators computes every aux component m in one
#: thread (the Rys roots and recurrences once instead of 2 lc + 1 times)
ALLM_MAX = 60


def allm(la, lb, lc):
    return ncart(la) * ncart(lb) * (2 * lc + 1) <= ALLM_MAX


def fmt(c):
    return repr(float(c))


def gen_class(la, lb, lc, two_c):
    name = f'int2c_{la}{lc}' if two_c else f'int3c_{la}{lb}{lc}'
    lab = la + lb
    nroots = (la + lb + lc) // 2 + 1
    ca, cb, cc = cart_powers(la), cart_powers(lb), cart_powers(lc)
    nca, ncb = len(ca), len(cb)
    am = (not two_c) and allm(la, lb, lc)
    nm = 2 * lc + 1 if am else 1
    Ca, Cb, Cc = cart2sph(la), cart2sph(lb), cart2sph(lc)
    L = []
    w = L.append
    if two_c:
        w(f'__global__ __launch_bounds__({THREADS}) void {name}(const BasisDev aux, '
          'const int *__restrict__ pl, int np, const int *__restrict__ ql, int nq, '
          'double *__restrict__ out, long ldo) {')
        w(f'    const long tid = (long)blockIdx.x * {THREADS} + threadIdx.x;')
        w(f'    if (tid >= (long)np * nq * {2 * lc + 1}) return;')
        w('    const int ip = (int)(tid % np);')
        w('    const long rest = tid / np;')
        w('    const int kq = (int)(rest % nq), m = (int)(rest / nq);')
        w('    const int ish = pl[ip], ksh = ql[kq];')
        w('    const BasisDev &ao = aux;')
        w('    const int jsh = ish;')
    else:
        w(f'__global__ __launch_bounds__({THREADS}) void {name}(const BasisDev ao, const BasisDev aux, '
          'const int *__restrict__ pairs, int np, const int *__restrict__ kl, int nk, '
          'double *__restrict__ out, long ldo, long p0, const int *__restrict__ cmap) {')
        w(f'    const long tid = (long)blockIdx.x * {THREADS} + threadIdx.x;')
        if am:
            w('    if (tid >= (long)np * nk) return;')
            w('    const int ip = (int)(tid % np), kq = (int)(tid / np);')
        else:
            w(f'    if (tid >= (long)np * nk * {2 * lc + 1}) return;')
            w('    const int ip = (int)(tid % np);')
            w('    const long rest = tid / np;')
            w('    const int kq = (int)(rest % nk), m = (int)(rest / nk);')
        w('    const int ish = pairs[2 * ip], jsh = pairs[2 * ip + 1], ksh = kl[kq];')
    w('    const double Ax = ao.xyz[3 * ish], Ay = ao.xyz[3 * ish + 1], Az = ao.xyz[3 * ish + 2];')
    if not two_c:
        w('    const double Bx = ao.xyz[3 * jsh], By = ao.xyz[3 * jsh + 1], Bz = ao.xyz[3 * jsh + 2];')
        w('    const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;')
        w('    const double rab2 = ABx * ABx + ABy * ABy + ABz * ABz;')
    w('    const double Cx = aux.xyz[3 * ksh], Cy = aux.xyz[3 * ksh + 1], Cz = aux.xyz[3 * ksh + 2];')
    w(f'    double acc[{nm * nca * ncb}];')
    w('#pragma unroll')
    w(f'    for (int t = 0; t < {nm * nca * ncb}; t++) acc[t] = 0.;')
    w('    const int npa = ao.nprim[ish], oa = ao.poff[ish];')
    if not two_c:
        w('    const int npb = ao.nprim[jsh], ob = ao.poff[jsh];')
    w('    const int npc = aux.nprim[ksh], oc = aux.poff[ksh];')
    w('    for (int ia = 0; ia < npa; ia++) {')
    w('        const double ea = ao.exps[oa + ia], da = ao.coefs[oa + ia];')
    if two_c:
        w('        {')
        w('        const double p = ea, fab = da;')
        w('        const double Px = Ax, Py = Ay, Pz = Az;')
    else:
        w('        for (int ib = 0; ib < npb; ib++) {')
        w('            const double eb = ao.exps[ob + ib];')
        w('            const double p = ea + eb, ip_ = 1.0 / p;')
        w('            const double fab = da * ao.coefs[ob + ib] * exp(-ea * eb * ip_ * rab2);')
        w('            if (fabs(fab) < PRIM_SCREEN) continue;')
        w('            const double Px = (ea * Ax + eb * Bx) * ip_, Py = (ea * Ay + eb * By) * ip_, '
          'Pz = (ea * Az + eb * Bz) * ip_;')
    w('            const double PAx = Px - Ax, PAy = Py - Ay, PAz = Pz - Az;')
    w('            const double PCx = Px - Cx, PCy = Py - Cy, PCz = Pz - Cz;')
    w('            const double rpc2 = PCx * PCx + PCy * PCy + PCz * PCz;')
    w('            for (int ic = 0; ic < npc; ic++) {')
    w('                const double q = aux.exps[oc + ic];')
    w('                const double pq = p + q, ipq = 1.0 / pq;')
    w('                const double x = p * q * ipq * rpc2;')
    w('                const double pref = TWO_PI_2P5 / (p * q * sqrt(pq)) * fab * aux.coefs[oc + ic];')
    w(f'                double ru[{nroots}], rw[{nroots}];')
    w(f'                rys_roots<{nroots}>(x, ru, rw);')
    w('#pragma unroll')
    w(f'                for (int r = 0; r < {nroots}; r++) {{')
    w('                    const double u = ru[r];')
    w('                    const double B00 = 0.5 * u * ipq;')
    if lab >= 2:
        w('                    const double B10 = 0.5 / p - 0.5 * q * u * ipq / p;')
    if lc >= 2:
        w('                    const double B01 = 0.5 / q - 0.5 * p * u * ipq / q;')
    w('                    const double qf = q * u * ipq, pf = p * u * ipq;')
    for d in 'xyz':
        w(f'                    const double C00{d} = PA{d} - qf * PC{d}, C0p{d} = pf * PC{d};')
    # VRR
    for d in 'xyz':
        init = 'rw[r] * pref' if d == 'z' else '1.0'
        w(f'                    const double g{d}_0_0 = {init};')
        for e in range(0, lab + 1):
            for f in range(0, lc + 1):
                if e == 0 and f == 0:
                    continue
                if f == 0:
                    terms = [f'C00{d} * g{d}_{e-1}_0']
                    if e - 1 >= 1:
                        terms.append(f'{e-1}.0 * B10 * g{d}_{e-2}_0')
                else:
                    terms = [f'C0p{d} * g{d}_{e}_{f-1}']
                    if f - 1 >= 1:
                        terms.append(f'{f-1}.0 * B01 * g{d}_{e}_{f-2}')
                    if e >= 1:
                        terms.append(f'{e}.0 * B00 * g{d}_{e-1}_{f-1}')
                w(f'                    const double g{d}_{e}_{f} = ' + ' + '.join(terms) + ';')
    # HRR: I{d}_{a}_{b}_{f}
    needed = {}
    for d in 'xyz':
        for f in range(lc + 1):
            for a in range(lab + 1):
                w(f'                    const double I{d}_{a}_0_{f} = g{d}_{a}_{f};')
            for b in range(1, lb + 1):
                for a in range(0, lab - b + 1):
                    w(f'                    const double I{d}_{a}_{b}_{f} = I{d}_{a+1}_{b-1}_{f} + '
                      f'AB{d} * I{d}_{a}_{b-1}_{f};')
    # products per m (every m in this thread, or the thread's m)
    def products(mi, base, indent):
        sup = [(k, Cc[mi, k]) for k in range(len(cc)) if abs(Cc[mi, k]) > 1e-15]
        for a_i, (ax, ay, az) in enumerate(ca):
            for b_i, (bx, by, bz) in enumerate(cb):
                terms = []
                for k, coef in sup:
                    kx, ky, kz = cc[k]
                    terms.append(f'{fmt(coef)} * Ix_{ax}_{bx}_{kx} * Iy_{ay}_{by}_{ky} * Iz_{az}_{bz}_{kz}')
                w(f'{indent}acc[{base + a_i * ncb + b_i}] += ' + ' + '.join(terms) + ';')
    if am:
        for mi in range(2 * lc + 1):
            products(mi, mi * nca * ncb, '                    ')
    else:
        w('                    switch (m) {')
        for mi in range(2 * lc + 1):
            w(f'                    case {mi}: {{')
            products(mi, 0, '                        ')
            w('                        break;')
            w('                    }')
        w('                    }')
    w('                }')          # roots
    w('            }')              # ic
    w('        }')                  # ib / block
    w('    }')                      # ia
    # bra to spherical and store
    nsa, nsb = 2 * la + 1, 2 * lb + 1
    if two_c:
        w('    const int prow0 = aux.aooff[ish], qcol = aux.aooff[ksh] + m;')
        for s in range(nsa):
            terms = [f'{fmt(Ca[s, a])} * acc[{a}]' for a in range(nca) if abs(Ca[s, a]) > 1e-15]
            w(f'    out[(long)(prow0 + {s}) * ldo + qcol] = ' + ' + '.join(terms) + ';')
    else:
        w('    const int i0 = ao.aooff[ish], j0 = ao.aooff[jsh];')
        w('    const bool diag = ish == jsh;')
        for mi in range(nm):
            mexpr = str(mi) if am else 'm'
            w(f'    {{ double *orow = out + (long)(aux.aooff[ksh] + {mexpr}) * ldo;')
            base = mi * nca * ncb
            for s in range(nsa):
                for t in range(nsb):
                    terms = []
                    for a in range(nca):
                        if abs(Ca[s, a]) < 1e-15:
                            continue
                        for b in range(ncb):
                            if abs(Cb[t, b]) < 1e-15:
                                continue
                            terms.append(f'{fmt(Ca[s, a] * Cb[t, b])} * acc[{base + a * ncb + b}]')
                    cond = f'if (!diag || {s} >= {t}) ' if la == lb else ''
                    w(f'    {cond}{{ const long i = i0 + {s}; const long f = i * (i + 1) / 2 + j0 + {t} - p0; '
                      f'orow[cmap ? cmap[f] : f] = ' + ' + '.join(terms) + '; }')
            w('    }')
    w('}')
    return name, '\n'.join(L)


def gen_int1e(la, lb):
    """S, T and V (all point charges) of one shell pair, one thread each."""
    name = f'int1e_{la}{lb}'
    ca, cb = cart_powers(la), cart_powers(lb)
    nca, ncb = len(ca), len(cb)
    Ca, Cb = cart2sph(la), cart2sph(lb)
    nroots = (la + lb) // 2 + 1
    L = []
    w = L.append
    w(f'__global__ __launch_bounds__({THREADS}) void {name}(const BasisDev ao, '
      'const int *__restrict__ pairs, int np, int natm, const double *__restrict__ charges, '
      'const double *__restrict__ atom_xyz, double *__restrict__ S, double *__restrict__ T, '
      'double *__restrict__ V, int nao) {')
    w(f'    const int ip = blockIdx.x * {THREADS} + threadIdx.x;')
    w('    if (ip >= np) return;')
    w('    const int ish = pairs[2 * ip], jsh = pairs[2 * ip + 1];')
    w('    const double Ax = ao.xyz[3 * ish], Ay = ao.xyz[3 * ish + 1], Az = ao.xyz[3 * ish + 2];')
    w('    const double Bx = ao.xyz[3 * jsh], By = ao.xyz[3 * jsh + 1], Bz = ao.xyz[3 * jsh + 2];')
    w('    const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;')
    w('    const double rab2 = ABx * ABx + ABy * ABy + ABz * ABz;')
    w(f'    double sa[{nca * ncb}], ta[{nca * ncb}], va[{nca * ncb}];')
    w('#pragma unroll')
    w(f'    for (int t = 0; t < {nca * ncb}; t++) {{ sa[t] = 0.; ta[t] = 0.; va[t] = 0.; }}')
    w('    const int npa = ao.nprim[ish], oa = ao.poff[ish], npb = ao.nprim[jsh], ob = ao.poff[jsh];')
    w('    for (int ia = 0; ia < npa; ia++) {')
    w('        const double ea = ao.exps[oa + ia], da = ao.coefs[oa + ia];')
    w('        for (int ib = 0; ib < npb; ib++) {')
    w('            const double eb = ao.exps[ob + ib];')
    w('            const double p = ea + eb, ip_ = 1.0 / p, h = 0.5 * ip_;')
    w('            const double fab = da * ao.coefs[ob + ib] * exp(-ea * eb * ip_ * rab2);')
    w('            if (fabs(fab) < PRIM_SCREEN) continue;')
    w('            const double Px = (ea * Ax + eb * Bx) * ip_, Py = (ea * Ay + eb * By) * ip_, '
      'Pz = (ea * Az + eb * Bz) * ip_;')
    w('            const double PAx = Px - Ax, PAy = Py - Ay, PAz = Pz - Az;')
    w('            const double PBx = Px - Bx, PBy = Py - By, PBz = Pz - Bz;')
    # 1D overlaps s(i, j) up to (la, lb + 2) by Obara-Saika, unit prefactor
    for d in 'xyz':
        for i in range(la + 1):
            for j in range(lb + 3):
                if i == 0 and j == 0:
                    w(f'            const double s{d}_0_0 = 1.0;')
                    continue
                if i > 0:
                    terms = [f'PA{d} * s{d}_{i-1}_{j}']
                    if i - 1 >= 1:
                        terms.append(f'{i-1}.0 * h * s{d}_{i-2}_{j}')
                    if j >= 1:
                        terms.append(f'{j}.0 * h * s{d}_{i-1}_{j-1}')
                else:
                    terms = [f'PB{d} * s{d}_{i}_{j-1}']
                    if j - 1 >= 1:
                        terms.append(f'{j-1}.0 * h * s{d}_{i}_{j-2}')
                w(f'            const double s{d}_{i}_{j} = ' + ' + '.join(terms) + ';')
        for i in range(la + 1):
            for j in range(lb + 1):
                terms = [f'-2.0 * eb * {2 * j + 1}.0 * s{d}_{i}_{j}', f'4.0 * eb * eb * s{d}_{i}_{j+2}']
                if j >= 2:
                    terms.append(f'{j * (j - 1)}.0 * s{d}_{i}_{j-2}')
                w(f'            const double k{d}_{i}_{j} = -0.5 * (' + ' + '.join(terms) + ');')
    w('            const double sp = fab * PI_1P5 * ip_ * sqrt(ip_);')
    for a_i, (ax, ay, az) in enumerate(ca):
        for b_i, (bx, by, bz) in enumerate(cb):
            k = a_i * ncb + b_i
            w(f'            sa[{k}] += sp * sx_{ax}_{bx} * sy_{ay}_{by} * sz_{az}_{bz};')
            w(f'            ta[{k}] += sp * (kx_{ax}_{bx} * sy_{ay}_{by} * sz_{az}_{bz} + '
              f'sx_{ax}_{bx} * ky_{ay}_{by} * sz_{az}_{bz} + sx_{ax}_{bx} * sy_{ay}_{by} * kz_{az}_{bz});')
    # nuclear attraction by Rys with point charges
    lab = la + lb
    w('            const double vp = -TWO_PI * ip_ * fab;')
    w('            for (int c = 0; c < natm; c++) {')
    w('                const double PCx = Px - atom_xyz[3 * c], PCy = Py - atom_xyz[3 * c + 1], '
      'PCz = Pz - atom_xyz[3 * c + 2];')
    w('                const double x = p * (PCx * PCx + PCy * PCy + PCz * PCz);')
    w(f'                double ru[{nroots}], rw[{nroots}];')
    w(f'                rys_roots<{nroots}>(x, ru, rw);')
    w('                const double zc = vp * charges[c];')
    w('#pragma unroll')
    w(f'                for (int r = 0; r < {nroots}; r++) {{')
    w('                    const double u = ru[r];')
    if lab >= 2:
        w('                    const double B10 = (1.0 - u) * h;')
    for d in 'xyz':
        w(f'                    const double C00{d} = PA{d} - u * PC{d};')
        init = 'zc * rw[r]' if d == 'z' else '1.0'
        w(f'                    const double g{d}_0 = {init};')
        for e in range(1, lab + 1):
            terms = [f'C00{d} * g{d}_{e-1}']
            if e - 1 >= 1:
                terms.append(f'{e-1}.0 * B10 * g{d}_{e-2}')
            w(f'                    const double g{d}_{e} = ' + ' + '.join(terms) + ';')
        for a in range(lab + 1):
            w(f'                    const double I{d}_{a}_0 = g{d}_{a};')
        for b in range(1, lb + 1):
            for a in range(0, lab - b + 1):
                w(f'                    const double I{d}_{a}_{b} = I{d}_{a+1}_{b-1} + AB{d} * I{d}_{a}_{b-1};')
    for a_i, (ax, ay, az) in enumerate(ca):
        for b_i, (bx, by, bz) in enumerate(cb):
            w(f'                    va[{a_i * ncb + b_i}] += Ix_{ax}_{bx} * Iy_{ay}_{by} * Iz_{az}_{bz};')
    w('                }')
    w('            }')
    w('        }')
    w('    }')
    w('    const int i0 = ao.aooff[ish], j0 = ao.aooff[jsh];')
    for s_ in range(2 * la + 1):
        for t in range(2 * lb + 1):
            idx = []
            for a in range(nca):
                if abs(Ca[s_, a]) < 1e-15:
                    continue
                for b in range(ncb):
                    if abs(Cb[t, b]) < 1e-15:
                        continue
                    idx.append((fmt(Ca[s_, a] * Cb[t, b]), a * ncb + b))
            for arr, outn in (('sa', 'S'), ('ta', 'T'), ('va', 'V')):
                expr = ' + '.join(f'{c} * {arr}[{k}]' for c, k in idx)
                w(f'    {{ const double v = {expr}; {outn}[(long)(i0 + {s_}) * nao + j0 + {t}] = v; '
                  f'{outn}[(long)(j0 + {t}) * nao + i0 + {s_}] = v; }}')
    w('}')
    return name, '\n'.join(L)


def main():
    parts = ['// Generated by electrolites_amd/codegen/gen_int.py; do not edit.',
             '// Rys-quadrature int3c2e / int2c2e kernels, one per angular class.',
             '#include "mdf_int.h"', '',
             'namespace mdf {', '']
    names3, names2 = {}, {}
    for la in range(LMAX_AO + 1):
        for lb in range(LMAX_AO + 1):
            for lc in range(LMAX_AUX + 1):
                n, src = gen_class(la, lb, lc, False)
                names3[(la, lb, lc)] = n
                parts.append(src)
                parts.append('')
    for la in range(LMAX_AUX + 1):
        for lc in range(LMAX_AUX + 1):
            n, src = gen_class(la, 0, lc, True)
            names2[(la, lc)] = n
            parts.append(src)
            parts.append('')
    names1 = {}
    for la in range(LMAX_AO + 1):
        for lb in range(LMAX_AO + 1):
            n, src = gen_int1e(la, lb)
            names1[(la, lb)] = n
            parts.append(src)
            parts.append('')
    # dispatch
    parts.append('int launch_int1e(int la, int lb, const BasisDev &ao, const int *pairs, int np, int natm, '
                 'const double *charges, const double *atom_xyz, double *S, double *T, double *V, int nao, '
                 'hipStream_t st) {')
    parts.append('    if (np == 0) return MDF_OK;')
    parts.append(f'    const dim3 g((unsigned)((np + {THREADS - 1}) / {THREADS})), b({THREADS});')
    parts.append(f'    switch (la * {LMAX_AO + 1} + lb) {{')
    for (la, lb), n in names1.items():
        parts.append(f'    case {la * (LMAX_AO + 1) + lb}: hipLaunchKernelGGL({n}, g, b, 0, st, ao, pairs, np, natm, '
                     'charges, atom_xyz, S, T, V, nao); break;')
    parts.append('    default: set_error("int1e: class (%d %d) not generated", la, lb); return MDF_EINVAL;')
    parts.append('    }')
    parts.append('    MDF_LAUNCHED("int1e");')
    parts.append('    return MDF_OK;')
    parts.append('}')
    parts.append('')
    cases = ' || '.join(f'(la == {a} && lb == {b} && lc == {c})' for (a, b, c) in names3 if allm(a, b, c))
    parts.append(f'static bool allm_class(int la, int lb, int lc) {{ return {cases}; }}')
    parts.append('')
    parts.append('int launch_int3c(int la, int lb, int lc, const BasisDev &ao, const BasisDev &aux, '
                 'const int *pairs, int np, const int *kl, int nk, double *out, long ldo, long p0, '
                 'const int *cmap, hipStream_t st) {')
    parts.append('    // one thread per (pair, aux shell) for the all-m classes, per (pair, aux shell, m) otherwise')
    parts.append('    const long n = (long)np * nk * (allm_class(la, lb, lc) ? 1 : 2 * lc + 1);')
    parts.append('    if (n == 0) return MDF_OK;')
    parts.append(f'    const dim3 g((unsigned)((n + {THREADS - 1}) / {THREADS})), b({THREADS});')
    parts.append(f'    switch ((la * {LMAX_AO + 1} + lb) * {LMAX_AUX + 1} + lc) {{')
    for (la, lb, lc), n in names3.items():
        key = (la * (LMAX_AO + 1) + lb) * (LMAX_AUX + 1) + lc
        parts.append(f'    case {key}: hipLaunchKernelGGL({n}, g, b, 0, st, ao, aux, pairs, np, kl, nk, out, ldo, p0, cmap); break;')
    parts.append('    default: set_error("int3c2e: class (%d %d|%d) not generated", la, lb, lc); return MDF_EINVAL;')
    parts.append('    }')
    parts.append('    MDF_LAUNCHED("int3c");')
    parts.append('    return MDF_OK;')
    parts.append('}')
    parts.append('')
    parts.append('int launch_int2c(int la, int lc, const BasisDev &aux, const int *pl, int np, const int *ql, '
                 'int nq, double *out, long ldo, hipStream_t st) {')
    parts.append('    const long n = (long)np * nq * (2 * lc + 1);')
    parts.append('    if (n == 0) return MDF_OK;')
    parts.append(f'    const dim3 g((unsigned)((n + {THREADS - 1}) / {THREADS})), b({THREADS});')
    parts.append(f'    switch (la * {LMAX_AUX + 1} + lc) {{')
    for (la, lc), n in names2.items():
        parts.append(f'    case {la * (LMAX_AUX + 1) + lc}: hipLaunchKernelGGL({n}, g, b, 0, st, aux, pl, np, ql, nq, out, ldo); break;')
    parts.append('    default: set_error("int2c2e: class (%d|%d) not generated", la, lc); return MDF_EINVAL;')
    parts.append('    }')
    parts.append('    MDF_LAUNCHED("int2c");')
    parts.append('    return MDF_OK;')
    parts.append('}')
    parts.append('')
    parts.append('}  // namespace mdf')
    with open(OUT, 'w') as f:
        f.write('\n'.join(parts) + '\n')
    print('wrote', OUT, len(names3), '3c and', len(names2), '2c kernels')


if __name__ == '__main__':
    main()
