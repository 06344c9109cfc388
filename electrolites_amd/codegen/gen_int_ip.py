"""Generate the nuclear-derivative Rys kernels (csrc/mdf_int_ip_gen.hip) of
the density-fitted gradient's two-electron term.

One kernel per angular class, as the integral kernels of gen_int.py, each
contracting its derivative integrals with a weight on the fly and adding the
result to a per-atom gradient:

* ``int3c_ip_<la><lb><lc>`` (``_m<m>``: one kernel per aux spherical
  component m, launched together, in the classes where a thread does not
  take all of them): one thread per (shell pair (I >= J), aux shell K, aux
  spherical component m).  The derivative of a primitive Cartesian
  Gaussian about its centre is ``d/dA_x x_A^a e^(-alpha r_A^2) = 2 alpha
  x_A^(a+1) e - a x_A^(a-1) e``, so with the Rys factors of gen_int.py
  ``dI/dA_x = 2 alpha I_x(a+1, b) - a I_x(a-1, b)`` (and the same about B with
  the second index); the aux centre follows from translational invariance,
  ``d/dC = -(d/dA + d/dB)``.  The thread reads the weight block
  ``W[P = (K, m), (i, j)]`` (per aux row one value per packed pair p = i(i+1)/2
  + j, i >= j, at column ``cmap[p]``, or p without a map; both orders of a
  pair take the one stored value), turns it into Cartesian
  weights, and adds ``sum W dI/dA`` to atom(I) and ``sum W dI/dB`` to atom(J)
  of ``grad``, and minus their sum to atom(K) of ``grad_aux`` (the
  auxiliary-basis response, which may be the same array); an off-diagonal
  shell pair counts twice
  (it stands for (I, J) and (J, I)).
* ``int2c_ip_<la><lc>``: one thread per ordered (aux shell P, aux shell Q,
  component m of Q), adding ``Z[P, Q] d(P|Q)/dA_P`` to atom(P) and its
  negative to atom(Q).

The per-atom sums are accumulated with FP64 vector-memory atomics (the order
of the additions is not fixed, so the last bits vary between runs).

Run: ``python -m electrolites_amd.codegen.gen_int_ip``.
"""
import os

from electrolites_amd.codegen.gen_int import LMAX_AO, LMAX_AUX, THREADS, fmt
from electrolites_amd.gto.mole import cart2sph, cart_powers

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, '..', 'csrc', 'mdf_int_ip_gen.hip')


def _vrr_hrr(w, lab_max, lc, lb_max, indent):
    """Rys 2-D integrals g(e, f) for e <= lab_max, f <= lc, then the
    horizontal transfer I(a, b, f) for b <= lb_max, a + b <= lab_max.  lc is
    one limit for all three directions or a dict of one per direction."""
    lcd = lc if isinstance(lc, dict) else {d: lc for d in 'xyz'}
    for d in 'xyz':
        lc = lcd[d]
        init = 'rw[r] * pref' if d == 'z' else '1.0'
        w(f'{indent}const double g{d}_0_0 = {init};')
        for e in range(0, lab_max + 1):
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
                w(f'{indent}const double g{d}_{e}_{f} = ' + ' + '.join(terms) + ';')
    for d in 'xyz':
        for f in range(lcd[d] + 1):
            for a in range(lab_max + 1):
                w(f'{indent}const double I{d}_{a}_0_{f} = g{d}_{a}_{f};')
            for b in range(1, lb_max + 1):
                for a in range(0, lab_max - b + 1):
                    w(f'{indent}const double I{d}_{a}_{b}_{f} = I{d}_{a+1}_{b-1}_{f} + AB{d} * I{d}_{a}_{b-1}_{f};')


def _deriv(d, a, b, k, wrt):
    """Expression of d I_d(a, b, k) / d(centre) for wrt 'A' (first index) or
    'B' (second index)."""
    if wrt == 'A':
        e = f'2.0 * ea * I{d}_{a+1}_{b}_{k}'
        if a:
            e = f'({e} - {a}.0 * I{d}_{a-1}_{b}_{k})'
    else:
        # I(a, b + 1) = I(a + 1, b) + AB I(a, b): the horizontal transfer one
        # level short, so no I(., lb + 1) is kept live
        e = f'2.0 * eb * (I{d}_{a+1}_{b}_{k} + AB{d} * I{d}_{a}_{b}_{k})'
        if b:
            e = f'({e} - {b}.0 * I{d}_{a}_{b-1}_{k})'
    return e


KCH = 4       # default aux shells one thread of a three-centre kernel loops over
QCH = 16      # aux shells one thread of a two-centre kernel loops over
LDS_MAX_ATOMS = 1024
#: a three-centre class whose Cartesian weight block times the aux components
#: holds at most this many values computes every aux component m in one
#: thread (the Rys roots and recurrences once instead of 2 lc + 1 times)
ALLM_IP_MAX = 15


def allm_ip(la, lb, lc):
    return len(cart_powers(la)) * len(cart_powers(lb)) * (2 * lc + 1) <= ALLM_IP_MAX


def _head(w, indent='    '):
    w(indent + 'extern __shared__ double sg[];')
    w(indent + 'if (use_lds) {')
    w(indent + f'    for (int i = threadIdx.x; i < nslot * natm; i += {THREADS}) sg[i] = 0.;')
    w(indent + '    __syncthreads();')
    w(indent + '}')


def _tail(w, indent='    '):
    w(indent + 'if (use_lds) {')
    w(indent + '    __syncthreads();')
    w(indent + f'    for (int i = threadIdx.x; i < nslot * natm; i += {THREADS}) {{')
    w(indent + '        const double v = sg[i];')
    w(indent + '        if (v != 0.) atomicAdd(i < 3 * natm ? grad + i : grad_aux + (i - 3 * natm), v);')
    w(indent + '    }')
    w(indent + '}')


def _component_limits(lc, mi):
    """Highest power per direction among the Cartesian functions of the
    spherical aux component mi (its Rys recurrences stop there)."""
    cc, Cc = cart_powers(lc), cart2sph(lc)
    sup = [cc[k] for k in range(len(cc)) if abs(Cc[mi, k]) > 1e-15]
    return {d: max(p[i] for p in sup) for i, d in enumerate('xyz')}


def gen3(la, lb, lc, m_fixed=None):
    """One three-centre derivative kernel: every aux component per thread
    (``allm_ip``), or the component ``m_fixed`` alone (one kernel per
    component, each with the recurrences that component needs: a kernel
    that switched on m kept every component's Rys factors live and spilled
    in the larger classes)."""
    name = f'int3c_ip_{la}{lb}{lc}' + ('' if m_fixed is None else f'_m{m_fixed}')
    lab = la + lb
    nroots = (lab + 1 + lc) // 2 + 1
    ca, cb, cc = cart_powers(la), cart_powers(lb), cart_powers(lc)
    nca, ncb = len(ca), len(cb)
    Ca, Cb, Cc = cart2sph(la), cart2sph(lb), cart2sph(lc)
    nsa, nsb = 2 * la + 1, 2 * lb + 1
    L = []
    w = L.append
    w(f'__global__ __launch_bounds__({THREADS}) void {name}(const BasisDev ao, const BasisDev aux, '
      'const int *__restrict__ pairs, int np, const int *__restrict__ kl, int nk, '
      'const double *__restrict__ W, long ldw, const int *__restrict__ cmap, '
      'const int *__restrict__ ao_atom, const int *__restrict__ aux_atom, int natm, int use_lds, '
      'int kch, double *grad, double *grad_aux) {')
    w('    const int nslot = 6;')
    _head(w)
    am = allm_ip(la, lb, lc)
    assert am == (m_fixed is None)
    ms = list(range(2 * lc + 1)) if am else [None]
    w('    const int nkc = (nk + kch - 1) / kch;')
    w(f'    const long tid = (long)blockIdx.x * {THREADS} + threadIdx.x;')
    w('    if (tid < (long)np * nkc) {')
    w('    const int ip = (int)(tid % np), kc = (int)(tid / np);')
    if not am:
        w(f'    const int m = {m_fixed};')
    w('    const int ish = pairs[2 * ip], jsh = pairs[2 * ip + 1];')
    w('    const double Ax = ao.xyz[3 * ish], Ay = ao.xyz[3 * ish + 1], Az = ao.xyz[3 * ish + 2];')
    w('    const double Bx = ao.xyz[3 * jsh], By = ao.xyz[3 * jsh + 1], Bz = ao.xyz[3 * jsh + 2];')
    w('    const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;')
    w('    const double rab2 = ABx * ABx + ABy * ABy + ABz * ABz;')
    w('    const long i0 = ao.aooff[ish], j0 = ao.aooff[jsh];')
    w('    const double fac = ish == jsh ? 1.0 : 2.0;')
    # the weight columns of the pair's function pairs (the same for every aux row)
    for s in range(nsa):
        for t in range(nsb):
            w(f'    long col_{s}_{t};')
            w(f'    {{ const long i = i0 + {s}, j = j0 + {t}; const long hi = i > j ? i : j, lo = i > j ? j : i; '
              f'const long pp = hi * (hi + 1) / 2 + lo; col_{s}_{t} = cmap ? cmap[pp] : pp; }}')
    w('    double gAx = 0., gAy = 0., gAz = 0., gBx = 0., gBy = 0., gBz = 0.;')
    w('    double cX = 0., cY = 0., cZ = 0.;')
    w('    int curC = -1;')
    w('    const int npa = ao.nprim[ish], oa = ao.poff[ish];')
    w('    const int npb = ao.nprim[jsh], ob = ao.poff[jsh];')
    w('    const int k1 = min(nk, (kc + 1) * kch);')
    w('    for (int kq = kc * kch; kq < k1; kq++) {')
    w('    const int ksh = kl[kq];')
    w('    const double Cx = aux.xyz[3 * ksh], Cy = aux.xyz[3 * ksh + 1], Cz = aux.xyz[3 * ksh + 2];')
    for mi in ms:
        sfx = '' if mi is None else f'{mi}_'
        mexpr = 'm' if mi is None else str(mi)
        w(f'    const double *wrow{sfx.rstrip("_")} = W + (long)(aux.aooff[ksh] + {mexpr}) * ldw;')
        for a in range(nca):
            for b in range(ncb):
                terms = []
                for s_ in range(nsa):
                    if abs(Ca[s_, a]) < 1e-15:
                        continue
                    for t in range(nsb):
                        if abs(Cb[t, b]) < 1e-15:
                            continue
                        terms.append(f'{fmt(Ca[s_, a] * Cb[t, b])} * wrow{sfx.rstrip("_")}[col_{s_}_{t}]')
                w(f'    const double wc_{sfx}{a}_{b} = fac * (' + (' + '.join(terms) if terms else '0.0') + ');')
    w('    double hAx = 0., hAy = 0., hAz = 0., hBx = 0., hBy = 0., hBz = 0.;')
    w('    const int npc = aux.nprim[ksh], oc = aux.poff[ksh];')
    w('    for (int ia = 0; ia < npa; ia++) {')
    w('        const double ea = ao.exps[oa + ia], da = ao.coefs[oa + ia];')
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
    # the roots one at a time in the classes whose terms keep every Rys factor
    # live across the per-component switch (unrolled, they spilled)
    w('#pragma unroll' if am else '#pragma unroll 1')
    w(f'                for (int r = 0; r < {nroots}; r++) {{')
    w('                    const double u = ru[r];')
    w('                    const double B00 = 0.5 * u * ipq;')
    w('                    const double B10 = 0.5 / p - 0.5 * q * u * ipq / p;')
    if lc >= 2:
        w('                    const double B01 = 0.5 / q - 0.5 * p * u * ipq / q;')
    w('                    const double qf = q * u * ipq, pf = p * u * ipq;')
    for d in 'xyz':
        w(f'                    const double C00{d} = PA{d} - qf * PC{d}, C0p{d} = pf * PC{d};')
    ind = '                    '
    _vrr_hrr(w, lab + 1, lc if am else _component_limits(lc, m_fixed), lb, ind)
    def products(mi, wname, indent):
        sup = [(k, Cc[mi, k]) for k in range(len(cc)) if abs(Cc[mi, k]) > 1e-15]
        for a_i, (ax, ay, az) in enumerate(ca):
            for b_i, (bx, by, bz) in enumerate(cb):
                for k, coef in sup:
                    kx, ky, kz = cc[k]
                    tw = f'{wname}{a_i}_{b_i} * {fmt(coef)}'
                    Ix, Iy, Iz = f'Ix_{ax}_{bx}_{kx}', f'Iy_{ay}_{by}_{ky}', f'Iz_{az}_{bz}_{kz}'
                    w(indent + f'{{ const double tw = {tw};')
                    w(indent + f'  hAx += tw * {_deriv("x", ax, bx, kx, "A")} * {Iy} * {Iz};')
                    w(indent + f'  hAy += tw * {Ix} * {_deriv("y", ay, by, ky, "A")} * {Iz};')
                    w(indent + f'  hAz += tw * {Ix} * {Iy} * {_deriv("z", az, bz, kz, "A")};')
                    w(indent + f'  hBx += tw * {_deriv("x", ax, bx, kx, "B")} * {Iy} * {Iz};')
                    w(indent + f'  hBy += tw * {Ix} * {_deriv("y", ay, by, ky, "B")} * {Iz};')
                    w(indent + f'  hBz += tw * {Ix} * {Iy} * {_deriv("z", az, bz, kz, "B")}; }}')
    if am:
        for mi in range(2 * lc + 1):
            products(mi, f'wc_{mi}_', ind)
    else:
        products(m_fixed, 'wc_', ind)
    w('                }')
    w('            }')
    w('        }')
    w('    }')
    w('    gAx += hAx; gAy += hAy; gAz += hAz; gBx += hBx; gBy += hBy; gBz += hBz;')
    w('    const int aC = aux_atom[ksh];')
    w('    if (aC != curC) {')
    w('        if (curC >= 0) ip_add(sg, use_lds, natm, grad_aux, curC, 1, cX, cY, cZ);')
    w('        curC = aC; cX = 0.; cY = 0.; cZ = 0.;')
    w('    }')
    w('    cX -= hAx + hBx; cY -= hAy + hBy; cZ -= hAz + hBz;')
    w('    }')          # kq
    w('    ip_add(sg, use_lds, natm, grad, ao_atom[ish], 0, gAx, gAy, gAz);')
    w('    ip_add(sg, use_lds, natm, grad, ao_atom[jsh], 0, gBx, gBy, gBz);')
    w('    if (curC >= 0) ip_add(sg, use_lds, natm, grad_aux, curC, 1, cX, cY, cZ);')
    w('    }')          # tid
    _tail(w)
    w('}')
    return name, '\n'.join(L)


def gen2(la, lc):
    name = f'int2c_ip_{la}{lc}'
    nroots = (la + 1 + lc) // 2 + 1
    ca, cc = cart_powers(la), cart_powers(lc)
    Ca, Cc = cart2sph(la), cart2sph(lc)
    nsa = 2 * la + 1
    L = []
    w = L.append
    w(f'__global__ __launch_bounds__({THREADS}) void {name}(const BasisDev aux, '
      'const int *__restrict__ pl, int np, const int *__restrict__ ql, int nq, '
      'const double *__restrict__ Z, long ldz, const int *__restrict__ aux_atom, int natm, int use_lds, '
      'double *grad) {')
    w('    const int nslot = 3;')
    w('    double *grad_aux = grad;')
    _head(w)
    w(f'    const int nqc = (nq + {QCH - 1}) / {QCH};')
    w(f'    const long tid = (long)blockIdx.x * {THREADS} + threadIdx.x;')
    w(f'    if (tid < (long)np * nqc * {2 * lc + 1}) {{')
    w('    const int ip = (int)(tid % np);')
    w('    const long rest = tid / np;')
    w('    const int qc = (int)(rest % nqc), m = (int)(rest / nqc);')
    w('    const int ish = pl[ip];')
    w('    const double Ax = aux.xyz[3 * ish], Ay = aux.xyz[3 * ish + 1], Az = aux.xyz[3 * ish + 2];')
    w('    const long r0 = aux.aooff[ish];')
    w('    const int npa = aux.nprim[ish], oa = aux.poff[ish];')
    w('    double gx = 0., gy = 0., gz = 0., cX = 0., cY = 0., cZ = 0.;')
    w('    int curQ = -1;')
    w(f'    const int q1 = min(nq, (qc + 1) * {QCH});')
    w(f'    for (int kq = qc * {QCH}; kq < q1; kq++) {{')
    w('    const int ksh = ql[kq];')
    w('    const double Cx = aux.xyz[3 * ksh], Cy = aux.xyz[3 * ksh + 1], Cz = aux.xyz[3 * ksh + 2];')
    w('    const long col = aux.aooff[ksh] + m;')
    for a in range(len(ca)):
        terms = [f'{fmt(Ca[s, a])} * Z[(r0 + {s}) * ldz + col]' for s in range(nsa) if abs(Ca[s, a]) > 1e-15]
        w(f'    const double wc_{a} = ' + (' + '.join(terms) if terms else '0.0') + ';')
    w('    double hx = 0., hy = 0., hz = 0.;')
    w('    const int npc = aux.nprim[ksh], oc = aux.poff[ksh];')
    w('    for (int ia = 0; ia < npa; ia++) {')
    w('        const double ea = aux.exps[oa + ia], fab = aux.coefs[oa + ia];')
    w('        const double p = ea;')
    w('        const double PAx = 0., PAy = 0., PAz = 0.;')
    w('        const double PCx = Ax - Cx, PCy = Ay - Cy, PCz = Az - Cz;')
    w('        const double rpc2 = PCx * PCx + PCy * PCy + PCz * PCz;')
    w('        for (int ic = 0; ic < npc; ic++) {')
    w('            const double q = aux.exps[oc + ic];')
    w('            const double pq = p + q, ipq = 1.0 / pq;')
    w('            const double x = p * q * ipq * rpc2;')
    w('            const double pref = TWO_PI_2P5 / (p * q * sqrt(pq)) * fab * aux.coefs[oc + ic];')
    w(f'            double ru[{nroots}], rw[{nroots}];')
    w(f'            rys_roots<{nroots}>(x, ru, rw);')
    w('#pragma unroll')
    w(f'            for (int r = 0; r < {nroots}; r++) {{')
    w('                const double u = ru[r];')
    w('                const double B00 = 0.5 * u * ipq;')
    w('                const double B10 = 0.5 / p - 0.5 * q * u * ipq / p;')
    if lc >= 2:
        w('                const double B01 = 0.5 / q - 0.5 * p * u * ipq / q;')
    w('                const double qf = q * u * ipq, pf = p * u * ipq;')
    for d in 'xyz':
        w(f'                const double C00{d} = PA{d} - qf * PC{d}, C0p{d} = pf * PC{d};')
    ind = '                '
    _vrr_hrr(w, la + 1, lc, 0, ind)
    w(ind + 'switch (m) {')
    for mi in range(2 * lc + 1):
        w(ind + f'case {mi}: {{')
        sup = [(k, Cc[mi, k]) for k in range(len(cc)) if abs(Cc[mi, k]) > 1e-15]
        for a_i, (ax, ay, az) in enumerate(ca):
            for k, coef in sup:
                kx, ky, kz = cc[k]
                Ix, Iy, Iz = f'Ix_{ax}_0_{kx}', f'Iy_{ay}_0_{ky}', f'Iz_{az}_0_{kz}'
                w(ind + f'    {{ const double tw = wc_{a_i} * {fmt(coef)};')
                w(ind + f'      hx += tw * {_deriv("x", ax, 0, kx, "A")} * {Iy} * {Iz};')
                w(ind + f'      hy += tw * {Ix} * {_deriv("y", ay, 0, ky, "A")} * {Iz};')
                w(ind + f'      hz += tw * {Ix} * {Iy} * {_deriv("z", az, 0, kz, "A")}; }}')
        w(ind + '    break;')
        w(ind + '}')
    w(ind + '}')
    w('            }')
    w('        }')
    w('    }')
    w('    gx += hx; gy += hy; gz += hz;')
    w('    const int aQ = aux_atom[ksh];')
    w('    if (aQ != curQ) {')
    w('        if (curQ >= 0) ip_add(sg, use_lds, natm, grad, curQ, 0, cX, cY, cZ);')
    w('        curQ = aQ; cX = 0.; cY = 0.; cZ = 0.;')
    w('    }')
    w('    cX -= hx; cY -= hy; cZ -= hz;')
    w('    }')          # kq
    w('    ip_add(sg, use_lds, natm, grad, aux_atom[ish], 0, gx, gy, gz);')
    w('    if (curQ >= 0) ip_add(sg, use_lds, natm, grad, curQ, 0, cX, cY, cZ);')
    w('    }')          # tid
    _tail(w)
    w('}')
    return name, '\n'.join(L)


def main():
    parts = ['// Generated by electrolites_amd/codegen/gen_int_ip.py; do not edit.',
             '// Nuclear-derivative Rys kernels of the DF gradient, one per angular class.',
             '#include <cstdlib>', '#include "mdf_int.h"', '', 'namespace mdf {', '',
             '// one atom\'s (x, y, z) into slot part (0 orbital centres, 1 aux centres) of the',
             '// block\'s LDS sums, or straight into the gradient g without them',
             '__device__ __forceinline__ void ip_add(double *sg, int use_lds, int natm, double *g, int atom, '
             'int part, double x, double y, double z) {',
             '    double *d = use_lds ? sg + 3 * (part * natm + atom) : g + 3 * atom;',
             '    atomicAdd(d, x);',
             '    atomicAdd(d + 1, y);',
             '    atomicAdd(d + 2, z);',
             '}', '',
             '// aux shells one thread of a three-centre kernel loops over (MDF_IP_KCH, default '
             f'{KCH})',
             'static int ip_kch() {',
             '    static int v = 0;',
             '    if (!v) {',
             '        const char *e = getenv("MDF_IP_KCH");',
             f'        const int x = e ? atoi(e) : {KCH};',
             f'        v = x >= 1 && x <= 4096 ? x : {KCH};',
             '    }',
             '    return v;',
             '}', '']
    n3, n2 = {}, {}
    for la in range(LMAX_AO + 1):
        for lb in range(LMAX_AO + 1):
            for lc in range(LMAX_AUX + 1):
                if allm_ip(la, lb, lc):
                    n, src = gen3(la, lb, lc)
                    n3[(la, lb, lc)] = [n]
                    parts += [src, '']
                else:
                    n3[(la, lb, lc)] = []
                    for mi in range(2 * lc + 1):
                        n, src = gen3(la, lb, lc, mi)
                        n3[(la, lb, lc)].append(n)
                        parts += [src, '']
    for la in range(LMAX_AUX + 1):
        for lc in range(LMAX_AUX + 1):
            n, src = gen2(la, lc)
            n2[(la, lc)] = n
            parts += [src, '']
    parts.append('int launch_int3c_ip(int la, int lb, int lc, const BasisDev &ao, const BasisDev &aux, '
                 'const int *pairs, int np, const int *kl, int nk, const double *W, long ldw, const int *cmap, '
                 'const int *ao_atom, const int *aux_atom, int natm, double *grad, double *grad_aux, hipStream_t st) {')
    parts.append('    const int kch = ip_kch();')
    parts.append('    const long nkc = (nk + kch - 1) / kch;')
    parts.append(f'    const int use_lds = natm <= {LDS_MAX_ATOMS};')
    parts.append('    const size_t shm = use_lds ? 6 * (size_t)natm * sizeof(double) : 0;')
    parts.append('    if ((long)np * nkc == 0) return MDF_OK;')
    parts.append(f'    const dim3 b({THREADS});')
    parts.append(f'    const dim3 g((unsigned)(((long)np * nkc + {THREADS - 1}) / {THREADS}));')
    parts.append(f'    switch ((la * {LMAX_AO + 1} + lb) * {LMAX_AUX + 1} + lc) {{')
    for (la, lb, lc), names in n3.items():
        key = (la * (LMAX_AO + 1) + lb) * (LMAX_AUX + 1) + lc
        parts.append(f'    case {key}:')
        for n in names:
            parts.append(f'        hipLaunchKernelGGL({n}, g, b, shm, st, ao, aux, pairs, np, kl, nk, W, ldw, '
                         'cmap, ao_atom, aux_atom, natm, use_lds, kch, grad, grad_aux);')
        parts.append('        break;')
    parts.append('    default: set_error("int3c2e_ip: class (%d %d|%d) not generated", la, lb, lc); return MDF_EINVAL;')
    parts.append('    }')
    parts.append('    MDF_LAUNCHED("int3c_ip");')
    parts.append('    return MDF_OK;')
    parts.append('}')
    parts.append('')
    parts.append('int launch_int2c_ip(int la, int lc, const BasisDev &aux, const int *pl, int np, const int *ql, '
                 'int nq, const double *Z, long ldz, const int *aux_atom, int natm, double *grad, hipStream_t st) {')
    parts.append(f'    const long n = (long)np * ((nq + {QCH - 1}) / {QCH}) * (2 * lc + 1);')
    parts.append(f'    const int use_lds = natm <= {LDS_MAX_ATOMS};')
    parts.append('    const size_t shm = use_lds ? 3 * (size_t)natm * sizeof(double) : 0;')
    parts.append('    if (n == 0) return MDF_OK;')
    parts.append(f'    const dim3 g((unsigned)((n + {THREADS - 1}) / {THREADS})), b({THREADS});')
    parts.append(f'    switch (la * {LMAX_AUX + 1} + lc) {{')
    for (la, lc), n in n2.items():
        parts.append(f'    case {la * (LMAX_AUX + 1) + lc}: hipLaunchKernelGGL({n}, g, b, shm, st, aux, pl, np, ql, nq, '
                     'Z, ldz, aux_atom, natm, use_lds, grad); break;')
    parts.append('    default: set_error("int2c2e_ip: class (%d|%d) not generated", la, lc); return MDF_EINVAL;')
    parts.append('    }')
    parts.append('    MDF_LAUNCHED("int2c_ip");')
    parts.append('    return MDF_OK;')
    parts.append('}')
    parts.append('')
    parts.append('}  // namespace mdf')
    with open(OUT, 'w') as f:
        f.write('\n'.join(parts) + '\n')
    print('wrote', OUT, sum(len(v) for v in n3.values()), '3c and', len(n2), '2c derivative kernels')


if __name__ == '__main__':
    main()
