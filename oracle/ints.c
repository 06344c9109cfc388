/*
 * CPU oracle for the Gaussian integrals of the density-fitted path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against; nothing in the shipped package links or calls it.
 *
 * Algorithm: McMurchie-Davidson (Hermite expansion + Boys function), an
 * independent route from the Rys quadrature the GPU kernels use.  It restates
 * what the reference obtains from GPU4PySCF/libcint:
 *   - int3c2e (ij|P): GPU4PySCF int3c2e evaluator, called at
 *     electrolites/fastdfcderi.py:604,667;
 *   - int2c2e (P|Q): electrolites/_dfj2c.py:234-262 (GPU4PySCF int2c2e);
 *   - overlap/kinetic/nuclear: PySCF int1e (hcore for the SCF driver).
 * Boys function: series + downward recursion, erf + upward recursion for
 * large T (the branch structure of electrolites/kernels/fastj_prologue.cu:91-145).
 * Conventions: libcint Cartesian order, PySCF real spherical harmonics,
 * primitive coefficients already normalised by the caller.
 *
 * Shell description (plain arrays, one entry per shell):
 *   l[s], nprim[s], off[s] (into exps/coefs), aooff[s] (spherical offset),
 *   xyz[3*s].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_LMAX 6
#define OR_LTOT (3 * OR_LMAX + 2)
#define NC(l) (((l) + 1) * ((l) + 2) / 2)

static const double PI = 3.14159265358979323846264338327950288;

/* ---------------------------------------------------------------- Boys */
void oracle_boys(int mmax, double t, double *F)
{
    int m;
    if (t < 1e-15) {
        for (m = 0; m <= mmax; m++)
            F[m] = 1. / (2 * m + 1) - t / (2 * m + 3);
        return;
    }
    if (t < 45.) {
        /* F_mmax(t) = e^-t sum_k (2t)^k / ((2mmax+1)(2mmax+3)...(2mmax+2k+1)) */
        double term = 1. / (2 * mmax + 1), sum = term;
        int k;
        for (k = 1; k < 400; k++) {
            term *= 2. * t / (2 * mmax + 2 * k + 1);
            sum += term;
            if (term < 1e-18 * sum)
                break;
        }
        double et = exp(-t);
        F[mmax] = et * sum;
        for (m = mmax - 1; m >= 0; m--)
            F[m] = (2. * t * F[m + 1] + et) / (2 * m + 1);
        return;
    }
    {
        double st = sqrt(t), et = exp(-t);
        F[0] = .5 * sqrt(PI / t) * erf(st);
        for (m = 0; m < mmax; m++)
            F[m + 1] = ((2 * m + 1) * F[m] - et) / (2. * t);
    }
}

/* ----------------------------------------------- spherical transforms */
static double fact(int n) { double r = 1; while (n > 1) r *= n--; return r; }
static double binom(int n, int k)
{
    if (k < 0 || k > n) return 0.;
    return fact(n) / (fact(k) * fact(n - k));
}

static int cart_index(int l, int lx, int ly)
{
    /* libcint order: lx descending, then ly descending */
    int idx = 0, x, y;
    for (x = l; x >= 0; x--)
        for (y = l - x; y >= 0; y--) {
            if (x == lx && y == ly) return idx;
            idx++;
        }
    return -1;
}

/* c[(2l+1) * ncart]: spherical = c . cartesian; Helgaker-Jorgensen-Olsen
 * eq. 6.4.47-50 rescaled to unit-normalised Y_lm; PySCF order (p: x,y,z). */
void oracle_cart2sph(int l, double *c)
{
    int nc = NC(l), i, t, u;
    memset(c, 0, sizeof(double) * (2 * l + 1) * nc);
    for (i = 0; i < 2 * l + 1; i++) {
        int m;
        if (l == 1) m = (i == 0) ? 1 : (i == 1 ? -1 : 0);
        else m = i - l;
        int am = abs(m);
        double vm = m >= 0 ? 0. : .5;
        double nlm = 1. / (pow(2., am) * fact(l)) *
                     sqrt(2. * fact(l + am) * fact(l - am) / (m == 0 ? 2. : 1.));
        for (t = 0; t <= (l - am) / 2; t++)
            for (u = 0; u <= t; u++) {
                double vmax = floor(am / 2. - vm) + vm, v;
                for (v = vm; v <= vmax + 1e-9; v += 1.) {
                    int two_v = (int)lround(2 * v);
                    int sgn = ((int)lround(t + v - vm)) % 2 ? -1 : 1;
                    double co = sgn * pow(.25, t) * binom(l, t) *
                                binom(l - t, am + t) * binom(t, u) * binom(am, two_v);
                    int lx = 2 * t + am - 2 * u - two_v;
                    int ly = 2 * u + two_v;
                    c[i * nc + cart_index(l, lx, ly)] += co * nlm;
                }
            }
        double scale = sqrt((2 * l + 1) / (4 * PI));
        int k;
        for (k = 0; k < nc; k++) c[i * nc + k] *= scale;
    }
}

/* ------------------------------------------------- Hermite expansions */
/* E[i][j][t], 1D, for i<=la, j<=lb (overlap distribution of two Gaussians) */
#define EIJ(E, i, j, t) E[((i) * (OR_LMAX + 3) + (j)) * (2 * OR_LMAX + 5) + (t)]
#define ESZ ((OR_LMAX + 1) * (OR_LMAX + 3) * (2 * OR_LMAX + 5))

static void hermite_e(int la, int lb, double a, double b, double xa, double xb,
                      double *E)
{
    double p = a + b, mu = a * b / p, xab = xa - xb;
    double xp = (a * xa + b * xb) / p, xpa = xp - xa, xpb = xp - xb;
    double oop = .5 / p;
    int i, j, t;
    memset(E, 0, sizeof(double) * ESZ);
    EIJ(E, 0, 0, 0) = exp(-mu * xab * xab);
    for (i = 0; i <= la; i++)
        for (j = 0; j <= lb; j++) {
            if (i == 0 && j == 0) continue;
            for (t = 0; t <= i + j; t++) {
                double v;
                if (i > 0) {
                    v = xpa * EIJ(E, i - 1, j, t);
                    if (t > 0) v += oop * EIJ(E, i - 1, j, t - 1);
                    if (t + 1 <= i - 1 + j) v += (t + 1) * EIJ(E, i - 1, j, t + 1);
                } else {
                    v = xpb * EIJ(E, i, j - 1, t);
                    if (t > 0) v += oop * EIJ(E, i, j - 1, t - 1);
                    if (t + 1 <= i + j - 1) v += (t + 1) * EIJ(E, i, j - 1, t + 1);
                }
                EIJ(E, i, j, t) = v;
            }
        }
}

/* R^0_{tuv}(alpha, X, Y, Z) for t+u+v <= L, returned in R0[t][u][v] */
#define RD (OR_LTOT + 1)
static void hermite_r(int L, double alpha, double X, double Y, double Z,
                      double *R0, double *work)
{
    /* work: (L+1) * RD^3 doubles, indexed [n][t][u][v] */
    double F[OR_LTOT + 2];
    double T = alpha * (X * X + Y * Y + Z * Z);
    int n, t, u, v, s;
    oracle_boys(L, T, F);
#define RW(n, t, u, v) work[(((n) * RD + (t)) * RD + (u)) * RD + (v)]
    double f = 1.;
    for (n = 0; n <= L; n++) {
        RW(n, 0, 0, 0) = f * F[n];
        f *= -2. * alpha;
    }
    for (s = 1; s <= L; s++)
        for (t = 0; t <= s; t++)
            for (u = 0; u <= s - t; u++) {
                v = s - t - u;
                for (n = 0; n <= L - s; n++) {
                    double val;
                    if (t > 0) {
                        val = X * RW(n + 1, t - 1, u, v);
                        if (t > 1) val += (t - 1) * RW(n + 1, t - 2, u, v);
                    } else if (u > 0) {
                        val = Y * RW(n + 1, t, u - 1, v);
                        if (u > 1) val += (u - 1) * RW(n + 1, t, u - 2, v);
                    } else {
                        val = Z * RW(n + 1, t, u, v - 1);
                        if (v > 1) val += (v - 1) * RW(n + 1, t, u, v - 2);
                    }
                    RW(n, t, u, v) = val;
                }
            }
    for (t = 0; t <= L; t++)
        for (u = 0; u <= L - t; u++)
            for (v = 0; v <= L - t - u; v++)
                R0[(t * RD + u) * RD + v] = RW(0, t, u, v);
#undef RW
}

/* ----------------------------------------- primitive Cartesian blocks */
/* (ab|c) over Cartesian monomials, out[ia][ib][ic] accumulated with scale */
static void prim3c(int la, const double *A, double a, int lb, const double *B,
                   double b, int lc, const double *C, double c, double scale,
                   double *out, double *Eab, double *Ec, double *R0, double *work)
{
    int L = la + lb + lc, d;
    double p = a + b, q = c;
    double P[3];
    for (d = 0; d < 3; d++) P[d] = (a * A[d] + b * B[d]) / p;
    double alpha = p * q / (p + q);
    double pref = 2. * pow(PI, 2.5) / (p * q * sqrt(p + q)) * scale;
    for (d = 0; d < 3; d++) {
        hermite_e(la, lb, a, b, A[d], B[d], Eab + d * ESZ);
        hermite_e(lc, 0, c, 0., C[d], C[d], Ec + d * ESZ);
    }
    hermite_r(L, alpha, P[0] - C[0], P[1] - C[1], P[2] - C[2], R0, work);
    int nca = NC(la), ncb = NC(lb), ncc = NC(lc);
    int ia, ib, ic, ax, ay, bx, by, cx, cy;
    ic = 0;
    for (cx = lc; cx >= 0; cx--)
        for (cy = lc - cx; cy >= 0; cy--, ic++) {
            int cz = lc - cx - cy;
            /* W[t][u][v] = sum_{tau,nu,phi} (-1)^(tau+nu+phi) Ec.. R[t+tau][u+nu][v+phi] */
            static __thread double W[RD * RD * RD];
            int lab = la + lb, t, u, v, ta, nu, ph;
            for (t = 0; t <= lab; t++)
                for (u = 0; u <= lab - t; u++)
                    for (v = 0; v <= lab - t - u; v++) {
                        double s = 0.;
                        for (ta = 0; ta <= cx; ta++) {
                            double ex = EIJ((Ec + 0 * ESZ), cx, 0, ta);
                            if (ex == 0.) continue;
                            for (nu = 0; nu <= cy; nu++) {
                                double ey = EIJ((Ec + 1 * ESZ), cy, 0, nu);
                                if (ey == 0.) continue;
                                for (ph = 0; ph <= cz; ph++) {
                                    double ez = EIJ((Ec + 2 * ESZ), cz, 0, ph);
                                    if (ez == 0.) continue;
                                    double sg = ((ta + nu + ph) & 1) ? -1. : 1.;
                                    s += sg * ex * ey * ez *
                                         R0[((t + ta) * RD + u + nu) * RD + v + ph];
                                }
                            }
                        }
                        W[(t * RD + u) * RD + v] = s;
                    }
            ia = 0;
            for (ax = la; ax >= 0; ax--)
                for (ay = la - ax; ay >= 0; ay--, ia++) {
                    int az = la - ax - ay;
                    ib = 0;
                    for (bx = lb; bx >= 0; bx--)
                        for (by = lb - bx; by >= 0; by--, ib++) {
                            int bz = lb - bx - by;
                            double s = 0.;
                            for (t = 0; t <= ax + bx; t++) {
                                double ex = EIJ((Eab + 0 * ESZ), ax, bx, t);
                                for (u = 0; u <= ay + by; u++) {
                                    double ey = EIJ((Eab + 1 * ESZ), ay, by, u);
                                    for (v = 0; v <= az + bz; v++) {
                                        double ez = EIJ((Eab + 2 * ESZ), az, bz, v);
                                        s += ex * ey * ez * W[(t * RD + u) * RD + v];
                                    }
                                }
                            }
                            out[(ia * ncb + ib) * ncc + ic] += pref * s;
                        }
                }
        }
    (void)L;
}

/* sph[i][j][k] = sum C_a[i][x] C_b[j][y] C_c[k][z] cart[x][y][z] */
static void cart_to_sph3(int la, int lb, int lc, const double *cart, double *sph)
{
    double ca[13 * 28], cb[13 * 28], cc[13 * 28];
    int nca = NC(la), ncb = NC(lb), ncc = NC(lc);
    int nsa = 2 * la + 1, nsb = 2 * lb + 1, nsc = 2 * lc + 1;
    int i, j, k, x, y, z;
    oracle_cart2sph(la, ca);
    oracle_cart2sph(lb, cb);
    oracle_cart2sph(lc, cc);
    double *t1 = malloc(sizeof(double) * nca * ncb * nsc);
    double *t2 = malloc(sizeof(double) * nca * nsb * nsc);
    for (x = 0; x < nca; x++)
        for (y = 0; y < ncb; y++)
            for (k = 0; k < nsc; k++) {
                double s = 0.;
                for (z = 0; z < ncc; z++) s += cc[k * ncc + z] * cart[(x * ncb + y) * ncc + z];
                t1[(x * ncb + y) * nsc + k] = s;
            }
    for (x = 0; x < nca; x++)
        for (j = 0; j < nsb; j++)
            for (k = 0; k < nsc; k++) {
                double s = 0.;
                for (y = 0; y < ncb; y++) s += cb[j * ncb + y] * t1[(x * ncb + y) * nsc + k];
                t2[(x * nsb + j) * nsc + k] = s;
            }
    for (i = 0; i < nsa; i++)
        for (j = 0; j < nsb; j++)
            for (k = 0; k < nsc; k++) {
                double s = 0.;
                for (x = 0; x < nca; x++) s += ca[i * nca + x] * t2[(x * nsb + j) * nsc + k];
                sph[(i * nsb + j) * nsc + k] = s;
            }
    free(t1);
    free(t2);
}

static int check_l(int nbas, const int *l)
{
    int s;
    for (s = 0; s < nbas; s++)
        if (l[s] < 0 || l[s] > OR_LMAX) return -1;
    return 0;
}

/* --------------------------------------------------------- int3c2e */
/* out[i][j][P], C order, i,j over nao (both triangles), P over naux */
int oracle_int3c2e(int nbas, const int *l, const int *nprim, const int *off,
                   const int *aooff, const double *xyz, const double *exps,
                   const double *coefs, int nao,
                   int nauxbas, const int *al, const int *anprim, const int *aoff,
                   const int *aaooff, const double *axyz, const double *aexps,
                   const double *acoefs, int naux, double *out)
{
    if (check_l(nbas, l) || check_l(nauxbas, al)) return -1;
    int ish;
#pragma omp parallel for schedule(dynamic)
    for (ish = 0; ish < nbas; ish++) {
        double *Eab = malloc(sizeof(double) * 3 * ESZ);
        double *Ec = malloc(sizeof(double) * 3 * ESZ);
        double *R0 = malloc(sizeof(double) * RD * RD * RD);
        double *work = malloc(sizeof(double) * (OR_LTOT + 1) * RD * RD * RD);
        double *cart = malloc(sizeof(double) * 28 * 28 * 28);
        double *sph = malloc(sizeof(double) * 13 * 13 * 13);
        int jsh, ksh;
        for (jsh = 0; jsh <= ish; jsh++)
            for (ksh = 0; ksh < nauxbas; ksh++) {
                int la = l[ish], lb = l[jsh], lc = al[ksh];
                int nca = NC(la), ncb = NC(lb), ncc = NC(lc);
                memset(cart, 0, sizeof(double) * nca * ncb * ncc);
                int ip, jp, kp;
                for (ip = 0; ip < nprim[ish]; ip++)
                    for (jp = 0; jp < nprim[jsh]; jp++)
                        for (kp = 0; kp < anprim[ksh]; kp++) {
                            double sc = coefs[off[ish] + ip] * coefs[off[jsh] + jp] *
                                        acoefs[aoff[ksh] + kp];
                            prim3c(la, xyz + 3 * ish, exps[off[ish] + ip],
                                   lb, xyz + 3 * jsh, exps[off[jsh] + jp],
                                   lc, axyz + 3 * ksh, aexps[aoff[ksh] + kp],
                                   sc, cart, Eab, Ec, R0, work);
                        }
                cart_to_sph3(la, lb, lc, cart, sph);
                int nsa = 2 * la + 1, nsb = 2 * lb + 1, nsc = 2 * lc + 1, i, j, k;
                for (i = 0; i < nsa; i++)
                    for (j = 0; j < nsb; j++)
                        for (k = 0; k < nsc; k++) {
                            double v = sph[(i * nsb + j) * nsc + k];
                            long I = aooff[ish] + i, J = aooff[jsh] + j, K = aaooff[ksh] + k;
                            out[(I * nao + J) * naux + K] = v;
                            out[(J * nao + I) * naux + K] = v;
                        }
            }
        free(Eab); free(Ec); free(R0); free(work); free(cart); free(sph);
    }
    return 0;
}

/* --------------------------------------------------------- int2c2e */
int oracle_int2c2e(int nauxbas, const int *al, const int *anprim, const int *aoff,
                   const int *aaooff, const double *axyz, const double *aexps,
                   const double *acoefs, int naux, double *out)
{
    if (check_l(nauxbas, al)) return -1;
    int ish;
#pragma omp parallel for schedule(dynamic)
    for (ish = 0; ish < nauxbas; ish++) {
        double *Eab = malloc(sizeof(double) * 3 * ESZ);
        double *Ec = malloc(sizeof(double) * 3 * ESZ);
        double *R0 = malloc(sizeof(double) * RD * RD * RD);
        double *work = malloc(sizeof(double) * (OR_LTOT + 1) * RD * RD * RD);
        double *cart = malloc(sizeof(double) * 28 * 28);
        double *sph = malloc(sizeof(double) * 13 * 13);
        int ksh;
        for (ksh = 0; ksh <= ish; ksh++) {
            int la = al[ish], lc = al[ksh];
            int nca = NC(la), ncc = NC(lc);
            memset(cart, 0, sizeof(double) * nca * ncc);
            int ip, kp;
            for (ip = 0; ip < anprim[ish]; ip++)
                for (kp = 0; kp < anprim[ksh]; kp++) {
                    double sc = acoefs[aoff[ish] + ip] * acoefs[aoff[ksh] + kp];
                    prim3c(la, axyz + 3 * ish, aexps[aoff[ish] + ip], 0,
                           axyz + 3 * ish, 0., lc, axyz + 3 * ksh,
                           aexps[aoff[ksh] + kp], sc, cart, Eab, Ec, R0, work);
                }
            cart_to_sph3(la, 0, lc, cart, sph);
            /* the middle index is a unit s "function" with exponent 0, not a
             * basis function: undo the Y_00 factor the transform applied */
            int i, k;
            for (i = 0; i < (2 * la + 1) * (2 * lc + 1); i++) sph[i] *= sqrt(4. * PI);
            for (i = 0; i < 2 * la + 1; i++)
                for (k = 0; k < 2 * lc + 1; k++) {
                    long I = aaooff[ish] + i, K = aaooff[ksh] + k;
                    out[I * naux + K] = sph[i * (2 * lc + 1) + k];
                    out[K * naux + I] = sph[i * (2 * lc + 1) + k];
                }
        }
        free(Eab); free(Ec); free(R0); free(work); free(cart); free(sph);
    }
    return 0;
}

/* ------------------------------------------------------------- int1e */
/* kind 0 overlap, 1 kinetic, 2 nuclear attraction (point charges) */
int oracle_int1e(int kind, int nbas, const int *l, const int *nprim, const int *off,
                 const int *aooff, const double *xyz, const double *exps,
                 const double *coefs, int nao, int natm, const double *charges,
                 const double *atom_xyz, double *out)
{
    if (check_l(nbas, l)) return -1;
    int ish;
#pragma omp parallel for schedule(dynamic)
    for (ish = 0; ish < nbas; ish++) {
        double *E = malloc(sizeof(double) * 3 * ESZ);
        double *R0 = malloc(sizeof(double) * RD * RD * RD);
        double *work = malloc(sizeof(double) * (OR_LTOT + 1) * RD * RD * RD);
        double cart[28 * 28], sph[13 * 13];
        int jsh;
        for (jsh = 0; jsh <= ish; jsh++) {
            int la = l[ish], lb = l[jsh];
            int nca = NC(la), ncb = NC(lb);
            memset(cart, 0, sizeof(cart));
            const double *A = xyz + 3 * ish, *B = xyz + 3 * jsh;
            int ip, jp, d;
            for (ip = 0; ip < nprim[ish]; ip++)
                for (jp = 0; jp < nprim[jsh]; jp++) {
                    double a = exps[off[ish] + ip], b = exps[off[jsh] + jp];
                    double sc = coefs[off[ish] + ip] * coefs[off[jsh] + jp];
                    double p = a + b;
                    for (d = 0; d < 3; d++)
                        hermite_e(la, lb + 2, a, b, A[d], B[d], E + d * ESZ);
                    int ia = 0, ax, ay, bx, by;
                    if (kind == 2) {
                        double P[3];
                        for (d = 0; d < 3; d++) P[d] = (a * A[d] + b * B[d]) / p;
                        int c;
                        for (c = 0; c < natm; c++) {
                            const double *Cc = atom_xyz + 3 * c;
                            hermite_r(la + lb, p, P[0] - Cc[0], P[1] - Cc[1],
                                      P[2] - Cc[2], R0, work);
                            double pref = -charges[c] * 2. * PI / p * sc;
                            ia = 0;
                            for (ax = la; ax >= 0; ax--)
                                for (ay = la - ax; ay >= 0; ay--, ia++) {
                                    int az = la - ax - ay, ib = 0;
                                    for (bx = lb; bx >= 0; bx--)
                                        for (by = lb - bx; by >= 0; by--, ib++) {
                                            int bz = lb - bx - by, t, u, v;
                                            double s = 0.;
                                            for (t = 0; t <= ax + bx; t++)
                                                for (u = 0; u <= ay + by; u++)
                                                    for (v = 0; v <= az + bz; v++)
                                                        s += EIJ((E + 0 * ESZ), ax, bx, t) *
                                                             EIJ((E + 1 * ESZ), ay, by, u) *
                                                             EIJ((E + 2 * ESZ), az, bz, v) *
                                                             R0[(t * RD + u) * RD + v];
                                            cart[ia * ncb + ib] += pref * s;
                                        }
                                }
                        }
                        continue;
                    }
                    double pp = pow(PI / p, 1.5) * sc;
                    ia = 0;
                    for (ax = la; ax >= 0; ax--)
                        for (ay = la - ax; ay >= 0; ay--, ia++) {
                            int az = la - ax - ay, ib = 0;
                            int ai[3] = {ax, ay, az};
                            for (bx = lb; bx >= 0; bx--)
                                for (by = lb - bx; by >= 0; by--, ib++) {
                                    int bz = lb - bx - by;
                                    int bi[3] = {bx, by, bz};
                                    double S[3], T[3];
                                    for (d = 0; d < 3; d++) {
                                        const double *Ed = E + d * ESZ;
                                        int i = ai[d], j = bi[d];
                                        S[d] = EIJ(Ed, i, j, 0);
                                        double t = -2. * b * (2 * j + 1) * EIJ(Ed, i, j, 0) +
                                                   4. * b * b * EIJ(Ed, i, j + 2, 0);
                                        if (j >= 2) t += j * (j - 1) * EIJ(Ed, i, j - 2, 0);
                                        T[d] = -.5 * t;
                                    }
                                    double v;
                                    if (kind == 0) v = S[0] * S[1] * S[2];
                                    else v = T[0] * S[1] * S[2] + S[0] * T[1] * S[2] +
                                             S[0] * S[1] * T[2];
                                    cart[ia * ncb + ib] += pp * v;
                                }
                        }
                }
            /* spherical */
            double ca[13 * 28], cb[13 * 28], tmp[28 * 13];
            oracle_cart2sph(la, ca);
            oracle_cart2sph(lb, cb);
            int nsa = 2 * la + 1, nsb = 2 * lb + 1, x, y, i, j;
            for (x = 0; x < nca; x++)
                for (j = 0; j < nsb; j++) {
                    double s = 0.;
                    for (y = 0; y < ncb; y++) s += cb[j * ncb + y] * cart[x * ncb + y];
                    tmp[x * nsb + j] = s;
                }
            for (i = 0; i < nsa; i++)
                for (j = 0; j < nsb; j++) {
                    double s = 0.;
                    for (x = 0; x < nca; x++) s += ca[i * nca + x] * tmp[x * nsb + j];
                    sph[i * nsb + j] = s;
                }
            for (i = 0; i < nsa; i++)
                for (j = 0; j < nsb; j++) {
                    long I = aooff[ish] + i, J = aooff[jsh] + j;
                    out[I * nao + J] = sph[i * nsb + j];
                    out[J * nao + I] = sph[i * nsb + j];
                }
        }
        free(E); free(R0); free(work);
    }
    return 0;
}
