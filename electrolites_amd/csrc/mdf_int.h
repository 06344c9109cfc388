// Shared declarations of the Rys integral kernels (hand-written entry points
// in mdf_int.hip, generated kernels in mdf_int_gen.hip).
#pragma once
#include "mdf_common.h"
#include "mdf_rys.h"

namespace mdf {

constexpr double TWO_PI_2P5 = 34.986836655249725;   // 2 pi^(5/2)
constexpr double TWO_PI = 6.283185307179586;        // 2 pi
constexpr double PI_1P5 = 5.568327996831708;        // pi^(3/2)
constexpr double PRIM_SCREEN = 1e-20;               // |c_a c_b exp(-ab/p R^2)| cutoff
constexpr int LMAX_AO = 2;
constexpr int LMAX_AUX = 4;

// Device view of one basis (all pointers are device memory).
struct BasisDev {
    const int *nprim, *poff, *aooff;
    const double *xyz, *exps, *coefs;
};

int launch_int3c(int la, int lb, int lc, const BasisDev &ao, const BasisDev &aux, const int *pairs, int np,
                 const int *kl, int nk, double *out, long ldo, long p0, const int *cmap, hipStream_t st);
int launch_int1e(int la, int lb, const BasisDev &ao, const int *pairs, int np, int natm, const double *charges,
                 const double *atom_xyz, double *S, double *T, double *V, int nao, hipStream_t st);
int launch_int2c(int la, int lc, const BasisDev &aux, const int *pl, int np, const int *ql, int nq,
                 double *out, long ldo, hipStream_t st);
int launch_int3c_ip(int la, int lb, int lc, const BasisDev &ao, const BasisDev &aux, const int *pairs, int np,
                    const int *kl, int nk, const double *W, long ldw, const int *cmap, const int *ao_atom,
                    const int *aux_atom, int natm, double *grad, double *grad_aux, hipStream_t st);
int launch_int2c_ip(int la, int lc, const BasisDev &aux, const int *pl, int np, const int *ql, int nq,
                    const double *Z, long ldz, const int *aux_atom, int natm, double *grad, hipStream_t st);

}  // namespace mdf
