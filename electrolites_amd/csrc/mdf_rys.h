// Rys quadrature roots and weights on the device.
//
// For n roots at argument x: u_i = t_i^2 in [0, 1) and w_i with
// sum_i w_i u_i^k = F_k(x) (k < 2n).  Below RYS_XMAX[n] each root and weight
// is a degree-RYS_DEG Chebyshev series on the unit interval of x holding it;
// above it the large-x (Gauss-Hermite) limit u = U/x, w = W/sqrt(x).  Tables
// come from codegen/gen_rys_tables.py (60-digit Golub-Welsch).
#pragma once
#include "mdf_common.h"
#include "rys_tables.inc"

namespace mdf {

template <int N>
__device__ __forceinline__ void rys_roots(double x, double (&u)[N], double (&w)[N]) {
    static_assert(N >= 1 && N <= RYS_NMAX, "unsupported number of Rys roots");
    if (x >= (double)RYS_XMAX[N]) {
        const double ix = 1.0 / x, isx = sqrt(ix);
#pragma unroll
        for (int i = 0; i < N; i++) {
            u[i] = RYS_LX_U[N * (N - 1) / 2 + i] * ix;
            w[i] = RYS_LX_W[N * (N - 1) / 2 + i] * isx;
        }
        return;
    }
    const int it = (int)x;
    const double s = 2.0 * (x - (double)it) - 1.0, s2 = 2.0 * s;
    const double *c = RYS_TABLE + RYS_OFFSET[N] + (long)it * (2 * N * (RYS_DEG + 1));
#pragma unroll
    for (int f = 0; f < 2 * N; f++) {
        const double *cf = c + f * (RYS_DEG + 1);
        double b1 = 0., b2 = 0.;
#pragma unroll
        for (int j = RYS_DEG; j >= 1; j--) {
            const double t = s2 * b1 - b2 + cf[j];
            b2 = b1;
            b1 = t;
        }
        const double v = s * b1 - b2 + cf[0];
        if (f < N) u[f] = v;
        else w[f - N] = v;
    }
}

}  // namespace mdf
