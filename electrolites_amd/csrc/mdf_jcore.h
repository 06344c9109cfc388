// Tile bodies of the Coulomb (J) contraction, shared by the single-item
// launches (mdf_j.hip) and the grouped persistent launch (mdf_grouped.hip),
// so an item gives bit-identical results alone and in a batch.
//
//     rho[n, P] = sum_p L[P, p] d[n, p]          (pass 1, column chunk x row block)
//     vj[n, p]  = sum_P rho[n, P] L[P, p]        (pass 2, column block x row group)
#pragma once
#include "mdf_common.h"

namespace mdf {
namespace jcore {

constexpr int J_THREADS = 256;
constexpr int J_LANE = 8;     // cderi values per lane per row (pass 1)
constexpr int J_RB = 32;      // rows per pass-1 tile

// The cderi is stored in S = double or float (the FP32 storage mode: every
// value is widened to FP64 before arithmetic).  A lane loads VEC values at a
// time (16 bytes when VEC * sizeof(S) == 16) and J_U = J_LANE / VEC loads per
// row, so a chunk is 4 * 64 * J_LANE = 2048 columns in either type.
template <typename S, int VEC>
struct vec_t;
template <>
struct vec_t<double, 1> { using T = double; };
template <>
struct vec_t<double, 2> { using T = double2; };
template <>
struct vec_t<float, 1> { using T = float; };
template <>
struct vec_t<float, 4> { using T = float4; };

template <typename S, int VEC>
__device__ __forceinline__ double lane_elem(const typename vec_t<S, VEC>::T &x, int e) {
    return (double)reinterpret_cast<const S *>(&x)[e];
}

constexpr long chunk_cols() { return 4L * 64 * J_LANE; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Pass-1 tile (chunk, rblock): the density slice of a chunk_cols() column
// chunk is held in registers and reused for J_RB rows; each row's partial dot
// product is reduced across the wavefront and the four wavefronts.
// part[(chunk * ndm + n) * naux + r]; ndm <= NDM densities are live.
template <int NDM, int VEC, typename S = double>
__device__ __forceinline__ void rho_partial_tile(const S *__restrict__ L, long ld, long naux, long npairs,
                                        const double *__restrict__ d, long ldd, int ndm,
                                        double *__restrict__ part, long chunk, long rblock) {
    constexpr int J_U = J_LANE / VEC;
    constexpr int WCOLS = 64 * VEC * J_U;
    constexpr int GCOLS = 4 * WCOLS;
    using V = typename vec_t<S, VEC>::T;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long cbase = chunk * GCOLS + wave * WCOLS + lane * VEC;
    const bool full = (chunk + 1) * GCOLS <= npairs;

    double dv[NDM][J_U * VEC];
#pragma unroll
    for (int n = 0; n < NDM; n++)
#pragma unroll
        for (int u = 0; u < J_U; u++)
#pragma unroll
            for (int v = 0; v < VEC; v++) {
                const long c = cbase + u * 64 * VEC + v;
                dv[n][u * VEC + v] = (n < ndm && c < npairs) ? d[n * ldd + c] : 0.;
            }

    __shared__ double red[4][NDM][J_RB];
    const long r0 = rblock * J_RB;
    const int nr = (int)min((long)J_RB, naux - r0);

    for (int rr = 0; rr < nr; rr++) {
        const S *row = L + (r0 + rr) * ld;
        double acc[NDM];
#pragma unroll
        for (int n = 0; n < NDM; n++) acc[n] = 0.;
        if (full) {
            V x[J_U];
#pragma unroll
            for (int u = 0; u < J_U; u++) x[u] = *reinterpret_cast<const V *>(row + cbase + u * 64 * VEC);
#pragma unroll
            for (int u = 0; u < J_U; u++) {
                if constexpr (VEC == 2 && sizeof(S) == 8) {
#pragma unroll
                    for (int n = 0; n < NDM; n++)
                        acc[n] += x[u].x * dv[n][2 * u] + x[u].y * dv[n][2 * u + 1];
                } else {
#pragma unroll
                    for (int v = 0; v < VEC; v++) {
                        const double xv = lane_elem<S, VEC>(x[u], v);
#pragma unroll
                        for (int n = 0; n < NDM; n++) acc[n] += xv * dv[n][u * VEC + v];
                    }
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < J_U; u++)
#pragma unroll
                for (int v = 0; v < VEC; v++) {
                    const long c = cbase + u * 64 * VEC + v;
                    const double xv = c < npairs ? (double)row[c] : 0.;
#pragma unroll
                    for (int n = 0; n < NDM; n++) acc[n] += xv * dv[n][u * VEC + v];
                }
        }
#pragma unroll
        for (int n = 0; n < NDM; n++) {
            const double s = wave_sum(acc[n]);
            if (lane == 0) red[wave][n][rr] = s;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ndm * nr; t += J_THREADS) {
        const int n = t / nr, rr = t % nr;
        const double s = red[0][n][rr] + red[1][n][rr] + red[2][n][rr] + red[3][n][rr];
        part[(chunk * ndm + n) * naux + r0 + rr] = s;
    }
    __syncthreads();
}

// rho[n, r] (+)= sum_c part[c, n, r] for the flat element range [e0, e1).
__device__ __forceinline__ void rho_reduce_range(const double *__restrict__ part, long nchunk, int ndm, long naux,
                                        double *__restrict__ rho, long ldrho, long e0, long e1,
                                        int accumulate = 0) {
    for (long t = e0 + threadIdx.x; t < e1; t += J_THREADS) {
        const int n = (int)(t / naux);
        const long r = t % naux;
        double s = accumulate ? rho[n * ldrho + r] : 0.;
        for (long c = 0; c < nchunk; c++) s += part[(c * ndm + n) * naux + r];
        rho[n * ldrho + r] = s;
    }
}

// Pass-2 tile (cblock, rows [r0, r1)): a thread owns VEC consecutive columns.
// out[n * ldo + c] for n < ndm.
template <int NDM, int VEC, typename S = double>
__device__ __forceinline__ void vj_pass_tile(const S *__restrict__ L, long ld, long npairs,
                                    const double *__restrict__ rho, long ldrho, int ndm, long r0, long r1,
                                    double *__restrict__ out, long ldo, long cblock) {
    using V = typename vec_t<S, VEC>::T;
    const long c = (cblock * J_THREADS + threadIdx.x) * VEC;
    double acc[NDM][VEC];
#pragma unroll
    for (int n = 0; n < NDM; n++)
#pragma unroll
        for (int v = 0; v < VEC; v++) acc[n][v] = 0.;
    if (c + VEC <= npairs) {
        const S *p = L + r0 * ld + c;
        long r = r0;
        for (; r + 8 <= r1; r += 8) {
            V x[8];
#pragma unroll
            for (int k = 0; k < 8; k++) x[k] = *reinterpret_cast<const V *>(p + k * ld);
            p += 8 * ld;
#pragma unroll
            for (int k = 0; k < 8; k++)
#pragma unroll
                for (int n = 0; n < NDM; n++) {
                    const double w = n < ndm ? rho[n * ldrho + r + k] : 0.;
#pragma unroll
                    for (int v = 0; v < VEC; v++) acc[n][v] += w * lane_elem<S, VEC>(x[k], v);
                }
        }
        for (; r < r1; r++) {
            const V x = *reinterpret_cast<const V *>(p);
            p += ld;
#pragma unroll
            for (int n = 0; n < NDM; n++) {
                const double w = n < ndm ? rho[n * ldrho + r] : 0.;
#pragma unroll
                for (int v = 0; v < VEC; v++) acc[n][v] += w * lane_elem<S, VEC>(x, v);
            }
        }
    } else if (c < npairs) {
        for (long r = r0; r < r1; r++)
#pragma unroll
            for (int v = 0; v < VEC; v++)
                if (c + v < npairs) {
                    const double x = (double)L[r * ld + c + v];
#pragma unroll
                    for (int n = 0; n < NDM; n++)
                        if (n < ndm) acc[n][v] += rho[n * ldrho + r] * x;
                }
    }
#pragma unroll
    for (int n = 0; n < NDM; n++)
#pragma unroll
        for (int v = 0; v < VEC; v++)
            if (n < ndm && c + v < npairs) out[n * ldo + c + v] = acc[n][v];
}

// vj[n, c] (+)= sum_g part[g, n, c] for the flat element range [e0, e1).
__device__ __forceinline__ void vj_reduce_range(const double *__restrict__ part, long ngroup, int ndm, long npairs,
                                       double *__restrict__ vj, long ldvj, int accumulate, long e0, long e1) {
    for (long t = e0 + threadIdx.x; t < e1; t += J_THREADS) {
        const int n = (int)(t / npairs);
        const long c = t % npairs;
        double s = accumulate ? vj[n * ldvj + c] : 0.;
        for (long g = 0; g < ngroup; g++) s += part[(g * ndm + n) * npairs + c];
        vj[n * ldvj + c] = s;
    }
}

// Decomposition of one item (a function of its shape only, so an item is cut
// the same way alone and in a batch).
struct JPlan {
    long nchunk;          // pass-1 column chunks
    long nrblock;         // pass-1 row blocks
    long ngroup;          // pass-2 row groups
    long rows_per_group;
    long ncblock;         // pass-2 column blocks
    size_t part1_words, part2_words;
};

// vec: cderi values per thread in pass 2 (1, 2 for double; 1, 4 for float)
inline JPlan plan_j(long naux, long npairs, int ndm, int vec) {
    JPlan p;
    const long gc = chunk_cols();
    p.nchunk = (npairs + gc - 1) / gc;
    p.nrblock = (naux + J_RB - 1) / J_RB;
    p.ncblock = (npairs + (long)J_THREADS * vec - 1) / ((long)J_THREADS * vec);
    long ng = 1;
    while (p.ncblock * ng < 4096 && naux / (ng * 2) >= 64) ng *= 2;
    p.ngroup = ng;
    p.rows_per_group = (naux + ng - 1) / ng;
    p.part1_words = (size_t)p.nchunk * ndm * naux;
    p.part2_words = (size_t)ng * ndm * npairs;
    return p;
}

}  // namespace jcore
}  // namespace mdf
