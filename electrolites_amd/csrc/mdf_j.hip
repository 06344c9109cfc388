// Coulomb (J) contraction over the packed cderi, FP64, HBM-bound.
//
// Restates electrolites/fastdf.py:348-386 (_vj_pass) and :688-745 (_get_j):
//     rho[n, P] = sum_p L[P, p] d[n, p]          (pass 1)
//     vj[n, p]  = sum_P rho[n, P] L[P, p]        (pass 2)
// L is the cderi, (naux, npairs) row-major with row stride ld (pair index
// contiguous, so both passes stream it with coalesced 16-byte loads).
//
// Pass 1: a workgroup owns a 2048-column chunk and a block of RB rows; the
// density slice of the chunk is held in registers and reused for every row,
// each row's partial dot product is reduced across the wavefront and the four
// wavefronts, and the per-chunk partials are summed by a second small kernel
// in a fixed order (deterministic, no atomics).
// Pass 2: a thread owns VEC consecutive columns and walks the rows of its
// row group with rho broadcast from the scalar cache; row groups (for
// parallelism at small npairs) are summed by a fixed-order reduction.
#include "mdf_common.h"
#include "../../include/mi355df.h"

namespace {

constexpr int J_THREADS = 256;
constexpr int J_U = 4;        // 16-byte loads per lane per row (pass 1)
constexpr int J_RB = 32;      // rows per workgroup (pass 1)

template <int VEC>
struct vec_t;
template <>
struct vec_t<1> { using T = double; };
template <>
struct vec_t<2> { using T = double2; };

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int NDM, int VEC>
__global__ __launch_bounds__(J_THREADS) void j_rho_partial(
    const double *__restrict__ L, long ld, long naux, long npairs,
    const double *__restrict__ d, long ldd, double *__restrict__ part) {
    constexpr int WCOLS = 64 * VEC * J_U;
    constexpr int GCOLS = 4 * WCOLS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long cbase = (long)blockIdx.x * GCOLS + wave * WCOLS + lane * VEC;
    const bool full = (long)(blockIdx.x + 1) * GCOLS <= npairs;

    double dv[NDM][J_U * VEC];
#pragma unroll
    for (int n = 0; n < NDM; n++)
#pragma unroll
        for (int u = 0; u < J_U; u++)
#pragma unroll
            for (int v = 0; v < VEC; v++) {
                long c = cbase + u * 64 * VEC + v;
                dv[n][u * VEC + v] = c < npairs ? d[n * ldd + c] : 0.;
            }

    __shared__ double red[4][NDM][J_RB];
    const long r0 = (long)blockIdx.y * J_RB;
    const int nr = (int)min((long)J_RB, naux - r0);

    for (int rr = 0; rr < nr; rr++) {
        const double *row = L + (r0 + rr) * ld;
        double acc[NDM];
#pragma unroll
        for (int n = 0; n < NDM; n++) acc[n] = 0.;
        if (full) {
            typename vec_t<VEC>::T x[J_U];
#pragma unroll
            for (int u = 0; u < J_U; u++)
                x[u] = *reinterpret_cast<const typename vec_t<VEC>::T *>(row + cbase + u * 64 * VEC);
#pragma unroll
            for (int u = 0; u < J_U; u++) {
                if constexpr (VEC == 2) {
#pragma unroll
                    for (int n = 0; n < NDM; n++)
                        acc[n] += x[u].x * dv[n][2 * u] + x[u].y * dv[n][2 * u + 1];
                } else {
#pragma unroll
                    for (int n = 0; n < NDM; n++) acc[n] += x[u] * dv[n][u];
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < J_U; u++)
#pragma unroll
                for (int v = 0; v < VEC; v++) {
                    long c = cbase + u * 64 * VEC + v;
                    double xv = c < npairs ? row[c] : 0.;
#pragma unroll
                    for (int n = 0; n < NDM; n++) acc[n] += xv * dv[n][u * VEC + v];
                }
        }
#pragma unroll
        for (int n = 0; n < NDM; n++) {
            double s = wave_sum(acc[n]);
            if (lane == 0) red[wave][n][rr] = s;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NDM * nr; t += J_THREADS) {
        int n = t / nr, rr = t % nr;
        double s = red[0][n][rr] + red[1][n][rr] + red[2][n][rr] + red[3][n][rr];
        part[((long)blockIdx.x * NDM + n) * naux + r0 + rr] = s;
    }
}

// rho[n, r] = sum_c part[c, n, r]
__global__ __launch_bounds__(J_THREADS) void j_rho_reduce(
    const double *__restrict__ part, long nchunk, int ndm, long naux,
    double *__restrict__ rho, long ldrho) {
    long t = (long)blockIdx.x * J_THREADS + threadIdx.x;
    if (t >= (long)ndm * naux) return;
    int n = (int)(t / naux);
    long r = t % naux;
    double s = 0.;
    for (long c = 0; c < nchunk; c++) s += part[(c * ndm + n) * naux + r];
    rho[n * ldrho + r] = s;
}

template <int NDM, int VEC>
__global__ __launch_bounds__(J_THREADS) void j_vj_pass(
    const double *__restrict__ L, long ld, long naux, long npairs,
    const double *__restrict__ rho, long ldrho, long rows_per_group,
    double *__restrict__ out, long ldo, long group_stride) {
    const long c = ((long)blockIdx.x * J_THREADS + threadIdx.x) * VEC;
    const long r0 = (long)blockIdx.y * rows_per_group;
    const long r1 = min(r0 + rows_per_group, naux);
    double acc[NDM][VEC];
#pragma unroll
    for (int n = 0; n < NDM; n++)
#pragma unroll
        for (int v = 0; v < VEC; v++) acc[n][v] = 0.;
    if (c + VEC <= npairs) {
        const double *p = L + r0 * ld + c;
        long r = r0;
        for (; r + 8 <= r1; r += 8) {
            typename vec_t<VEC>::T x[8];
#pragma unroll
            for (int k = 0; k < 8; k++)
                x[k] = *reinterpret_cast<const typename vec_t<VEC>::T *>(p + k * ld);
            p += 8 * ld;
#pragma unroll
            for (int k = 0; k < 8; k++)
#pragma unroll
                for (int n = 0; n < NDM; n++) {
                    double w = rho[n * ldrho + r + k];
                    if constexpr (VEC == 2) {
                        acc[n][0] += w * x[k].x;
                        acc[n][1] += w * x[k].y;
                    } else {
                        acc[n][0] += w * x[k];
                    }
                }
        }
        for (; r < r1; r++) {
            typename vec_t<VEC>::T x = *reinterpret_cast<const typename vec_t<VEC>::T *>(p);
            p += ld;
#pragma unroll
            for (int n = 0; n < NDM; n++) {
                double w = rho[n * ldrho + r];
                if constexpr (VEC == 2) {
                    acc[n][0] += w * x.x;
                    acc[n][1] += w * x.y;
                } else {
                    acc[n][0] += w * x;
                }
            }
        }
    } else if (c < npairs) {
        for (long r = r0; r < r1; r++)
#pragma unroll
            for (int v = 0; v < VEC; v++)
                if (c + v < npairs) {
                    double x = L[r * ld + c + v];
#pragma unroll
                    for (int n = 0; n < NDM; n++) acc[n][v] += rho[n * ldrho + r] * x;
                }
    }
    double *o = out + (long)blockIdx.y * group_stride;
#pragma unroll
    for (int n = 0; n < NDM; n++)
#pragma unroll
        for (int v = 0; v < VEC; v++)
            if (c + v < npairs) o[n * ldo + c + v] = acc[n][v];
}

// vj[n, c] (+)= sum_g part[g, n, c]
__global__ __launch_bounds__(J_THREADS) void j_vj_reduce(
    const double *__restrict__ part, long ngroup, int ndm, long npairs,
    double *__restrict__ vj, long ldvj, int accumulate) {
    long t = (long)blockIdx.x * J_THREADS + threadIdx.x;
    if (t >= (long)ndm * npairs) return;
    int n = (int)(t / npairs);
    long c = t % npairs;
    double s = accumulate ? vj[n * ldvj + c] : 0.;
    for (long g = 0; g < ngroup; g++) s += part[(g * ndm + n) * npairs + c];
    vj[n * ldvj + c] = s;
}

struct JPlan {
    long nchunk;      // pass-1 column chunks
    long ngroup;      // pass-2 row groups
    long rows_per_group;
    size_t part1_words, part2_words;
};

JPlan plan_j(long naux, long npairs, int ndm_step, int vec) {
    JPlan p;
    const long gcols = 4L * 64 * vec * J_U;
    p.nchunk = mdf::cdiv(npairs, gcols);
    const long wg2 = mdf::cdiv(npairs, (long)J_THREADS * vec);
    // enough row groups to put ~4096 workgroups in flight, >= 64 rows each
    long ng = 1;
    while (wg2 * ng < 4096 && naux / (ng * 2) >= 64) ng *= 2;
    p.ngroup = ng;
    p.rows_per_group = mdf::cdiv(naux, ng);
    p.part1_words = (size_t)p.nchunk * ndm_step * naux;
    p.part2_words = (size_t)ng * ndm_step * npairs;
    return p;
}

template <int NDM, int VEC>
int run_j_group(const double *L, long ld, long naux, long npairs, const double *d, long ldd,
                double *rho, long ldrho, double *vj, long ldvj, int accumulate,
                double *work, const JPlan &pl, hipStream_t st) {
    double *part1 = work;
    double *part2 = work + mdf::align_up(pl.part1_words, 64);
    dim3 g1((unsigned)pl.nchunk, (unsigned)mdf::cdiv(naux, (long)J_RB));
    hipLaunchKernelGGL((j_rho_partial<NDM, VEC>), g1, dim3(J_THREADS), 0, st,
                       L, ld, naux, npairs, d, ldd, part1);
    MDF_LAUNCHED("j_rho_partial");
    long nt = (long)NDM * naux;
    hipLaunchKernelGGL(j_rho_reduce, dim3((unsigned)mdf::cdiv(nt, (long)J_THREADS)),
                       dim3(J_THREADS), 0, st, part1, pl.nchunk, NDM, naux, rho, ldrho);
    MDF_LAUNCHED("j_rho_reduce");
    dim3 g2((unsigned)mdf::cdiv(npairs, (long)J_THREADS * VEC), (unsigned)pl.ngroup);
    if (pl.ngroup == 1 && !accumulate) {
        hipLaunchKernelGGL((j_vj_pass<NDM, VEC>), g2, dim3(J_THREADS), 0, st,
                           L, ld, naux, npairs, rho, ldrho, pl.rows_per_group, vj, ldvj, 0L);
        MDF_LAUNCHED("j_vj_pass");
        return MDF_OK;
    }
    double *tgt = part2;
    hipLaunchKernelGGL((j_vj_pass<NDM, VEC>), g2, dim3(J_THREADS), 0, st,
                       L, ld, naux, npairs, rho, ldrho, pl.rows_per_group, tgt, npairs,
                       (long)NDM * npairs);
    MDF_LAUNCHED("j_vj_pass");
    nt = (long)NDM * npairs;
    hipLaunchKernelGGL(j_vj_reduce, dim3((unsigned)mdf::cdiv(nt, (long)J_THREADS)),
                       dim3(J_THREADS), 0, st, tgt, pl.ngroup, NDM, npairs, vj, ldvj, accumulate);
    MDF_LAUNCHED("j_vj_reduce");
    return MDF_OK;
}

template <int VEC>
int dispatch_j(int nd, const double *L, long ld, long naux, long npairs, const double *d, long ldd,
               double *rho, long ldrho, double *vj, long ldvj, int accumulate, double *work,
               const JPlan &pl, hipStream_t st) {
    switch (nd) {
    case 1: return run_j_group<1, VEC>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st);
    case 2: return run_j_group<2, VEC>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st);
    case 3: return run_j_group<3, VEC>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st);
    default: return run_j_group<4, VEC>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st);
    }
}

}  // namespace

static int j_vec(const double *L, long ld, const double *d, long ldd) {
    bool ok = (ld % 2 == 0) && (ldd % 2 == 0) && ((uintptr_t)L % 16 == 0) && ((uintptr_t)d % 16 == 0);
    return ok ? 2 : 1;
}

extern "C" size_t mdf_get_j_workspace(long naux, long npairs, int n_dm) {
    (void)n_dm;
    JPlan p1 = plan_j(naux, npairs, 4, 1), p2 = plan_j(naux, npairs, 4, 2);
    size_t w1 = mdf::align_up(p1.part1_words, 64) + p1.part2_words;
    size_t w2 = mdf::align_up(p2.part1_words, 64) + p2.part2_words;
    return (w1 > w2 ? w1 : w2) * sizeof(double);
}

extern "C" int mdf_get_j(const double *cderi, long naux, long npairs, long ld,
                         const double *dm_pairs, int n_dm, long ldd,
                         double *rho, long ldrho, double *vj_pairs, long ldvj, int accumulate,
                         void *work, size_t work_bytes, void *stream) {
    MDF_CHECK_ARG(naux >= 0 && npairs >= 0 && n_dm >= 1, "mdf_get_j: bad sizes naux=%ld npairs=%ld n_dm=%d", naux, npairs, n_dm);
    MDF_CHECK_ARG(ld >= npairs && ldd >= npairs && ldvj >= npairs && ldrho >= naux,
                  "mdf_get_j: leading dimensions too small");
    MDF_CHECK_ARG(cderi && dm_pairs && rho && vj_pairs, "mdf_get_j: null pointer");
    hipStream_t st = mdf::as_stream(stream);
    if (npairs == 0) return MDF_OK;
    if (naux == 0) {
        if (!accumulate)
            for (int n = 0; n < n_dm; n++)
                MDF_HIP(hipMemsetAsync(vj_pairs + n * ldvj, 0, npairs * sizeof(double), st));
        return MDF_OK;
    }
    size_t need = mdf_get_j_workspace(naux, npairs, n_dm);
    MDF_CHECK_ARG(work != nullptr, "mdf_get_j: workspace is null");
    if (work_bytes < need) {
        mdf::set_error("mdf_get_j: workspace %zu bytes < %zu required", work_bytes, need);
        return MDF_EWORKSPACE;
    }
    int vec = j_vec(cderi, ld, dm_pairs, ldd);
    for (int n0 = 0; n0 < n_dm; n0 += 4) {
        int nd = n_dm - n0 < 4 ? n_dm - n0 : 4;
        JPlan pl = plan_j(naux, npairs, nd, vec);
        int rc = vec == 2
            ? dispatch_j<2>(nd, cderi, ld, naux, npairs, dm_pairs + n0 * ldd, ldd, rho + n0 * ldrho, ldrho,
                            vj_pairs + n0 * ldvj, ldvj, accumulate, (double *)work, pl, st)
            : dispatch_j<1>(nd, cderi, ld, naux, npairs, dm_pairs + n0 * ldd, ldd, rho + n0 * ldrho, ldrho,
                            vj_pairs + n0 * ldvj, ldvj, accumulate, (double *)work, pl, st);
        if (rc) return rc;
    }
    return MDF_OK;
}
