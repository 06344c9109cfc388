// Coulomb (J) contraction over the packed cderi, FP64, HBM-bound.
//
// Restates electrolites/fastdf.py:348-386 (_vj_pass) and :688-745 (_get_j):
//     rho[n, P] = sum_p L[P, p] d[n, p]          (pass 1)
//     vj[n, p]  = sum_P rho[n, P] L[P, p]        (pass 2)
// L is the cderi, (naux, npairs) row-major with row stride ld (pair index
// contiguous, so both passes stream it with coalesced 16-byte loads).
//
// Pass 1: a workgroup owns a 2048-column chunk and a block of 32 rows; the
// density slice of the chunk is held in registers and reused for every row,
// each row's partial dot product is reduced across the wavefront and the four
// wavefronts, and the per-chunk partials are summed by a second small kernel
// in a fixed order (deterministic, no atomics).
// Pass 2: a thread owns two consecutive columns and walks the rows of its
// row group with rho broadcast from the scalar cache; row groups (for
// parallelism at small npairs) are summed by a fixed-order reduction.
// The tile bodies live in mdf_jcore.h (shared with the grouped launch).
#include "mdf_common.h"
#include "mdf_jcore.h"
#include "../../include/mi355df.h"

using namespace mdf;
using namespace mdf::jcore;

namespace {

template <int NDM, int VEC, typename S>
__global__ __launch_bounds__(J_THREADS) void j_rho_partial(const S *__restrict__ L, long ld, long naux,
                                                           long npairs, const double *__restrict__ d, long ldd,
                                                           double *__restrict__ part) {
    rho_partial_tile<NDM, VEC, S>(L, ld, naux, npairs, d, ldd, NDM, part, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(J_THREADS) void j_rho_reduce(const double *__restrict__ part, long nchunk, int ndm,
                                                          long naux, double *__restrict__ rho, long ldrho) {
    const long e0 = (long)blockIdx.x * J_THREADS;
    rho_reduce_range(part, nchunk, ndm, naux, rho, ldrho, e0, min(e0 + J_THREADS, (long)ndm * naux));
}

template <int NDM, int VEC, typename S>
__global__ __launch_bounds__(J_THREADS) void j_vj_pass(const S *__restrict__ L, long ld, long naux,
                                                       long npairs, const double *__restrict__ rho, long ldrho,
                                                       long rows_per_group, double *__restrict__ out, long ldo,
                                                       long group_stride) {
    const long r0 = (long)blockIdx.y * rows_per_group;
    const long r1 = min(r0 + rows_per_group, naux);
    vj_pass_tile<NDM, VEC, S>(L, ld, npairs, rho, ldrho, NDM, r0, r1, out + (long)blockIdx.y * group_stride, ldo,
                              blockIdx.x);
}

__global__ __launch_bounds__(J_THREADS) void j_vj_reduce(const double *__restrict__ part, long ngroup, int ndm,
                                                         long npairs, double *__restrict__ vj, long ldvj,
                                                         int accumulate) {
    const long e0 = (long)blockIdx.x * J_THREADS;
    vj_reduce_range(part, ngroup, ndm, npairs, vj, ldvj, accumulate, e0, min(e0 + J_THREADS, (long)ndm * npairs));
}

template <int NDM, int VEC, typename S>
int run_j_group(const S *L, long ld, long naux, long npairs, const double *d, long ldd,
                double *rho, long ldrho, double *vj, long ldvj, int accumulate,
                double *work, const JPlan &pl, hipStream_t st) {
    double *part1 = work;
    double *part2 = work + align_up(pl.part1_words, 64);
    dim3 g1((unsigned)pl.nchunk, (unsigned)pl.nrblock);
    hipLaunchKernelGGL((j_rho_partial<NDM, VEC, S>), g1, dim3(J_THREADS), 0, st, L, ld, naux, npairs, d, ldd, part1);
    MDF_LAUNCHED("j_rho_partial");
    long nt = (long)NDM * naux;
    hipLaunchKernelGGL(j_rho_reduce, dim3((unsigned)cdiv(nt, (long)J_THREADS)), dim3(J_THREADS), 0, st,
                       part1, pl.nchunk, NDM, naux, rho, ldrho);
    MDF_LAUNCHED("j_rho_reduce");
    dim3 g2((unsigned)pl.ncblock, (unsigned)pl.ngroup);
    if (pl.ngroup == 1 && !accumulate) {
        hipLaunchKernelGGL((j_vj_pass<NDM, VEC, S>), g2, dim3(J_THREADS), 0, st,
                           L, ld, naux, npairs, rho, ldrho, pl.rows_per_group, vj, ldvj, 0L);
        MDF_LAUNCHED("j_vj_pass");
        return MDF_OK;
    }
    hipLaunchKernelGGL((j_vj_pass<NDM, VEC, S>), g2, dim3(J_THREADS), 0, st,
                       L, ld, naux, npairs, rho, ldrho, pl.rows_per_group, part2, npairs, (long)NDM * npairs);
    MDF_LAUNCHED("j_vj_pass");
    nt = (long)NDM * npairs;
    hipLaunchKernelGGL(j_vj_reduce, dim3((unsigned)cdiv(nt, (long)J_THREADS)), dim3(J_THREADS), 0, st,
                       part2, pl.ngroup, NDM, npairs, vj, ldvj, accumulate);
    MDF_LAUNCHED("j_vj_reduce");
    return MDF_OK;
}

template <int VEC, typename S>
int dispatch_j(int nd, const S *L, long ld, long naux, long npairs, const double *d, long ldd,
               double *rho, long ldrho, double *vj, long ldvj, int accumulate, double *work,
               const JPlan &pl, hipStream_t st) {
    switch (nd) {
    case 1: return run_j_group<1, VEC, S>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st);
    case 2: return run_j_group<2, VEC, S>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st);
    case 3: return run_j_group<3, VEC, S>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st);
    default: return run_j_group<4, VEC, S>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st);
    }
}

}  // namespace

// cderi values per thread in pass 2: a 16-byte vector when the rows allow it
static int j_vec(const void *L, long ld, int dtype) {
    const int w = dtype == MDF_DTYPE_F32 ? 4 : 2;
    return (ld % w == 0) && ((uintptr_t)L % 16 == 0) ? w : 1;
}

extern "C" size_t mdf_get_j_workspace(long naux, long npairs, int n_dm) {
    (void)n_dm;
    size_t w = 0;
    for (int vec : {1, 2, 4}) {
        JPlan p = plan_j(naux, npairs, 4, vec);
        const size_t x = align_up(p.part1_words, 64) + p.part2_words;
        w = x > w ? x : w;
    }
    return w * sizeof(double);
}

template <typename S>
static int get_j_typed(const S *cderi, int vec, long naux, long npairs, long ld, const double *dm_pairs, int n_dm,
                       long ldd, double *rho, long ldrho, double *vj_pairs, long ldvj, int accumulate, void *work,
                       hipStream_t st) {
    for (int n0 = 0; n0 < n_dm; n0 += 4) {
        const int nd = n_dm - n0 < 4 ? n_dm - n0 : 4;
        JPlan pl = plan_j(naux, npairs, nd, vec);
        int rc;
        const double *d = dm_pairs + n0 * ldd;
        double *r = rho + n0 * ldrho, *v = vj_pairs + n0 * ldvj;
        if (vec == 1)
            rc = dispatch_j<1, S>(nd, cderi, ld, naux, npairs, d, ldd, r, ldrho, v, ldvj, accumulate, (double *)work,
                                  pl, st);
        else if constexpr (sizeof(S) == 8)
            rc = dispatch_j<2, S>(nd, cderi, ld, naux, npairs, d, ldd, r, ldrho, v, ldvj, accumulate, (double *)work,
                                  pl, st);
        else
            rc = dispatch_j<4, S>(nd, cderi, ld, naux, npairs, d, ldd, r, ldrho, v, ldvj, accumulate, (double *)work,
                                  pl, st);
        if (rc) return rc;
    }
    return MDF_OK;
}

extern "C" int mdf_get_j(const void *cderi, int cderi_dtype, long naux, long npairs, long ld,
                         const double *dm_pairs, int n_dm, long ldd,
                         double *rho, long ldrho, double *vj_pairs, long ldvj, int accumulate,
                         void *work, size_t work_bytes, void *stream) {
    MDF_CHECK_ARG(naux >= 0 && npairs >= 0 && n_dm >= 1, "mdf_get_j: bad sizes naux=%ld npairs=%ld n_dm=%d", naux, npairs, n_dm);
    MDF_CHECK_ARG(ld >= npairs && ldd >= npairs && ldvj >= npairs && ldrho >= naux,
                  "mdf_get_j: leading dimensions too small");
    MDF_CHECK_ARG(cderi && dm_pairs && rho && vj_pairs, "mdf_get_j: null pointer");
    MDF_CHECK_ARG(cderi_dtype == MDF_DTYPE_F64 || cderi_dtype == MDF_DTYPE_F32, "mdf_get_j: cderi dtype %d",
                  cderi_dtype);
    hipStream_t st = as_stream(stream);
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
        set_error("mdf_get_j: workspace %zu bytes < %zu required", work_bytes, need);
        return MDF_EWORKSPACE;
    }
    const int vec = j_vec(cderi, ld, cderi_dtype);
    if (cderi_dtype == MDF_DTYPE_F32)
        return get_j_typed((const float *)cderi, vec, naux, npairs, ld, dm_pairs, n_dm, ldd, rho, ldrho, vj_pairs,
                           ldvj, accumulate, work, st);
    return get_j_typed((const double *)cderi, vec, naux, npairs, ld, dm_pairs, n_dm, ldd, rho, ldrho, vj_pairs, ldvj,
                       accumulate, work, st);
}
