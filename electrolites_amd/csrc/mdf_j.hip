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
                                                          long naux, double *__restrict__ rho, long ldrho,
                                                          int accumulate) {
    const long e0 = (long)blockIdx.x * J_THREADS;
    rho_reduce_range(part, nchunk, ndm, naux, rho, ldrho, e0, min(e0 + J_THREADS, (long)ndm * naux), accumulate);
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

// The same tiles under their own names for the two halves called one
// chunk at a time (mdf_j_rho / mdf_j_vj: the J build without a stored cderi
// and the DF gradient's Coulomb density), so a kernel trace separates them
// from the full J of mdf_get_j.
template <int NDM, int VEC, typename S>
__global__ __launch_bounds__(J_THREADS) void jh_rho_partial(const S *__restrict__ L, long ld, long naux,
                                                            long npairs, const double *__restrict__ d, long ldd,
                                                            double *__restrict__ part) {
    rho_partial_tile<NDM, VEC, S>(L, ld, naux, npairs, d, ldd, NDM, part, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(J_THREADS) void jh_rho_reduce(const double *__restrict__ part, long nchunk, int ndm,
                                                           long naux, double *__restrict__ rho, long ldrho,
                                                           int accumulate) {
    const long e0 = (long)blockIdx.x * J_THREADS;
    rho_reduce_range(part, nchunk, ndm, naux, rho, ldrho, e0, min(e0 + J_THREADS, (long)ndm * naux), accumulate);
}

template <int NDM, int VEC, typename S>
__global__ __launch_bounds__(J_THREADS) void jh_vj_pass(const S *__restrict__ L, long ld, long naux,
                                                        long npairs, const double *__restrict__ rho, long ldrho,
                                                        long rows_per_group, double *__restrict__ out, long ldo,
                                                        long group_stride) {
    const long r0 = (long)blockIdx.y * rows_per_group;
    const long r1 = min(r0 + rows_per_group, naux);
    vj_pass_tile<NDM, VEC, S>(L, ld, npairs, rho, ldrho, NDM, r0, r1, out + (long)blockIdx.y * group_stride, ldo,
                              blockIdx.x);
}

__global__ __launch_bounds__(J_THREADS) void jh_vj_reduce(const double *__restrict__ part, long ngroup, int ndm,
                                                          long npairs, double *__restrict__ vj, long ldvj,
                                                          int accumulate) {
    const long e0 = (long)blockIdx.x * J_THREADS;
    vj_reduce_range(part, ngroup, ndm, npairs, vj, ldvj, accumulate, e0, min(e0 + J_THREADS, (long)ndm * npairs));
}

// passes: 1 = rho only (rho += with accumulate_rho), 2 = vj only (from the
// given rho), 3 = both (rho written, then vj); halves: launch the jh_* names
template <int NDM, int VEC, typename S>
int run_j_group(const S *L, long ld, long naux, long npairs, const double *d, long ldd,
                double *rho, long ldrho, double *vj, long ldvj, int accumulate,
                double *work, const JPlan &pl, hipStream_t st, int passes = 3, int accumulate_rho = 0,
                bool halves = false) {
    double *part1 = work;
    double *part2 = work + align_up(pl.part1_words, 64);
    long nt;
    if (passes & 1) {
        dim3 g1((unsigned)pl.nchunk, (unsigned)pl.nrblock);
        auto rho_partial = halves ? jh_rho_partial<NDM, VEC, S> : j_rho_partial<NDM, VEC, S>;
        hipLaunchKernelGGL(rho_partial, g1, dim3(J_THREADS), 0, st, L, ld, naux, npairs, d, ldd, part1);
        MDF_LAUNCHED("j_rho_partial");
        nt = (long)NDM * naux;
        auto rho_reduce = halves ? jh_rho_reduce : j_rho_reduce;
        hipLaunchKernelGGL(rho_reduce, dim3((unsigned)cdiv(nt, (long)J_THREADS)), dim3(J_THREADS), 0, st, part1,
                           pl.nchunk, NDM, naux, rho, ldrho, accumulate_rho);
        MDF_LAUNCHED("j_rho_reduce");
    }
    if (!(passes & 2)) return MDF_OK;
    dim3 g2((unsigned)pl.ncblock, (unsigned)pl.ngroup);
    auto vj_pass = halves ? jh_vj_pass<NDM, VEC, S> : j_vj_pass<NDM, VEC, S>;
    if (pl.ngroup == 1 && !accumulate) {
        hipLaunchKernelGGL(vj_pass, g2, dim3(J_THREADS), 0, st,
                           L, ld, naux, npairs, rho, ldrho, pl.rows_per_group, vj, ldvj, 0L);
        MDF_LAUNCHED("j_vj_pass");
        return MDF_OK;
    }
    hipLaunchKernelGGL(vj_pass, g2, dim3(J_THREADS), 0, st,
                       L, ld, naux, npairs, rho, ldrho, pl.rows_per_group, part2, npairs, (long)NDM * npairs);
    MDF_LAUNCHED("j_vj_pass");
    nt = (long)NDM * npairs;
    auto vj_reduce = halves ? jh_vj_reduce : j_vj_reduce;
    hipLaunchKernelGGL(vj_reduce, dim3((unsigned)cdiv(nt, (long)J_THREADS)), dim3(J_THREADS), 0, st, part2,
                       pl.ngroup, NDM, npairs, vj, ldvj, accumulate);
    MDF_LAUNCHED("j_vj_reduce");
    return MDF_OK;
}

template <int VEC, typename S>
int dispatch_j(int nd, const S *L, long ld, long naux, long npairs, const double *d, long ldd,
               double *rho, long ldrho, double *vj, long ldvj, int accumulate, double *work,
               const JPlan &pl, hipStream_t st, int passes = 3, int acc_rho = 0, bool halves = false) {
    switch (nd) {
    case 1: return run_j_group<1, VEC, S>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st, passes, acc_rho, halves);
    case 2: return run_j_group<2, VEC, S>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st, passes, acc_rho, halves);
    case 3: return run_j_group<3, VEC, S>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st, passes, acc_rho, halves);
    default: return run_j_group<4, VEC, S>(L, ld, naux, npairs, d, ldd, rho, ldrho, vj, ldvj, accumulate, work, pl, st, passes, acc_rho, halves);
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
                       hipStream_t st, int passes = 3, int acc_rho = 0) {
    const bool halves = passes != 3;
    for (int n0 = 0; n0 < n_dm; n0 += 4) {
        const int nd = n_dm - n0 < 4 ? n_dm - n0 : 4;
        JPlan pl = plan_j(naux, npairs, nd, vec);
        int rc;
        const double *d = dm_pairs + n0 * ldd;
        double *r = rho + n0 * ldrho, *v = vj_pairs + n0 * ldvj;
        if (vec == 1)
            rc = dispatch_j<1, S>(nd, cderi, ld, naux, npairs, d, ldd, r, ldrho, v, ldvj, accumulate, (double *)work,
                                  pl, st, passes, acc_rho, halves);
        else if constexpr (sizeof(S) == 8)
            rc = dispatch_j<2, S>(nd, cderi, ld, naux, npairs, d, ldd, r, ldrho, v, ldvj, accumulate, (double *)work,
                                  pl, st, passes, acc_rho, halves);
        else
            rc = dispatch_j<4, S>(nd, cderi, ld, naux, npairs, d, ldd, r, ldrho, v, ldvj, accumulate, (double *)work,
                                  pl, st, passes, acc_rho, halves);
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

// The two halves of mdf_get_j on their own, for a three-centre block that is
// produced and consumed chunk by chunk (the J build without a stored cderi):
//   mdf_j_rho: rho[n, P] (+)= sum_p B[P, p] d[n, p]
//   mdf_j_vj:  vj[n, p] (+)= sum_P c[n, P] B[P, p]
// Replace GPU4PySCF's Int3c2eOpt.contract_dm / contract_auxvec as the
// reference's RI-J path calls them (electrolites/fastdfj3c.py:388-530,
// :726-784).  Same workspace as mdf_get_j.
static int j_half(const void *B, int dtype, long naux, long width, long ld, const double *x, int n_dm, long ldx,
                  double *y, long ldy, int accumulate, void *work, size_t work_bytes, void *stream, int passes) {
    MDF_CHECK_ARG(naux >= 0 && width >= 0 && n_dm >= 1 && ld >= width, "mdf_j_%s: bad sizes",
                  passes == 1 ? "rho" : "vj");
    MDF_CHECK_ARG(B && x && y, "mdf_j_%s: null pointer", passes == 1 ? "rho" : "vj");
    MDF_CHECK_ARG(passes == 1 ? (ldx >= width && ldy >= naux) : (ldx >= naux && ldy >= width),
                  "mdf_j_%s: leading dimensions too small", passes == 1 ? "rho" : "vj");
    MDF_CHECK_ARG(dtype == MDF_DTYPE_F64 || dtype == MDF_DTYPE_F32, "mdf_j: dtype %d", dtype);
    hipStream_t st = as_stream(stream);
    if (width == 0 || naux == 0) {
        if (!accumulate) {
            const long n = passes == 1 ? naux : width;
            for (int i = 0; i < n_dm && n > 0; i++) MDF_HIP(hipMemsetAsync(y + i * ldy, 0, n * sizeof(double), st));
        }
        return MDF_OK;
    }
    const size_t need = mdf_get_j_workspace(naux, width, n_dm);
    MDF_CHECK_ARG(work != nullptr && work_bytes >= need, "mdf_j: workspace %zu bytes < %zu required", work_bytes,
                  need);
    const int vec = j_vec(B, ld, dtype);
    // pass 1: d = x, rho = y; pass 2: rho = x (the fitted coefficients), vj = y
    if (dtype == MDF_DTYPE_F32)
        return passes == 1
            ? get_j_typed((const float *)B, vec, naux, width, ld, x, n_dm, ldx, y, ldy, y, ldy, 0, work, st, 1,
                          accumulate)
            : get_j_typed((const float *)B, vec, naux, width, ld, x, n_dm, ldx, (double *)x, ldx, y, ldy, accumulate,
                          work, st, 2, 0);
    return passes == 1
        ? get_j_typed((const double *)B, vec, naux, width, ld, x, n_dm, ldx, y, ldy, y, ldy, 0, work, st, 1,
                      accumulate)
        : get_j_typed((const double *)B, vec, naux, width, ld, x, n_dm, ldx, (double *)x, ldx, y, ldy, accumulate,
                      work, st, 2, 0);
}

extern "C" int mdf_j_rho(const void *B, int dtype, long naux, long width, long ld, const double *d, int n_dm,
                         long ldd, double *rho, long ldrho, int accumulate, void *work, size_t work_bytes,
                         void *stream) {
    return j_half(B, dtype, naux, width, ld, d, n_dm, ldd, rho, ldrho, accumulate, work, work_bytes, stream, 1);
}

extern "C" int mdf_j_vj(const void *B, int dtype, long naux, long width, long ld, const double *c, int n_dm,
                        long ldc, double *vj, long ldvj, int accumulate, void *work, size_t work_bytes,
                        void *stream) {
    return j_half(B, dtype, naux, width, ld, c, n_dm, ldc, vj, ldvj, accumulate, work, work_bytes, stream, 2);
}
// ---- J with no allocation: the partial sums in a file-scope device array ----
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The
// partials of both passes (the only scratch mdf_get_j uses) live in
// `scratch`, MDF_NOALLOC_SCRATCH_BYTES per device, held from module load to
// process end.  Its device address is looked up once per device with that
// device current (hipGetSymbolAddress) and only that address is handed to
// kernels.  Every partial is written before it is read, so nothing needs
// clearing between calls.  Calls that overlap on one device race on the
// array.
#include <mutex>

static __device__ __align__(256) double scratch[MDF_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

int scratch_of_current_device(void **out) {
    int dev = 0;
    MDF_HIP(hipGetDevice(&dev));
    MDF_CHECK_ARG(dev >= 0 && dev < MAX_DEVICES, "mdf_get_j_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        MDF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        MDF_CHECK_ARG(p != nullptr, "mdf_get_j_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MDF_OK;
}

// bytes of partials mdf_get_j's plan needs for these sizes (its first,
// largest group of at most four densities)
size_t j_need_bytes(long naux, long npairs, int n_dm, int vec) {
    const int nd = n_dm < 4 ? n_dm : 4;
    JPlan p = plan_j(naux, npairs, nd, vec);
    return (align_up(p.part1_words, 64) + p.part2_words) * sizeof(double);
}
}  // namespace

extern "C" int mdf_get_j_noalloc(const void *cderi, int cderi_dtype, long naux, long npairs, long ld,
                                 const double *dm_pairs, int n_dm, long ldd, double *rho, long ldrho,
                                 double *vj_pairs, long ldvj, int accumulate, void *stream) {
    MDF_CHECK_ARG(naux >= 0 && npairs >= 0 && n_dm >= 1, "mdf_get_j_noalloc: bad sizes naux=%ld npairs=%ld n_dm=%d",
                  naux, npairs, n_dm);
    MDF_CHECK_ARG(ld >= npairs && ldd >= npairs && ldvj >= npairs && ldrho >= naux,
                  "mdf_get_j_noalloc: leading dimensions too small");
    MDF_CHECK_ARG(cderi && dm_pairs && rho && vj_pairs, "mdf_get_j_noalloc: null pointer");
    MDF_CHECK_ARG(cderi_dtype == MDF_DTYPE_F64 || cderi_dtype == MDF_DTYPE_F32,
                  "mdf_get_j_noalloc: cderi dtype %d", cderi_dtype);
    if (npairs == 0 || naux == 0)
        return mdf_get_j(cderi, cderi_dtype, naux, npairs, ld, dm_pairs, n_dm, ldd, rho, ldrho, vj_pairs, ldvj,
                         accumulate, nullptr, 0, stream);
    const int vec = j_vec(cderi, ld, cderi_dtype);
    const size_t need = j_need_bytes(naux, npairs, n_dm, vec);
    if (need > (size_t)MDF_NOALLOC_SCRATCH_BYTES) {
        set_error("mdf_get_j_noalloc: naux %ld x npairs %ld with %d densities needs %zu bytes of partials, "
                  "above the %zu-byte scratch array; use mdf_get_j with a workspace",
                  naux, npairs, n_dm, need, (size_t)MDF_NOALLOC_SCRATCH_BYTES);
        return MDF_EINVAL;
    }
    void *work = nullptr;
    if (int rc = scratch_of_current_device(&work)) return rc;
    hipStream_t st = as_stream(stream);
    if (cderi_dtype == MDF_DTYPE_F32)
        return get_j_typed((const float *)cderi, vec, naux, npairs, ld, dm_pairs, n_dm, ldd, rho, ldrho, vj_pairs,
                           ldvj, accumulate, work, st);
    return get_j_typed((const double *)cderi, vec, naux, npairs, ld, dm_pairs, n_dm, ldd, rho, ldrho, vj_pairs, ldvj,
                       accumulate, work, st);
}
