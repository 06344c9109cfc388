// Exchange (K) contraction over the packed cderi on FP64 MFMA.
//
// Restates electrolites/fastdf.py:974-1005 (unpack + step 1 + step 2) and
// :330-335 (_syrk):
//     R[L, i, k] = sum_j A_L[i, j] F[j, k]                 (step 1)
//     K[i, j]   += sum_{L,k} R_l[L, i, k] R_r[L, j, k]     (step 2)
// where A_L is row L of the cderi unpacked to a symmetric nao x nao matrix.
// The reference materialises A_L with GPU4PySCF's fill_symmetric before the
// GEMM; here the unpack is fused into the step-1 A-operand load: each 128x16
// A tile is read straight from the packed triangle (row-contiguous below the
// diagonal, column-contiguous above it, so both halves load coalesced).
//
// Mode 0 (one factor, fr == NULL): step 2 is a symmetric rank-k update and
// only lower-triangle 128x128 output tiles are computed.  Mode 1: general
// GEMM.  Step 2 splits the (L, k) reduction across workgroups (ksplit); the
// per-split partials are summed in a fixed order by k_finalize, so results do
// not depend on scheduling.
//
// Precision: prec = 0 runs both steps on FP64 MFMA; prec = 1 converts the
// packed cderi and the factors to FP32 on the way into LDS, runs both steps
// on FP32 MFMA with FP32 accumulation and stores R in FP32 (the mixed-
// precision SCF cycles); the cross-split sum and the output stay FP64.
#include "mdf_common.h"
#include "mdf_gemm.h"
#include "mdf_kcore.h"
#include "../../include/mi355df.h"

using namespace mdf;
using namespace mdf::kcore;

namespace {

// 1-D grid of round_xcd(nkt * nit * nL) workgroups in XCD order; tile
// (k-tile fastest, then i-tile, then aux row) so the k-tiles of one A tile
// share an L2.
template <int BN, bool MAPPED, typename T, typename Src, bool TRI = false>
__global__ __launch_bounds__(Step1Shape<BN>::THREADS) void k_step1(
    const Src *__restrict__ cderi, long ld, const int *__restrict__ pmap, int nao,
    const double *__restrict__ F, long ldf, int nocc, T *__restrict__ R, long ldr,
    long strideL, int nkt, int nit, long ntiles) {
    __shared__ T lds[Step1Shape<BN>::S::LDS_ELEMS];
    const long x = xcd_tile(blockIdx.x, gridDim.x);
    if (x >= ntiles) return;
    const int kt = (int)(x % nkt), it = (int)((x / nkt) % nit);
    const long L = x / ((long)nkt * nit);
    step1_tile<BN, MAPPED, T, Src, TRI>(cderi + L * ld, pmap, nao, F, ldf, nocc, R + L * strideL, ldr, it * 128,
                                        kt * BN, lds);
}

// 1-D grid of ntiles_grid * ksplit workgroups, (tile, split) by step2_coords.
template <bool SYM, bool VEC, typename T>
__global__ __launch_bounds__(G_THREADS, 2) void k_step2(
    const T *__restrict__ Rl, const T *__restrict__ Rr, int nao, int nocc,
    long ldr, long strideL, int nL, int ntile, int ntiles_grid, int ksplit, double *__restrict__ part,
    long part_stride, int beta) {
    __shared__ T lds[TileShape<S2_B, S2_B>::LDS_ELEMS];
    int tile, split, ti, tj;
    step2_coords(blockIdx.x, ntiles_grid, tile, split);
    if (SYM) {
        tri_tile(tile, ti, tj);
    } else {
        ti = tile / ntile;
        tj = tile % ntile;
    }
    const int kchunks = (nocc + G_BK - 1) / G_BK;
    const long nsteps = (long)nL * kchunks;
    const long s0 = nsteps * split / ksplit, s1 = nsteps * (split + 1) / ksplit;
    step2_tile<VEC, T>(Rl, Rr, nao, nocc, ldr, strideL, ti * S2_B, tj * S2_B, s0, s1,
                       part + (long)split * part_stride, beta, lds);
}

__global__ __launch_bounds__(256) void k_finalize(
    const double *__restrict__ part, int ksplit, long part_stride, int nao, int sym,
    double *__restrict__ vk, long ldvk, int accumulate) {
    const long e0 = (long)blockIdx.x * 256;
    finalize_range(part, ksplit, part_stride, nao, sym, vk, ldvk, accumulate, e0,
                   min(e0 + 256, (long)nao * nao));
}

// ---- split-BF16 (prec 2)
__global__ __launch_bounds__(256) void k_split_ft(const double *__restrict__ F, long ldf, int nao, int nocc,
                                                  b3::u16 *hi, b3::u16 *lo, long ldft) {
    const long e0 = (long)blockIdx.x * 1024;
    kb3::split_ft_range(F, ldf, nao, nocc, hi, lo, ldft, e0, min(e0 + 1024, (long)nocc * ldft));
}

template <bool MAPPED, int BN, typename Src>
__global__ __launch_bounds__(kb3::Step1Cfg<BN>::THREADS) void k_step1_b3(
    const Src *__restrict__ cderi, long ld, const int *__restrict__ pmap, int nao, const b3::u16 *fthi,
    const b3::u16 *ftlo, long ldft, int nocc, b3::u16 *Rhi, b3::u16 *Rlo, long ldr, long strideL, int nkt,
    int nit, long ntiles) {
    __shared__ b3::u16 lds[kb3::Step1Cfg<BN>::S::LDS_ELEMS];
    const long x = xcd_tile(blockIdx.x, gridDim.x);
    if (x >= ntiles) return;
    const int kt = (int)(x % nkt), it = (int)((x / nkt) % nit);
    const long z = x / ((long)nkt * nit);
    kb3::step1_tile<MAPPED, BN, Src>(cderi + z * ld, pmap, nao, fthi, ftlo, ldft, nocc, Rhi + z * strideL,
                                Rlo + z * strideL, ldr, it * kb3::BM, kt * BN, lds);
}

template <int BN, typename Src>
void launch_step1_b3(int nL, hipStream_t st, const Src *blk, long ld, const int *pmap, int nao,
                     const b3::u16 *fthi, const b3::u16 *ftlo, long ldft, int nocc, b3::u16 *rhi, b3::u16 *rlo,
                     long ldr, long strideL) {
    const int nkt = (int)cdiv(ldr, (long)BN), nit = (int)cdiv(nao, kb3::BM);
    const long ntiles = (long)nkt * nit * nL;
    const dim3 g((unsigned)round_xcd(ntiles)), b(kb3::Step1Cfg<BN>::THREADS);
    if (pmap)
        hipLaunchKernelGGL((k_step1_b3<true, BN, Src>), g, b, 0, st, blk, ld, pmap, nao, fthi, ftlo, ldft, nocc, rhi,
                           rlo, ldr, strideL, nkt, nit, ntiles);
    else
        hipLaunchKernelGGL((k_step1_b3<false, BN, Src>), g, b, 0, st, blk, ld, pmap, nao, fthi, ftlo, ldft, nocc, rhi,
                           rlo, ldr, strideL, nkt, nit, ntiles);
}

template <bool SYM>
__global__ __launch_bounds__(G_THREADS, 2) void k_step2_b3(
    const b3::u16 *lhi, const b3::u16 *llo, const b3::u16 *rhi, const b3::u16 *rlo, int nao, long ldr,
    long strideL, int nL, int ntile, int ntiles_grid, int ksplit, double *__restrict__ part, long part_stride,
    int beta) {
    __shared__ b3::u16 lds[b3::Shape<S2_B, S2_B>::LDS_ELEMS];
    int tile, split, ti, tj;
    step2_coords(blockIdx.x, ntiles_grid, tile, split);
    if (SYM) {
        tri_tile(tile, ti, tj);
    } else {
        ti = tile / ntile;
        tj = tile % ntile;
    }
    const long nsteps = (long)nL * (ldr / b3::BK);
    const long s0 = nsteps * split / ksplit, s1 = nsteps * (split + 1) / ksplit;
    kb3::step2_tile(lhi, llo, rhi, rlo, nao, ldr, strideL, ti * S2_B, tj * S2_B, s0, s1,
                    part + (long)split * part_stride, beta, lds);
}

template <int BN, typename T, typename Src, bool TRI = false>
void launch_step1(int nL, hipStream_t st, const Src *blk, long ld, const int *pmap, int nao,
                  const double *f, long ldf, int nocc, T *r, long ldr, long strideL) {
    const int nkt = (int)cdiv(nocc, BN), nit = (int)cdiv(nao, 128);
    const long ntiles = (long)nkt * nit * nL;
    const dim3 g((unsigned)round_xcd(ntiles)), b(Step1Shape<BN>::THREADS);
    if (pmap)
        hipLaunchKernelGGL((k_step1<BN, true, T, Src, TRI>), g, b, 0, st, blk, ld, pmap, nao, f, ldf,
                           nocc, r, ldr, strideL, nkt, nit, ntiles);
    else
        hipLaunchKernelGGL((k_step1<BN, false, T, Src, TRI>), g, b, 0, st, blk, ld, pmap, nao, f, ldf,
                           nocc, r, ldr, strideL, nkt, nit, ntiles);
}

template <typename T, typename Src, bool TRI = false>
int step1(const KPlan &p, int nL, hipStream_t st, const Src *blk, long ld, const int *pmap, int nao,
          const double *f, long ldf, int nocc, T *r, long ldr, long strideL) {
    switch (p.bn) {
    case 320: launch_step1<320, T, Src, TRI>(nL, st, blk, ld, pmap, nao, f, ldf, nocc, r, ldr, strideL); break;
    case 256: launch_step1<256, T, Src, TRI>(nL, st, blk, ld, pmap, nao, f, ldf, nocc, r, ldr, strideL); break;
    case 192: launch_step1<192, T, Src, TRI>(nL, st, blk, ld, pmap, nao, f, ldf, nocc, r, ldr, strideL); break;
    case 160: launch_step1<160, T, Src, TRI>(nL, st, blk, ld, pmap, nao, f, ldf, nocc, r, ldr, strideL); break;
    case 128: launch_step1<128, T, Src, TRI>(nL, st, blk, ld, pmap, nao, f, ldf, nocc, r, ldr, strideL); break;
    default: launch_step1<64, T, Src, TRI>(nL, st, blk, ld, pmap, nao, f, ldf, nocc, r, ldr, strideL); break;
    }
    MDF_LAUNCHED("k_step1");
    return MDF_OK;
}

template <typename T>
int step2(const KPlan &p, int nL, bool two, hipStream_t st, const T *rl, const T *rr, int nao, int nocc,
          long ldr, long strideL, double *part, long part_stride, int beta) {
    const dim3 g2((unsigned)((long)p.ntiles_grid * p.ksplit));
    const bool vec = (ldr % (16 / (long)sizeof(T)) == 0) && ((uintptr_t)rl % 16 == 0) && ((uintptr_t)rr % 16 == 0);
    if (two && vec)
        hipLaunchKernelGGL((k_step2<false, true, T>), g2, dim3(G_THREADS), 0, st, rl, rr, nao, nocc, ldr,
                           strideL, nL, p.ntile, p.ntiles_grid, p.ksplit, part, part_stride, beta);
    else if (two)
        hipLaunchKernelGGL((k_step2<false, false, T>), g2, dim3(G_THREADS), 0, st, rl, rr, nao, nocc, ldr,
                           strideL, nL, p.ntile, p.ntiles_grid, p.ksplit, part, part_stride, beta);
    else if (vec)
        hipLaunchKernelGGL((k_step2<true, true, T>), g2, dim3(G_THREADS), 0, st, rl, rl, nao, nocc, ldr,
                           strideL, nL, p.ntile, p.ntiles_grid, p.ksplit, part, part_stride, beta);
    else
        hipLaunchKernelGGL((k_step2<true, false, T>), g2, dim3(G_THREADS), 0, st, rl, rl, nao, nocc, ldr,
                           strideL, nL, p.ntile, p.ntiles_grid, p.ksplit, part, part_stride, beta);
    MDF_LAUNCHED("k_step2");
    return MDF_OK;
}

// The block loop of mdf_get_k for element type T.
template <typename T, typename Src>
int run_k(const KPlan &p, const Src *cderi, long naux, long ld, const int *pair_map, int nao,
          const double *fl, const double *fr, int nocc, long ldf, double *vk, long ldvk, int accumulate,
          void *work, hipStream_t st) {
    const bool two = fr != nullptr;
    double *part = (double *)work;
    T *rl = (T *)(part + align_up(p.part_words, 64));
    T *rr = two ? rl + align_up(p.r_words, 64) : rl;
    const long ldr = r_ld(nocc, sizeof(T)), strideL = (long)nao * ldr;
    const long part_stride = (long)nao * nao;
    for (long L0 = 0; L0 < naux; L0 += p.nL) {
        const int nL = (int)(naux - L0 < p.nL ? naux - L0 : p.nL);
        const Src *blk = cderi + L0 * ld;
        int rc = step1<T, Src>(p, nL, st, blk, ld, pair_map, nao, fl, ldf, nocc, rl, ldr, strideL);
        if (rc) return rc;
        if (two) {
            rc = step1<T, Src>(p, nL, st, blk, ld, pair_map, nao, fr, ldf, nocc, rr, ldr, strideL);
            if (rc) return rc;
        }
        rc = step2<T>(p, nL, two, st, rl, rr, nao, nocc, ldr, strideL, part, part_stride, L0 > 0);
        if (rc) return rc;
    }
    long nt = (long)nao * nao;
    hipLaunchKernelGGL(k_finalize, dim3((unsigned)cdiv(nt, 256L)), dim3(256), 0, st,
                       part, p.ksplit, part_stride, nao, two ? 0 : 1, vk, ldvk, accumulate);
    MDF_LAUNCHED("k_finalize");
    return MDF_OK;
}

// The block loop of mdf_get_k on the split-BF16 engine.
template <typename Src>
int run_k_b3(const KPlan &p, const Src *cderi, long naux, long ld, const int *pair_map, int nao,
             const double *fl, const double *fr, int nocc, long ldf, double *vk, long ldvk, int accumulate,
             void *work, hipStream_t st) {
    const bool two = fr != nullptr;
    const kb3::Layout y = kb3::layout(nao, nocc, p.nL, p.ksplit, two);
    double *part = (double *)work;
    kb3::Side sd[2] = {kb3::side_ptrs((char *)work, y, 0), kb3::side_ptrs((char *)work, y, two ? 1 : 0)};
    const long strideL = (long)nao * y.ldr, part_stride = (long)nao * nao;
    const long nft = (long)nocc * y.ldft;
    for (int side = 0; side < (two ? 2 : 1); side++) {
        hipLaunchKernelGGL(k_split_ft, dim3((unsigned)cdiv(nft, 1024L)), dim3(256), 0, st, side ? fr : fl, ldf, nao,
                           nocc, sd[side].fthi, sd[side].ftlo, y.ldft);
        MDF_LAUNCHED("k_split_ft");
    }
    for (long L0 = 0; L0 < naux; L0 += p.nL) {
        const int nL = (int)(naux - L0 < p.nL ? naux - L0 : p.nL);
        const Src *blk = cderi + L0 * ld;
        const int bn = kb3::pick_bn(y.ldr);
        for (int side = 0; side < (two ? 2 : 1); side++) {
            const kb3::Side &q = sd[side];
            switch (bn) {
            case 320:
                launch_step1_b3<320>(nL, st, blk, ld, pair_map, nao, q.fthi, q.ftlo, y.ldft, nocc, q.rhi, q.rlo,
                                     y.ldr, strideL);
                break;
            case 256:
                launch_step1_b3<256>(nL, st, blk, ld, pair_map, nao, q.fthi, q.ftlo, y.ldft, nocc, q.rhi, q.rlo,
                                     y.ldr, strideL);
                break;
            case 192:
                launch_step1_b3<192>(nL, st, blk, ld, pair_map, nao, q.fthi, q.ftlo, y.ldft, nocc, q.rhi, q.rlo,
                                     y.ldr, strideL);
                break;
            default:
                launch_step1_b3<128>(nL, st, blk, ld, pair_map, nao, q.fthi, q.ftlo, y.ldft, nocc, q.rhi, q.rlo,
                                     y.ldr, strideL);
                break;
            }
            MDF_LAUNCHED("k_step1_b3");
        }
        const dim3 g2((unsigned)((long)p.ntiles_grid * p.ksplit));
        const kb3::Side &a = sd[0], &b = sd[two ? 1 : 0];
        if (two)
            hipLaunchKernelGGL(k_step2_b3<false>, g2, dim3(G_THREADS), 0, st, a.rhi, a.rlo, b.rhi, b.rlo, nao, y.ldr,
                               strideL, nL, p.ntile, p.ntiles_grid, p.ksplit, part, part_stride, (int)(L0 > 0));
        else
            hipLaunchKernelGGL(k_step2_b3<true>, g2, dim3(G_THREADS), 0, st, a.rhi, a.rlo, b.rhi, b.rlo, nao, y.ldr,
                               strideL, nL, p.ntile, p.ntiles_grid, p.ksplit, part, part_stride, (int)(L0 > 0));
        MDF_LAUNCHED("k_step2_b3");
    }
    long nt = (long)nao * nao;
    hipLaunchKernelGGL(k_finalize, dim3((unsigned)cdiv(nt, 256L)), dim3(256), 0, st,
                       part, p.ksplit, part_stride, nao, two ? 0 : 1, vk, ldvk, accumulate);
    MDF_LAUNCHED("k_finalize");
    return MDF_OK;
}

}  // namespace

static size_t k_workspace(const KPlan &p, bool two, size_t elem) {
    return align_up(p.part_words, 64) * sizeof(double) + (two ? 2 : 1) * align_up(p.r_words, 64) * elem;
}

// Aux rows per pass and workspace of one call for any precision.
static long k_rows_rule(long naux, int nao, int nocc, bool two, int prec) {
    if (prec == MDF_PREC_BF16X3) return block_rows(naux, nao, (int)kb3::pad32(nocc), two, 4);
    return block_rows(naux, nao, nocc, two, prec == MDF_PREC_FP32 ? sizeof(float) : sizeof(double));
}
static size_t k_bytes(long naux, int nao, int nocc, bool two, int prec, long rows) {
    KPlan p = plan_k(naux, nao, nocc, two, rows, prec == MDF_PREC_FP32 ? sizeof(float) : sizeof(double));
    if (prec == MDF_PREC_BF16X3) return kb3::layout(nao, nocc, p.nL, p.ksplit, two).total;
    return k_workspace(p, two, prec == MDF_PREC_FP32 ? sizeof(float) : sizeof(double));
}

extern "C" size_t mdf_get_k_workspace(long naux, int nao, int nocc, int two_factors, int prec,
                                      long max_block_rows) {
    long rows = k_rows_rule(naux, nao, nocc, two_factors != 0, prec);
    if (max_block_rows > 0 && max_block_rows < rows) rows = max_block_rows;
    return k_bytes(naux, nao, nocc, two_factors != 0, prec, rows);
}

extern "C" int mdf_get_k(const void *cderi, int cderi_dtype, long naux, long npairs, long ld, const int *pair_map,
                         int nao, const double *fl, const double *fr, int nocc, long ldf, int prec,
                         double *vk, long ldvk, int accumulate, void *work, size_t work_bytes,
                         void *stream) {
    MDF_CHECK_ARG(naux >= 0 && nao >= 0 && nocc >= 0, "mdf_get_k: bad sizes");
    MDF_CHECK_ARG(cderi && fl && vk, "mdf_get_k: null pointer");
    MDF_CHECK_ARG(ld >= npairs && ldf >= nocc && ldvk >= nao, "mdf_get_k: leading dimension too small");
    MDF_CHECK_ARG(prec == MDF_PREC_FP64 || prec == MDF_PREC_FP32 || prec == MDF_PREC_BF16X3,
                  "mdf_get_k: unknown precision %d", prec);
    MDF_CHECK_ARG(cderi_dtype == MDF_DTYPE_F64 || cderi_dtype == MDF_DTYPE_F32, "mdf_get_k: cderi dtype %d",
                  cderi_dtype);
    MDF_CHECK_ARG(pair_map || npairs == (long)nao * (nao + 1) / 2,
                  "mdf_get_k: without a pair map the cderi must hold the full packed triangle "
                  "(npairs=%ld, nao=%d)", npairs, nao);
    hipStream_t st = as_stream(stream);
    if (nao == 0) return MDF_OK;
    const bool two = fr != nullptr;
    if (nocc == 0 || naux == 0) {
        if (!accumulate)
            for (int i = 0; i < nao; i++) MDF_HIP(hipMemsetAsync(vk + (long)i * ldvk, 0, nao * sizeof(double), st));
        return MDF_OK;
    }
    MDF_CHECK_ARG(work != nullptr, "mdf_get_k: workspace is null");
    size_t need1 = k_bytes(naux, nao, nocc, two, prec, 1);
    if (work_bytes < need1) {
        set_error("mdf_get_k: workspace %zu bytes < %zu required", work_bytes, need1);
        return MDF_EWORKSPACE;
    }
    // passes of the shape rule's aux rows (the grouped launch cuts the same
    // way); shorter when the workspace is smaller
    long rows = k_rows_rule(naux, nao, nocc, two, prec);
    if (k_bytes(naux, nao, nocc, two, prec, rows) > work_bytes) {
        long lo = 1, hi = rows;          // largest rows that fit
        while (lo < hi) {
            const long mid = (lo + hi + 1) / 2;
            if (k_bytes(naux, nao, nocc, two, prec, mid) <= work_bytes) lo = mid; else hi = mid - 1;
        }
        rows = lo;
    }
    KPlan p = plan_k(naux, nao, nocc, two, rows, prec == MDF_PREC_FP32 ? sizeof(float) : sizeof(double));
    if (cderi_dtype == MDF_DTYPE_F32) {
        const float *c = (const float *)cderi;
        if (prec == MDF_PREC_BF16X3)
            return run_k_b3(p, c, naux, ld, pair_map, nao, fl, fr, nocc, ldf, vk, ldvk, accumulate, work, st);
        if (prec == MDF_PREC_FP32)
            return run_k<float>(p, c, naux, ld, pair_map, nao, fl, fr, nocc, ldf, vk, ldvk, accumulate, work, st);
        return run_k<double>(p, c, naux, ld, pair_map, nao, fl, fr, nocc, ldf, vk, ldvk, accumulate, work, st);
    }
    const double *c = (const double *)cderi;
    if (prec == MDF_PREC_BF16X3)
        return run_k_b3(p, c, naux, ld, pair_map, nao, fl, fr, nocc, ldf, vk, ldvk, accumulate, work, st);
    if (prec == MDF_PREC_FP32)
        return run_k<float>(p, c, naux, ld, pair_map, nao, fl, fr, nocc, ldf, vk, ldvk, accumulate, work, st);
    return run_k<double>(p, c, naux, ld, pair_map, nao, fl, fr, nocc, ldf, vk, ldvk, accumulate, work, st);
}

// The first step of mdf_get_k alone, FP64: R[L] = A_L F for the nL aux rows
// of one block (A_L the unpacked symmetric row L), R[L * strideL + i * ldr +
// k] for i < nao, k < nocc.  The DF gradient's rhok = F_r^T A_L F_l takes its
// inner product from here.
namespace {
template <bool TRI>
int k_step1_entry(const char *name, const void *cderi, int cderi_dtype, long nL, long npairs, long ld,
                  const int *pair_map, int nao, const double *f, long ldf, int nocc, double *R, long ldr,
                  long strideL, void *stream) {
    MDF_CHECK_ARG(nL >= 0 && nao >= 0 && nocc >= 0 && nL < (1L << 31), "%s: bad sizes", name);
    MDF_CHECK_ARG(cderi && f && R, "%s: null pointer", name);
    MDF_CHECK_ARG(ld >= npairs && ldf >= nocc && ldr >= nocc && strideL >= (long)nao * ldr,
                  "%s: leading dimension too small", name);
    MDF_CHECK_ARG(cderi_dtype == MDF_DTYPE_F64 || cderi_dtype == MDF_DTYPE_F32, "%s: cderi dtype %d", name,
                  cderi_dtype);
    MDF_CHECK_ARG(pair_map || npairs == (long)nao * (nao + 1) / 2,
                  "%s: without a pair map the cderi must hold the full packed triangle "
                  "(npairs=%ld, nao=%d)", name, npairs, nao);
    if (nL == 0 || nao == 0 || nocc == 0) return MDF_OK;
    KPlan p{};
    p.bn = pick_bn(nocc);
    hipStream_t st = as_stream(stream);
    if (cderi_dtype == MDF_DTYPE_F32)
        return step1<double, float, TRI>(p, (int)nL, st, (const float *)cderi, ld, pair_map, nao, f, ldf, nocc, R,
                                         ldr, strideL);
    return step1<double, double, TRI>(p, (int)nL, st, (const double *)cderi, ld, pair_map, nao, f, ldf, nocc, R,
                                      ldr, strideL);
}
}  // namespace

extern "C" int mdf_k_step1_lower(const void *cderi, int cderi_dtype, long nL, long npairs, long ld,
                                 const int *pair_map, int nao, const double *f, long ldf, int nocc, double *R,
                                 long ldr, long strideL, void *stream) {
    return k_step1_entry<true>("mdf_k_step1_lower", cderi, cderi_dtype, nL, npairs, ld, pair_map, nao, f, ldf,
                               nocc, R, ldr, strideL, stream);
}

extern "C" int mdf_k_step1(const void *cderi, int cderi_dtype, long nL, long npairs, long ld, const int *pair_map,
                           int nao, const double *f, long ldf, int nocc, double *R, long ldr, long strideL,
                           void *stream) {
    return k_step1_entry<false>("mdf_k_step1", cderi, cderi_dtype, nL, npairs, ld, pair_map, nao, f, ldf, nocc, R,
                                ldr, strideL, stream);
}
