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
#include <type_traits>

#include "mdf_common.h"
#include "mdf_gemm.h"
#include "../../include/mi355df.h"

using namespace mdf;

namespace {

__device__ inline long tri(long i) { return i * (i + 1) / 2; }

// ---------------------------------------------------------------- step 1
// A operand of step 1: rows i0..i0+127 of the unpacked A_L, K-step kt covers
// j = 16 kt .. 16 kt + 15.
template <bool MAPPED, typename T = double>
struct PackedALoader {
    static constexpr int N = 128 * G_BK / G_THREADS;     // 8
    static constexpr int TPJ = G_THREADS / G_BK;         // threads per j row (upper)
    const double *row;   // cderi row L
    const int *pmap;     // MAPPED: pair index of (i, j), -1 if not stored
    int nao, i0;
    int upper;           // layout used by the last load()
    __device__ void load(int kt, T (&r)[N]) {
        const int t = threadIdx.x, j0 = kt * G_BK;
        if (!MAPPED && j0 >= i0 + 127) {
            // every j > every i: (i, j) at tri(j) + i, contiguous in i
            upper = 1;
            const int j = j0 + t / TPJ, ib = i0 + (t % TPJ) * N;
            const double *src = row + tri(j) + ib;
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = (j < nao && ib + v < nao) ? (T)src[v] : (T)0;
            return;
        }
        upper = 0;
        const int i = i0 + (t >> 1), jb = j0 + (t & 1) * N;
        if (!MAPPED && i0 >= j0 + G_BK - 1) {
            // every i >= every j: contiguous in j
            const double *src = row + tri(i) + jb;
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = (i < nao) ? (T)src[v] : (T)0;
            return;
        }
#pragma unroll
        for (int v = 0; v < N; v++) {
            const int j = jb + v;
            double x = 0.;
            if (i < nao && j < nao) {
                if (MAPPED) {
                    const int p = pmap[(long)i * nao + j];
                    x = p >= 0 ? row[p] : 0.;
                } else {
                    x = row[i >= j ? tri(i) + j : tri(j) + i];
                }
            }
            r[v] = (T)x;
        }
    }
    __device__ void store(const T (&r)[N], T *lds, int pitch) const {
        const int t = threadIdx.x;
        if (upper) {
            const int jj = t / TPJ, ib = (t % TPJ) * N;
#pragma unroll
            for (int v = 0; v < N; v++) lds[jj * pitch + ib + v] = r[v];
        } else {
            const int ii = t >> 1, jb = (t & 1) * N;
#pragma unroll
            for (int v = 0; v < N; v++) lds[(jb + v) * pitch + ii] = r[v];
        }
    }
};

template <int BN, bool MAPPED, typename T>
__global__ __launch_bounds__(G_THREADS) void k_step1(
    const double *__restrict__ cderi, long ld, const int *__restrict__ pmap, int nao,
    const double *__restrict__ F, long ldf, int nocc, T *__restrict__ R, long ldr,
    long strideL) {
    using S = TileShape<128, BN>;
    __shared__ T lds[S::LDS_ELEMS];
    const int k0 = blockIdx.x * BN, i0 = blockIdx.y * 128;
    PackedALoader<MAPPED, T> la{cderi + (long)blockIdx.z * ld, pmap, nao, i0, 0};
    ColsLoader<BN, T, double> lb{F, ldf, k0, nocc, 0, nao};
    Acc<S::MI, S::NI, T> acc;
    acc.zero();
    gemm_mainloop<128, BN>(la, lb, (nao + G_BK - 1) / G_BK, lds, acc);
    T *Rl = R + (long)blockIdx.z * strideL;
    gemm_epilogue<128, BN, T>(acc, [&](int r, int c, T v) {
        const int i = i0 + r, k = k0 + c;
        if (i < nao && k < nocc) Rl[(long)i * ldr + k] = v;
    });
}

// ---------------------------------------------------------------- step 2
// Rows r0..r0+127 of R[L] (k contiguous); K-step kt of this split is
// (L, kc) = divmod(s0 + kt, kchunks), columns kc*16 .. kc*16 + 15.
template <bool VEC, typename T>
struct TransformedLoader {
    static constexpr int N = 128 * G_BK / G_THREADS;     // 8
    static constexpr int EPV = 16 / (int)sizeof(T);       // elements per 16-byte load
    using V = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    const T *R;
    long strideL, ldr;
    int r0, nao, nocc, kchunks;
    long L;              // current aux row of the next load()
    int kc;              // current k chunk of the next load()
    // thread t owns row t % 128 and k half t / 128, so a wavefront's LDS
    // stores cover 64 consecutive rows of one k (no bank conflicts)
    __device__ void load(int, T (&r)[N]) {
        // called for kt = 0, 1, 2, ... in order: advance (L, kc) in place
        const int t = threadIdx.x;
        const int kb = kc * G_BK + (t >> 7) * N;
        const int row = r0 + (t & 127);
        const T *src = R + L * strideL + (long)row * ldr + kb;
        if (VEC && row < nao && kb + N <= nocc) {
#pragma unroll
            for (int v = 0; v < N; v += EPV) {
                const V x = *reinterpret_cast<const V *>(src + v);
                const T *xe = reinterpret_cast<const T *>(&x);
#pragma unroll
                for (int e = 0; e < EPV; e++) r[v + e] = xe[e];
            }
        } else {
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = (row < nao && kb + v < nocc) ? src[v] : (T)0;
        }
        if (++kc == kchunks) {
            kc = 0;
            ++L;
        }
    }
    __device__ void store(const T (&r)[N], T *lds, int pitch) const {
        const int t = threadIdx.x, ii = t & 127, kb = (t >> 7) * N;
#pragma unroll
        for (int v = 0; v < N; v++) lds[(kb + v) * pitch + ii] = r[v];
    }
};

constexpr int S2_B = 128;

template <bool SYM, bool VEC, typename T>
__global__ __launch_bounds__(G_THREADS, 2) void k_step2(
    const T *__restrict__ Rl, const T *__restrict__ Rr, int nao, int nocc,
    long ldr, long strideL, int nL, int ntile, double *__restrict__ part, long part_stride,
    int beta) {
    using S = TileShape<S2_B, S2_B>;
    __shared__ T lds[S::LDS_ELEMS];
    int ti, tj;
    if (SYM) {
        const int x = blockIdx.x;
        ti = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
        while (ti * (ti + 1) / 2 > x) ti--;
        while ((ti + 1) * (ti + 2) / 2 <= x) ti++;
        tj = x - ti * (ti + 1) / 2;
    } else {
        ti = blockIdx.x / ntile;
        tj = blockIdx.x % ntile;
    }
    const int i0 = ti * S2_B, j0 = tj * S2_B;
    const int kchunks = (nocc + G_BK - 1) / G_BK;
    const long nsteps = (long)nL * kchunks;
    const long s0 = nsteps * blockIdx.y / gridDim.y, s1 = nsteps * (blockIdx.y + 1) / gridDim.y;
    const long Ls = s0 / kchunks;
    const int kcs = (int)(s0 - Ls * kchunks);
    TransformedLoader<VEC, T> la{Rl, strideL, ldr, i0, nao, nocc, kchunks, Ls, kcs};
    TransformedLoader<VEC, T> lb{Rr, strideL, ldr, j0, nao, nocc, kchunks, Ls, kcs};
    Acc<S::MI, S::NI, T> acc;
    acc.zero();
    gemm_mainloop<S2_B, S2_B>(la, lb, (int)(s1 - s0), lds, acc);
    double *o = part + (long)blockIdx.y * part_stride;
    gemm_epilogue<S2_B, S2_B, T>(acc, [&](int r, int c, T v) {
        const int i = i0 + r, j = j0 + c;
        if (i < nao && j < nao) {
            double *dst = o + (long)i * nao + j;
            *dst = beta ? *dst + (double)v : (double)v;
        }
    });
}

// vk[i, j] (+)= sum_s part[s, I, J]; SYM reads the lower-tile copy for
// upper-tile entries.
__global__ __launch_bounds__(256) void k_finalize(
    const double *__restrict__ part, int ksplit, long part_stride, int nao, int sym,
    double *__restrict__ vk, long ldvk, int accumulate) {
    long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)nao * nao) return;
    int i = (int)(t / nao), j = (int)(t % nao);
    int I = i, J = j;
    if (sym && (i / S2_B) < (j / S2_B)) { I = j; J = i; }
    double s = 0.;
    for (int k = 0; k < ksplit; k++) s += part[k * part_stride + (long)I * nao + J];
    double *dst = vk + (long)i * ldvk + j;
    *dst = accumulate ? *dst + s : s;
}

struct KPlan {
    int ntile, ntiles_grid, ksplit, bn;
    long nL;            // aux rows per block
    size_t r_words;     // one R buffer
    size_t part_words;
};

// step-1 N tile: the width in {64, 128, 160} that pads nocc least (ties: wider)
int pick_bn(int nocc) {
    const int cand[3] = {160, 128, 64};
    int best = 64;
    long best_pad = 1L << 40;
    for (int b : cand) {
        long pad = (long)((nocc + b - 1) / b) * b - nocc;
        if (pad < best_pad) { best_pad = pad; best = b; }
    }
    return best;
}

// Resident step-2 workgroups on this device: 2 per CU (LDS and registers).
int resident_step2_wgs() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 512;
    if (!cached[dev]) {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
            cu = 256;
        cached[dev] = 2 * cu;
    }
    return cached[dev];
}

// Split of the (L, k) reduction: the ksplit in 1..32 whose grid fills whole
// rounds of resident workgroups best (ties: the larger split).
int pick_ksplit(int ntiles_grid) {
    const int res = resident_step2_wgs();
    int best = 1;
    double best_eff = -1.;
    for (int ks = 1; ks <= 32; ks++) {
        const long wgs = (long)ntiles_grid * ks;
        const long rounds = (wgs + res - 1) / res;
        const double eff = (double)wgs / (double)(rounds * res);
        if (eff >= best_eff - 1e-9) { best_eff = eff; best = ks; }
    }
    return best;
}

KPlan plan_k(long naux, int nao, int nocc, bool two, long max_rows) {
    KPlan p;
    p.ntile = (nao + S2_B - 1) / S2_B;
    p.ntiles_grid = two ? p.ntile * p.ntile : p.ntile * (p.ntile + 1) / 2;
    p.ksplit = pick_ksplit(p.ntiles_grid);
    p.bn = pick_bn(nocc);
    p.nL = max_rows < naux ? max_rows : naux;
    if (p.nL < 1) p.nL = 1;
    p.r_words = (size_t)p.nL * nao * nocc;
    p.part_words = (size_t)p.ksplit * nao * nao;
    return p;
}

template <int BN, typename T>
void launch_step1(dim3 g, hipStream_t st, const double *blk, long ld, const int *pmap, int nao,
                  const double *f, long ldf, int nocc, T *r, long ldr, long strideL) {
    if (pmap)
        hipLaunchKernelGGL((k_step1<BN, true, T>), g, dim3(G_THREADS), 0, st, blk, ld, pmap, nao, f, ldf,
                           nocc, r, ldr, strideL);
    else
        hipLaunchKernelGGL((k_step1<BN, false, T>), g, dim3(G_THREADS), 0, st, blk, ld, pmap, nao, f, ldf,
                           nocc, r, ldr, strideL);
}

template <typename T>
int step1(const KPlan &p, int nL, hipStream_t st, const double *blk, long ld, const int *pmap, int nao,
          const double *f, long ldf, int nocc, T *r, long ldr, long strideL) {
    dim3 g((unsigned)cdiv(nocc, p.bn), (unsigned)cdiv(nao, 128), (unsigned)nL);
    switch (p.bn) {
    case 160: launch_step1<160, T>(g, st, blk, ld, pmap, nao, f, ldf, nocc, r, ldr, strideL); break;
    case 128: launch_step1<128, T>(g, st, blk, ld, pmap, nao, f, ldf, nocc, r, ldr, strideL); break;
    default: launch_step1<64, T>(g, st, blk, ld, pmap, nao, f, ldf, nocc, r, ldr, strideL); break;
    }
    MDF_LAUNCHED("k_step1");
    return MDF_OK;
}

template <typename T>
int step2(const KPlan &p, int nL, bool two, hipStream_t st, const T *rl, const T *rr, int nao, int nocc,
          long ldr, long strideL, double *part, long part_stride, int beta) {
    dim3 g2((unsigned)p.ntiles_grid, (unsigned)p.ksplit);
    const bool vec = (ldr % (16 / (long)sizeof(T)) == 0) && ((uintptr_t)rl % 16 == 0) && ((uintptr_t)rr % 16 == 0);
    if (two && vec)
        hipLaunchKernelGGL((k_step2<false, true, T>), g2, dim3(G_THREADS), 0, st, rl, rr, nao, nocc, ldr,
                           strideL, nL, p.ntile, part, part_stride, beta);
    else if (two)
        hipLaunchKernelGGL((k_step2<false, false, T>), g2, dim3(G_THREADS), 0, st, rl, rr, nao, nocc, ldr,
                           strideL, nL, p.ntile, part, part_stride, beta);
    else if (vec)
        hipLaunchKernelGGL((k_step2<true, true, T>), g2, dim3(G_THREADS), 0, st, rl, rl, nao, nocc, ldr,
                           strideL, nL, p.ntile, part, part_stride, beta);
    else
        hipLaunchKernelGGL((k_step2<true, false, T>), g2, dim3(G_THREADS), 0, st, rl, rl, nao, nocc, ldr,
                           strideL, nL, p.ntile, part, part_stride, beta);
    MDF_LAUNCHED("k_step2");
    return MDF_OK;
}

// The block loop of mdf_get_k for element type T.
template <typename T>
int run_k(const KPlan &p, const double *cderi, long naux, long ld, const int *pair_map, int nao,
          const double *fl, const double *fr, int nocc, long ldf, double *vk, long ldvk, int accumulate,
          void *work, hipStream_t st) {
    const bool two = fr != nullptr;
    double *part = (double *)work;
    T *rl = (T *)(part + align_up(p.part_words, 64));
    T *rr = two ? rl + align_up(p.r_words, 64) : rl;
    const long ldr = nocc, strideL = (long)nao * nocc;
    const long part_stride = (long)nao * nao;
    for (long L0 = 0; L0 < naux; L0 += p.nL) {
        const int nL = (int)(naux - L0 < p.nL ? naux - L0 : p.nL);
        const double *blk = cderi + L0 * ld;
        int rc = step1<T>(p, nL, st, blk, ld, pair_map, nao, fl, ldf, nocc, rl, ldr, strideL);
        if (rc) return rc;
        if (two) {
            rc = step1<T>(p, nL, st, blk, ld, pair_map, nao, fr, ldf, nocc, rr, ldr, strideL);
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

}  // namespace

static size_t k_workspace(const KPlan &p, bool two, size_t elem) {
    return align_up(p.part_words, 64) * sizeof(double) + (two ? 2 : 1) * align_up(p.r_words, 64) * elem;
}

extern "C" size_t mdf_get_k_workspace(long naux, int nao, int nocc, int two_factors, int prec,
                                      long max_block_rows) {
    KPlan p = plan_k(naux, nao, nocc, two_factors != 0, max_block_rows);
    return k_workspace(p, two_factors != 0, prec == MDF_PREC_FP32 ? sizeof(float) : sizeof(double));
}

extern "C" int mdf_get_k(const double *cderi, long naux, long npairs, long ld, const int *pair_map,
                         int nao, const double *fl, const double *fr, int nocc, long ldf, int prec,
                         double *vk, long ldvk, int accumulate, void *work, size_t work_bytes,
                         void *stream) {
    MDF_CHECK_ARG(naux >= 0 && nao >= 0 && nocc >= 0, "mdf_get_k: bad sizes");
    MDF_CHECK_ARG(cderi && fl && vk, "mdf_get_k: null pointer");
    MDF_CHECK_ARG(ld >= npairs && ldf >= nocc && ldvk >= nao, "mdf_get_k: leading dimension too small");
    MDF_CHECK_ARG(prec == MDF_PREC_FP64 || prec == MDF_PREC_FP32, "mdf_get_k: unknown precision %d", prec);
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
    const size_t elem = prec == MDF_PREC_FP32 ? sizeof(float) : sizeof(double);
    KPlan p0 = plan_k(naux, nao, nocc, two, 1);
    size_t need1 = k_workspace(p0, two, elem);
    if (work_bytes < need1) {
        set_error("mdf_get_k: workspace %zu bytes < %zu required", work_bytes, need1);
        return MDF_EWORKSPACE;
    }
    size_t per_row = (size_t)nao * nocc * elem * (two ? 2 : 1);
    size_t part_bytes = align_up(p0.part_words, 64) * sizeof(double);
    long rows = (long)((work_bytes - part_bytes) / per_row);
    KPlan p = plan_k(naux, nao, nocc, two, rows);
    while (p.nL > 1 && k_workspace(p, two, elem) > work_bytes) p = plan_k(naux, nao, nocc, two, p.nL - 1);
    if (prec == MDF_PREC_FP32)
        return run_k<float>(p, cderi, naux, ld, pair_map, nao, fl, fr, nocc, ldf, vk, ldvk, accumulate, work, st);
    return run_k<double>(p, cderi, naux, ld, pair_map, nao, fl, fr, nocc, ldf, vk, ldvk, accumulate, work, st);
}
