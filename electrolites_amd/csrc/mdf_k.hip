// Exchange (K) contraction over the packed cderi on FP64 MFMA.
//
// Restates electrolites/fastdf.py:974-1005 (unpack + step 1 + step 2) and
// :330-335 (_syrk):
//     R[L, i, k] = sum_j A_L[i, j] F[j, k]          (step 1, occupied transform)
//     K[i, j]   += sum_{L,k} R_l[L, i, k] R_r[L, j, k]   (step 2)
// where A_L is row L of the cderi unpacked to a symmetric nao x nao matrix.
// The reference materialises A_L with GPU4PySCF's fill_symmetric before the
// GEMM; here the unpack is fused into the step-1 A-operand load: each 128x16
// A tile is read straight from the packed triangle (row-contiguous below the
// diagonal, column-contiguous above it, so both halves load coalesced).
//
// Mode 0 (one factor, fr == NULL): step 2 is a symmetric rank-k update and
// only lower-triangle output tiles are computed.  Mode 1: general GEMM.
// Step 2 splits the (L, k) reduction across workgroups (ksplit) and the
// per-split partials are summed in a fixed order by k_finalize.
//
// MFMA: v_mfma_f64_16x16x4_f64 (A: lane = i + 16 k; B: lane = j + 16 k;
// D: lane = j + 16 (i % 4), element i / 4 -- row i = lane/16 + 4*element,
// measured on gfx950 with an in-bounds one-hot probe).
#include "mdf_common.h"
#include "../../include/mi355df.h"

namespace {

constexpr int K_THREADS = 256;

__device__ inline double4_t mfma_f64(double a, double b, double4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ inline long tri(long i) { return i * (i + 1) / 2; }

// ---------------------------------------------------------------- step 1
constexpr int S1_BM = 128, S1_BN = 64, S1_BK = 16;
constexpr int S1_PA = S1_BM + 2, S1_PB = S1_BN + 2;   // LDS pitches (doubles)

template <bool MAPPED>
__global__ __launch_bounds__(K_THREADS) void k_step1(
    const double *__restrict__ cderi, long ld, const int *__restrict__ pmap, int nao,
    const double *__restrict__ F, long ldf, int nocc,
    double *__restrict__ R, long ldr, long strideL) {
    __shared__ double As[S1_BK][S1_PA];
    __shared__ double Bs[S1_BK][S1_PB];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int k0 = blockIdx.x * S1_BN, i0 = blockIdx.y * S1_BM;
    const double *Arow = cderi + (long)blockIdx.z * ld;

    double4_t acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = double4_t{0., 0., 0., 0.};

    for (int j0 = 0; j0 < nao; j0 += S1_BK) {
        // ---- A tile (i0..i0+127, j0..j0+15) from the packed triangle
        if (!MAPPED && j0 >= i0 + S1_BM - 1) {
            // every j > every i: element (i, j) at tri(j) + i, contiguous in i
            const int jj = t >> 4, ib = (t & 15) * 8;
            const int j = j0 + jj;
            const double *src = Arow + tri(j) + i0 + ib;
#pragma unroll
            for (int v = 0; v < 8; v++) {
                int i = i0 + ib + v;
                As[jj][ib + v] = (j < nao && i < nao) ? src[v] : 0.;
            }
        } else {
            const int ii = t >> 1, jb = (t & 1) * 8;
            const int i = i0 + ii;
#pragma unroll
            for (int v = 0; v < 8; v++) {
                int j = j0 + jb + v;
                double x = 0.;
                if (i < nao && j < nao) {
                    if (MAPPED) {
                        int p = pmap[(long)i * nao + j];
                        x = p >= 0 ? Arow[p] : 0.;
                    } else {
                        x = Arow[i >= j ? tri(i) + j : tri(j) + i];
                    }
                }
                As[jb + v][ii] = x;
            }
        }
        // ---- B tile: F[j0..j0+15][k0..k0+63]
        {
            const int jj = t >> 4, kb = (t & 15) * 4;
            const int j = j0 + jj;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int k = k0 + kb + v;
                Bs[jj][kb + v] = (j < nao && k < nocc) ? F[(long)j * ldf + k] : 0.;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kq = 0; kq < S1_BK / 4; kq++) {
            const int kk = kq * 4 + (lane >> 4);
            double a[4], b[2];
#pragma unroll
            for (int mi = 0; mi < 4; mi++) a[mi] = As[kk][wm * 64 + mi * 16 + (lane & 15)];
#pragma unroll
            for (int ni = 0; ni < 2; ni++) b[ni] = Bs[kk][wn * 32 + ni * 16 + (lane & 15)];
#pragma unroll
            for (int mi = 0; mi < 4; mi++)
#pragma unroll
                for (int ni = 0; ni < 2; ni++) acc[mi][ni] = mfma_f64(a[mi], b[ni], acc[mi][ni]);
        }
        __syncthreads();
    }
    double *Rl = R + (long)blockIdx.z * strideL;
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++) {
            const int k = k0 + wn * 32 + ni * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = i0 + wm * 64 + mi * 16 + (lane >> 4) + 4 * v;
                if (i < nao && k < nocc) Rl[(long)i * ldr + k] = acc[mi][ni][v];
            }
        }
}

// ---------------------------------------------------------------- step 2
constexpr int S2_B = 64, S2_BK = 16, S2_P = S2_B + 2;

template <bool SYM>
__global__ __launch_bounds__(K_THREADS) void k_step2(
    const double *__restrict__ Rl, const double *__restrict__ Rr, int nao, int nocc,
    long ldr, long strideL, int nL, int ntile, double *__restrict__ part, long part_stride,
    int beta) {
    __shared__ double As[S2_BK][S2_P];
    __shared__ double Bs[S2_BK][S2_P];
    int ti, tj;
    if (SYM) {
        // blockIdx.x -> (ti >= tj) in row-major lower-triangle order
        int x = blockIdx.x;
        ti = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
        while (ti * (ti + 1) / 2 > x) ti--;
        while ((ti + 1) * (ti + 2) / 2 <= x) ti++;
        tj = x - ti * (ti + 1) / 2;
    } else {
        ti = blockIdx.x / ntile;
        tj = blockIdx.x % ntile;
    }
    const int i0 = ti * S2_B, j0 = tj * S2_B;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    const int kchunks = (nocc + S2_BK - 1) / S2_BK;
    const long nsteps = (long)nL * kchunks;
    const long s0 = nsteps * blockIdx.y / gridDim.y, s1 = nsteps * (blockIdx.y + 1) / gridDim.y;

    double4_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = double4_t{0., 0., 0., 0.};

    const int ii = t >> 2, kb = (t & 3) * 4;
    for (long s = s0; s < s1; s++) {
        const long L = s / kchunks;
        const int kc = (int)(s % kchunks) * S2_BK;
        const double *a = Rl + L * strideL;
        const double *b = Rr + L * strideL;
        const int ia = i0 + ii, jb = j0 + ii;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k = kc + kb + v;
            As[kb + v][ii] = (ia < nao && k < nocc) ? a[(long)ia * ldr + k] : 0.;
            Bs[kb + v][ii] = (jb < nao && k < nocc) ? b[(long)jb * ldr + k] : 0.;
        }
        __syncthreads();
#pragma unroll
        for (int kq = 0; kq < S2_BK / 4; kq++) {
            const int kk = kq * 4 + (lane >> 4);
            double av[2], bv[2];
#pragma unroll
            for (int mi = 0; mi < 2; mi++) av[mi] = As[kk][wm * 32 + mi * 16 + (lane & 15)];
#pragma unroll
            for (int ni = 0; ni < 2; ni++) bv[ni] = Bs[kk][wn * 32 + ni * 16 + (lane & 15)];
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
#pragma unroll
                for (int ni = 0; ni < 2; ni++) acc[mi][ni] = mfma_f64(av[mi], bv[ni], acc[mi][ni]);
        }
        __syncthreads();
    }
    double *o = part + (long)blockIdx.y * part_stride;
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int ni = 0; ni < 2; ni++) {
            const int j = j0 + wn * 32 + ni * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = i0 + wm * 32 + mi * 16 + (lane >> 4) + 4 * v;
                if (i < nao && j < nao) {
                    double *dst = o + (long)i * nao + j;
                    *dst = beta ? *dst + acc[mi][ni][v] : acc[mi][ni][v];
                }
            }
        }
}

// vk[i, j] (+)= sum_s part[s, I, J]; SYM reads the lower-tile copy for
// upper-tile entries.
__global__ __launch_bounds__(K_THREADS) void k_finalize(
    const double *__restrict__ part, int ksplit, long part_stride, int nao, int sym,
    double *__restrict__ vk, long ldvk, int accumulate) {
    long t = (long)blockIdx.x * K_THREADS + threadIdx.x;
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
    int ntile, ntiles_grid, ksplit;
    long nL;            // aux rows per block
    size_t r_words;     // one R buffer
    size_t part_words;
};

KPlan plan_k(long naux, int nao, int nocc, bool two, long max_rows) {
    KPlan p;
    p.ntile = (nao + S2_B - 1) / S2_B;
    p.ntiles_grid = two ? p.ntile * p.ntile : p.ntile * (p.ntile + 1) / 2;
    p.ksplit = 1;
    while (p.ntiles_grid * p.ksplit < 1024 && p.ksplit < 16) p.ksplit *= 2;
    p.nL = max_rows < naux ? max_rows : naux;
    if (p.nL < 1) p.nL = 1;
    p.r_words = (size_t)p.nL * nao * nocc;
    p.part_words = (size_t)p.ksplit * nao * nao;
    return p;
}

}  // namespace

extern "C" size_t mdf_get_k_workspace(long naux, int nao, int nocc, int two_factors, long max_block_rows) {
    KPlan p = plan_k(naux, nao, nocc, two_factors != 0, max_block_rows);
    return (mdf::align_up(p.part_words, 64) + (two_factors ? 2 : 1) * mdf::align_up(p.r_words, 64)) *
           sizeof(double);
}

extern "C" int mdf_get_k(const double *cderi, long naux, long npairs, long ld, const int *pair_map,
                         int nao, const double *fl, const double *fr, int nocc, long ldf,
                         double *vk, long ldvk, int accumulate, void *work, size_t work_bytes,
                         void *stream) {
    MDF_CHECK_ARG(naux >= 0 && nao >= 0 && nocc >= 0, "mdf_get_k: bad sizes");
    MDF_CHECK_ARG(cderi && fl && vk, "mdf_get_k: null pointer");
    MDF_CHECK_ARG(ld >= npairs && ldf >= nocc && ldvk >= nao, "mdf_get_k: leading dimension too small");
    MDF_CHECK_ARG(pair_map || npairs == (long)nao * (nao + 1) / 2,
                  "mdf_get_k: without a pair map the cderi must hold the full packed triangle "
                  "(npairs=%ld, nao=%d)", npairs, nao);
    hipStream_t st = mdf::as_stream(stream);
    if (nao == 0) return MDF_OK;
    const bool two = fr != nullptr;
    if (nocc == 0 || naux == 0) {
        if (!accumulate)
            for (int i = 0; i < nao; i++) MDF_HIP(hipMemsetAsync(vk + (long)i * ldvk, 0, nao * sizeof(double), st));
        return MDF_OK;
    }
    MDF_CHECK_ARG(work != nullptr, "mdf_get_k: workspace is null");
    // largest aux block that fits the workspace
    KPlan p0 = plan_k(naux, nao, nocc, two, 1);
    size_t part_bytes = mdf::align_up(p0.part_words, 64) * sizeof(double);
    size_t per_row = (size_t)nao * nocc * sizeof(double) * (two ? 2 : 1);
    if (work_bytes < part_bytes + per_row + 1024) {
        mdf::set_error("mdf_get_k: workspace %zu bytes < %zu required", work_bytes,
                       part_bytes + per_row + 1024);
        return MDF_EWORKSPACE;
    }
    long rows = (long)((work_bytes - part_bytes - 1024) / per_row);
    KPlan p = plan_k(naux, nao, nocc, two, rows);
    while (p.nL > 1 && mdf_get_k_workspace(naux, nao, nocc, two, p.nL) > work_bytes) p = plan_k(naux, nao, nocc, two, p.nL - 1);

    double *part = (double *)work;
    double *rl = part + mdf::align_up(p.part_words, 64);
    double *rr = two ? rl + mdf::align_up(p.r_words, 64) : rl;
    const long ldr = nocc, strideL = (long)nao * nocc;
    const long part_stride = (long)nao * nao;

    for (long L0 = 0; L0 < naux; L0 += p.nL) {
        const int nL = (int)(naux - L0 < p.nL ? naux - L0 : p.nL);
        dim3 g1((unsigned)mdf::cdiv(nocc, S1_BN), (unsigned)mdf::cdiv(nao, S1_BM), (unsigned)nL);
        const double *blk = cderi + L0 * ld;
        if (pair_map)
            hipLaunchKernelGGL(k_step1<true>, g1, dim3(K_THREADS), 0, st, blk, ld, pair_map, nao, fl, ldf,
                               nocc, rl, ldr, strideL);
        else
            hipLaunchKernelGGL(k_step1<false>, g1, dim3(K_THREADS), 0, st, blk, ld, pair_map, nao, fl, ldf,
                               nocc, rl, ldr, strideL);
        MDF_LAUNCHED("k_step1");
        if (two) {
            if (pair_map)
                hipLaunchKernelGGL(k_step1<true>, g1, dim3(K_THREADS), 0, st, blk, ld, pair_map, nao, fr, ldf,
                                   nocc, rr, ldr, strideL);
            else
                hipLaunchKernelGGL(k_step1<false>, g1, dim3(K_THREADS), 0, st, blk, ld, pair_map, nao, fr, ldf,
                                   nocc, rr, ldr, strideL);
            MDF_LAUNCHED("k_step1");
        }
        dim3 g2((unsigned)p.ntiles_grid, (unsigned)p.ksplit);
        const int beta = L0 > 0;
        if (two)
            hipLaunchKernelGGL(k_step2<false>, g2, dim3(K_THREADS), 0, st, rl, rr, nao, nocc, ldr, strideL, nL,
                               p.ntile, part, part_stride, beta);
        else
            hipLaunchKernelGGL(k_step2<true>, g2, dim3(K_THREADS), 0, st, rl, rl, nao, nocc, ldr, strideL, nL,
                               p.ntile, part, part_stride, beta);
        MDF_LAUNCHED("k_step2");
    }
    long nt = (long)nao * nao;
    hipLaunchKernelGGL(k_finalize, dim3((unsigned)mdf::cdiv(nt, (long)K_THREADS)), dim3(K_THREADS), 0, st,
                       part, p.ksplit, part_stride, nao, two ? 0 : 1, vk, ldvk, accumulate);
    MDF_LAUNCHED("k_finalize");
    return MDF_OK;
}
