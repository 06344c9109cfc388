// MFMA tile engine shared by the K contraction and the metric solve.
//
// One workgroup (256 threads = 4 wavefronts as 2 x 2) computes a BM x BN
// tile of C = sum_k A[m, k] B[k, n] over nk K-steps of BK = 16.  Operands are
// staged through double-buffered LDS (A as As[k][m], B as Bs[k][n]) with the
// next K-step's global loads held in registers while the current step's MFMAs
// run, so there is one barrier per K-step.  Loaders decide where a tile comes
// from (packed triangle, row-major matrix, transformed block ...) and convert
// to the engine's element type T on the way into LDS:
//
//   struct Loader {
//     using reg_t = ...;                      // register type (the source type)
//     static constexpr int N;                 // elements per thread per K-step
//     void load(int kt, reg_t (&r)[N]);       // global -> registers
//     void store(const reg_t (&r)[N], T *lds, int pitch);   // convert -> LDS
//   };
//
// Prefetched values stay in the source type until the store, so converting
// them never makes the current K-step's MFMAs wait on the prefetch.
//
// T = double: v_mfma_f64_16x16x4_f64; T = float: v_mfma_f32_16x16x4_f32.
// Lane layouts measured on gfx950 with one-hot probes: A lane = m + 16 k,
// B lane = n + 16 k for both; the result element v of lane l is row
// l/16 + 4 v (f64) or 4 (l/16) + v (f32), column l % 16.
#pragma once
#include "mdf_common.h"

typedef float float4_t __attribute__((ext_vector_type(4)));

// LDS (address space 3) pointers: operand reads through them compile to
// ds_read; through a generic pointer they become flat loads that wait on the
// in-flight global prefetch (s_waitcnt vmcnt(0)) and break the pipeline.
#define MDF_LDS __attribute__((address_space(3)))

namespace mdf {

constexpr int G_THREADS = 256;
constexpr int G_BK = 16;

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
    using acc_t = double4_t;
    __device__ __forceinline__ static acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ __forceinline__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

template <>
struct Mfma<float> {
    using acc_t = float4_t;
    __device__ __forceinline__ static acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ __forceinline__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

__device__ __forceinline__ double4_t mfma16(double a, double b, double4_t c) { return Mfma<double>::mma(a, b, c); }

// A workgroup of WGM x WGN wavefronts (64 WGM WGN threads) computes a
// BM x BN tile; wavefront w owns sub-tile (w % WGM, w / WGM).
template <int BM, int BN, int WGM = 2, int WGN = 2>
struct TileShape {
    static constexpr int THREADS = 64 * WGM * WGN;
    static constexpr int WM = BM / WGM, WN = BN / WGN;   // per-wave tile
    static constexpr int MI = WM / 16, NI = WN / 16;     // MFMA tiles per wave
    static_assert(MI * 16 == WM && NI * 16 == WN, "wave tile must be a multiple of 16");
    static constexpr int PA = BM + 16, PB = BN + 16;     // LDS pitches (elements)
    static constexpr int LDS_ELEMS = 2 * G_BK * (PA + PB);
    static constexpr int LDS_DOUBLES = LDS_ELEMS;
};

// Accumulator tile of one wavefront.
template <int MI, int NI, typename T = double>
struct Acc {
    typename Mfma<T>::acc_t v[MI][NI];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int a = 0; a < MI; a++)
#pragma unroll
            for (int b = 0; b < NI; b++) v[a][b] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    }
};

// Runs the K loop.  `lds` must hold TileShape<BM,BN,WGM,WGN>::LDS_ELEMS
// elements of T.
template <int BM, int BN, class LA, class LB, typename T = double, int WGM = 2, int WGN = 2>
__device__ __forceinline__ void gemm_mainloop(LA &la, LB &lb, int nk, T *lds,
                                     Acc<TileShape<BM, BN, WGM, WGN>::MI, TileShape<BM, BN, WGM, WGN>::NI, T> &acc) {
    using S = TileShape<BM, BN, WGM, WGN>;
    constexpr int SA = G_BK * S::PA, SB = G_BK * S::PB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    T *const Bbase = lds + 2 * SA;
    const MDF_LDS T *Ar = (const MDF_LDS T *)(lds + wm * S::WM + (lane & 15));
    const MDF_LDS T *Br = (const MDF_LDS T *)(Bbase + wn * S::WN + (lane & 15));
    typename LA::reg_t ra[LA::N];
    typename LB::reg_t rb[LB::N];
    if (nk <= 0) return;
    la.load(0, ra);
    lb.load(0, rb);
    la.store(ra, lds, S::PA);
    lb.store(rb, Bbase, S::PB);
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            la.load(kt + 1, ra);
            lb.load(kt + 1, rb);
        }
        const MDF_LDS T *A = Ar + cur * SA;
        const MDF_LDS T *B = Br + cur * SB;
#pragma unroll
        for (int kq = 0; kq < G_BK / 4; kq++) {
            const int kk = kq * 4 + (lane >> 4);
            T a[S::MI], b[S::NI];
#pragma unroll
            for (int mi = 0; mi < S::MI; mi++) a[mi] = A[kk * S::PA + mi * 16];
#pragma unroll
            for (int ni = 0; ni < S::NI; ni++) b[ni] = B[kk * S::PB + ni * 16];
#pragma unroll
            for (int mi = 0; mi < S::MI; mi++)
#pragma unroll
                for (int ni = 0; ni < S::NI; ni++) acc.v[mi][ni] = Mfma<T>::mma(a[mi], b[ni], acc.v[mi][ni]);
            // the next K-step's tile goes into the other buffer (read by no
            // wave since the last barrier) halfway through this step's MFMAs,
            // so the stores overlap the matrix pipe instead of following it
            if (kq == G_BK / 8 - 1 && more) {
                la.store(ra, lds + (cur ^ 1) * SA, S::PA);
                lb.store(rb, Bbase + (cur ^ 1) * SB, S::PB);
            }
        }
        __syncthreads();
    }
}

// Visits every accumulator element with its tile-local (row, col); the
// element-to-row layout is that of the MFMA type LT (an FP64 shadow of an
// FP32 accumulator keeps the FP32 layout).
template <int BM, int BN, typename LT, class AccT, class F, int WGM = 2, int WGN = 2>
__device__ __forceinline__ void gemm_epilogue_layout(const AccT &acc, F &&f) {
    using S = TileShape<BM, BN, WGM, WGN>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
#pragma unroll
    for (int mi = 0; mi < S::MI; mi++)
#pragma unroll
        for (int ni = 0; ni < S::NI; ni++) {
            const int c = wn * S::WN + ni * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = wm * S::WM + mi * 16 + Mfma<LT>::row(lane, v);
                f(r, c, acc.v[mi][ni][v]);
            }
        }
}

template <int BM, int BN, typename T = double, class F>
__device__ __forceinline__ void gemm_epilogue(const Acc<TileShape<BM, BN>::MI, TileShape<BM, BN>::NI, T> &acc, F &&f) {
    gemm_epilogue_layout<BM, BN, T>(acc, f);
}

// ---------------------------------------------------------------- loaders
// Row-major matrix M[row][k] (k contiguous, element type S) as the A or B
// operand: tile rows r0 .. r0+R-1, K-step kt covers k = kt*BK .. +BK-1.
// Out-of-range reads 0.
template <int R, typename T = double, typename Src = double, int THREADS = G_THREADS>
struct RowsLoader {
    using reg_t = Src;
    static constexpr int N = R * G_BK / THREADS;
    static_assert(N >= 1 && G_BK % N == 0 && (THREADS / (G_BK / N)) == R, "RowsLoader shape");
    const Src *base;      // &M[0][0]
    long ld;              // row stride
    int r0, nrows, k0, kmax;
    __device__ __forceinline__ void load(int kt, Src (&r)[N]) const {
        const int t = threadIdx.x;
        const int row = r0 + t / (G_BK / N);
        const int kb = k0 + kt * G_BK + (t % (G_BK / N)) * N;
        if (row < nrows && kb + N <= kmax) {
            // the whole chunk is in range: plain loads, no per-element select
            const Src *src = base + (long)row * ld + kb;
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = src[v];
            return;
        }
        // edge: unconditional loads at clamped (in-bounds) addresses, the
        // out-of-range elements selected to zero
        const Src *src = base + (long)min(row, nrows - 1) * ld;
#pragma unroll
        for (int v = 0; v < N; v++) {
            const Src x = src[min(kb + v, kmax - 1)];
            r[v] = (row < nrows && kb + v < kmax) ? x : (Src)0;
        }
    }
    __device__ __forceinline__ void store(const Src (&r)[N], T *lds, int pitch) const {
        const int t = threadIdx.x;
        const int row = t / (G_BK / N);
        const int kb = (t % (G_BK / N)) * N;
#pragma unroll
        for (int v = 0; v < N; v++) lds[(kb + v) * pitch + row] = (T)r[v];
    }
};

// Row-major matrix M[k][col] (col contiguous) as the B operand: tile columns
// c0 .. c0+C-1.
template <int C, typename T = double, typename Src = double, int THREADS = G_THREADS>
struct ColsLoader {
    using reg_t = Src;
    static constexpr int N = C * G_BK / THREADS;
    const Src *base;
    long ld;
    int c0, ncols, k0, kmax;
    __device__ __forceinline__ void load(int kt, Src (&r)[N]) const {
        const int t = threadIdx.x;
        // each thread takes N consecutive columns of one k row
        constexpr int TPR = C / N;              // threads per k row
        const int kk = t / TPR;
        const int cb = c0 + (t % TPR) * N;
        const int k = k0 + kt * G_BK + kk;
        if (kk < G_BK && k < kmax && cb + N <= ncols) {
            // the whole chunk is in range: plain loads, no per-element select
            const Src *src = base + (long)k * ld + cb;
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = src[v];
            return;
        }
        // edge: unconditional loads at clamped (in-bounds) addresses, the
        // out-of-range elements selected to zero
        const Src *src = base + (long)min(k, kmax - 1) * ld;
#pragma unroll
        for (int v = 0; v < N; v++) {
            const Src x = src[min(cb + v, ncols - 1)];
            r[v] = (kk < G_BK && k < kmax && cb + v < ncols) ? x : (Src)0;
        }
    }
    __device__ __forceinline__ void store(const Src (&r)[N], T *lds, int pitch) const {
        const int t = threadIdx.x;
        constexpr int TPR = C / N;
        const int kk = t / TPR;
        const int cb = (t % TPR) * N;
        if (kk < G_BK) {
#pragma unroll
            for (int v = 0; v < N; v++) lds[kk * pitch + cb + v] = (T)r[v];
        }
    }
};

}  // namespace mdf
