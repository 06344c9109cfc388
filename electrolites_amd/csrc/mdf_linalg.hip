// FP64 dense linear algebra of the cderi build on the MFMA tile engine.
//
// The reference decomposes the metric and applies its inverse factor with
// cuSOLVER/cuBLAS (electrolites/fastdfcderi.py:173-215 _decompose_j2c,
// :479-481 explicit L^-1, :237-261 _trmm_right_lowerT).  Here:
//   mdf_gemm        C = alpha op(A) op(B) + beta C (row-major; B may be
//                   transposed; A may be lower triangular; C may be
//                   restricted to its lower 128x128 tiles)
//   mdf_potrf       blocked right-looking Cholesky A = L L^T (lower, in
//                   place), 128-wide panels: one-workgroup diagonal block
//                   factorisation + inverse, panel = A21 inv(L11)^T, trailing
//                   SYRK on the engine.  A non-positive pivot sets *info.
//   mdf_trtri       X = L^-1 by block rows from the diagonal inverses
//   mdf_trmm        C = L^-1 B with the triangular A bound folded into the
//                   K loop (the cderi = L^-1 (P|pq) product).
#include <cmath>

#include "mdf_common.h"
#include "mdf_gemm.h"
#include "../../include/mi355df.h"

using namespace mdf;

namespace {

constexpr int LB = 128;   // panel / tile width

// Lower 128-tile x of a square result: (ti, tj) with tj <= ti, row-major
// over the lower triangle of tiles.
__device__ __forceinline__ void lower_tile(int x, int &ti, int &tj) {
    ti = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
    while (ti * (ti + 1) / 2 > x) ti--;
    while ((ti + 1) * (ti + 2) / 2 <= x) ti++;
    tj = x - ti * (ti + 1) / 2;
}

// TRI: 0 = A general; 1 = A lower triangular (row block m reads k < m0 + 128);
// 2 = A upper triangular (row block m reads k >= m0).  blockIdx.z is the
// batch index (operand strides sA, sB, sC).
template <int BN, bool BT, int TRI, bool LOWER_C>
__global__ __launch_bounds__(G_THREADS, 2) void gemm_f64_kernel(
    int M, int N, int K, double alpha, const double *__restrict__ A, long lda, long sA,
    const double *__restrict__ B, long ldb, long sB, double beta, double *C, long ldc, long sC) {
    using S = TileShape<128, BN>;
    __shared__ double lds[S::LDS_DOUBLES];
    A += (long)blockIdx.z * sA;
    B += (long)blockIdx.z * sB;
    C += (long)blockIdx.z * sC;
    int ti, tj;
    if (LOWER_C) {
        lower_tile(blockIdx.x, ti, tj);
    } else {
        ti = blockIdx.y;
        tj = blockIdx.x;
    }
    const int m0 = ti * 128, n0 = tj * BN;
    const int k0 = TRI == 2 ? min(K, m0) : 0;
    const int kmax = TRI == 1 ? min(K, m0 + 128) : K;
    RowsLoader<128> la{A, lda, m0, M, k0, kmax};
    Acc<S::MI, S::NI> acc;
    acc.zero();
    const int nk = (kmax - k0 + G_BK - 1) / G_BK;
    if (BT) {
        RowsLoader<BN> lb{B, ldb, n0, N, k0, kmax};
        gemm_mainloop<128, BN>(la, lb, nk, lds, acc);
    } else {
        ColsLoader<BN> lb{B, ldb, n0, N, k0, kmax};
        gemm_mainloop<128, BN>(la, lb, nk, lds, acc);
    }
    gemm_epilogue<128, BN>(acc, [&](int r, int c, double v) {
        const int m = m0 + r, n = n0 + c;
        if (m < M && n < N) {
            double *dst = C + (long)m * ldc + n;
            *dst = beta == 0. ? alpha * v : alpha * v + beta * *dst;
        }
    });
}

// C = alpha A B + beta C with 128 x 320 tiles on 2 x 4 wavefronts (B not
// transposed, A general): for N in (256, 320] one column tile covers C's
// width, where 128-wide tiles would compute 384 columns and read every A
// tile three times (the gradient's F_r^T R_L, N = nocc).
__global__ __launch_bounds__(512) void gemm_f64_wide(int M, int N, int K, double alpha, const double *__restrict__ A,
                                                     long lda, long sA, const double *__restrict__ B, long ldb,
                                                     long sB, double beta, double *C, long ldc, long sC) {
    using S = TileShape<128, 320, 2, 4>;
    __shared__ double lds[S::LDS_DOUBLES];
    A += (long)blockIdx.z * sA;
    B += (long)blockIdx.z * sB;
    C += (long)blockIdx.z * sC;
    const int m0 = blockIdx.y * 128;
    RowsLoader<128, double, double, 512> la{A, lda, m0, M, 0, K};
    ColsLoader<320, double, double, 512> lb{B, ldb, 0, N, 0, K};
    Acc<S::MI, S::NI> acc;
    acc.zero();
    gemm_mainloop<128, 320, decltype(la), decltype(lb), double, 2, 4>(la, lb, (K + G_BK - 1) / G_BK, lds, acc);
    auto emit = [&](int r, int c, double v) {
        const int m = m0 + r;
        if (m < M && c < N) {
            double *dst = C + (long)m * ldc + c;
            *dst = beta == 0. ? alpha * v : alpha * v + beta * *dst;
        }
    };
    gemm_epilogue_layout<128, 320, double, decltype(acc), decltype(emit) &, 2, 4>(acc, emit);
}

// The lower 128-tiles of P_z = A_z op(B_z) (n x n), put at the stored
// pairs: W[z ldw + c] = base + alpha P_z[i][j] with c = pcol[i (i + 1) / 2 + j]
// for n > i >= j with c >= 0, base = iscale[z] ivals[c] when iscale is given
// (W need not hold anything), else W's own value.  No P_z is written.  Every
// column of the tile is looked up first, then every base value is loaded
// (all loads in flight together: the columns of one tile are distinct, so
// no load depends on a store of the same tile), then every store issues.
template <bool BT>
__global__ __launch_bounds__(G_THREADS, 2) void gemm_pairs_kernel(int n, int K, double alpha,
                                                                  const double *__restrict__ A, long lda, long sA,
                                                                  const double *__restrict__ B, long ldb, long sB,
                                                                  const int *__restrict__ pcol, double *W, long ldw,
                                                                  const double *__restrict__ iscale,
                                                                  const double *__restrict__ ivals) {
    using S = TileShape<128, 128>;
    __shared__ double lds[S::LDS_DOUBLES];
    A += (long)blockIdx.z * sA;
    B += (long)blockIdx.z * sB;
    W += (long)blockIdx.z * ldw;
    int ti, tj;
    lower_tile(blockIdx.x, ti, tj);
    const int m0 = ti * 128, n0 = tj * 128;
    RowsLoader<128> la{A, lda, m0, n, 0, K};
    Acc<S::MI, S::NI> acc;
    acc.zero();
    const int nk = (K + G_BK - 1) / G_BK;
    if (BT) {
        RowsLoader<128> lb{B, ldb, n0, n, 0, K};
        gemm_mainloop<128, 128>(la, lb, nk, lds, acc);
    } else {
        ColsLoader<128> lb{B, ldb, n0, n, 0, K};
        gemm_mainloop<128, 128>(la, lb, nk, lds, acc);
    }
    // per row group of 16 x 64 elements: every column's byte offset
    // (32-bit, so the loads and stores take the uniform row base plus one
    // register each; ~0 for pairs not stored), then every base value (the
    // loads in flight together: the columns of one tile are distinct), then
    // every store
    constexpr unsigned NONE = ~0u;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % 2, wn = wave / 2;
    const double s = iscale ? iscale[blockIdx.z] : 0.;
    const char *vb = (const char *)ivals;
    char *wb = (char *)W;
#pragma unroll
    for (int mi = 0; mi < S::MI; mi++) {
        unsigned off[S::NI][4];
#pragma unroll
        for (int ni = 0; ni < S::NI; ni++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = m0 + wm * S::WM + mi * 16 + Mfma<double>::row(lane, v);
                const int j = n0 + wn * S::WN + ni * 16 + (lane & 15);
                // the packed index's byte offset fits 32 bits (checked at entry)
                const unsigned po = (unsigned)(i * (i + 1) / 2 + j) * 4u;
                const int c = (i < n && j <= i) ? *(const int *)((const char *)pcol + po) : -1;
                off[ni][v] = c >= 0 ? (unsigned)c * 8u : NONE;
            }
#pragma unroll
        for (int ni = 0; ni < S::NI; ni++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const unsigned o = off[ni][v];
                double base = 0.;
                if (o != NONE) base = iscale ? s * *(const double *)(vb + o) : *(const double *)(wb + o);
                acc.v[mi][ni][v] = base + alpha * acc.v[mi][ni][v];
            }
#pragma unroll
        for (int ni = 0; ni < S::NI; ni++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (off[ni][v] != NONE) *(double *)(wb + off[ni][v]) = acc.v[mi][ni][v];
    }
}

// Packed upper triangles (row-major over a <= b: the offset of (a, b) is
// a n - a (a - 1) / 2 + b - a) of n x n matrices, by 32 x 32 tile pairs
// (ta <= tb) staged in LDS so that both the rows and the columns of G are
// read along rows.  blockIdx.x is the tile pair, blockIdx.y strides the
// batch.
constexpr int SY_T = 32;

__device__ __forceinline__ long upper_offset(long n, long a, long b) { return a * n - a * (a - 1) / 2 + (b - a); }

__device__ __forceinline__ void upper_tile(int x, int &ta, int &tb) {
    lower_tile(x, tb, ta);   // ta <= tb
}

// P_z[(a, b)] = s_ab (G_z[a][b] + G_z[b][a]), s = dscale on the diagonal,
// oscale off it.
__global__ __launch_bounds__(256) void sym_pack_kernel(int n, const double *__restrict__ G, long sG,
                                                       double *__restrict__ P, long sP, long batch, double dscale,
                                                       double oscale) {
    __shared__ double tab[SY_T][SY_T + 1], tba[SY_T][SY_T + 1];
    int ta, tb;
    upper_tile(blockIdx.x, ta, tb);
    const int a0 = ta * SY_T, b0 = tb * SY_T;
    for (long z = blockIdx.y; z < batch; z += gridDim.y) {
        const double *g = G + z * sG;
        double *p = P + z * sP;
#pragma unroll
        for (int e = threadIdx.x; e < SY_T * SY_T; e += 256) {
            const int r = e / SY_T, c = e % SY_T;
            tab[r][c] = (a0 + r < n && b0 + c < n) ? g[(long)(a0 + r) * n + b0 + c] : 0.;
            tba[r][c] = (b0 + r < n && a0 + c < n) ? g[(long)(b0 + r) * n + a0 + c] : 0.;
        }
        __syncthreads();
#pragma unroll
        for (int e = threadIdx.x; e < SY_T * SY_T; e += 256) {
            const int r = e / SY_T, c = e % SY_T;
            const int a = a0 + r, b = b0 + c;
            if (a <= b && b < n) p[upper_offset(n, a, b)] = (a == b ? dscale : oscale) * (tab[r][c] + tba[c][r]);
        }
        __syncthreads();
    }
}

// G_z[a][b] = G_z[b][a] = s_ab P_z[(a, b)] for a <= b.
__global__ __launch_bounds__(256) void sym_unpack_kernel(int n, const double *__restrict__ P, long sP,
                                                         double *__restrict__ G, long sG, long batch, double dscale,
                                                         double oscale) {
    __shared__ double t[SY_T][SY_T + 1];
    int ta, tb;
    upper_tile(blockIdx.x, ta, tb);
    const int a0 = ta * SY_T, b0 = tb * SY_T;
    for (long z = blockIdx.y; z < batch; z += gridDim.y) {
        const double *p = P + z * sP;
        double *g = G + z * sG;
#pragma unroll
        for (int e = threadIdx.x; e < SY_T * SY_T; e += 256) {
            const int r = e / SY_T, c = e % SY_T;
            const int a = a0 + r, b = b0 + c;
            t[r][c] = (a <= b && b < n) ? (a == b ? dscale : oscale) * p[upper_offset(n, a, b)] : 0.;
        }
        __syncthreads();
#pragma unroll
        for (int e = threadIdx.x; e < SY_T * SY_T; e += 256) {
            const int r = e / SY_T, c = e % SY_T;
            // rows of the a-block: (a0 + r, b0 + c), from the tile or, on a
            // diagonal tile below its diagonal, from the mirrored entry
            if (a0 + r < n && b0 + c < n) g[(long)(a0 + r) * n + b0 + c] = (ta < tb || r <= c) ? t[r][c] : t[c][r];
            // rows of the b-block: (b0 + r, a0 + c) = (a0 + c, b0 + r) mirrored
            if (ta < tb && b0 + r < n && a0 + c < n) g[(long)(b0 + r) * n + a0 + c] = t[c][r];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ SP2
// One step of SP2 purification of a symmetric n x n X on 64-tiles of its
// lower triangle: Y = X^2 or Y = 2X - X^2, whichever trace lies closer to
// nocc, with tr(X^2) = ||X||_F^2 (X symmetric) known before the product, so
// the choice, the product, the update and the next step's sums are one
// kernel.  Each lower tile writes its block of Y and the mirrored block, and
// its share of tr(Y) and ||Y||_F^2 into part_out[2 tile]; the next step's
// blocks all sum part_in in one fixed order (no atomics: every block takes
// the same branch and the sums are reproducible).
constexpr int SP2_T = 64;

// sum of part[2 t + q] over t < ntiles, the same order in every block
__device__ __forceinline__ double2 sp2_sums(const double *part, int ntiles, double *red) {
    double a = 0., b = 0.;
    for (int t = threadIdx.x; t < ntiles; t += G_THREADS) {
        a += part[2 * t];
        b += part[2 * t + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[2 * w] = a;
        red[2 * w + 1] = b;
    }
    __syncthreads();
    double2 r{0., 0.};
    for (int i = 0; i < G_THREADS / 64; i++) {
        r.x += red[2 * i];
        r.y += red[2 * i + 1];
    }
    __syncthreads();
    return r;
}

// the block's (tr, ||.||_F^2 share) into part[2 tile]
__device__ __forceinline__ void sp2_put(double tr, double fro, double *part, int tile, double *red) {
    for (int off = 32; off > 0; off >>= 1) {
        tr += __shfl_xor(tr, off, 64);
        fro += __shfl_xor(fro, off, 64);
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[2 * w] = tr;
        red[2 * w + 1] = fro;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0., b = 0.;
        for (int i = 0; i < G_THREADS / 64; i++) {
            a += red[2 * i];
            b += red[2 * i + 1];
        }
        part[2 * tile] = a;
        part[2 * tile + 1] = b;
    }
}

// the partial sums of X itself (the first step's input)
__global__ __launch_bounds__(G_THREADS) void sp2_norms_kernel(int n, const double *__restrict__ X, long ldx,
                                                              double *__restrict__ part) {
    __shared__ double red[2 * G_THREADS / 64];
    int ti, tj;
    lower_tile(blockIdx.x, ti, tj);
    const int m0 = ti * SP2_T, n0 = tj * SP2_T;
    const double wgt = ti == tj ? 1. : 2.;
    double tr = 0., fro = 0.;
    for (int e = threadIdx.x; e < SP2_T * SP2_T; e += G_THREADS) {
        const int i = m0 + e / SP2_T, j = n0 + e % SP2_T;
        if (i < n && j < n) {
            const double x = X[(long)i * ldx + j];
            fro += wgt * x * x;
            if (i == j) tr += x;
        }
    }
    sp2_put(tr, fro, part, blockIdx.x, red);
}

__global__ __launch_bounds__(G_THREADS, 2) void sp2_step_kernel(int n, double nocc, const double *__restrict__ X,
                                                                long ldx, const double *__restrict__ part_in,
                                                                double *__restrict__ Y, long ldy,
                                                                double *__restrict__ part_out, int ntiles) {
    using S = TileShape<SP2_T, SP2_T>;
    __shared__ double lds[S::LDS_DOUBLES];
    int ti, tj;
    lower_tile(blockIdx.x, ti, tj);
    const int m0 = ti * SP2_T, n0 = tj * SP2_T;
    const double2 sums = sp2_sums(part_in, ntiles, lds);
    const double trX = sums.x, trX2 = sums.y;
    const bool up = fabs(trX2 - nocc) >= fabs(2. * trX - trX2 - nocc);
    // X X = X X^T: both operands are rows of X
    RowsLoader<SP2_T> la{X, ldx, m0, n, 0, n};
    RowsLoader<SP2_T> lb{X, ldx, n0, n, 0, n};
    Acc<S::MI, S::NI> acc;
    acc.zero();
    gemm_mainloop<SP2_T, SP2_T>(la, lb, (n + G_BK - 1) / G_BK, lds, acc);
    const double wgt = ti == tj ? 1. : 2.;
    double tr = 0., fro = 0.;
    gemm_epilogue<SP2_T, SP2_T>(acc, [&](int r, int c, double v) {
        const int i = m0 + r, j = n0 + c;
        if (i < n && j < n) {
            const double y = up ? 2. * X[(long)i * ldx + j] - v : v;
            Y[(long)i * ldy + j] = y;
            if (ti != tj) Y[(long)j * ldy + i] = y;
            fro += wgt * y * y;
            if (i == j) tr += y;
        }
    });
    sp2_put(tr, fro, part_out, blockIdx.x, lds);
}

constexpr long MAX_GRID_Z = 65535;

template <int BN, bool BT, int TRI, bool LOWER_C>
int launch_gemm(int M, int N, int K, double alpha, const double *A, long lda, long sA, const double *B, long ldb,
                long sB, double beta, double *C, long ldc, long sC, long batch, hipStream_t st) {
    if (M <= 0 || N <= 0 || batch <= 0) return MDF_OK;
    for (long b0 = 0; b0 < batch; b0 += MAX_GRID_Z) {
        const unsigned nz = (unsigned)(batch - b0 < MAX_GRID_Z ? batch - b0 : MAX_GRID_Z);
        dim3 g;
        if (LOWER_C) {
            const int t = (M + 127) / 128;
            g = dim3((unsigned)(t * (t + 1) / 2), 1, nz);
        } else {
            g = dim3((unsigned)((N + BN - 1) / BN), (unsigned)((M + 127) / 128), nz);
        }
        hipLaunchKernelGGL((gemm_f64_kernel<BN, BT, TRI, LOWER_C>), g, dim3(G_THREADS), 0, st, M, N, K, alpha,
                           A + b0 * sA, lda, sA, B + b0 * sB, ldb, sB, beta, C + b0 * sC, ldc, sC);
        MDF_LAUNCHED("gemm_f64_kernel");
    }
    return MDF_OK;
}

int launch_wide(int M, int N, int K, double alpha, const double *A, long lda, long sA, const double *B, long ldb,
                long sB, double beta, double *C, long ldc, long sC, long batch, hipStream_t st) {
    if (M <= 0 || N <= 0 || batch <= 0) return MDF_OK;
    for (long b0 = 0; b0 < batch; b0 += MAX_GRID_Z) {
        const unsigned nz = (unsigned)(batch - b0 < MAX_GRID_Z ? batch - b0 : MAX_GRID_Z);
        hipLaunchKernelGGL(gemm_f64_wide, dim3(1, (unsigned)((M + 127) / 128), nz), dim3(512), 0, st, M, N, K, alpha,
                           A + b0 * sA, lda, sA, B + b0 * sB, ldb, sB, beta, C + b0 * sC, ldc, sC);
        MDF_LAUNCHED("gemm_f64_wide");
    }
    return MDF_OK;
}

template <bool BT, bool LOWER_C>
int gemm_tri(int tri, int m, int n, int k, double alpha, const double *A, long lda, long sA, const double *B,
             long ldb, long sB, double beta, double *C, long ldc, long sC, long batch, hipStream_t st) {
    if (tri == 1)
        return launch_gemm<128, BT, 1, LOWER_C>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, st);
    if (tri == 2)
        return launch_gemm<128, BT, 2, LOWER_C>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, st);
    return launch_gemm<128, BT, 0, LOWER_C>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, st);
}

// Cholesky of one diagonal block (n <= 128, lower, in place) by one
// workgroup of 1024 threads, then its inverse into D (ldd = 128).  Both
// triangles live in LDS, packed (i(i+1)/2 + j, 2 x 66 KB), so the 128
// dependent steps of each phase touch LDS only.  A non-positive pivot
// records (global index + 1) in *info and continues with pivot 1.
constexpr int DIAG_THREADS = 1024;
constexpr int DIAG_PARTS = DIAG_THREADS / LB;     // partial sums per inverse entry

__device__ __forceinline__ int ptri(int i) { return i * (i + 1) / 2; }

__global__ __launch_bounds__(DIAG_THREADS) void potrf_diag_kernel(double *A, long lda, int n, int k0,
                                                                  double *D, int *info) {
    __shared__ double Lp[LB * (LB + 1) / 2];
    __shared__ double Xp[LB * (LB + 1) / 2];
    __shared__ double red[DIAG_PARTS][LB];
    const int t = threadIdx.x;
    for (int e = t; e < n * n; e += DIAG_THREADS) {
        const int i = e / n, j = e % n;
        if (j <= i) Lp[ptri(i) + j] = A[(long)i * lda + j];
    }
    // right-looking factorisation
    for (int k = 0; k < n; k++) {
        __syncthreads();
        const double d = Lp[ptri(k) + k];
        const bool bad = !(d > 0.);
        const double piv = bad ? 1. : sqrt(d);
        __syncthreads();
        if (t == 0) {
            Lp[ptri(k) + k] = piv;
            if (bad && info) atomicCAS(info, 0, k0 + k + 1);
        }
        for (int i = k + 1 + t; i < n; i += DIAG_THREADS) Lp[ptri(i) + k] /= piv;
        __syncthreads();
        const int m = n - k - 1;
        for (int e = t; e < m * m; e += DIAG_THREADS) {
            const int i = k + 1 + e / m, j = k + 1 + e % m;
            if (j <= i) Lp[ptri(i) + j] -= Lp[ptri(i) + k] * Lp[ptri(j) + k];
        }
    }
    __syncthreads();
    // inverse by forward substitution, one row per step:
    // X[i][j] = (delta_ij - sum_{j<=k<i} L[i][k] X[k][j]) / L[i][i], the sum
    // split over DIAG_PARTS threads per column and added in a fixed order
    const int j = t % LB, part = t / LB;
    for (int i = 0; i < n; i++) {
        double s = 0.;
        if (j < i)
            for (int k = j + part; k < i; k += DIAG_PARTS) s += Lp[ptri(i) + k] * Xp[ptri(k) + j];
        red[part][j] = s;
        __syncthreads();
        if (part == 0 && j <= i) {
            double acc = 0.;
#pragma unroll
            for (int q = 0; q < DIAG_PARTS; q++) acc += red[q][j];
            Xp[ptri(i) + j] = ((i == j ? 1. : 0.) - acc) / Lp[ptri(i) + i];
        }
        __syncthreads();
    }
    // the factor back (strict upper triangle of the block zeroed) and D
    for (int e = t; e < n * n; e += DIAG_THREADS) {
        const int i = e / n, c = e % n;
        A[(long)i * lda + c] = c <= i ? Lp[ptri(i) + c] : 0.;
    }
    for (int e = t; e < LB * LB; e += DIAG_THREADS) {
        const int i = e / LB, c = e % LB;
        D[(long)i * LB + c] = (i < n && c <= i) ? Xp[ptri(i) + c] : 0.;
    }
}

}  // namespace

extern "C" int mdf_gemm_batched(int transb, int tri_a, int lower_c, long M, long N, long K, double alpha,
                                const double *A, long lda, long strideA, const double *B, long ldb, long strideB,
                                double beta, double *C, long ldc, long strideC, long batch, void *stream) {
    MDF_CHECK_ARG(M >= 0 && N >= 0 && K >= 0 && M < (1L << 31) && N < (1L << 31) && K < (1L << 31) && batch >= 0,
                  "mdf_gemm: bad sizes");
    MDF_CHECK_ARG((A && B && C) || batch == 0, "mdf_gemm: null pointer");
    MDF_CHECK_ARG(!lower_c || M == N, "mdf_gemm: lower_c needs a square C");
    MDF_CHECK_ARG(tri_a >= 0 && tri_a <= 2 && (tri_a == 0 || M == K), "mdf_gemm: tri_a %d needs a square A", tri_a);
    MDF_CHECK_ARG(lda >= K && ldc >= N && ldb >= (transb ? K : N), "mdf_gemm: leading dimension too small");
    MDF_CHECK_ARG(strideA >= 0 && strideB >= 0 && strideC >= 0, "mdf_gemm: negative batch stride");
    hipStream_t st = as_stream(stream);
    const int m = (int)M, n = (int)N, k = (int)K;
    if (!transb && tri_a == 0 && !lower_c && N > 256 && N <= 320)
        return launch_wide(m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc, strideC, batch, st);
    if (lower_c)
        return transb ? gemm_tri<true, true>(tri_a, m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                             strideC, batch, st)
                      : gemm_tri<false, true>(tri_a, m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                              strideC, batch, st);
    return transb ? gemm_tri<true, false>(tri_a, m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                          strideC, batch, st)
                  : gemm_tri<false, false>(tri_a, m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                                           strideC, batch, st);
}

extern "C" int mdf_gemm_pairs_batched(int transb, long n, long K, double alpha, const double *A, long lda,
                                      long strideA, const double *B, long ldb, long strideB, const int *pair_col,
                                      double *W, long ldw, const double *init_scale, const double *init_vals,
                                      long batch, void *stream) {
    MDF_CHECK_ARG(n >= 0 && K >= 0 && n < (1L << 31) && K < (1L << 31) && batch >= 0 &&
                  n * (n + 1) / 2 < (1L << 30) && ldw < (1L << 29), "mdf_gemm_pairs_batched: bad sizes");
    MDF_CHECK_ARG((A && B && pair_col && W) || batch == 0 || n == 0, "mdf_gemm_pairs_batched: null pointer");
    MDF_CHECK_ARG(!init_scale == !init_vals, "mdf_gemm_pairs_batched: init_scale and init_vals go together");
    MDF_CHECK_ARG(lda >= K && ldb >= (transb ? K : n) && strideA >= 0 && strideB >= 0 && ldw >= 0,
                  "mdf_gemm_pairs_batched: leading dimension or stride");
    if (n == 0 || batch == 0) return MDF_OK;
    hipStream_t st = as_stream(stream);
    const int t = (int)((n + 127) / 128);
    for (long b0 = 0; b0 < batch; b0 += MAX_GRID_Z) {
        const unsigned nz = (unsigned)(batch - b0 < MAX_GRID_Z ? batch - b0 : MAX_GRID_Z);
        const dim3 g((unsigned)(t * (t + 1) / 2), 1, nz);
        const double *isc = init_scale ? init_scale + b0 : nullptr;
        if (transb)
            hipLaunchKernelGGL(gemm_pairs_kernel<true>, g, dim3(G_THREADS), 0, st, (int)n, (int)K, alpha,
                               A + b0 * strideA, lda, strideA, B + b0 * strideB, ldb, strideB, pair_col, W + b0 * ldw,
                               ldw, isc, init_vals);
        else
            hipLaunchKernelGGL(gemm_pairs_kernel<false>, g, dim3(G_THREADS), 0, st, (int)n, (int)K, alpha,
                               A + b0 * strideA, lda, strideA, B + b0 * strideB, ldb, strideB, pair_col, W + b0 * ldw,
                               ldw, isc, init_vals);
        MDF_LAUNCHED("gemm_pairs_kernel");
    }
    return MDF_OK;
}

// Sizes of mdf_sym_pack / mdf_sym_unpack: n x n matrices of src_min / dst_min
// elements, one per batch stride.
static int sym_check(const char *name, long n, const void *src, long s_src, long src_min, const void *dst,
                     long s_dst, long dst_min, long batch) {
    MDF_CHECK_ARG(n >= 0 && batch >= 0 && n < (1L << 20), "%s: bad sizes", name);
    if (n == 0 || batch == 0) return MDF_OK;
    MDF_CHECK_ARG(src && dst, "%s: null pointer", name);
    MDF_CHECK_ARG(s_src >= 0 && s_dst >= 0 && (batch == 1 || (s_src >= src_min && s_dst >= dst_min)),
                  "%s: batch stride below one matrix", name);
    return MDF_OK;
}

static dim3 sym_grid(long n, long batch) {
    const long t = (n + SY_T - 1) / SY_T;
    return dim3((unsigned)(t * (t + 1) / 2), (unsigned)(batch < 65535 ? batch : 65535));
}

extern "C" int mdf_sym_pack(long n, const double *G, long strideG, double *P, long strideP, double diag_scale,
                            double off_scale, long batch, void *stream) {
    const long np = n * (n + 1) / 2;
    int rc = sym_check("mdf_sym_pack", n, G, strideG, n * n, P, strideP, np, batch);
    if (rc || n == 0 || batch == 0) return rc;
    hipLaunchKernelGGL(sym_pack_kernel, sym_grid(n, batch), dim3(256), 0, as_stream(stream), (int)n, G, strideG, P,
                       strideP, batch, diag_scale, off_scale);
    MDF_LAUNCHED("sym_pack_kernel");
    return MDF_OK;
}

extern "C" int mdf_sym_unpack(long n, const double *P, long strideP, double *G, long strideG, double diag_scale,
                              double off_scale, long batch, void *stream) {
    const long np = n * (n + 1) / 2;
    int rc = sym_check("mdf_sym_unpack", n, P, strideP, np, G, strideG, n * n, batch);
    if (rc || n == 0 || batch == 0) return rc;
    hipLaunchKernelGGL(sym_unpack_kernel, sym_grid(n, batch), dim3(256), 0, as_stream(stream), (int)n, P, strideP, G,
                       strideG, batch, diag_scale, off_scale);
    MDF_LAUNCHED("sym_unpack_kernel");
    return MDF_OK;
}

extern "C" long mdf_sp2_tiles(long n) {
    const long t = (n + SP2_T - 1) / SP2_T;
    return n > 0 ? t * (t + 1) / 2 : 0;
}

extern "C" int mdf_sp2_begin(long n, const double *X, long ldx, double *part, void *stream) {
    MDF_CHECK_ARG(n >= 0 && n < (1L << 20) && ldx >= n, "mdf_sp2_begin: bad sizes");
    if (n == 0) return MDF_OK;
    MDF_CHECK_ARG(X && part, "mdf_sp2_begin: null pointer");
    hipLaunchKernelGGL(sp2_norms_kernel, dim3((unsigned)mdf_sp2_tiles(n)), dim3(G_THREADS), 0, as_stream(stream),
                       (int)n, X, ldx, part);
    MDF_LAUNCHED("sp2_norms_kernel");
    return MDF_OK;
}

extern "C" int mdf_sp2_step(long n, double nocc, const double *X, long ldx, const double *part_in, double *Y,
                            long ldy, double *part_out, void *stream) {
    MDF_CHECK_ARG(n >= 0 && n < (1L << 20) && ldx >= n && ldy >= n, "mdf_sp2_step: bad sizes");
    if (n == 0) return MDF_OK;
    MDF_CHECK_ARG(X && Y && part_in && part_out, "mdf_sp2_step: null pointer");
    MDF_CHECK_ARG(X != Y && part_in != part_out, "mdf_sp2_step: Y and part_out must not alias X and part_in");
    const long nt = mdf_sp2_tiles(n);
    hipLaunchKernelGGL(sp2_step_kernel, dim3((unsigned)nt), dim3(G_THREADS), 0, as_stream(stream), (int)n, nocc, X,
                       ldx, part_in, Y, ldy, part_out, (int)nt);
    MDF_LAUNCHED("sp2_step_kernel");
    return MDF_OK;
}

extern "C" int mdf_gemm(int transb, int tri_a, int lower_c, long M, long N, long K, double alpha,
                        const double *A, long lda, const double *B, long ldb, double beta, double *C, long ldc,
                        void *stream) {
    return mdf_gemm_batched(transb, tri_a, lower_c, M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1,
                            stream);
}

// workspace: the diagonal-block inverses D (nblocks x 128 x 128, kept for
// mdf_trtri) followed by an n x 128 panel buffer
extern "C" size_t mdf_potrf_workspace(long n) {
    return ((size_t)((n + LB - 1) / LB) * LB * LB + (size_t)n * LB) * sizeof(double) + 256;
}

extern "C" int mdf_potrf(double *A, long n, long lda, int *info, void *work, size_t work_bytes, void *stream) {
    MDF_CHECK_ARG(A && n >= 0 && lda >= n && n < (1L << 31), "mdf_potrf: bad arguments");
    MDF_CHECK_ARG(work && work_bytes >= mdf_potrf_workspace(n), "mdf_potrf: workspace too small");
    hipStream_t st = as_stream(stream);
    double *D = (double *)work;
    double *P = D + (size_t)((n + LB - 1) / LB) * LB * LB;
    for (long k0 = 0; k0 < n; k0 += LB) {
        const int kb = (int)(n - k0 < LB ? n - k0 : LB);
        double *Akk = A + k0 * lda + k0;
        double *Dk = D + (k0 / LB) * LB * LB;
        hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(DIAG_THREADS), 0, st, Akk, lda, kb, (int)k0, Dk, info);
        MDF_LAUNCHED("potrf_diag_kernel");
        const long m2 = n - k0 - kb;
        if (m2 <= 0) break;
        double *A21 = A + (k0 + kb) * lda + k0;
        // A21 <- A21 inv(L11)^T, through the panel buffer
        int rc = launch_gemm<128, true, 0, false>((int)m2, kb, kb, 1., A21, lda, 0, Dk, LB, 0, 0., P, LB, 0, 1, st);
        if (rc) return rc;
        MDF_HIP(hipMemcpy2DAsync(A21, lda * sizeof(double), P, LB * sizeof(double), kb * sizeof(double), m2,
                                 hipMemcpyDeviceToDevice, st));
        // A22 <- A22 - A21 A21^T (lower tiles)
        double *A22 = A + (k0 + kb) * lda + k0 + kb;
        rc = launch_gemm<128, true, 0, true>((int)m2, (int)m2, kb, -1., A21, lda, 0, A21, lda, 0, 1., A22, lda, 0, 1, st);
        if (rc) return rc;
    }
    return MDF_OK;
}

extern "C" size_t mdf_trtri_workspace(long n) { return (size_t)LB * n * sizeof(double) + 256; }

// X = L^-1 for the lower factor L left by mdf_potrf; D holds its diagonal
// block inverses (mdf_potrf's workspace).  X must not alias L.
extern "C" int mdf_trtri(const double *L, long n, long ldl, const double *D, double *X, long ldx, void *work,
                         size_t work_bytes, void *stream) {
    MDF_CHECK_ARG(L && D && X && n >= 0 && ldl >= n && ldx >= n, "mdf_trtri: bad arguments");
    MDF_CHECK_ARG(work && work_bytes >= mdf_trtri_workspace(n), "mdf_trtri: workspace too small");
    hipStream_t st = as_stream(stream);
    double *T = (double *)work;      // LB x n
    // zero X, then place the diagonal inverses
    {
        const long tot = n * ldx;
        MDF_HIP(hipMemsetAsync(X, 0, tot * sizeof(double), st));
    }
    for (long b0 = 0; b0 < n; b0 += LB) {
        const int kb = (int)(n - b0 < LB ? n - b0 : LB);
        const double *Db = D + (b0 / LB) * LB * LB;
        MDF_HIP(hipMemcpy2DAsync(X + b0 * ldx + b0, ldx * sizeof(double), Db, LB * sizeof(double),
                                 kb * sizeof(double), kb, hipMemcpyDeviceToDevice, st));
        if (b0 == 0) continue;
        // T = L[b, 0:b0] X[0:b0, 0:b0];  X[b, 0:b0] = -D_b T
        int rc = launch_gemm<128, false, 0, false>(kb, (int)b0, (int)b0, 1., L + b0 * ldl, ldl, 0, X, ldx, 0, 0.,
                                                   T, b0, 0, 1, st);
        if (rc) return rc;
        rc = launch_gemm<128, false, 0, false>(kb, (int)b0, kb, -1., Db, LB, 0, T, b0, 0, 0., X + b0 * ldx, ldx, 0,
                                               1, st);
        if (rc) return rc;
    }
    return MDF_OK;
}

// C[0:n, 0:ncols] = Linv[0:n, 0:n] B[0:n, 0:ncols], Linv lower triangular.
extern "C" int mdf_trmm(const double *Linv, long n, long ldl, const double *B, long ldb, long ncols, double *C,
                        long ldc, void *stream) {
    MDF_CHECK_ARG(Linv && B && C && n >= 0 && ncols >= 0 && ldl >= n && ldb >= ncols && ldc >= ncols,
                  "mdf_trmm: bad arguments");
    MDF_CHECK_ARG(n < (1L << 31) && ncols < (1L << 31), "mdf_trmm: sizes beyond int range");
    return launch_gemm<128, false, 1, false>((int)n, (int)ncols, (int)n, 1., Linv, ldl, 0, B, ldb, 0, 0., C, ldc, 0,
                                             1, as_stream(stream));
}
