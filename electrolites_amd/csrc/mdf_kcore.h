// Tile bodies of the exchange (K) contraction, shared by the single-item
// launches (mdf_k.hip) and the grouped persistent launch (mdf_grouped.hip),
// so an item gives bit-identical results alone and in a batch.
//
//     R[L, i, k] = sum_j A_L[i, j] F[j, k]                 (step 1)
//     K[i, j]   += sum_{L,k} R_l[L, i, k] R_r[L, j, k]     (step 2)
#pragma once
#include <cstdlib>
#include <type_traits>

#include "mdf_common.h"
#include "mdf_gemm.h"
#include "mdf_b3.h"

namespace mdf {
namespace kcore {

__device__ __forceinline__ long tri(long i) { return i * (i + 1) / 2; }

// ---------------------------------------------------------------- step 1
// A operand of step 1: rows i0..i0+127 of the unpacked A_L, K-step kt covers
// j = 16 kt .. 16 kt + 15.  TRI: the lower triangle of A_L with its
// diagonal halved (A_L = T + T^T), zero above the diagonal.
template <bool MAPPED, typename T = double, typename Src = double, int THREADS = G_THREADS, bool TRI = false>
struct PackedALoader {
    using reg_t = Src;                                   // the cderi storage type
    static constexpr int N = 128 * G_BK / THREADS;       // 8 (256 threads) or 4 (512)
    static constexpr int TPJ = THREADS / G_BK;           // threads per j row (upper)
    static constexpr int TPR = G_BK / N;                 // threads per i row (lower)
    const Src *row;      // cderi row L
    const int *pmap;     // MAPPED: pair index of (i, j), -1 if not stored
    int nao, i0;
    int upper;           // layout used by the last load()
    // MAPPED: the pair indices of the next K-step, loaded one call ahead so
    // the value gathers do not wait on the map (calls come with kt = 0, 1, ...)
    int pm[N];
    __device__ __forceinline__ void map_step(int kt, int (&p)[N]) const {
        const int t = threadIdx.x, i = i0 + t / TPR, jb = kt * G_BK + (t % TPR) * N;
        if ((nao & 3) == 0 && i < nao && jb + N <= nao) {
            // a thread's N map entries as 16-byte loads (N is 4 or 8, jb a
            // multiple of 4, so the row start i nao + jb is one too)
            const int4 *src = reinterpret_cast<const int4 *>(pmap + (long)i * nao + jb);
#pragma unroll
            for (int q = 0; q < N / 4; q++) {
                const int4 m = src[q];
                p[4 * q] = m.x;
                p[4 * q + 1] = m.y;
                p[4 * q + 2] = m.z;
                p[4 * q + 3] = m.w;
            }
            if (TRI) {
#pragma unroll
                for (int v = 0; v < N; v++)
                    if (jb + v > i) p[v] = -1;
            }
            return;
        }
#pragma unroll
        for (int v = 0; v < N; v++)
            p[v] = (i < nao && jb + v < nao && (!TRI || jb + v <= i)) ? pmap[(long)i * nao + jb + v] : -1;
    }
    __device__ __forceinline__ void load(int kt, Src (&r)[N]) {
        const int t = threadIdx.x, j0 = kt * G_BK;
        if (MAPPED) {
            // lanes run along j of one row i in every K-step (the transposed
            // layout above the diagonal measured no faster here)
            upper = 0;
            if (kt == 0) map_step(0, pm);
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = pm[v] >= 0 ? row[pm[v]] : (Src)0;
            if (TRI) {
                const int i = i0 + t / TPR, jb = j0 + (t % TPR) * N;
#pragma unroll
                for (int v = 0; v < N; v++)
                    if (jb + v == i) r[v] *= (Src)0.5;
            }
            map_step(kt + 1, pm);
            return;
        }
        if (TRI && j0 >= i0 + 127) {
            // above the diagonal: zero (the K-loop of a TRI tile stops at
            // the diagonal block, so this is only a guard)
            upper = 0;
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = (Src)0;
            return;
        }
        if (!MAPPED && j0 >= i0 + 127) {
            // every j > every i: (i, j) at tri(j) + i, contiguous in i
            upper = 1;
            const int j = j0 + t / TPJ, ib = i0 + (t % TPJ) * N;
            const Src *src = row + tri(j) + ib;
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = (j < nao && ib + v < nao) ? src[v] : (Src)0;
            return;
        }
        upper = 0;
        const int i = i0 + t / TPR, jb = j0 + (t % TPR) * N;
        if (!MAPPED && i0 >= j0 + G_BK - 1) {
            // every i >= every j: contiguous in j
            const Src *src = row + tri(i) + jb;
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = (i < nao) ? src[v] : (Src)0;
            return;
        }
#pragma unroll
        for (int v = 0; v < N; v++) {
            const int j = jb + v;
            Src x = 0;
            if (i < nao && j < nao) {
                if (MAPPED) {
                    const int p = pmap[(long)i * nao + j];
                    x = p >= 0 ? row[p] : 0.;
                } else if (TRI) {
                    x = i > j ? row[tri(i) + j] : i == j ? (Src)0.5 * row[tri(i) + i] : (Src)0;
                } else {
                    x = row[i >= j ? tri(i) + j : tri(j) + i];
                }
            }
            r[v] = x;
        }
    }
    __device__ __forceinline__ void store(const Src (&r)[N], T *lds, int pitch) const {
        const int t = threadIdx.x;
        if (upper) {
            const int jj = t / TPJ, ib = (t % TPJ) * N;
#pragma unroll
            for (int v = 0; v < N; v++) lds[jj * pitch + ib + v] = (T)r[v];
        } else {
            const int ii = t / TPR, jb = (t % TPR) * N;
#pragma unroll
            for (int v = 0; v < N; v++) lds[(jb + v) * pitch + ii] = (T)r[v];
        }
    }
};

// Step-1 tile widths: 64, 128 and 160 columns on 2 x 2 wavefronts, 256 and
// 320 on 2 x 4 (a workgroup of 512 threads then covers up to 320 occupied
// columns, so each cderi tile is read once per 320 columns instead of once
// per 128 or 160).
template <int BN>
struct Step1Shape {
    static constexpr int WGN = BN >= 192 ? 4 : 2;
    static constexpr int THREADS = 128 * WGN;
    using S = TileShape<128, BN, 2, WGN>;
};

// One 128 x BN tile of step 1 for aux row `row` (the cderi row L): rows
// i0.., columns (occupied index) k0.. of R_L.
// TRI: A_L replaced by its lower triangle with the diagonal halved, and the
// K loop stops after the tile's diagonal block (about half the work; for a
// symmetric bilinear form F^T A_L F = G + G^T with G = F^T (TRI product)).
template <int BN, bool MAPPED, typename T, typename Src = double, bool TRI = false>
__device__ __forceinline__ void step1_tile(const Src *__restrict__ row, const int *__restrict__ pmap, int nao,
                                  const double *__restrict__ F, long ldf, int nocc, T *__restrict__ Rl, long ldr,
                                  int i0, int k0, T *lds) {
    using C = Step1Shape<BN>;
    using S = typename C::S;
    PackedALoader<MAPPED, T, Src, C::THREADS, TRI> la{row, pmap, nao, i0, 0};
    ColsLoader<BN, T, double, C::THREADS> lb{F, ldf, k0, nocc, 0, nao};
    Acc<S::MI, S::NI, T> acc;
    acc.zero();
    const int jend = TRI ? min(nao, i0 + 128) : nao;
    gemm_mainloop<128, BN, decltype(la), decltype(lb), T, 2, C::WGN>(la, lb, (jend + G_BK - 1) / G_BK, lds, acc);
    auto emit = [&](int r, int c, T v) {
        const int i = i0 + r, k = k0 + c;
        if (i < nao && k < nocc) Rl[(long)i * ldr + k] = v;
    };
    gemm_epilogue_layout<128, BN, T, decltype(acc), decltype(emit) &, 2, C::WGN>(acc, emit);
}

// ---------------------------------------------------------------- step 2
// Rows r0..r0+127 of R[L] (k contiguous); K-step kt of this split is
// (L, kc) = divmod(s0 + kt, kchunks), columns kc*16 .. kc*16 + 15.
template <bool VEC, typename T>
struct TransformedLoader {
    using reg_t = T;
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
    __device__ __forceinline__ void load(int, T (&r)[N]) {
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
    __device__ __forceinline__ void store(const T (&r)[N], T *lds, int pitch) const {
        const int t = threadIdx.x, ii = t & 127, kb = (t >> 7) * N;
#pragma unroll
        for (int v = 0; v < N; v++) lds[(kb + v) * pitch + ii] = r[v];
    }
};

constexpr int S2_B = 128;
// FP32 step 2 adds its MFMA accumulators into an FP64 shadow every S2_FLUSH
// K-steps (S2_FLUSH * 16 products per element), so the rounding error of the
// long (L, k) reduction grows with the flush length, not with naux * nocc.
constexpr int S2_FLUSH = 32;

// One 128 x 128 output tile (i0, j0) of step 2 over the K-steps [s0, s1) of
// the (L, k-chunk) reduction; writes (beta = 0) or adds (beta = 1) into o.
template <bool VEC, typename T>
__device__ __forceinline__ void step2_tile(const T *__restrict__ Rl, const T *__restrict__ Rr, int nao, int nocc, long ldr,
                                  long strideL, int i0, int j0, long s0, long s1, double *__restrict__ o, int beta,
                                  T *lds) {
    using S = TileShape<S2_B, S2_B>;
    const int kchunks = (nocc + G_BK - 1) / G_BK;
    const long Ls = s0 / kchunks;
    const int kcs = (int)(s0 - Ls * kchunks);
    TransformedLoader<VEC, T> la{Rl, strideL, ldr, i0, nao, nocc, kchunks, Ls, kcs};
    TransformedLoader<VEC, T> lb{Rr, strideL, ldr, j0, nao, nocc, kchunks, Ls, kcs};
    Acc<S::MI, S::NI, T> acc;
    auto emit = [&](int r, int c, double v) {
        const int i = i0 + r, j = j0 + c;
        if (i < nao && j < nao) {
            double *dst = o + (long)i * nao + j;
            *dst = beta ? *dst + v : v;
        }
    };
    if constexpr (sizeof(T) == 8) {
        acc.zero();
        gemm_mainloop<S2_B, S2_B>(la, lb, (int)(s1 - s0), lds, acc);
        gemm_epilogue<S2_B, S2_B, T>(acc, emit);
    } else {
        Acc<S::MI, S::NI, double> sum;
        sum.zero();
        for (long s = s0; s < s1; s += S2_FLUSH) {
            acc.zero();
            // the loaders advance (L, k-chunk) in place, so consecutive calls
            // continue the reduction where the previous one stopped
            gemm_mainloop<S2_B, S2_B>(la, lb, (int)min((long)S2_FLUSH, s1 - s), lds, acc);
#pragma unroll
            for (int a = 0; a < S::MI; a++)
#pragma unroll
                for (int b = 0; b < S::NI; b++)
#pragma unroll
                    for (int v = 0; v < 4; v++) sum.v[a][b][v] += (double)acc.v[a][b][v];
        }
        gemm_epilogue_layout<S2_B, S2_B, T>(sum, emit);
    }
}

// Lower-triangle tile index x -> (ti >= tj), row-major.
__device__ __forceinline__ void tri_tile(int x, int &ti, int &tj) {
    ti = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
    while (ti * (ti + 1) / 2 > x) ti--;
    while ((ti + 1) * (ti + 2) / 2 <= x) ti++;
    tj = x - ti * (ti + 1) / 2;
}

// vk[i, j] (+)= sum_s part[s, I, J] for the flat range [e0, e1); SYM reads the
// lower-tile copy for upper-tile entries.
__device__ __forceinline__ void finalize_range(const double *__restrict__ part, int ksplit, long part_stride, int nao,
                                      int sym, double *__restrict__ vk, long ldvk, int accumulate, long e0, long e1) {
    for (long t = e0 + threadIdx.x; t < e1; t += 256) {
        const int i = (int)(t / nao), j = (int)(t % nao);
        int I = i, J = j;
        if (sym && (i / S2_B) < (j / S2_B)) { I = j; J = i; }
        double s = 0.;
        for (int k = 0; k < ksplit; k++) s += part[k * part_stride + (long)I * nao + J];
        double *dst = vk + (long)i * ldvk + j;
        *dst = accumulate ? *dst + s : s;
    }
}

// step-1 N tile: the width in {64, 128, 160} that pads nocc least (ties: wider)
// step-1 tile width: each k-tile re-reads the cderi tile (worth about 64
// columns of MFMA work), so minimise nkt * (64 + BN) over the widths; ties
// go to the wider tile
inline int pick_bn(int nocc) {
    const int cand[6] = {64, 128, 160, 192, 256, 320};
    // MDF_K_STEP1_BN forces one of the widths (for tuning runs)
    if (const char *e = getenv("MDF_K_STEP1_BN")) {
        const int v = atoi(e);
        for (int b : cand)
            if (b == v) return b;
    }
    int best = 64;
    long best_cost = 1L << 40;
    for (int b : cand) {
        const long cost = (long)((nocc + b - 1) / b) * (64 + b);
        if (cost <= best_cost) { best_cost = cost; best = b; }
    }
    return best;
}

struct KPlan {
    int ntile, ntiles_grid, ksplit, bn;
    long nL;            // aux rows per block
    size_t r_words;     // one R buffer
    size_t part_words;
};

// Resident step-2 workgroups on this device: 2 per CU (LDS and registers).
inline int resident_step2_wgs() {
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

// Split of the (L, k) reduction: a multiple of the XCD count, so that each
// XCD takes whole splits (step2_coords) and the output tiles running on one
// XCD at a time read the same R panels through its L2; among 8, 16, 24, 32
// the one whose per-XCD share fills whole rounds of resident workgroups
// best (ties: the larger split).
inline int pick_ksplit(int ntiles_grid) {
    const int slots = resident_step2_wgs() / N_XCD;
    int best = N_XCD;
    double best_eff = -1.;
    for (int ks = N_XCD; ks <= 4 * N_XCD; ks += N_XCD) {
        const long items = (long)ntiles_grid * ks / N_XCD;
        const long rounds = (items + slots - 1) / slots;
        const double eff = (double)items / (double)(rounds * slots);
        if (eff >= best_eff - 1e-9) { best_eff = eff; best = ks; }
    }
    return best;
}

// (output tile, split) of workgroup w of a 1-D step-2 grid of ntiles_grid *
// ksplit workgroups (ksplit a multiple of N_XCD).  Workgroups reach the XCDs
// round-robin (w % N_XCD); XCD x takes splits x, x + N_XCD, ..., every
// output tile of one split before the next, so the tiles resident on an XCD
// together mostly share one split's K range: each R panel of that range is
// fetched into the XCD's L2 once and read there by every tile of its row
// and column.
__device__ __forceinline__ void step2_coords(long w, int ntiles_grid, int &tile, int &split) {
    const int xcd = (int)(w % N_XCD);
    const long l = w / N_XCD;
    split = xcd + N_XCD * (int)(l / ntiles_grid);
    tile = (int)(l % ntiles_grid);
}

// Aux rows per pass: the transformed blocks of one pass (R_l and R_r) stay
// within K_BLOCK_BYTES.  A function of the item's shape only, so the single
// entry and the grouped launch cut an item into the same passes.
// The environment variable MDF_K_BLOCK_BYTES overrides the bound (read on
// every plan, so tests can force several passes on small items).
constexpr size_t K_BLOCK_BYTES = 8ull << 30;

inline size_t k_block_bytes() {
    const char *e = getenv("MDF_K_BLOCK_BYTES");
    if (e && *e) {
        const long long v = atoll(e);
        if (v > 0) return (size_t)v;
    }
    return K_BLOCK_BYTES;
}

inline long block_rows(long naux, int nao, int nocc, bool two, size_t elem) {
    const size_t per_row = (size_t)nao * nocc * elem * (two ? 2 : 1);
    long r = per_row ? (long)(k_block_bytes() / per_row) : naux;
    if (r > naux) r = naux;
    return r < 1 ? 1 : r;
}

// Row stride of a transformed block R[L][i][k]: nocc rounded up to whole
// 16-byte vectors, so step 2 streams every row with vector loads whatever
// the parity of nocc (columns nocc..ldr-1 are never written or read).
inline long r_ld(int nocc, size_t elem) {
    const long e = 16 / (long)elem;
    return (nocc + e - 1) / e * e;
}

inline KPlan plan_k(long naux, int nao, int nocc, bool two, long max_rows, size_t elem = 8) {
    KPlan p;
    p.ntile = (nao + S2_B - 1) / S2_B;
    p.ntiles_grid = two ? p.ntile * p.ntile : p.ntile * (p.ntile + 1) / 2;
    p.ksplit = pick_ksplit(p.ntiles_grid);
    p.bn = pick_bn(nocc);
    p.nL = max_rows < naux ? max_rows : naux;
    if (p.nL < 1) p.nL = 1;
    p.r_words = (size_t)p.nL * nao * r_ld(nocc, elem);
    p.part_words = (size_t)p.ksplit * nao * nao;
    return p;
}

// ------------------------------------------------- split-BF16 (prec 2)
// The same two steps on the three-term split-BF16 engine (mdf_b3.h).  The
// factors are pre-split into transposed planes FT[k][j] (j zero-padded to
// ldft = a multiple of 32 >= nao); R is stored split, as two planes of
// [L][i][k] bfloat16 with k zero-padded to ldr = a multiple of 32 >= nocc,
// so step 2 copies it into LDS without conversion.  Both are stored in
// chunks of 32 along their contiguous index (FT: (j / 32, k, j % 32); R:
// r_index), so each K-step reads its tile of a plane as one contiguous run.
namespace kb3 {

using b3::u16;
constexpr int BM = 128;               // step-1 rows per tile (step 2: S2_B x S2_B)
// step 2 adds its FP32 accumulators into the FP64 partial in global memory
// every S2_FLUSH K-steps (S2_FLUSH * 32 reduction terms)
constexpr int S2_FLUSH = 64;

__host__ __device__ inline long pad32(long n) { return (n + 31) / 32 * 32; }

// Element (i, k) of one aux row's split R plane (nao rows, ldr = pad32(nocc)
// columns): stored as ldr / 32 chunks of 32 columns, each chunk nao rows of
// 32 contiguous values, so the 128 rows x 32 columns a step-2 K-step reads
// are one contiguous 8 KB run per plane (row-major rows would give 128 runs
// of 64 bytes, half a cache line each).  The plane's size is nao * ldr, as
// row-major.
__host__ __device__ inline long r_index(int nao, int i, int k) {
    return ((long)(k >> 5) * nao + i) * 32 + (k & 31);
}

// A operand: rows i0..i0+127 of the unpacked A_L; K-step kt covers
// j = 32 kt .. 32 kt + 31.  THREADS threads, N = 4096 / THREADS values each.
template <bool MAPPED, int THREADS, typename Src = double>
struct PackedA32 {
    using reg_t = Src;
    static constexpr int N = 128 * b3::BK / THREADS;
    static constexpr int TPR = b3::BK / N;           // threads per row (lower layout)
    const Src *row;
    const int *pmap;
    int nao, i0;
    int upper;
    int pm[N];           // MAPPED: the next K-step's pair indices (one call ahead)
    int pm_upper;        // and their layout
    // Above the diagonal (every j > every i) the transposed layout of the
    // packed loader: thread t takes i = i0 + t % 128 and j = j0 + N (t / 128)
    // + v, so a load instruction reads consecutive i of the map row j and of
    // the stored pairs (j, i).
    __device__ __forceinline__ void map_step(int kt, int (&p)[N], int &up) const {
        const int t = threadIdx.x, j0 = kt * b3::BK;
        up = j0 >= i0 + 127;
        if (up) {
            const int i = i0 + (t & 127), jb = j0 + (t >> 7) * N;
#pragma unroll
            for (int v = 0; v < N; v++) p[v] = (i < nao && jb + v < nao) ? pmap[(long)(jb + v) * nao + i] : -1;
        } else {
            const int i = i0 + t / TPR, jb = j0 + (t % TPR) * N;
            if ((nao & 3) == 0 && i < nao && jb + N <= nao) {
                // N (a multiple of 4) map entries as 16-byte loads
                const int4 *src = reinterpret_cast<const int4 *>(pmap + (long)i * nao + jb);
#pragma unroll
                for (int q = 0; q < N / 4; q++) {
                    const int4 m = src[q];
                    p[4 * q] = m.x;
                    p[4 * q + 1] = m.y;
                    p[4 * q + 2] = m.z;
                    p[4 * q + 3] = m.w;
                }
                return;
            }
#pragma unroll
            for (int v = 0; v < N; v++) p[v] = (i < nao && jb + v < nao) ? pmap[(long)i * nao + jb + v] : -1;
        }
    }
    __device__ __forceinline__ void load(int kt, Src (&r)[N]) {
        const int t = threadIdx.x, j0 = kt * b3::BK;
        if (MAPPED) {
            if (kt == 0) map_step(0, pm, pm_upper);
            upper = pm_upper;
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = pm[v] >= 0 ? row[pm[v]] : (Src)0;
            map_step(kt + 1, pm, pm_upper);
            return;
        }
        if (!MAPPED && j0 >= i0 + 127) {
            // every j > every i: (i, j) at tri(j) + i, contiguous in i.  Thread
            // t takes row i = i0 + t % 128 and j = j0 + N (t / 128) + v: each
            // load instruction reads 64 consecutive i of one j (coalesced), and
            // the row lands in LDS as N contiguous k (16-byte stores).
            upper = 1;
            const int i = i0 + (t & 127), jb = j0 + (t >> 7) * N;
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = (i < nao && jb + v < nao) ? row[tri(jb + v) + i] : (Src)0;
            return;
        }
        upper = 0;
        const int i = i0 + t / TPR, jb = j0 + (t % TPR) * N;
        if (!MAPPED && i0 >= j0 + b3::BK - 1) {
            // every i >= every j: contiguous in j
            const Src *src = row + tri(i) + jb;
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = (i < nao) ? src[v] : (Src)0;
            return;
        }
#pragma unroll
        for (int v = 0; v < N; v++) {
            const int j = jb + v;
            Src x = 0;
            if (i < nao && j < nao) {
                if (MAPPED) {
                    const int p = pmap[(long)i * nao + j];
                    x = p >= 0 ? row[p] : 0.;
                } else {
                    x = row[i >= j ? tri(i) + j : tri(j) + i];
                }
            }
            r[v] = x;
        }
    }
    __device__ __forceinline__ void store(const Src (&r)[N], u16 *hi, u16 *lo) const {
        const int t = threadIdx.x;
        if (upper)
            b3::put<N>(r, hi, lo, t & 127, (t >> 7) * N);
        else
            b3::put<N>(r, hi, lo, t / TPR, (t % TPR) * N);
    }
};

// B operand of step 1: rows k0..k0+BN-1 of the split FT planes.
template <int BN, int THREADS>
struct FTLoader {
    using P = b3::PlaneRowsN<BN, THREADS>;
    using reg_t = uint4;
    static constexpr int N = P::N;
    const u16 *hi, *lo;
    long ld;
    int k0, nocc;
    // the planes are chunk-major (split_ft_range): K-step kt is chunk kt,
    // nocc rows of 32 values
    __device__ __forceinline__ void load(int kt, uint4 (&r)[N]) const {
        P::fetch(hi, lo, b3::BK, (long)kt * b3::BK * nocc, k0, nocc, r);
    }
    __device__ __forceinline__ void store(const uint4 (&r)[N], u16 *h, u16 *l) const { P::put(r, h, l); }
};

// Step-1 tile widths: 128 columns on 2 x 2 wavefronts, 192, 256 or 320 on
// 2 x 4 (one workgroup then covers up to 320 occupied columns, so each cderi tile
// is read once per factor instead of once per 128 columns).
template <int BN>
struct Step1Cfg {
    static constexpr int WGN = BN == 128 ? 2 : 4;
    static constexpr int THREADS = 128 * WGN;
    using S = b3::Shape<BM, BN, 2, WGN>;
};

inline int pick_bn(long ldr) {
    int best = 128;
    long pad = -1;
    for (int bn : {128, 192, 256, 320}) {
        const long p = (ldr + bn - 1) / bn * bn - ldr;
        if (pad < 0 || p < pad || (p == pad && bn > best)) {
            best = bn;
            pad = p;
        }
    }
    return best;
}

// One 128 x BN tile of step 1 for the cderi row `row`: rows i0.., columns
// k0.. of the split R planes (row stride ldr; columns nocc..ldr-1 get zeros).
template <bool MAPPED, int BN, typename Src = double>
__device__ __forceinline__ void step1_tile(const Src *__restrict__ row, const int *__restrict__ pmap, int nao,
                                           const u16 *fthi, const u16 *ftlo, long ldft, int nocc, u16 *Rhi,
                                           u16 *Rlo, long ldr, int i0, int k0, u16 *lds) {
    using C = Step1Cfg<BN>;
    PackedA32<MAPPED, C::THREADS, Src> la{row, pmap, nao, i0, 0};
    FTLoader<BN, C::THREADS> lb{fthi, ftlo, ldft, k0, nocc};
    b3::Acc32<C::S::MI, C::S::NI> acc;
    acc.zero();
    b3::mainloop<BM, BN, 2, C::WGN>(la, lb, (nao + b3::BK - 1) / b3::BK, lds, acc);
    // The split tile leaves through LDS one plane at a time (the mainloop's
    // last barrier has freed it): each lane puts its MFMA results at their
    // (row, column), then every thread writes whole 16-byte runs of 8
    // columns.  ldr and k0 are multiples of 32, so a run lies entirely inside
    // or outside ldr.
    constexpr int P = BN + 8;                  // row pitch (u16), 16-byte multiple
    constexpr int CH = BN / 8;                 // 16-byte runs per row
    static_assert(BM * P <= C::S::LDS_ELEMS, "one output plane must fit the tile's LDS");
    for (int plane = 0; plane < 2; plane++) {
        u16 *R = plane ? Rlo : Rhi;
        b3::epilogue<BM, BN, 2, C::WGN>(acc, [&](int r, int c, float v) {
            u16 h, l;
            b3::split(v, h, l);
            lds[r * P + c] = plane ? l : h;
        });
        __syncthreads();
        for (int u = threadIdx.x; u < BM * CH; u += C::THREADS) {
            const int r = u / CH, c = (u % CH) * 8;
            const int i = i0 + r, k = k0 + c;
            if (i < nao && k < ldr)
                *reinterpret_cast<uint4 *>(R + r_index(nao, i, k)) = *reinterpret_cast<const uint4 *>(lds + r * P + c);
        }
        __syncthreads();
    }
}

// Rows r0..r0+127 of the R planes; K-step kt of this split is
// (L, kc) = divmod(s0 + kt, ldr / 32), advanced in place by load().  A
// K-step's 128 rows x 32 k of one plane are 8 KB in one piece (r_index).
struct RLoader {
    using reg_t = uint4;
    static constexpr int N = b3::PlaneRows128::N;
    const u16 *hi, *lo;
    long strideL, ldr;
    int r0, nao, kchunks;
    long L;
    int kc;
    __device__ __forceinline__ void load(int, uint4 (&r)[N]) {
        const int t = threadIdx.x, row = r0 + (t >> 1);
        const bool ok = row < nao;
        const long off = L * strideL + r_index(nao, row, kc * b3::BK) + (t & 1) * 16;
        b3::PlaneRows128::fetch(ok ? hi + off : hi, ok ? lo + off : lo, ok, r);
        if (++kc == kchunks) {
            kc = 0;
            ++L;
        }
    }
    __device__ __forceinline__ void store(const uint4 (&r)[N], u16 *h, u16 *l) const {
        b3::PlaneRows128::put(r, h, l);
    }
};

// One 128 x 128 output tile (i0, j0) of step 2 over the K-steps [s0, s1);
// writes (beta = 0) or adds (beta = 1) into the FP64 partial o.
__device__ __forceinline__ void step2_tile(const u16 *lhi, const u16 *llo, const u16 *rhi, const u16 *rlo, int nao,
                                           long ldr, long strideL, int i0, int j0, long s0, long s1,
                                           double *__restrict__ o, int beta, u16 *lds) {
    const int kchunks = (int)(ldr / b3::BK);
    const long Ls = s0 / kchunks;
    const int kcs = (int)(s0 - Ls * kchunks);
    RLoader la{lhi, llo, strideL, ldr, i0, nao, kchunks, Ls, kcs};
    RLoader lb{rhi, rlo, strideL, ldr, j0, nao, kchunks, Ls, kcs};
    b3::Acc32<b3::Shape<S2_B, S2_B>::MI, b3::Shape<S2_B, S2_B>::NI> acc;
    bool first = true;
    long s = s0;
    do {
        const int n = (int)min((long)S2_FLUSH, s1 - s);
        acc.zero();
        b3::mainloop<S2_B, S2_B, 2, 2>(la, lb, n, lds, acc);
        const bool add = beta || !first;
        b3::epilogue<S2_B, S2_B, 2, 2>(acc, [&](int r, int c, float v) {
            const int i = i0 + r, j = j0 + c;
            if (i < nao && j < nao) {
                double *dst = o + (long)i * nao + j;
                *dst = add ? *dst + (double)v : (double)v;
            }
        });
        first = false;
        s += n;
    } while (s < s1);
}

// FT[k][j] = F[j][k] split into two planes, j zero-padded to ldft, stored
// chunk-major: (j / 32, k, j % 32).
__device__ __forceinline__ void split_ft_range(const double *__restrict__ F, long ldf, int nao, int nocc, u16 *hi,
                                               u16 *lo, long ldft, long e0, long e1) {
    // e runs over the planes in storage order: chunks of 32 j, each nocc
    // rows k of 32 contiguous j (a step-1 K-step then reads BN rows x 32 j
    // of a plane as one contiguous run)
    for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const long chunk = e / ((long)nocc * 32), rem = e % ((long)nocc * 32);
        const int k = (int)(rem / 32), j = (int)(chunk * 32 + rem % 32);
        const double x = j < nao ? F[(long)j * ldf + k] : 0.;
        u16 h, l;
        b3::split((float)x, h, l);
        hi[e] = h;
        lo[e] = l;
    }
}

// Workspace layout of one item (bytes from its base): step-2 partials, then
// per factor side the FT planes and the R planes.
struct Layout {
    long ldr, ldft;
    size_t ft_elems, r_elems;        // per plane
    size_t part_bytes, side_bytes, total;
};

inline Layout layout(int nao, int nocc, long nL, int ksplit, bool two) {
    Layout y;
    y.ldr = pad32(nocc);
    y.ldft = pad32(nao);
    y.ft_elems = (size_t)nocc * y.ldft;
    y.r_elems = (size_t)nL * nao * y.ldr;
    y.part_bytes = align_up((size_t)ksplit * nao * nao * sizeof(double), 256);
    y.side_bytes = align_up(2 * y.ft_elems * sizeof(u16), 256) + align_up(2 * y.r_elems * sizeof(u16), 256);
    y.total = y.part_bytes + (two ? 2 : 1) * y.side_bytes;
    return y;
}

// pointers of one side (0 left, 1 right) inside an item's workspace
struct Side {
    u16 *fthi, *ftlo, *rhi, *rlo;
};
inline Side side_ptrs(char *base, const Layout &y, int side) {
    char *b = base + y.part_bytes + side * y.side_bytes;
    Side s;
    s.fthi = (u16 *)b;
    s.ftlo = s.fthi + y.ft_elems;
    s.rhi = (u16 *)(b + align_up(2 * y.ft_elems * sizeof(u16), 256));
    s.rlo = s.rhi + y.r_elems;
    return s;
}

}  // namespace kb3

}  // namespace kcore
}  // namespace mdf
