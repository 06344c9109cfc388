// Split-BF16 three-term MFMA engine (v_mfma_f32_16x16x32_bf16, gfx950).
//
// Every operand value x is split into two bfloat16 values, hi = bf16(x) and
// lo = bf16(x - hi) (round to nearest even), so x = hi + lo to about 2^-17
// relative.  A product is taken as three MFMA terms,
//     a b ~ lo_a hi_b + hi_a lo_b + hi_a hi_b,
// accumulated in FP32 (the dropped lo_a lo_b term is below 2^-16 relative).
// One MFMA covers 32 reduction steps, so a K-step here is BK = 32.
//
// LDS holds each operand tile as two planes (hi, lo) of [row][k] bfloat16
// with k contiguous and a row pitch of PK = 40 elements (80 bytes): a lane's
// eight k values are one 16-byte ds_read, and eight consecutive rows start in
// distinct bank groups.  Double-buffered; the next K-step's global loads are
// held in registers (in the source type) while the current step's MFMAs run.
//
// MFMA lane layouts (measured with one-hot probes): lane l holds
// A[m = l % 16][k = 8 (l / 16) + s] and B[k = 8 (l / 16) + s][n = l % 16]
// for slot s = 0..7; result element v of lane l is row 4 (l / 16) + v,
// column l % 16 (the FP32 layout, so gemm_epilogue_layout<.., float> applies).
#pragma once
#include "mdf_common.h"
#include "mdf_gemm.h"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

namespace mdf {
namespace b3 {

typedef unsigned short u16;
constexpr int BK = 32;
constexpr int PK = BK + 8;

__device__ __forceinline__ u16 bf16_rne(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__device__ __forceinline__ float bf16_f(u16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ void split(float x, u16 &hi, u16 &lo) {
    hi = bf16_rne(x);
    lo = bf16_rne(x - bf16_f(hi));
}

// A workgroup of WGM x WGN wavefronts computes a BM x BN tile; each
// wavefront a (BM / WGM) x (BN / WGN) sub-tile of MI x NI MFMA blocks.
template <int BM, int BN, int WGM = 2, int WGN = 2>
struct Shape {
    static constexpr int THREADS = 64 * WGM * WGN;
    static constexpr int WM = BM / WGM, WN = BN / WGN;
    static constexpr int MI = WM / 16, NI = WN / 16;
    static_assert(MI * 16 == WM && NI * 16 == WN, "wave tile must be a multiple of 16");
    static constexpr int A_PLANE = BM * PK, B_PLANE = BN * PK;    // u16 elements
    static constexpr int STAGE = 2 * A_PLANE + 2 * B_PLANE;
    static constexpr int LDS_ELEMS = 2 * STAGE;
};

template <int MI, int NI>
using Acc32 = Acc<MI, NI, float>;

__device__ __forceinline__ float4_t mma(bf16x8_t a, bf16x8_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Loader interface:
//   using reg_t; static constexpr int N;
//   void load(int kt, reg_t (&r)[N]);                 // global -> registers
//   void store(const reg_t (&r)[N], u16 *hi, u16 *lo) const;   // -> the two planes
template <int BM, int BN, int WGM, int WGN, class LA, class LB>
__device__ __forceinline__ void mainloop(LA &la, LB &lb, int nk, u16 *lds,
                                         Acc32<Shape<BM, BN, WGM, WGN>::MI, Shape<BM, BN, WGM, WGN>::NI> &acc) {
    using S = Shape<BM, BN, WGM, WGN>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    const int aoff = (wm * S::WM + (lane & 15)) * PK + 8 * (lane >> 4);
    const int boff = 2 * S::A_PLANE + (wn * S::WN + (lane & 15)) * PK + 8 * (lane >> 4);
    typename LA::reg_t ra[LA::N];
    typename LB::reg_t rb[LB::N];
    if (nk <= 0) return;
    la.load(0, ra);
    lb.load(0, rb);
    la.store(ra, lds, lds + S::A_PLANE);
    lb.store(rb, lds + 2 * S::A_PLANE, lds + 2 * S::A_PLANE + S::B_PLANE);
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            la.load(kt + 1, ra);
            lb.load(kt + 1, rb);
        }
        const MDF_LDS u16 *base = (const MDF_LDS u16 *)(lds + cur * S::STAGE);
        bf16x8_t ahi[S::MI], alo[S::MI], bhi[S::NI], blo[S::NI];
#pragma unroll
        for (int mi = 0; mi < S::MI; mi++) {
            ahi[mi] = *(const MDF_LDS bf16x8_t *)(base + aoff + mi * 16 * PK);
            alo[mi] = *(const MDF_LDS bf16x8_t *)(base + aoff + mi * 16 * PK + S::A_PLANE);
        }
#pragma unroll
        for (int ni = 0; ni < S::NI; ni++) {
            bhi[ni] = *(const MDF_LDS bf16x8_t *)(base + boff + ni * 16 * PK);
            blo[ni] = *(const MDF_LDS bf16x8_t *)(base + boff + ni * 16 * PK + S::B_PLANE);
        }
#pragma unroll
        for (int mi = 0; mi < S::MI; mi++)
#pragma unroll
            for (int ni = 0; ni < S::NI; ni++) {
                acc.v[mi][ni] = mma(alo[mi], bhi[ni], acc.v[mi][ni]);
                acc.v[mi][ni] = mma(ahi[mi], blo[ni], acc.v[mi][ni]);
                acc.v[mi][ni] = mma(ahi[mi], bhi[ni], acc.v[mi][ni]);
            }
        if (more) {
            u16 *nxt = lds + (cur ^ 1) * S::STAGE;
            la.store(ra, nxt, nxt + S::A_PLANE);
            lb.store(rb, nxt + 2 * S::A_PLANE, nxt + 2 * S::A_PLANE + S::B_PLANE);
        }
        __syncthreads();
    }
}

// Visits every accumulator element with its tile-local (row, col) (the FP32
// MFMA result layout).
template <int BM, int BN, int WGM, int WGN, class AccT, class F>
__device__ __forceinline__ void epilogue(const AccT &acc, F &&f) {
    using S = Shape<BM, BN, WGM, WGN>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
#pragma unroll
    for (int mi = 0; mi < S::MI; mi++)
#pragma unroll
        for (int ni = 0; ni < S::NI; ni++) {
            const int c = wn * S::WN + ni * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; v++) f(wm * S::WM + mi * 16 + 4 * (lane >> 4) + v, c, acc.v[mi][ni][v]);
        }
}

// Writes NV (8 or 16) values (k = kb .. kb + NV - 1 of LDS row `row`) into
// both planes as 16-byte stores.
template <int NV, typename V>
__device__ __forceinline__ void put(const V (&x)[NV], u16 *hi, u16 *lo, int row, int kb) {
    static_assert(NV % 8 == 0, "put: whole 16-byte groups");
    uint4 h[NV / 8], l[NV / 8];
    u16 *hs = reinterpret_cast<u16 *>(h), *ls = reinterpret_cast<u16 *>(l);
#pragma unroll
    for (int v = 0; v < NV; v++) split((float)x[v], hs[v], ls[v]);
    uint4 *dh = reinterpret_cast<uint4 *>(hi + row * PK + kb);
    uint4 *dl = reinterpret_cast<uint4 *>(lo + row * PK + kb);
#pragma unroll
    for (int g = 0; g < NV / 8; g++) {
        dh[g] = h[g];
        dl[g] = l[g];
    }
}
template <typename V>
__device__ __forceinline__ void put16(const V (&x)[16], u16 *hi, u16 *lo, int row, int kb) {
    put<16>(x, hi, lo, row, kb);
}

// Two already-split planes P[row][k] (row stride ld elements, k contiguous,
// padded with zeros to a multiple of BK) as an operand with 128 rows: thread
// t copies row t / 2, k half t % 2 (16 values = 32 bytes per plane).
// Rows at or beyond nrows read zero.  The K-step index maps to a k offset
// through `koff(kt)` supplied by the owner.
struct PlaneRows128 {
    using reg_t = uint4;
    static constexpr int N = 4;       // hi[0..1], lo[0..1]
    __device__ __forceinline__ static void fetch(const u16 *hi, const u16 *lo, bool ok, uint4 (&r)[N]) {
        if (ok) {
            const uint4 *h = reinterpret_cast<const uint4 *>(hi);
            const uint4 *l = reinterpret_cast<const uint4 *>(lo);
            r[0] = h[0];
            r[1] = h[1];
            r[2] = l[0];
            r[3] = l[1];
        } else {
#pragma unroll
            for (int v = 0; v < N; v++) r[v] = make_uint4(0, 0, 0, 0);
        }
    }
    __device__ __forceinline__ static void put(const uint4 (&r)[N], u16 *hi, u16 *lo) {
        const int t = threadIdx.x, row = t >> 1, kb = (t & 1) * 16;
        uint4 *dh = reinterpret_cast<uint4 *>(hi + row * PK + kb);
        uint4 *dl = reinterpret_cast<uint4 *>(lo + row * PK + kb);
        dh[0] = r[0];
        dh[1] = r[1];
        dl[0] = r[2];
        dl[1] = r[3];
    }
};

// ROWS rows of two split planes (row r0 + r at hi/lo + r * ld + koff, 32
// values per row) for a workgroup of THREADS threads: unit u = t + THREADS q
// copies half u % 2 (16 values) of row u / 2.  Rows at or beyond nrows read
// zero.
template <int ROWS, int THREADS>
struct PlaneRowsN {
    static constexpr int UNITS = 2 * ROWS;
    static constexpr int Q = (UNITS + THREADS - 1) / THREADS;
    using reg_t = uint4;
    static constexpr int N = 4 * Q;
    __device__ __forceinline__ static void fetch(const u16 *hi, const u16 *lo, long ld, long koff, int r0,
                                                 int nrows, uint4 (&r)[N]) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int u = threadIdx.x + THREADS * q, row = u >> 1;
            const bool ok = u < UNITS && r0 + row < nrows;
            const long off = (long)(r0 + row) * ld + koff + (u & 1) * 16;
            if (ok) {
                const uint4 *h = reinterpret_cast<const uint4 *>(hi + off);
                const uint4 *l = reinterpret_cast<const uint4 *>(lo + off);
                r[4 * q] = h[0];
                r[4 * q + 1] = h[1];
                r[4 * q + 2] = l[0];
                r[4 * q + 3] = l[1];
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++) r[4 * q + v] = make_uint4(0, 0, 0, 0);
            }
        }
    }
    __device__ __forceinline__ static void put(const uint4 (&r)[N], u16 *hi, u16 *lo) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int u = threadIdx.x + THREADS * q, row = u >> 1, kb = (u & 1) * 16;
            if (u < UNITS) {
                uint4 *dh = reinterpret_cast<uint4 *>(hi + row * PK + kb);
                uint4 *dl = reinterpret_cast<uint4 *>(lo + row * PK + kb);
                dh[0] = r[4 * q];
                dh[1] = r[4 * q + 1];
                dl[0] = r[4 * q + 2];
                dl[1] = r[4 * q + 3];
            }
        }
    }
};

}  // namespace b3
}  // namespace mdf
