// Grouped J/K over many (molecule, density) items in persistent launches.
//
// The reference has no batched path: each get_jk call contracts one
// molecule's cderi (electrolites/fastdf.py:805-1026).  Many small molecules
// leave most of an MI355X idle, so this entry contracts a whole batch at once:
// a plan flattens every item's work into tile tables (J pass-1 tiles, J pass-2
// tiles, K step-1 tiles, K step-2 tiles and the fixed-order reductions), each
// phase is ONE persistent grid-stride launch over its table, and a run is
// kernel launches only (tables and item descriptors live in device memory the
// plan owns), so a run can be captured in a HIP graph and replayed.
//
// Each item is cut exactly as the single-item entries cut it (mdf_jcore.h /
// mdf_kcore.h plans depend on the item's own shape only: J chunks and row
// groups, K passes of block_rows() aux rows, step-2 splits) and runs the same
// tile bodies, so an item's J and K are bit-identical alone and in a batch.
//
// K runs in waves: consecutive items whose transformed blocks and step-2
// partials fit the workspace limit together share one launch per pass
// (step 1 of pass b for every item of the wave, then step 2 of pass b), and
// the wave ends with its finalize launch; the next wave reuses the same
// workspace.  J runs before K, so K's workspace overlaps J's.
#include <algorithm>
#include <cstring>
#include <vector>

#include "mdf_common.h"
#include "mdf_gemm.h"
#include "mdf_jcore.h"
#include "mdf_kcore.h"
#include "../../include/mi355df.h"

using namespace mdf;

namespace {

constexpr int GJ_NDM = 8;   // densities per molecule in one grouped J pass
// J tiles are launched in classes by the item's density count (at most 1, 2,
// 4 or 8), each with the matching register width: an item with one density
// does not pay for eight accumulators and eight wavefront sums per row.
constexpr int GJ_CLASSES = 4;
inline int gj_class(int ndm) { return ndm <= 1 ? 0 : ndm <= 2 ? 1 : ndm <= 4 ? 2 : 3; }

// Per-molecule J work on the device.
struct GJItem {
    const void *cderi;      // double or float (the plan's storage type)
    long ld, naux, npairs;
    const double *d;        // (ndm, npairs)
    double *rho, *vj;       // (ndm, naux), (ndm, npairs)
    double *part1, *part2;  // workspace
    long nchunk, ngroup, rows_per_group, ncblock;
    int ndm, vec;
};

// Per-(molecule, density) K work on the device.
struct GKItem {
    const void *cderi;
    long ld, naux;
    const int *pmap;
    const double *fl, *fr;  // (nao, nocc); fr NULL -> symmetric
    double *vk;             // (nao, nao)
    void *rl, *rr;          // transformed blocks (T), nL aux rows each
    double *part;           // ksplit x nao x nao
    long nL;                // aux rows per pass
    int nao, nocc, ksplit, ntile, vec;
    // split-BF16 (prec 2): workspace offsets of [side][FT hi, FT lo, R hi, R lo]
    void *b3[2][4];
    long ldr, ldft;         // R row stride (FP64/FP32: kcore::r_ld; split-BF16: a multiple of 32), FT row stride
    int bn3;                // split-BF16 step-1 tile width (kb3::pick_bn(ldr))
    int bn1;                // FP64/FP32 step-1 tile width (the plan's bn, or kcore::pick_bn(nocc))
};

struct Tile {
    int item, a, b, c;
};

// A launch of the K sequence: kind 0 = factor split (prec 2), 1 = step 1,
// 2 = step 2, 3 = finalize, over tiles [start, start + count) of that kind's
// table.  Split-BF16 step-1 launches hold items of one tile width `bn` and
// one `mapped` flag each.
struct Phase {
    int kind;
    long start, count;
    int bn, mapped;
};

}  // namespace

struct mdf_grouped_plan {
    int with_j, with_k, prec, bn;
    int dtype;                          // cderi storage type of every item
    int nj, nk;                         // items
    long t_j1, t_jr1, t_j2, t_jr2;      // tile counts
    long jc1[GJ_CLASSES + 1], jc2[GJ_CLASSES + 1];   // pass-1/2 tile offsets per density class
    long t_k1, t_k2, t_kf;
    void *dev;                          // owned device memory: items + tables
    GJItem *jitems;
    GKItem *kitems;
    Tile *j1, *jr1, *j2, *jr2, *k0, *k1, *k2, *kf;
    long t_k0;
    std::vector<Phase> kphases;
    int kwaves;
    size_t work_bytes;
    // host copies of the per-item workspace offsets (relative to work)
    std::vector<GJItem> hj;
    std::vector<GKItem> hk;
};

namespace {

constexpr long RED_CHUNK = 4096;    // elements per reduction tile
constexpr size_t MDF_GROUPED_DEFAULT_LIMIT = 32ull << 30;   // K wave workspace

template <int NDM, typename S>
__global__ __launch_bounds__(jcore::J_THREADS) void gj_pass1(const GJItem *items, const Tile *tiles, long n,
                                                             char *work) {
    constexpr int W = 16 / sizeof(S);       // values per 16-byte vector
    for (long t = blockIdx.x; t < n; t += gridDim.x) {
        const Tile e = tiles[t];
        const GJItem &it = items[e.item];
        double *part = (double *)(work + (size_t)it.part1);
        const S *L = (const S *)it.cderi;
        if (it.vec == W)
            jcore::rho_partial_tile<NDM, W, S>(L, it.ld, it.naux, it.npairs, it.d, it.npairs, it.ndm, part, e.a, e.b);
        else
            jcore::rho_partial_tile<NDM, 1, S>(L, it.ld, it.naux, it.npairs, it.d, it.npairs, it.ndm, part, e.a, e.b);
    }
}

__global__ __launch_bounds__(jcore::J_THREADS) void gj_reduce1(const GJItem *items, const Tile *tiles, long n,
                                                               char *work) {
    for (long t = blockIdx.x; t < n; t += gridDim.x) {
        const Tile e = tiles[t];
        const GJItem &it = items[e.item];
        const long e0 = (long)e.a * RED_CHUNK, e1 = min(e0 + RED_CHUNK, (long)it.ndm * it.naux);
        jcore::rho_reduce_range((const double *)(work + (size_t)it.part1), it.nchunk, it.ndm, it.naux, it.rho,
                                it.naux, e0, e1);
    }
}

template <int NDM, typename S>
__global__ __launch_bounds__(jcore::J_THREADS) void gj_pass2(const GJItem *items, const Tile *tiles, long n,
                                                             char *work) {
    constexpr int W = 16 / sizeof(S);
    for (long t = blockIdx.x; t < n; t += gridDim.x) {
        const Tile e = tiles[t];
        const GJItem &it = items[e.item];
        const long r0 = (long)e.b * it.rows_per_group, r1 = min(r0 + it.rows_per_group, it.naux);
        double *out = (double *)(work + (size_t)it.part2) + (long)e.b * it.ndm * it.npairs;
        const S *L = (const S *)it.cderi;
        if (it.vec == W)
            jcore::vj_pass_tile<NDM, W, S>(L, it.ld, it.npairs, it.rho, it.naux, it.ndm, r0, r1, out, it.npairs,
                                           e.a);
        else
            jcore::vj_pass_tile<NDM, 1, S>(L, it.ld, it.npairs, it.rho, it.naux, it.ndm, r0, r1, out, it.npairs,
                                           e.a);
    }
}

__global__ __launch_bounds__(jcore::J_THREADS) void gj_reduce2(const GJItem *items, const Tile *tiles, long n,
                                                               char *work) {
    for (long t = blockIdx.x; t < n; t += gridDim.x) {
        const Tile e = tiles[t];
        const GJItem &it = items[e.item];
        const long e0 = (long)e.a * RED_CHUNK, e1 = min(e0 + RED_CHUNK, (long)it.ndm * it.npairs);
        jcore::vj_reduce_range((const double *)(work + (size_t)it.part2), it.ngroup, it.ndm, it.npairs, it.vj,
                               it.npairs, 0, e0, e1);
    }
}

// Step-1 tiles are visited in XCD runs (xcd_tile order): each XCD takes a
// contiguous eighth of the table, so the k-tiles of one A tile share an L2.
#define XCD_RUN_LOOP(t, n)                                                                   \
    const long xr_per = gridDim.x / N_XCD, xr_x = blockIdx.x % N_XCD, xr_j = blockIdx.x / N_XCD; \
    for (long t = (n) * xr_x / N_XCD + xr_j; t < (n) * (xr_x + 1) / N_XCD; t += xr_per)

template <int BN, typename T, typename Src>
__global__ __launch_bounds__(kcore::Step1Shape<BN>::THREADS) void gk_step1(const GKItem *items, const Tile *tiles,
                                                                          long n, char *work) {
    __shared__ T lds[kcore::Step1Shape<BN>::S::LDS_ELEMS];
    XCD_RUN_LOOP(t, n) {
        const Tile e = tiles[t];            // (item, L, i-tile, k-tile | right-factor flag << 16)
        const GKItem &it = items[e.item];
        const bool right = (e.c >> 16) != 0;
        const int ktile = e.c & 0xffff;
        const double *F = right ? it.fr : it.fl;
        const long lb = e.a % it.nL;        // row within the pass
        T *R = (T *)(work + (size_t)(right ? it.rr : it.rl)) + lb * it.nao * it.ldr;
        const Src *row = (const Src *)it.cderi + (long)e.a * it.ld;
        if (it.pmap)
            kcore::step1_tile<BN, true, T, Src>(row, it.pmap, it.nao, F, it.nocc, it.nocc, R, it.ldr, e.b * 128,
                                                ktile * BN, lds);
        else
            kcore::step1_tile<BN, false, T, Src>(row, it.pmap, it.nao, F, it.nocc, it.nocc, R, it.ldr, e.b * 128,
                                                 ktile * BN, lds);
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(G_THREADS, 2) void gk_step2(const GKItem *items, const Tile *tiles, long n,
                                                         char *work) {
    __shared__ T lds[TileShape<kcore::S2_B, kcore::S2_B>::LDS_ELEMS];
    for (long t = blockIdx.x; t < n; t += gridDim.x) {
        const Tile e = tiles[t];            // (item, tile index, split, pass)
        const GKItem &it = items[e.item];
        const long L0 = (long)e.c * it.nL;
        const long nLb = min(it.nL, it.naux - L0);
        int ti, tj;
        if (it.fr == nullptr) {
            kcore::tri_tile(e.a, ti, tj);
        } else {
            ti = e.a / it.ntile;
            tj = e.a % it.ntile;
        }
        const int kchunks = (it.nocc + G_BK - 1) / G_BK;
        const long nsteps = nLb * kchunks;
        const long s0 = nsteps * e.b / it.ksplit, s1 = nsteps * (e.b + 1) / it.ksplit;
        const T *rl = (const T *)(work + (size_t)it.rl);
        const T *rr = it.fr ? (const T *)(work + (size_t)it.rr) : rl;
        double *o = (double *)(work + (size_t)it.part) + (long)e.b * it.nao * it.nao;
        if (it.vec)
            kcore::step2_tile<true, T>(rl, rr, it.nao, it.nocc, it.ldr, (long)it.nao * it.ldr, ti * kcore::S2_B,
                                       tj * kcore::S2_B, s0, s1, o, e.c > 0, lds);
        else
            kcore::step2_tile<false, T>(rl, rr, it.nao, it.nocc, it.ldr, (long)it.nao * it.ldr, ti * kcore::S2_B,
                                        tj * kcore::S2_B, s0, s1, o, e.c > 0, lds);
        __syncthreads();
    }
}

// ---- split-BF16 (prec 2) counterparts
__global__ __launch_bounds__(256) void gk_split_ft(const GKItem *items, const Tile *tiles, long n, char *work) {
    for (long t = blockIdx.x; t < n; t += gridDim.x) {
        const Tile e = tiles[t];            // (item, chunk, side)
        const GKItem &it = items[e.item];
        using kcore::kb3::u16;
        const long e0 = (long)e.a * 1024, e1 = min(e0 + 1024, (long)it.nocc * it.ldft);
        kcore::kb3::split_ft_range(e.b ? it.fr : it.fl, it.nocc, it.nao, it.nocc,
                                   (u16 *)(work + (size_t)it.b3[e.b][0]), (u16 *)(work + (size_t)it.b3[e.b][1]),
                                   it.ldft, e0, e1);
    }
}

template <bool MAPPED, int BN, typename Src>
__global__ __launch_bounds__(kcore::kb3::Step1Cfg<BN>::THREADS) void gk_step1_b3(const GKItem *items,
                                                                                 const Tile *tiles, long n,
                                                                                 char *work) {
    using kcore::kb3::u16;
    __shared__ u16 lds[kcore::kb3::Step1Cfg<BN>::S::LDS_ELEMS];
    XCD_RUN_LOOP(t, n) {
        const Tile e = tiles[t];            // (item, L, i-tile, k-tile | side << 16)
        const GKItem &it = items[e.item];
        const int side = e.c >> 16, ktile = e.c & 0xffff;
        const long lb = e.a % it.nL;
        const long off = lb * it.nao * it.ldr;
        const u16 *fthi = (const u16 *)(work + (size_t)it.b3[side][0]);
        const u16 *ftlo = (const u16 *)(work + (size_t)it.b3[side][1]);
        u16 *rhi = (u16 *)(work + (size_t)it.b3[side][2]) + off;
        u16 *rlo = (u16 *)(work + (size_t)it.b3[side][3]) + off;
        const Src *row = (const Src *)it.cderi + (long)e.a * it.ld;
        // a phase holds only mapped or only unmapped items (one instance
        // each keeps the register budget of the single-item kernel)
        kcore::kb3::step1_tile<MAPPED, BN, Src>(row, it.pmap, it.nao, fthi, ftlo, it.ldft, it.nocc, rhi, rlo, it.ldr,
                                           e.b * kcore::kb3::BM, ktile * BN, lds);
        __syncthreads();
    }
}

__global__ __launch_bounds__(G_THREADS, 2) void gk_step2_b3(const GKItem *items, const Tile *tiles, long n,
                                                             char *work) {
    using kcore::kb3::u16;
    __shared__ u16 lds[b3::Shape<kcore::S2_B, kcore::S2_B>::LDS_ELEMS];
    for (long t = blockIdx.x; t < n; t += gridDim.x) {
        const Tile e = tiles[t];            // (item, tile index, split, pass)
        const GKItem &it = items[e.item];
        const long L0 = (long)e.c * it.nL;
        const long nLb = min(it.nL, it.naux - L0);
        int ti, tj;
        if (it.fr == nullptr) {
            kcore::tri_tile(e.a, ti, tj);
        } else {
            ti = e.a / it.ntile;
            tj = e.a % it.ntile;
        }
        const long nsteps = nLb * (it.ldr / b3::BK);
        const long s0 = nsteps * e.b / it.ksplit, s1 = nsteps * (e.b + 1) / it.ksplit;
        const int rs = it.fr ? 1 : 0;
        double *o = (double *)(work + (size_t)it.part) + (long)e.b * it.nao * it.nao;
        kcore::kb3::step2_tile((const u16 *)(work + (size_t)it.b3[0][2]), (const u16 *)(work + (size_t)it.b3[0][3]),
                               (const u16 *)(work + (size_t)it.b3[rs][2]), (const u16 *)(work + (size_t)it.b3[rs][3]),
                               it.nao, it.ldr, (long)it.nao * it.ldr, ti * kcore::S2_B, tj * kcore::S2_B, s0, s1, o,
                               e.c > 0, lds);
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void gk_finalize(const GKItem *items, const Tile *tiles, long n, char *work) {
    for (long t = blockIdx.x; t < n; t += gridDim.x) {
        const Tile e = tiles[t];
        const GKItem &it = items[e.item];
        const long nn = (long)it.nao * it.nao;
        const long e0 = (long)e.a * RED_CHUNK, e1 = min(e0 + RED_CHUNK, nn);
        kcore::finalize_range((const double *)(work + (size_t)it.part), it.ksplit, nn, it.nao, it.fr == nullptr,
                              it.vk, it.nao, 0, e0, e1);
    }
}

int persistent_grid(long ntiles, int per_cu) {
    int dev = 0, cu = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev);
    long g = (long)cu * per_cu;
    if (g > ntiles) g = ntiles;
    return (int)round_xcd(g > 0 ? g : 1);
}

}  // namespace

extern "C" int mdf_grouped_plan_create(const mdf_jk_item *items, int nitems, int with_j, int with_k, int prec,
                                       int bn, size_t work_limit, mdf_grouped_plan **out, size_t *work_bytes) {
    MDF_CHECK_ARG(items && nitems >= 0 && out && work_bytes, "mdf_grouped_plan_create: null argument");
    MDF_CHECK_ARG(prec == MDF_PREC_FP64 || prec == MDF_PREC_FP32 || prec == MDF_PREC_BF16X3,
                  "mdf_grouped_plan_create: precision %d", prec);
    MDF_CHECK_ARG(bn == 0 || bn == 64 || bn == 128 || bn == 160 || bn == 192 || bn == 256 || bn == 320,
                  "mdf_grouped_plan_create: bn %d not in {0,64,128,160,192,256,320}", bn);
    for (int m = 0; m < nitems; m++)
        MDF_CHECK_ARG(items[m].cderi_dtype == items[0].cderi_dtype &&
                          (items[m].cderi_dtype == MDF_DTYPE_F64 || items[m].cderi_dtype == MDF_DTYPE_F32),
                      "mdf_grouped_plan_create: item %d cderi dtype %d (one storage type per plan)", m,
                      items[m].cderi_dtype);
    const size_t elem = prec == MDF_PREC_FP32 ? sizeof(float) : sizeof(double);
    mdf_grouped_plan *p = new mdf_grouped_plan();
    p->with_j = with_j;
    p->with_k = with_k;
    p->prec = prec;
    p->dtype = nitems ? items[0].cderi_dtype : MDF_DTYPE_F64;
    const int wvec = p->dtype == MDF_DTYPE_F32 ? 4 : 2;     // values per 16-byte vector
    std::vector<Tile> j1, jr1, j2, jr2, k0, k1, k2, kf;
    std::vector<Tile> j1c[GJ_CLASSES], j2c[GJ_CLASSES];
    std::vector<int> kgrid;             // step-2 output tiles per K item
    if (work_limit == 0) work_limit = MDF_GROUPED_DEFAULT_LIMIT;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = align_up(off + bytes, 256); return o; };
    int maxocc = 0;
    for (int m = 0; m < nitems; m++) maxocc = items[m].nocc > maxocc ? items[m].nocc : maxocc;
    p->bn = bn ? bn : kcore::pick_bn(maxocc > 0 ? maxocc : 1);
    for (int m = 0; m < nitems; m++) {
        const mdf_jk_item &x = items[m];
        if (x.n_dm < 1 || x.n_dm > GJ_NDM || x.naux < 0 || x.nao < 0 || x.ld < x.npairs || !x.cderi ||
            (!x.pair_map && x.npairs != (long)x.nao * (x.nao + 1) / 2)) {
            set_error("mdf_grouped_plan_create: item %d invalid (n_dm %d, shapes, or pair map)", m, x.n_dm);
            delete p;
            return MDF_EINVAL;
        }
        if (with_j) {
            if (!x.dm_pairs || !x.rho || !x.vj_pairs) {
                set_error("mdf_grouped_plan_create: item %d lacks J buffers", m);
                delete p;
                return MDF_EINVAL;
            }
            GJItem g{};
            g.cderi = x.cderi; g.ld = x.ld; g.naux = x.naux; g.npairs = x.npairs;
            g.d = x.dm_pairs; g.rho = x.rho; g.vj = x.vj_pairs; g.ndm = x.n_dm;
            g.vec = ((x.ld % wvec == 0) && ((uintptr_t)x.cderi % 16 == 0)) ? wvec : 1;
            jcore::JPlan pl = jcore::plan_j(x.naux, x.npairs, x.n_dm, g.vec);
            g.nchunk = pl.nchunk; g.ngroup = pl.ngroup; g.rows_per_group = pl.rows_per_group; g.ncblock = pl.ncblock;
            g.part1 = (double *)take(pl.part1_words * sizeof(double));
            g.part2 = (double *)take(pl.part2_words * sizeof(double));
            const int ji = (int)p->hj.size();
            const int jc = gj_class(x.n_dm);
            for (long c = 0; c < pl.nchunk; c++)
                for (long r = 0; r < pl.nrblock; r++) j1c[jc].push_back({ji, (int)c, (int)r, 0});
            for (long e = 0; e < cdiv((long)x.n_dm * x.naux, RED_CHUNK); e++) jr1.push_back({ji, (int)e, 0, 0});
            for (long g2 = 0; g2 < pl.ngroup; g2++)
                for (long c = 0; c < pl.ncblock; c++) j2c[jc].push_back({ji, (int)c, (int)g2, 0});
            for (long e = 0; e < cdiv((long)x.n_dm * x.npairs, RED_CHUNK); e++) jr2.push_back({ji, (int)e, 0, 0});
            p->hj.push_back(g);
        }
        if (with_k && x.nocc > 0 && x.naux > 0) {
            if (!x.fl || !x.vk) {
                set_error("mdf_grouped_plan_create: item %d lacks K buffers", m);
                delete p;
                return MDF_EINVAL;
            }
            const bool two = x.fr != nullptr;
            const bool b3 = prec == MDF_PREC_BF16X3;
            const long nL = b3 ? kcore::block_rows(x.naux, x.nao, (int)kcore::kb3::pad32(x.nocc), two, 4)
                               : kcore::block_rows(x.naux, x.nao, x.nocc, two, elem);
            kcore::KPlan kp = kcore::plan_k(x.naux, x.nao, x.nocc, two, nL, b3 ? sizeof(double) : elem);
            for (int d = 0; d < x.n_dm; d++) {
                GKItem g{};
                g.cderi = x.cderi; g.ld = x.ld; g.naux = x.naux; g.pmap = x.pair_map;
                g.fl = x.fl + (long)d * x.nao * x.nocc;
                g.fr = two ? x.fr + (long)d * x.nao * x.nocc : nullptr;
                g.vk = x.vk + (long)d * x.nao * x.nao;
                g.nao = x.nao; g.nocc = x.nocc; g.ksplit = kp.ksplit; g.ntile = kp.ntile;
                g.nL = kp.nL;
                g.bn1 = bn ? bn : kcore::pick_bn(x.nocc);
                g.vec = 1;             // R rows padded to whole 16-byte vectors (kcore::r_ld)
                // workspace offsets are assigned per wave below; keep the sizes
                if (b3) {
                    const kcore::kb3::Layout y = kcore::kb3::layout(x.nao, x.nocc, kp.nL, kp.ksplit, two);
                    g.ldr = y.ldr;
                    g.ldft = y.ldft;
                    g.bn3 = kcore::kb3::pick_bn(y.ldr);
                    g.rl = (void *)0;
                    g.part = (double *)y.total;
                } else {
                    g.ldr = kcore::r_ld(x.nocc, elem);
                    g.rl = (void *)(align_up(kp.r_words * elem, 256));
                    g.part = (double *)(align_up(kp.part_words * sizeof(double), 256));
                }
                g.rr = nullptr;
                p->hk.push_back(g);
                kgrid.push_back(kp.ntiles_grid);
            }
        }
    }
    // K waves: consecutive items whose blocks and partials fit work_limit
    // (at least one item per wave); every wave's workspace starts at 0
    size_t kmax = 0;
    p->kwaves = 0;
    for (size_t w0 = 0; w0 < p->hk.size();) {
        size_t used = 0, w1 = w0;
        while (w1 < p->hk.size()) {
            const GKItem &g = p->hk[w1];
            const size_t need = (g.fr ? 2 : 1) * (size_t)g.rl + (size_t)g.part;
            if (w1 > w0 && used + need > work_limit) break;
            used += need;
            w1++;
        }
        size_t o = 0;
        long maxpass = 0;
        for (size_t i = w0; i < w1; i++) {
            GKItem &g = p->hk[i];
            if (prec == MDF_PREC_BF16X3) {
                const kcore::kb3::Layout y = kcore::kb3::layout(g.nao, g.nocc, g.nL, g.ksplit, g.fr != nullptr);
                for (int side = 0; side < 2; side++) {
                    const kcore::kb3::Side q = kcore::kb3::side_ptrs((char *)o, y, g.fr ? side : 0);
                    g.b3[side][0] = q.fthi; g.b3[side][1] = q.ftlo; g.b3[side][2] = q.rhi; g.b3[side][3] = q.rlo;
                }
                g.part = (double *)o;
                o += y.total;
            } else {
                const size_t rb = (size_t)g.rl, pb = (size_t)g.part;
                g.rl = (void *)o; o += rb;
                if (g.fr) { g.rr = (void *)o; o += rb; }
                g.part = (double *)o; o += pb;
            }
            maxpass = std::max(maxpass, cdiv(g.naux, g.nL));
        }
        kmax = std::max(kmax, o);
        if (prec == MDF_PREC_BF16X3) {
            const long s0 = k0.size();
            for (size_t i = w0; i < w1; i++) {
                const GKItem &g = p->hk[i];
                for (int side = 0; side < (g.fr ? 2 : 1); side++)
                    for (long e = 0; e < cdiv((long)g.nocc * g.ldft, 1024L); e++) k0.push_back({(int)i, (int)e, side, 0});
            }
            p->kphases.push_back({0, s0, (long)k0.size() - s0, 0, 0});
        }
        const bool b3 = prec == MDF_PREC_BF16X3;
        for (long b = 0; b < maxpass; b++) {
          // one step-1 launch per tile width (and, for split-BF16, per pair-map flag)
          for (int cls = 0; cls < (b3 ? 8 : 6); cls++) {
            static const int kBN3[4] = {128, 192, 256, 320};
            static const int kBN1[6] = {64, 128, 160, 192, 256, 320};
            const int mp = b3 ? cls / 4 : 0, cbn = b3 ? kBN3[cls % 4] : kBN1[cls];
            const long s1 = k1.size();
            for (size_t i = w0; i < w1; i++) {
                const GKItem &g = p->hk[i];
                const long L0 = b * g.nL;
                if (L0 >= g.naux) continue;
                if (b3 && ((g.pmap != nullptr) != (mp == 1) || g.bn3 != cbn)) continue;
                if (!b3 && g.bn1 != cbn) continue;
                const long L1 = std::min(L0 + g.nL, g.naux);
                const int nkt = b3 ? (int)cdiv(g.ldr, (long)cbn) : (int)cdiv(g.nocc, cbn);
                const int nit = (int)cdiv(g.nao, 128);
                for (int side = 0; side < (g.fr ? 2 : 1); side++)
                    for (long L = L0; L < L1; L++)
                        for (int it = 0; it < nit; it++)
                            for (int kt = 0; kt < nkt; kt++) k1.push_back({(int)i, (int)L, it, kt | (side << 16)});
            }
            // longest tiles first (cost ~ the item's nao K-steps): the
            // grid-stride loop then ends on short tiles, not on one long one;
            // each item's tiles stay contiguous for the XCD runs
            std::stable_sort(k1.begin() + s1, k1.end(), [&](const Tile &x, const Tile &y) {
                return p->hk[x.item].nao > p->hk[y.item].nao;
            });
            p->kphases.push_back({1, s1, (long)k1.size() - s1, cbn, mp});
          }
            const long s2 = k2.size();
            for (size_t i = w0; i < w1; i++) {
                const GKItem &g = p->hk[i];
                if (b * g.nL >= g.naux) continue;
                for (int s = 0; s < g.ksplit; s++)
                    for (int tl = 0; tl < kgrid[i]; tl++) k2.push_back({(int)i, tl, s, (int)b});
            }
            // longest first: a step-2 tile's cost is its share of the pass's
            // (L, k) steps
            auto s2cost = [&](const Tile &x) {
                const GKItem &g = p->hk[x.item];
                const long nLb = std::min(g.nL, g.naux - (long)x.c * g.nL);
                const long kc = prec == MDF_PREC_BF16X3 ? g.ldr / b3::BK : cdiv((long)g.nocc, (long)G_BK);
                return nLb * kc / g.ksplit;
            };
            std::stable_sort(k2.begin() + s2, k2.end(),
                             [&](const Tile &x, const Tile &y) { return s2cost(x) > s2cost(y); });
            p->kphases.push_back({2, s2, (long)k2.size() - s2, 0, 0});
        }
        const long sf = kf.size();
        for (size_t i = w0; i < w1; i++) {
            const GKItem &g = p->hk[i];
            for (long e = 0; e < cdiv((long)g.nao * g.nao, RED_CHUNK); e++) kf.push_back({(int)i, (int)e, 0, 0});
        }
        p->kphases.push_back({3, sf, (long)kf.size() - sf, 0, 0});
        p->kwaves++;
        w0 = w1;
    }
    off = std::max(off, kmax);
    p->nj = (int)p->hj.size();
    p->nk = (int)p->hk.size();
    for (int c = 0; c < GJ_CLASSES; c++) {
        p->jc1[c] = (long)j1.size();
        p->jc2[c] = (long)j2.size();
        j1.insert(j1.end(), j1c[c].begin(), j1c[c].end());
        j2.insert(j2.end(), j2c[c].begin(), j2c[c].end());
    }
    p->jc1[GJ_CLASSES] = (long)j1.size();
    p->jc2[GJ_CLASSES] = (long)j2.size();
    p->t_j1 = j1.size(); p->t_jr1 = jr1.size(); p->t_j2 = j2.size(); p->t_jr2 = jr2.size();
    p->t_k0 = k0.size(); p->t_k1 = k1.size(); p->t_k2 = k2.size(); p->t_kf = kf.size();
    p->work_bytes = off + 256;
    // device block: items then tables
    size_t dsz = 0;
    auto dtake = [&](size_t bytes) { size_t o = dsz; dsz = align_up(dsz + bytes, 256); return o; };
    const size_t o_j = dtake(p->hj.size() * sizeof(GJItem)), o_k = dtake(p->hk.size() * sizeof(GKItem));
    const size_t o_j1 = dtake(j1.size() * sizeof(Tile)), o_jr1 = dtake(jr1.size() * sizeof(Tile));
    const size_t o_j2 = dtake(j2.size() * sizeof(Tile)), o_jr2 = dtake(jr2.size() * sizeof(Tile));
    const size_t o_k1 = dtake(k1.size() * sizeof(Tile)), o_k2 = dtake(k2.size() * sizeof(Tile));
    const size_t o_kf = dtake(kf.size() * sizeof(Tile));
    const size_t o_k0 = dtake(k0.size() * sizeof(Tile));
    std::vector<char> host(dsz + 256, 0);
    auto put = [&](size_t o, const void *src, size_t bytes) { if (bytes) memcpy(host.data() + o, src, bytes); };
    put(o_j, p->hj.data(), p->hj.size() * sizeof(GJItem));
    put(o_k, p->hk.data(), p->hk.size() * sizeof(GKItem));
    put(o_j1, j1.data(), j1.size() * sizeof(Tile));
    put(o_jr1, jr1.data(), jr1.size() * sizeof(Tile));
    put(o_j2, j2.data(), j2.size() * sizeof(Tile));
    put(o_jr2, jr2.data(), jr2.size() * sizeof(Tile));
    put(o_k1, k1.data(), k1.size() * sizeof(Tile));
    put(o_k2, k2.data(), k2.size() * sizeof(Tile));
    put(o_kf, kf.data(), kf.size() * sizeof(Tile));
    put(o_k0, k0.data(), k0.size() * sizeof(Tile));
    hipError_t e = hipMalloc(&p->dev, host.size());
    if (e != hipSuccess) {
        set_error("mdf_grouped_plan_create: hipMalloc(%zu) failed: %s", host.size(), hipGetErrorString(e));
        delete p;
        return MDF_EHIP;
    }
    // the upload goes through a private non-blocking stream: a plain
    // hipMemcpy would also wait for the work queued on the null stream (a
    // serving loop creates plans while its previous runs are in flight)
    hipStream_t up = nullptr;
    e = hipStreamCreateWithFlags(&up, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemcpyAsync(p->dev, host.data(), host.size(), hipMemcpyHostToDevice, up);
    if (e == hipSuccess) e = hipStreamSynchronize(up);
    if (up) (void)hipStreamDestroy(up);
    if (e != hipSuccess) {
        set_error("mdf_grouped_plan_create: copy failed: %s", hipGetErrorString(e));
        (void)hipFree(p->dev);
        delete p;
        return MDF_EHIP;
    }
    char *d = (char *)p->dev;
    p->jitems = (GJItem *)(d + o_j); p->kitems = (GKItem *)(d + o_k);
    p->j1 = (Tile *)(d + o_j1); p->jr1 = (Tile *)(d + o_jr1); p->j2 = (Tile *)(d + o_j2);
    p->jr2 = (Tile *)(d + o_jr2); p->k1 = (Tile *)(d + o_k1); p->k2 = (Tile *)(d + o_k2);
    p->kf = (Tile *)(d + o_kf);
    p->k0 = (Tile *)(d + o_k0);
    *out = p;
    *work_bytes = p->work_bytes;
    return MDF_OK;
}

// Pass PASS (1 or 2) of the grouped J over n tiles of density class c.
// Persistent workgroups, eight per CU for one or two densities (few
// registers; the loads need many waves in flight), four otherwise.
template <int PASS, int NDM, typename S>
static void launch_gj_t(const GJItem *items, const Tile *tiles, long n, char *w, hipStream_t st) {
    const dim3 g(persistent_grid(n, NDM <= 2 ? 8 : 4)), b(jcore::J_THREADS);
    if (PASS == 1)
        hipLaunchKernelGGL((gj_pass1<NDM, S>), g, b, 0, st, items, tiles, n, w);
    else
        hipLaunchKernelGGL((gj_pass2<NDM, S>), g, b, 0, st, items, tiles, n, w);
}

template <int PASS, typename S>
static void launch_gj_s(int c, const GJItem *items, const Tile *tiles, long n, char *w, hipStream_t st) {
    switch (c) {
    case 0: launch_gj_t<PASS, 1, S>(items, tiles, n, w, st); break;
    case 1: launch_gj_t<PASS, 2, S>(items, tiles, n, w, st); break;
    case 2: launch_gj_t<PASS, 4, S>(items, tiles, n, w, st); break;
    default: launch_gj_t<PASS, GJ_NDM, S>(items, tiles, n, w, st); break;
    }
}

template <int PASS>
static void launch_gj(int c, int dtype, const GJItem *items, const Tile *tiles, long n, char *w, hipStream_t st) {
    if (dtype == MDF_DTYPE_F32)
        launch_gj_s<PASS, float>(c, items, tiles, n, w, st);
    else
        launch_gj_s<PASS, double>(c, items, tiles, n, w, st);
}

template <int BN, typename Src>
static void launch_gk1_s(const mdf_grouped_plan *p, const Tile *tiles, long n, char *w, hipStream_t st) {
    const int g = persistent_grid(n, BN >= 192 ? 1 : 2);
    const dim3 b(kcore::Step1Shape<BN>::THREADS);
    if (p->prec == MDF_PREC_FP32)
        hipLaunchKernelGGL((gk_step1<BN, float, Src>), dim3(g), b, 0, st, p->kitems, tiles, n, w);
    else
        hipLaunchKernelGGL((gk_step1<BN, double, Src>), dim3(g), b, 0, st, p->kitems, tiles, n, w);
}

template <int BN>
static void launch_gk1(const mdf_grouped_plan *p, const Tile *tiles, long n, char *w, hipStream_t st) {
    if (p->dtype == MDF_DTYPE_F32)
        launch_gk1_s<BN, float>(p, tiles, n, w, st);
    else
        launch_gk1_s<BN, double>(p, tiles, n, w, st);
}

template <bool MAPPED, int BN>
static void launch_gk1_b3(const mdf_grouped_plan *p, const Tile *tiles, long n, char *w, hipStream_t st) {
    // one workgroup per CU (the 192-, 256- and 320-wide tiles hold 100-140 KB of LDS)
    const int per_cu = BN == 128 ? 2 : 1;
    const dim3 g(persistent_grid(n, per_cu)), b(kcore::kb3::Step1Cfg<BN>::THREADS);
    if (p->dtype == MDF_DTYPE_F32)
        hipLaunchKernelGGL((gk_step1_b3<MAPPED, BN, float>), g, b, 0, st, p->kitems, tiles, n, w);
    else
        hipLaunchKernelGGL((gk_step1_b3<MAPPED, BN, double>), g, b, 0, st, p->kitems, tiles, n, w);
}

extern "C" int mdf_grouped_jk(const mdf_grouped_plan *p, void *work, size_t work_bytes, void *stream) {
    MDF_CHECK_ARG(p && (work || p->work_bytes <= 256), "mdf_grouped_jk: null argument");
    if (work_bytes < p->work_bytes) {
        set_error("mdf_grouped_jk: workspace %zu bytes < %zu required", work_bytes, p->work_bytes);
        return MDF_EWORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    char *w = (char *)work;
    if (p->with_j && p->nj) {
        for (int c = 0; c < GJ_CLASSES; c++) {
            const long n = p->jc1[c + 1] - p->jc1[c];
            if (n == 0) continue;
            launch_gj<1>(c, p->dtype, p->jitems, p->j1 + p->jc1[c], n, w, st);
            MDF_LAUNCHED("gj_pass1");
        }
        hipLaunchKernelGGL(gj_reduce1, dim3(persistent_grid(p->t_jr1, 4)), dim3(jcore::J_THREADS), 0, st, p->jitems,
                           p->jr1, p->t_jr1, w);
        MDF_LAUNCHED("gj_reduce1");
        for (int c = 0; c < GJ_CLASSES; c++) {
            const long n = p->jc2[c + 1] - p->jc2[c];
            if (n == 0) continue;
            launch_gj<2>(c, p->dtype, p->jitems, p->j2 + p->jc2[c], n, w, st);
            MDF_LAUNCHED("gj_pass2");
        }
        hipLaunchKernelGGL(gj_reduce2, dim3(persistent_grid(p->t_jr2, 4)), dim3(jcore::J_THREADS), 0, st, p->jitems,
                           p->jr2, p->t_jr2, w);
        MDF_LAUNCHED("gj_reduce2");
    }
    if (p->with_k && p->nk) {
        const bool b3 = p->prec == MDF_PREC_BF16X3;
        for (const Phase &ph : p->kphases) {
            if (ph.count == 0) continue;
            if (ph.kind == 0) {
                hipLaunchKernelGGL(gk_split_ft, dim3(persistent_grid(ph.count, 4)), dim3(256), 0, st, p->kitems,
                                   p->k0 + ph.start, ph.count, w);
                MDF_LAUNCHED("gk_split_ft");
            } else if (ph.kind == 1 && b3) {
                const Tile *t1 = p->k1 + ph.start;
                switch (ph.bn * 2 + ph.mapped) {
                case 640: launch_gk1_b3<false, 320>(p, t1, ph.count, w, st); break;
                case 641: launch_gk1_b3<true, 320>(p, t1, ph.count, w, st); break;
                case 512: launch_gk1_b3<false, 256>(p, t1, ph.count, w, st); break;
                case 513: launch_gk1_b3<true, 256>(p, t1, ph.count, w, st); break;
                case 384: launch_gk1_b3<false, 192>(p, t1, ph.count, w, st); break;
                case 385: launch_gk1_b3<true, 192>(p, t1, ph.count, w, st); break;
                case 257: launch_gk1_b3<true, 128>(p, t1, ph.count, w, st); break;
                default: launch_gk1_b3<false, 128>(p, t1, ph.count, w, st); break;
                }
                MDF_LAUNCHED("gk_step1_b3");
            } else if (ph.kind == 2 && b3) {
                hipLaunchKernelGGL(gk_step2_b3, dim3(persistent_grid(ph.count, 2)), dim3(G_THREADS), 0, st,
                                   p->kitems, p->k2 + ph.start, ph.count, w);
                MDF_LAUNCHED("gk_step2_b3");
            } else if (ph.kind == 1) {
                const Tile *t = p->k1 + ph.start;
                switch (ph.bn) {
                case 320: launch_gk1<320>(p, t, ph.count, w, st); break;
                case 256: launch_gk1<256>(p, t, ph.count, w, st); break;
                case 192: launch_gk1<192>(p, t, ph.count, w, st); break;
                case 160: launch_gk1<160>(p, t, ph.count, w, st); break;
                case 128: launch_gk1<128>(p, t, ph.count, w, st); break;
                default: launch_gk1<64>(p, t, ph.count, w, st); break;
                }
                MDF_LAUNCHED("gk_step1");
            } else if (ph.kind == 2) {
                const int g2 = persistent_grid(ph.count, 2);
                if (p->prec == MDF_PREC_FP32)
                    hipLaunchKernelGGL(gk_step2<float>, dim3(g2), dim3(G_THREADS), 0, st, p->kitems,
                                       p->k2 + ph.start, ph.count, w);
                else
                    hipLaunchKernelGGL(gk_step2<double>, dim3(g2), dim3(G_THREADS), 0, st, p->kitems,
                                       p->k2 + ph.start, ph.count, w);
                MDF_LAUNCHED("gk_step2");
            } else {
                hipLaunchKernelGGL(gk_finalize, dim3(persistent_grid(ph.count, 4)), dim3(256), 0, st, p->kitems,
                                   p->kf + ph.start, ph.count, w);
                MDF_LAUNCHED("gk_finalize");
            }
        }
    }
    return MDF_OK;
}

extern "C" int mdf_grouped_plan_destroy(mdf_grouped_plan *p) {
    if (!p) return MDF_OK;
    if (p->dev) MDF_HIP(hipFree(p->dev));
    delete p;
    return MDF_OK;
}

extern "C" int mdf_grouped_plan_info(const mdf_grouped_plan *p, long *tiles, int ntiles) {
    MDF_CHECK_ARG(p && tiles && ntiles >= 7, "mdf_grouped_plan_info: bad arguments");
    tiles[0] = p->t_j1; tiles[1] = p->t_jr1; tiles[2] = p->t_j2; tiles[3] = p->t_jr2;
    tiles[4] = p->t_k1; tiles[5] = p->t_k2; tiles[6] = p->t_kf;
    if (ntiles >= 9) {
        tiles[7] = p->kwaves;
        tiles[8] = (long)p->kphases.size();
    }
    return MDF_OK;
}
