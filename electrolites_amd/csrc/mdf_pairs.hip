// Pair-vector gather/scatter around the J/K contractions.
//
// mdf_pack_dm restates electrolites/fastdf.py:674-686 (_sparse_dm):
//     d[n, p] = 2 D[i, j]            (hermi = 1)
//             = D[i, j] + D[j, i]    (hermi = 0), diagonal pairs halved.
// mdf_unpack_pairs restates the vj scatter of fastdf.py:1014-1015 and
// :736-738: vj[:, i, j] = vj[:, j, i] = vj_pairs, zero where no pair is
// stored.
// Pairs are given by an address list (addr = i*nao + j, i >= j) or, when the
// list is NULL, are the full lower triangle p = i(i+1)/2 + j.
#include <cmath>

#include "mdf_common.h"
#include "../../include/mi355df.h"

namespace {

__device__ inline void pair_ij(long p, const int *addr, int nao, int &i, int &j) {
    if (addr) {
        int a = addr[p];
        i = a / nao;
        j = a % nao;
    } else {
        long r = (long)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
        while (r * (r + 1) / 2 > p) r--;
        while ((r + 1) * (r + 2) / 2 <= p) r++;
        i = (int)r;
        j = (int)(p - r * (r + 1) / 2);
    }
}

__global__ __launch_bounds__(256) void pack_dm_kernel(const double *__restrict__ dms, int ndm, int nao,
                                                      long ldm, const int *__restrict__ addr, long npairs,
                                                      int hermi, double *__restrict__ d, long ldd) {
    long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= npairs) return;
    int i, j;
    pair_ij(p, addr, nao, i, j);
    for (int n = 0; n < ndm; n++) {
        const double *D = dms + n * ldm;
        double v = hermi ? 2. * D[(long)i * nao + j] : D[(long)i * nao + j] + D[(long)j * nao + i];
        if (i == j) v *= .5;
        d[n * ldd + p] = v;
    }
}

__global__ __launch_bounds__(256) void fill_zero_kernel(double *__restrict__ x, long n) {
    long t = (long)blockIdx.x * 256 + threadIdx.x;
    long stride = (long)gridDim.x * 256;
    for (; t < n; t += stride) x[t] = 0.;
}

__global__ __launch_bounds__(256) void unpack_pairs_kernel(const double *__restrict__ vp, int ndm, long ldv,
                                                           const int *__restrict__ addr, long npairs, int nao,
                                                           double *__restrict__ out, long ldo) {
    long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= npairs) return;
    int i, j;
    pair_ij(p, addr, nao, i, j);
    for (int n = 0; n < ndm; n++) {
        double v = vp[n * ldv + p];
        out[n * ldo + (long)i * nao + j] = v;
        out[n * ldo + (long)j * nao + i] = v;
    }
}

}  // namespace

extern "C" int mdf_pack_dm(const double *dms, int n_dm, int nao, long ldm, const int *pair_addr,
                           long npairs, int hermi, double *dm_pairs, long ldd, void *stream) {
    MDF_CHECK_ARG(dms && dm_pairs && n_dm >= 1 && nao >= 0 && ldm >= (long)nao * nao && ldd >= npairs,
                  "mdf_pack_dm: bad arguments");
    MDF_CHECK_ARG(pair_addr || npairs == (long)nao * (nao + 1) / 2,
                  "mdf_pack_dm: without pair addresses npairs must be nao(nao+1)/2");
    if (npairs == 0) return MDF_OK;
    hipLaunchKernelGGL(pack_dm_kernel, dim3((unsigned)mdf::cdiv(npairs, 256L)), dim3(256), 0,
                       mdf::as_stream(stream), dms, n_dm, nao, ldm, pair_addr, npairs, hermi, dm_pairs, ldd);
    MDF_LAUNCHED("pack_dm_kernel");
    return MDF_OK;
}

extern "C" int mdf_unpack_pairs(const double *vj_pairs, int n_dm, long ldv, const int *pair_addr,
                                long npairs, int nao, double *out, long ldo, void *stream) {
    MDF_CHECK_ARG(vj_pairs && out && n_dm >= 1 && ldo >= (long)nao * nao && ldv >= npairs,
                  "mdf_unpack_pairs: bad arguments");
    MDF_CHECK_ARG(pair_addr || npairs == (long)nao * (nao + 1) / 2,
                  "mdf_unpack_pairs: without pair addresses npairs must be nao(nao+1)/2");
    hipStream_t st = mdf::as_stream(stream);
    if (pair_addr) {
        // screened pairs leave holes that must read as zero
        long n = (long)n_dm * ldo;
        unsigned g = (unsigned)std::min(mdf::cdiv(n, 256L), 65536L);
        hipLaunchKernelGGL(fill_zero_kernel, dim3(g), dim3(256), 0, st, out, n);
        MDF_LAUNCHED("fill_zero_kernel");
    }
    if (npairs == 0) return MDF_OK;
    hipLaunchKernelGGL(unpack_pairs_kernel, dim3((unsigned)mdf::cdiv(npairs, 256L)), dim3(256), 0, st,
                       vj_pairs, n_dm, ldv, pair_addr, npairs, nao, out, ldo);
    MDF_LAUNCHED("unpack_pairs_kernel");
    return MDF_OK;
}

__global__ __launch_bounds__(256) void cast_f32_kernel(const double *__restrict__ src, long lds, long cols,
                                                       float *__restrict__ dst, long ldd, long total) {
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long r = e / cols, c = e % cols;
        dst[r * ldd + c] = (float)src[r * lds + c];
    }
}

extern "C" int mdf_cast_f32(const double *src, long lds, long rows, long cols, float *dst, long ldd,
                            void *stream) {
    MDF_CHECK_ARG(rows >= 0 && cols >= 0 && lds >= cols && ldd >= cols, "mdf_cast_f32: bad shape");
    MDF_CHECK_ARG((src && dst) || rows * cols == 0, "mdf_cast_f32: null pointer");
    const long total = rows * cols;
    if (total == 0) return MDF_OK;
    const long blocks = mdf::cdiv(total, 256L);
    hipLaunchKernelGGL(cast_f32_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0,
                       mdf::as_stream(stream), src, lds, cols, dst, ldd, total);
    MDF_LAUNCHED("cast_f32_kernel");
    return MDF_OK;
}

namespace {
// one row per blockIdx.y (grid-strided over rows), columns grid-strided over
// blockIdx.x: no per-element division, coalesced on both sides
__global__ __launch_bounds__(256) void widen_f32_kernel(const float *__restrict__ src, long lds, long rows, long cols,
                                                        double *__restrict__ dst, long ldd) {
    for (long r = blockIdx.y; r < rows; r += gridDim.y) {
        const float *s = src + r * lds;
        double *d = dst + r * ldd;
        for (long c = (long)blockIdx.x * 256 + threadIdx.x; c < cols; c += (long)gridDim.x * 256) d[c] = (double)s[c];
    }
}
}  // namespace

extern "C" int mdf_widen_f32(const float *src, long lds, long rows, long cols, double *dst, long ldd,
                             void *stream) {
    MDF_CHECK_ARG(rows >= 0 && cols >= 0 && lds >= cols && ldd >= cols, "mdf_widen_f32: bad shape");
    MDF_CHECK_ARG((src && dst) || rows * cols == 0, "mdf_widen_f32: null pointer");
    if (rows * cols == 0) return MDF_OK;
    const dim3 g((unsigned)std::min(mdf::cdiv(cols, 256L), 64L), (unsigned)std::min(rows, 4096L));
    hipLaunchKernelGGL(widen_f32_kernel, g, dim3(256), 0, mdf::as_stream(stream), src, lds, rows, cols, dst, ldd);
    MDF_LAUNCHED("widen_f32_kernel");
    return MDF_OK;
}

extern "C" int mdf_copy_2d(void *dst, long dpitch, const void *src, long spitch, long width, long rows,
                           void *stream) {
    MDF_CHECK_ARG(rows >= 0 && width >= 0 && dpitch >= width && spitch >= width, "mdf_copy_2d: bad shape");
    MDF_CHECK_ARG((src && dst) || rows * width == 0, "mdf_copy_2d: null pointer");
    if (rows * width == 0) return MDF_OK;
    MDF_HIP(hipMemcpy2DAsync(dst, (size_t)dpitch, src, (size_t)spitch, (size_t)width, (size_t)rows, hipMemcpyDefault,
                             mdf::as_stream(stream)));
    return MDF_OK;
}

extern "C" int mdf_fill_zero(double *x, long n, void *stream) {
    MDF_CHECK_ARG(x || n == 0, "mdf_fill_zero: null pointer");
    if (n == 0) return MDF_OK;
    unsigned g = (unsigned)std::min(mdf::cdiv(n, 256L), 65536L);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(g), dim3(256), 0, mdf::as_stream(stream), x, n);
    MDF_LAUNCHED("fill_zero_kernel");
    return MDF_OK;
}

namespace {
__global__ __launch_bounds__(256) void fill_int_kernel(int *__restrict__ x, long n, int v) {
    long t = (long)blockIdx.x * 256 + threadIdx.x;
    long stride = (long)gridDim.x * 256;
    for (; t < n; t += stride) x[t] = v;
}

__global__ __launch_bounds__(256) void pair_map_kernel(const int *__restrict__ addr, long npairs, int nao,
                                                       int *__restrict__ map) {
    long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= npairs) return;
    int a = addr[p], i = a / nao, j = a % nao;
    map[(long)i * nao + j] = (int)p;
    map[(long)j * nao + i] = (int)p;
}
}  // namespace

// map[i*nao + j] = pair index of (i, j) (either order), -1 where not stored.
extern "C" int mdf_build_pair_map(const int *pair_addr, long npairs, int nao, int *map, void *stream) {
    MDF_CHECK_ARG(pair_addr && map && nao > 0 && npairs < (1L << 31), "mdf_build_pair_map: bad arguments");
    hipStream_t st = mdf::as_stream(stream);
    long n = (long)nao * nao;
    hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)std::min(mdf::cdiv(n, 256L), 65536L)), dim3(256), 0, st,
                       map, n, -1);
    MDF_LAUNCHED("fill_int_kernel");
    if (npairs == 0) return MDF_OK;
    hipLaunchKernelGGL(pair_map_kernel, dim3((unsigned)mdf::cdiv(npairs, 256L)), dim3(256), 0, st,
                       pair_addr, npairs, nao, map);
    MDF_LAUNCHED("pair_map_kernel");
    return MDF_OK;
}

namespace {
// dst[c][r] = src[r][c] through a 64 x 64 LDS tile (one pad column keeps
// the column reads off a single bank); both sides read and write whole
// 64-element row segments.  Grid-strided over tiles in both directions.
template <typename T>
__global__ __launch_bounds__(256) void transpose_kernel(const T *__restrict__ src, long rows, long cols, long lds,
                                                        T *__restrict__ dst, long ldd) {
    __shared__ T tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (long rt = blockIdx.y; rt * 64 < rows; rt += gridDim.y) {
        for (long ct = blockIdx.x; ct * 64 < cols; ct += gridDim.x) {
            const long r0 = rt * 64, c0 = ct * 64;
            for (int k = ty; k < 64; k += 4) {
                const long r = r0 + k, c = c0 + tx;
                if (r < rows && c < cols) tile[k][tx] = src[r * lds + c];
            }
            __syncthreads();
            for (int k = ty; k < 64; k += 4) {
                const long c = c0 + k, r = r0 + tx;
                if (c < cols && r < rows) dst[c * ldd + r] = tile[tx][k];
            }
            __syncthreads();
        }
    }
}
}  // namespace

extern "C" int mdf_transpose(const void *src, int dtype, long rows, long cols, long lds, void *dst, long ldd,
                             void *stream) {
    MDF_CHECK_ARG(rows >= 0 && cols >= 0 && lds >= cols && ldd >= rows, "mdf_transpose: bad shape");
    MDF_CHECK_ARG((src && dst) || rows * cols == 0, "mdf_transpose: null pointer");
    MDF_CHECK_ARG(dtype == MDF_DTYPE_F64 || dtype == MDF_DTYPE_F32, "mdf_transpose: dtype %d", dtype);
    if (rows * cols == 0) return MDF_OK;
    const dim3 g((unsigned)std::min(mdf::cdiv(cols, 64L), 4096L), (unsigned)std::min(mdf::cdiv(rows, 64L), 4096L));
    if (dtype == MDF_DTYPE_F64)
        hipLaunchKernelGGL(transpose_kernel<double>, g, dim3(256), 0, mdf::as_stream(stream),
                           static_cast<const double *>(src), rows, cols, lds, static_cast<double *>(dst), ldd);
    else
        hipLaunchKernelGGL(transpose_kernel<float>, g, dim3(256), 0, mdf::as_stream(stream),
                           static_cast<const float *>(src), rows, cols, lds, static_cast<float *>(dst), ldd);
    MDF_LAUNCHED("transpose_kernel");
    return MDF_OK;
}
