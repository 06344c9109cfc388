// FP64 matrix-core ceiling of one MI355X (gfx950): register-only loops of
// v_mfma_f64_16x16x4f64 with NACC independent accumulators per wavefront at
// 1 and 2 wavefronts per SIMD, plus a vector FMA loop for comparison.  Every
// block also reads the shader clock (clock64) and the 100 MHz wall clock
// (wall_clock64) around its loop, so the rate can be set against the clock
// the loop actually ran at.
//
// Build it twice, with the compiler's own accumulator placement (AGPRs for
// these loops) and with accumulators forced into VGPRs:
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -o mfma_agpr electrolites_amd/probes/mfma_ceiling.hip
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -mllvm -amdgpu-mfma-vgpr-form -o mfma_vgpr electrolites_amd/probes/mfma_ceiling.hip
// Not part of the library: the package build compiles csrc/ only.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

typedef double double4_t __attribute__((ext_vector_type(4)));

template <int NACC>
__global__ __launch_bounds__(256) void mfma_loop(double *out, long long *clk, int iters, double seed) {
    const int lane = threadIdx.x & 63;
    const long long c0 = clock64(), w0 = wall_clock64();
    double a[4], b[5];
#pragma unroll
    for (int j = 0; j < 4; j++) a[j] = seed + lane * 1e-3 + j;
#pragma unroll
    for (int j = 0; j < 5; j++) b[j] = seed * 0.5 - lane * 1e-3 + j;
    double4_t acc[NACC];
#pragma unroll
    for (int j = 0; j < NACC; j++) acc[j] = double4_t{0, 0, 0, 0};
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int j = 0; j < NACC; j++)
            acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[j % 4], b[(j / 4) % 5], acc[j], 0, 0, 0);
    }
    double s = 0;
#pragma unroll
    for (int j = 0; j < NACC; j++) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    out[blockIdx.x * 256 + threadIdx.x] = s;
    const long long c1 = clock64(), w1 = wall_clock64();
    if (threadIdx.x == 0) {
        clk[2 * blockIdx.x] = c1 - c0;
        clk[2 * blockIdx.x + 1] = w1 - w0;
    }
}

__global__ __launch_bounds__(256) void fma_loop(double *out, int iters, double seed) {
    double x[16];
#pragma unroll
    for (int j = 0; j < 16; j++) x[j] = seed + j * 1e-3 + threadIdx.x * 1e-6;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int j = 0; j < 16; j++) x[j] = __builtin_fma(x[j], 0.999999, 1e-9);
    }
    double s = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) s += x[j];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

#define CK(x)                                                                          \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);     \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

template <int NACC>
static int run(double *out, long long *clk, int ncu, int wave_per_simd, int iters, int wall_khz) {
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int grid = ncu * wave_per_simd;       // 4 wavefronts per block: one per SIMD
    mfma_loop<NACC><<<grid, 256>>>(out, clk, 16, 1.0);
    CK(hipDeviceSynchronize());
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
        CK(hipEventRecord(e0));
        mfma_loop<NACC><<<grid, 256>>>(out, clk, iters, 1.0);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms;
        CK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) best = ms;
    }
    std::vector<long long> h(2 * grid);
    CK(hipMemcpy(h.data(), clk, h.size() * sizeof(long long), hipMemcpyDeviceToHost));
    double cyc = 0, wall = 0;
    for (int i = 0; i < grid; i++) {
        cyc += (double)h[2 * i];
        wall += (double)h[2 * i + 1];
    }
    const double mhz = cyc / wall * wall_khz * 1e-3;
    const double tflops = (double)grid * 4 * iters * NACC * 2048.0 / best * 1e-9;
    // published peak 78.6 TF/s is 1024 SIMDs x 2.4 GHz x 32 flop per cycle
    const double peak_at_clock = 1024.0 * mhz * 1e6 * 32 * 1e-12;
    printf("v_mfma_f64_16x16x4  %2d accumulators  %d wave/SIMD  %7.3f ms  %6.2f TF/s  clock %5.0f MHz  "
           "(peak at that clock %5.1f TF/s)\n",
           NACC, wave_per_simd, best, tflops, mhz, peak_at_clock);
    CK(hipEventDestroy(e0));
    CK(hipEventDestroy(e1));
    return 0;
}

int main() {
    hipDeviceProp_t p;
    CK(hipGetDeviceProperties(&p, 0));
    int wall_khz = 0;
    CK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, 0));
    const int ncu = p.multiProcessorCount;
    printf("%s, %d CUs, peak clock %d MHz\n", p.gcnArchName, ncu, p.clockRate / 1000);
    double *out;
    long long *clk;
    CK(hipMalloc(&out, (size_t)ncu * 4 * 256 * sizeof(double)));
    CK(hipMalloc(&clk, (size_t)ncu * 4 * 2 * sizeof(long long)));
    for (int w = 1; w <= 2; w++) {
        if (run<8>(out, clk, ncu, w, 4000, wall_khz)) return 1;
        if (run<20>(out, clk, ncu, w, 1600, wall_khz)) return 1;
    }
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int grid = ncu * 4, iters = 20000;
    fma_loop<<<grid, 256>>>(out, 16, 1.0);
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0));
    fma_loop<<<grid, 256>>>(out, iters, 1.0);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms;
    CK(hipEventElapsedTime(&ms, e0, e1));
    printf("v_fma_f64 (16 chains, 4 wave/SIMD)  %7.3f ms  %6.2f TF/s\n", ms,
           (double)grid * 256 * iters * 16 * 2 / ms * 1e-9);
    CK(hipFree(out));
    CK(hipFree(clk));
    return 0;
}
