// Shared definitions of the mi355df HIP library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#define MDF_OK 0
#define MDF_EINVAL 1      // bad argument (shape, pointer, unsupported case)
#define MDF_EHIP 2        // a HIP runtime call or kernel launch failed
#define MDF_EWORKSPACE 3  // workspace too small for the call

namespace mdf {

// Records a formatted message for mdf_last_error() (thread-local).
void set_error(const char *fmt, ...);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

template <typename T>
inline T cdiv(T a, T b) { return (a + b - 1) / b; }

// XCD-aware tile order.  The 8 XCDs of an MI355X each have their own L2 and
// receive workgroups round-robin by workgroup id.  xcd_tile() maps the id w
// of a grid of g workgroups (g a multiple of 8) to a tile index so that each
// XCD takes a contiguous run of tiles: neighbouring tiles, which read the
// same operand rows, then meet in one L2 instead of eight.
constexpr int N_XCD = 8;
__host__ __device__ inline long xcd_tile(long w, long g) { return (w % N_XCD) * (g / N_XCD) + w / N_XCD; }
__host__ __device__ inline long round_xcd(long g) { return (g + N_XCD - 1) / N_XCD * N_XCD; }

}  // namespace mdf

#define MDF_CHECK_ARG(cond, ...)            \
    do {                                    \
        if (!(cond)) {                      \
            mdf::set_error(__VA_ARGS__);    \
            return MDF_EINVAL;              \
        }                                   \
    } while (0)

#define MDF_HIP(call)                                                           \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            mdf::set_error("%s failed: %s", #call, hipGetErrorString(e_));      \
            return MDF_EHIP;                                                    \
        }                                                                       \
    } while (0)

#define MDF_LAUNCHED(name)                                                      \
    do {                                                                        \
        hipError_t e_ = hipGetLastError();                                      \
        if (e_ != hipSuccess) {                                                 \
            mdf::set_error("launch of %s failed: %s", name, hipGetErrorString(e_)); \
            return MDF_EHIP;                                                    \
        }                                                                       \
    } while (0)

typedef double double4_t __attribute__((ext_vector_type(4)));
