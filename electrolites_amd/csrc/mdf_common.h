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
