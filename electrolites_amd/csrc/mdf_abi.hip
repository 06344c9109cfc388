// ABI version, error reporting and device queries of the mi355df library.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "mdf_common.h"
#include "../../include/mi355df.h"

namespace mdf {
static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
}  // namespace mdf

extern "C" int mdf_abi_version(void) { return MDF_ABI_VERSION; }

extern "C" const char *mdf_last_error(void) { return mdf::g_err; }

extern "C" int mdf_device_info(int *cu_count, long long *hbm_bytes, char *arch, int arch_len) {
    int dev = 0;
    MDF_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MDF_HIP(hipGetDeviceProperties(&prop, dev));
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (long long)prop.totalGlobalMem;
    if (arch && arch_len > 0) {
        strncpy(arch, prop.gcnArchName, arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return MDF_OK;
}
