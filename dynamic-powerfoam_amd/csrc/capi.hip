// capi.hip — library-level C-ABI entry points (device query, messages).
#include <hip/hip_runtime.h>
#include "pf_hip.h"

extern "C" int pf_device_count(int32_t *n) {
    if (!n) return PF_ERR_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        *n = 0;
        return PF_ERR_HIP;
    }
    *n = c;
    return PF_OK;
}

extern "C" const char *pf_status_string(int32_t code) {
    switch (code) {
        case PF_OK: return "ok";
        case PF_ERR_ARG: return "invalid argument (null pointer, size or enum)";
        case PF_ERR_DTYPE: return "unsupported attribute or face dtype";
        case PF_ERR_HIP: return "HIP runtime error";
        case PF_ERR_OVERFLOW: return "caller-provided buffer too small";
        default: return "unknown status";
    }
}
