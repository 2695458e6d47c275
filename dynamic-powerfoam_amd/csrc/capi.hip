// capi.hip — library-level C-ABI entry points (device query, messages).
#include <hip/hip_runtime.h>
#include "pf_hip.h"

// The ABI's struct sizes, which bindings mirror field for field.
static_assert(sizeof(pf_camera) == 48, "pf_camera layout");
static_assert(sizeof(pf_scene) == 104, "pf_scene layout");
static_assert(sizeof(pf_walk_out) == 88, "pf_walk_out layout");
static_assert(sizeof(pf_walk_record) == 40, "pf_walk_record layout");
static_assert(sizeof(pf_grads) == 48, "pf_grads layout");
static_assert(sizeof(pf_serve_scene) == 120, "pf_serve_scene layout");
static_assert(sizeof(pf_serve_request) == 168, "pf_serve_request layout");
static_assert(sizeof(pf_cech_grid) == 32, "pf_cech_grid layout");

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
