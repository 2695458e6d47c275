// order.hip — Morton (Z-order) codes of the sites, the key by which the
// reference orders cells (morton_code_kernel, geometry.py:18-47, sorted by
// morton_sort, geometry.py:50-85, from sort_points, scene.py:197-233).
// With cells in this order, cells near in space are near in memory, so the
// gathers of neighbouring rays share cache lines and pages.
#include "pf_device.h"
#include "pf_hip.h"

namespace pf {

// 21 low bits of v spread to every third bit (bit k -> bit 3k).
__device__ __forceinline__ uint64_t spread3(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__device__ __forceinline__ uint64_t quantise(float p, float lo, float extent) {
    const float res = 2097152.0f;  // 2^21 steps per axis
    const float q = (p - lo) / fmaxf(extent, 1e-6f) * res;
    return (uint64_t)fminf(fmaxf(q, 0.0f), res - 1.0f);
}

__global__ __launch_bounds__(256) void morton_kernel(const float *pts, int64_t n, float lx, float ly, float lz,
                                                     float ex, float ey, float ez, int64_t *codes) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t x = spread3(quantise(pts[3 * i], lx, ex));
    const uint64_t y = spread3(quantise(pts[3 * i + 1], ly, ey));
    const uint64_t z = spread3(quantise(pts[3 * i + 2], lz, ez));
    codes[i] = (int64_t)((x << 2) | (y << 1) | z);
}

}  // namespace pf

using namespace pf;

extern "C" int pf_morton_codes(const float *points, int64_t num_points, const float *lower, const float *upper,
                               int64_t *codes, void *stream) {
    if (num_points < 0 || (num_points > 0 && (!points || !codes)) || !lower || !upper) return PF_ERR_ARG;
    if (num_points == 0) return PF_OK;
    const unsigned blocks = (unsigned)((num_points + 255) / 256);
    hipLaunchKernelGGL(morton_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, points, num_points, lower[0],
                       lower[1], lower[2], upper[0] - lower[0], upper[1] - lower[1], upper[2] - lower[2], codes);
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}
