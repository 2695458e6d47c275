// sh.hip — view-dependent colour pre-pass: real spherical harmonics up to
// degree 3 evaluated per texel site, producing the walk's texel_rgb.
//
// Dataflow follows the reference's colour pre-pass (powerfoam/color_fn.py:
// per texel site, direction = normalize(site - eye) :71, value + 0.5 then
// clamped at 0 :95-98) with spherical harmonics in place of the spherical
// Voronoi lobes; basis constants and signs are those of
// powerfoam/sh_utils.py:4-74 (the standard real-SH ordering).
#include "pf_device.h"
#include "pf_hip.h"
#include "pf_sh.h"

namespace pf {

// One thread per (cell, texel): coefficients (N, (D+1)^2, 3) f32, sites
// (N, K, 3) f32 -> rgb (N, K, 3) in the attribute dtype.
template <typename A, int DEG>
__global__ __launch_bounds__(256) void sh_fwd_kernel(const float *coeffs, const float *sites, int64_t NK, int K,
                                                     const float *cams, int B, A *rgb, int64_t rgb_cam_stride) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NK) return;
    const int b = blockIdx.y;
    constexpr int NC = (DEG + 1) * (DEG + 1);
    const f3 eye = mk(cams[12 * b], cams[12 * b + 1], cams[12 * b + 2]);
    const f3 s = mk(sites[3 * t], sites[3 * t + 1], sites[3 * t + 2]);
    float Y[16];
    const f3 v = sh_value<DEG>(coeffs + (t / K) * NC * 3, s, eye, Y);
    const f3 c = sh_color(v);
    A *o = rgb + (int64_t)b * rgb_cam_stride + 3 * t;
    o[0] = (A)c.x;
    o[1] = (A)c.y;
    o[2] = (A)c.z;
}

// One thread per cell: sums over its K texels and all B cameras, writes
// (accumulates) dL/dcoeffs once, no atomics.  The cell's coefficients are
// read once into registers (read per texel, they were fetched from L2 and
// HBM K times over).  The clamp at 0 passes the gradient only where the
// unclamped value is positive.
template <int DEG>
__global__ __launch_bounds__(256) void sh_bwd_kernel(const float *coeffs, const float *sites, int64_t N, int K,
                                                     const float *cams, int B, const float *grad_rgb,
                                                     int64_t grad_cam_stride, float *grad_coeffs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    constexpr int NC = (DEG + 1) * (DEG + 1);
    float k[NC * 3];
#pragma unroll
    for (int l = 0; l < NC * 3; ++l) k[l] = coeffs[i * NC * 3 + l];
    float acc[NC * 3];
#pragma unroll
    for (int l = 0; l < NC * 3; ++l) acc[l] = 0.0f;
    for (int b = 0; b < B; ++b) {
        const f3 eye = mk(cams[12 * b], cams[12 * b + 1], cams[12 * b + 2]);
        for (int j = 0; j < K; ++j) {
            const int64_t t = i * K + j;
            const float *gr = grad_rgb + (int64_t)b * grad_cam_stride + 3 * t;
            const float g0 = gr[0], g1 = gr[1], g2 = gr[2];
            if (g0 == 0.0f && g1 == 0.0f && g2 == 0.0f) continue;
            const f3 s = mk(sites[3 * t], sites[3 * t + 1], sites[3 * t + 2]);
            float Y[16];
            const f3 v = sh_value<DEG>(k, s, eye, Y);
            const float v0 = v.x, v1 = v.y, v2 = v.z;
            const float m0 = (v0 + 0.5f) > 0.0f ? g0 : 0.0f;
            const float m1 = (v1 + 0.5f) > 0.0f ? g1 : 0.0f;
            const float m2 = (v2 + 0.5f) > 0.0f ? g2 : 0.0f;
#pragma unroll
            for (int l = 0; l < NC; ++l) {
                acc[3 * l + 0] = __builtin_fmaf(Y[l], m0, acc[3 * l + 0]);
                acc[3 * l + 1] = __builtin_fmaf(Y[l], m1, acc[3 * l + 1]);
                acc[3 * l + 2] = __builtin_fmaf(Y[l], m2, acc[3 * l + 2]);
            }
        }
    }
    float *o = grad_coeffs + i * NC * 3;
#pragma unroll
    for (int l = 0; l < NC * 3; ++l) o[l] += acc[l];
}

}  // namespace pf

using namespace pf;

extern "C" int pf_sh_forward(const float *coeffs, int32_t degree, const float *texel_sites, int64_t num_cells,
                             int32_t num_texel_sites, const pf_camera *cams, int32_t num_cams, void *texel_rgb,
                             int32_t rgb_dtype, int64_t rgb_cam_stride, void *stream) {
    if (!coeffs || !texel_sites || num_cells <= 0 || num_texel_sites <= 0 || !cams || num_cams <= 0 || !texel_rgb)
        return PF_ERR_ARG;
    if (degree < 0 || degree > 3) return PF_ERR_ARG;
    if (num_cams > 1 && rgb_cam_stride < num_cells * num_texel_sites * 3) return PF_ERR_ARG;
    const int64_t NK = num_cells * num_texel_sites;
    dim3 grid((unsigned)((NK + 255) / 256), (unsigned)num_cams);
    hipStream_t st = (hipStream_t)stream;
    const float *cf = (const float *)cams;
#define PF_SH_F(A, D)                                                                                              \
    hipLaunchKernelGGL((sh_fwd_kernel<A, D>), grid, dim3(256), 0, st, coeffs, texel_sites, NK, num_texel_sites, cf, \
                       num_cams, (A *)texel_rgb, rgb_cam_stride)
    if (rgb_dtype == PF_F32) {
        switch (degree) {
            case 0: PF_SH_F(float, 0); break;
            case 1: PF_SH_F(float, 1); break;
            case 2: PF_SH_F(float, 2); break;
            default: PF_SH_F(float, 3); break;
        }
    } else if (rgb_dtype == PF_F16) {
        switch (degree) {
            case 0: PF_SH_F(_Float16, 0); break;
            case 1: PF_SH_F(_Float16, 1); break;
            case 2: PF_SH_F(_Float16, 2); break;
            default: PF_SH_F(_Float16, 3); break;
        }
    } else return PF_ERR_DTYPE;
#undef PF_SH_F
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}

extern "C" int pf_sh_backward(const float *coeffs, int32_t degree, const float *texel_sites, int64_t num_cells,
                              int32_t num_texel_sites, const pf_camera *cams, int32_t num_cams,
                              const float *grad_texel_rgb, int64_t grad_cam_stride, float *grad_coeffs,
                              void *stream) {
    if (!coeffs || !texel_sites || num_cells <= 0 || num_texel_sites <= 0 || !cams || num_cams <= 0 ||
        !grad_texel_rgb || !grad_coeffs)
        return PF_ERR_ARG;
    if (degree < 0 || degree > 3) return PF_ERR_ARG;
    dim3 grid((unsigned)((num_cells + 255) / 256));
    hipStream_t st = (hipStream_t)stream;
    const float *cf = (const float *)cams;
#define PF_SH_B(D)                                                                                                  \
    hipLaunchKernelGGL((sh_bwd_kernel<D>), grid, dim3(256), 0, st, coeffs, texel_sites, num_cells, num_texel_sites, \
                       cf, num_cams, grad_texel_rgb, grad_cam_stride, grad_coeffs)
    switch (degree) {
        case 0: PF_SH_B(0); break;
        case 1: PF_SH_B(1); break;
        case 2: PF_SH_B(2); break;
        default: PF_SH_B(3); break;
    }
#undef PF_SH_B
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}
