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

namespace pf {

constexpr float SH_C0 = 0.28209479177387814f;
constexpr float SH_C1 = 0.4886025119029199f;
constexpr float SH_C2_0 = 1.0925484305920792f, SH_C2_1 = -1.0925484305920792f, SH_C2_2 = 0.31539156525252005f,
                SH_C2_3 = -1.0925484305920792f, SH_C2_4 = 0.5462742152960396f;
constexpr float SH_C3_0 = -0.5900435899266435f, SH_C3_1 = 2.890611442640554f, SH_C3_2 = -0.4570457994644658f,
                SH_C3_3 = 0.3731763325901154f, SH_C3_4 = -0.4570457994644658f, SH_C3_5 = 1.445305721320277f,
                SH_C3_6 = -0.5900435899266435f;

__device__ __forceinline__ void sh_basis(f3 dir, int degree, float *Y) {
    const float x = dir.x, y = dir.y, z = dir.z;
    Y[0] = SH_C0;
    if (degree < 1) return;
    Y[1] = -SH_C1 * y;
    Y[2] = SH_C1 * z;
    Y[3] = -SH_C1 * x;
    if (degree < 2) return;
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    Y[4] = SH_C2_0 * xy;
    Y[5] = SH_C2_1 * yz;
    Y[6] = SH_C2_2 * (2.0f * zz - xx - yy);
    Y[7] = SH_C2_3 * xz;
    Y[8] = SH_C2_4 * (xx - yy);
    if (degree < 3) return;
    Y[9] = SH_C3_0 * y * (3.0f * xx - yy);
    Y[10] = SH_C3_1 * xy * z;
    Y[11] = SH_C3_2 * y * (4.0f * zz - xx - yy);
    Y[12] = SH_C3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    Y[13] = SH_C3_4 * x * (4.0f * zz - xx - yy);
    Y[14] = SH_C3_5 * z * (xx - yy);
    Y[15] = SH_C3_6 * x * (xx - 3.0f * yy);
}

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
    f3 dir = s - eye;
    dir = divs(dir, length(dir));
    float Y[16];
    sh_basis(dir, DEG, Y);
    const float *k = coeffs + (t / K) * NC * 3;
    float r = 0.0f, g = 0.0f, bl = 0.0f;
#pragma unroll
    for (int l = 0; l < NC; ++l) {
        r = __builtin_fmaf(Y[l], k[3 * l + 0], r);
        g = __builtin_fmaf(Y[l], k[3 * l + 1], g);
        bl = __builtin_fmaf(Y[l], k[3 * l + 2], bl);
    }
    A *o = rgb + (int64_t)b * rgb_cam_stride + 3 * t;
    o[0] = (A)fmax_(r + 0.5f, 0.0f);
    o[1] = (A)fmax_(g + 0.5f, 0.0f);
    o[2] = (A)fmax_(bl + 0.5f, 0.0f);
}

// One thread per cell: sums over its K texels and all B cameras, writes
// (accumulates) dL/dcoeffs once, no atomics.  The clamp at 0 passes the
// gradient only where the unclamped value is positive.
template <int DEG>
__global__ __launch_bounds__(256) void sh_bwd_kernel(const float *coeffs, const float *sites, int64_t N, int K,
                                                     const float *cams, int B, const float *grad_rgb,
                                                     int64_t grad_cam_stride, float *grad_coeffs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    constexpr int NC = (DEG + 1) * (DEG + 1);
    const float *k = coeffs + i * NC * 3;
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
            f3 dir = s - eye;
            dir = divs(dir, length(dir));
            float Y[16];
            sh_basis(dir, DEG, Y);
            float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
#pragma unroll
            for (int l = 0; l < NC; ++l) {
                v0 = __builtin_fmaf(Y[l], k[3 * l + 0], v0);
                v1 = __builtin_fmaf(Y[l], k[3 * l + 1], v1);
                v2 = __builtin_fmaf(Y[l], k[3 * l + 2], v2);
            }
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
