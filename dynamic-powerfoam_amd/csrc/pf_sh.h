// pf_sh.h — real spherical harmonics up to degree 3 (constants and signs of
// powerfoam/sh_utils.py:4-74) and the colour rule of the reference's colour
// pre-pass (value + 0.5, clamped at 0: powerfoam/color_fn.py:95-98).  Shared
// by the SH pre-pass kernels and the walk's in-kernel colour evaluation so
// both produce identical bits.
#pragma once
#include "pf_device.h"

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

// Unclamped value sum_l Y_l(normalize(site - eye)) k_l for one texel site.
template <int DEG>
__device__ __forceinline__ f3 sh_value(const float *k, f3 site, f3 eye, float *Y) {
    constexpr int NC = (DEG + 1) * (DEG + 1);
    f3 dir = site - eye;
    dir = divs(dir, length(dir));
    sh_basis(dir, DEG, Y);
    float r = 0.0f, g = 0.0f, b = 0.0f;
#pragma unroll
    for (int l = 0; l < NC; ++l) {
        r = __builtin_fmaf(Y[l], k[3 * l + 0], r);
        g = __builtin_fmaf(Y[l], k[3 * l + 1], g);
        b = __builtin_fmaf(Y[l], k[3 * l + 2], b);
    }
    return mk(r, g, b);
}

__device__ __forceinline__ f3 sh_color(f3 v) {
    return mk(fmax_(v.x + 0.5f, 0.0f), fmax_(v.y + 0.5f, 0.0f), fmax_(v.z + 0.5f, 0.0f));
}

}  // namespace pf
