// pf_texel.h — texel lookup of the walk (raytrace.py:42-113), shared by
// the walk kernels and the rasterizer's benchmark / visualize kernels
// (rasterize.py:67-138), which use the same r^-2-multiply variant.
#pragma once
#include "pf_device.h"

namespace pf {

constexpr float kTemp = 10.0f;  // texel weight temperature (raytrace.py:39)

template <typename A> __device__ __forceinline__ f3 ld3(const A *p) {
    return mk((float)p[0], (float)p[1], (float)p[2]);
}

// Texel lookup of the walk: raytrace.py:42-113.  Two passes over the K texel
// sites: height at the base-plane query point, then colour at the displaced
// plane's query point.  Weights use the walk's r^-2 multiply (:61, :68).
struct TexFwd {
    float t0, dp0, height, wsum0, t_surf, dp, wsum;
    f3 rgb;
};

template <typename A, int KT>
__device__ __forceinline__ TexFwd texel_fwd(f3 o, f3 d, float t_near, f3 c, f3 n, float r,
                                            const A *sites, const A *rgbs, const A *heights, int K) {
    TexFwd out;
    if constexpr (KT > 0) K = KT;
    ray_plane(o, d, c, n, 0.0f, out.t0, out.dp0);
    const float tq0 = out.dp0 >= 0.0f ? t_near : fmax_(t_near, out.t0);
    const f3 q0 = ray_at(o, tq0, d);
    const float inv_r2 = 1.0f / (r * r);
    float hs = 0.0f, ws = 0.0f;
#pragma unroll
    for (int i = 0; i < (KT > 0 ? KT : 16); ++i) {
        if (KT == 0 && i >= K) break;
        const f3 s = ld3(sites + 3 * i);
        const f3 dv = q0 - s;
        const float dsq = dot(dv, dv) * inv_r2;
        const float w = pf_expf(-kTemp * dsq);
        hs = __builtin_fmaf(w, (float)heights[i], hs);
        ws = ws + w;
    }
    ws = fmax_(ws, 1e-20f);
    out.wsum0 = ws;
    out.height = hs / ws;
    ray_plane(o, d, c, n, out.height, out.t_surf, out.dp);
    const float tq = out.dp >= 0.0f ? t_near : fmax_(t_near, out.t_surf);
    const f3 q = ray_at(o, tq, d);
    f3 acc = mk(0.0f, 0.0f, 0.0f);
    float ws2 = 0.0f;
#pragma unroll
    for (int i = 0; i < (KT > 0 ? KT : 16); ++i) {
        if (KT == 0 && i >= K) break;
        const f3 s = ld3(sites + 3 * i);
        const f3 dv = q - s;
        const float dsq = dot(dv, dv) * inv_r2;
        const float w = pf_expf(-kTemp * dsq);
        acc = macc(acc, w, ld3(rgbs + 3 * i));
        ws2 = ws2 + w;
    }
    ws2 = fmax_(ws2, 1e-20f);
    out.wsum = ws2;
    out.rgb = divs(acc, ws2);
    return out;
}


}  // namespace pf
