// pf_device.h — gfx950 device math for the power-diagram ray tracer.
//
// Arithmetic contract (shared with the CPU restatement in oracle/, which is
// written separately to the same contract so that visited-cell sequences
// agree bit for bit):
//   * built with -ffp-contract=off; fused multiply-adds appear only where
//     written as fmaf: in dot products, in multiply-accumulates (s += a*b)
//     and in ray points o + t*d;
//   * division and sqrt are the correctly rounded IEEE operations (hipcc's
//     default for f32);
//   * min/max are the ternaries a<b?a:b and a>b?a:b;
//   * every exp that can change a transmittance, and hence where a walk
//     stops, goes through pf_expf below, a fixed polynomial evaluated the
//     same way on the CPU.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pf {

struct f3 { float x, y, z; };

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) {
    return __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x));
}
// o + t*d
__device__ __forceinline__ f3 ray_at(f3 o, float t, f3 d) {
    return f3{__builtin_fmaf(t, d.x, o.x), __builtin_fmaf(t, d.y, o.y), __builtin_fmaf(t, d.z, o.z)};
}
// acc + w*v
__device__ __forceinline__ f3 macc(f3 acc, float w, f3 v) {
    return f3{__builtin_fmaf(w, v.x, acc.x), __builtin_fmaf(w, v.y, acc.y), __builtin_fmaf(w, v.z, acc.z)};
}
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float fmin_(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmax_(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float length(f3 a) { return __builtin_sqrtf(dot(a, a)); }

// exp(x) for f32, evaluated identically on host and device:
// x = k ln2 + r (Cody-Waite, fma), exp(r) by its degree-7 Taylor polynomial
// in Horner form (truncation < 0.1 ulp on |r| <= ln2/2), scaled by ldexp.
__device__ __forceinline__ float pf_expf(float x) {
    if (!(x > -104.0f)) return (x != x) ? x : 0.0f;
    if (x > 88.72283935546875f) return __builtin_inff();
    const float k = __builtin_rintf(x * 1.44269502162933349609375f);
    float r = __builtin_fmaf(k, -0.693145751953125f, x);
    r = __builtin_fmaf(k, -1.428606765330187045e-06f, r);
    float p = 1.98412698412698412698e-04f;              // 1/5040
    p = __builtin_fmaf(p, r, 1.38888888888888888889e-03f);  // 1/720
    p = __builtin_fmaf(p, r, 8.33333333333333333333e-03f);  // 1/120
    p = __builtin_fmaf(p, r, 4.16666666666666666667e-02f);  // 1/24
    p = __builtin_fmaf(p, r, 1.66666666666666666667e-01f);  // 1/6
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    return __builtin_ldexpf(p, (int)k);
}

// log(x) for f32 with the same host/device contract as pf_expf:
// x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
// s = (m-1)/(m+1), odd series to s^11 (|s| <= 0.1716).
__device__ __forceinline__ float pf_logf(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    int e;
    float m = __builtin_frexpf(x, &e);  // m in [0.5, 1)
    if (m < 0.70710678118654752f) { m = m * 2.0f; e -= 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 0.0909090909090909091f;             // 1/11
    p = __builtin_fmaf(p, s2, 0.111111111111111111f);  // 1/9
    p = __builtin_fmaf(p, s2, 0.142857142857142857f);  // 1/7
    p = __builtin_fmaf(p, s2, 0.2f);
    p = __builtin_fmaf(p, s2, 0.333333333333333333f);
    p = __builtin_fmaf(p, s2, 1.0f);
    const float lm = 2.0f * s * p;
    const float ef = (float)e;
    return __builtin_fmaf(ef, 0.693145751953125f, __builtin_fmaf(ef, 1.428606765330187045e-06f, lm));
}

template <typename T> __device__ __forceinline__ float ld(const T *p) { return (float)(*p); }

// ---------------------------------------------------------------- camera --
// Pinhole ray through pixel (i, j): powerfoam/camera.py:19-24, normalised
// again at raytrace.py:153-154.
struct Cam {
    f3 eye, right, up, fwd;
    int W, H, start;
};

__device__ __forceinline__ Cam load_cam(const float *c /* pf_camera */) {
    Cam k;
    k.eye = mk(c[0], c[1], c[2]);
    k.right = mk(c[3], c[4], c[5]);
    k.up = mk(c[6], c[7], c[8]);
    const int *ci = reinterpret_cast<const int *>(c);
    k.start = ci[9];
    k.W = ci[10];
    k.H = ci[11];
    f3 f = cross(k.up, k.right);
    k.fwd = divs(f, length(f));
    return k;
}

__device__ __forceinline__ f3 ray_dir(const Cam &c, int pix_i, int pix_j) {
    const float x = 2.0f * (float)pix_j / ((float)c.W - 1.0f) - 1.0f;
    const float y = 1.0f - 2.0f * (float)pix_i / ((float)c.H - 1.0f);
    f3 d;
    d.x = __builtin_fmaf(y, c.up.x, x * c.right.x) + c.fwd.x;
    d.y = __builtin_fmaf(y, c.up.y, x * c.right.y) + c.fwd.y;
    d.z = __builtin_fmaf(y, c.up.z, x * c.right.z) + c.fwd.z;
    return divs(d, length(d));
}

// --------------------------------------------------------- intersections --
// rendering_math.py:4-27 — chord from the perpendicular distance.
struct SphereHit { bool hit; float tn, tf; };
__device__ __forceinline__ SphereHit ray_sphere(f3 o, f3 d, f3 c, float r) {
    const f3 oc = o - c;
    const float qa = dot(d, d);
    const float b = dot(oc, d) / qa;
    const f3 perp = oc - b * d;
    const float h = (r * r - dot(perp, perp)) / qa;
    if (h < 0.0f) return SphereHit{false, 0.0f, 0.0f};
    const float sq = __builtin_sqrtf(h);
    const float tf = -b + sq;
    const float tn = -b - sq;
    if (tn < 0.0f && tf < 0.0f) return SphereHit{false, 0.0f, 0.0f};
    if (tn < 0.0f) return SphereHit{true, 0.0f, tf};
    return SphereHit{true, tn, tf};
}

// rendering_math.py:30-65.
struct SphereGrad { f3 dn_dc, df_dc; float dn_dr, df_dr; };
__device__ __forceinline__ SphereGrad ray_sphere_bwd(f3 o, f3 d, f3 c, float r) {
    SphereGrad g{mk(0, 0, 0), mk(0, 0, 0), 0.0f, 0.0f};
    const f3 oc = o - c;
    const float qa = dot(d, d);
    const float b = dot(oc, d) / qa;
    const f3 perp = oc - b * d;
    const float h = (r * r - dot(perp, perp)) / qa;
    if (h < 0.0f) return g;
    const float sq = __builtin_sqrtf(h);
    const float tf = -b + sq, tn = -b - sq;
    if (tn < 0.0f && tf < 0.0f) return g;
    const f3 a = divs(d, qa);
    const f3 q = divs(perp, qa * sq);
    g.df_dc = a + q;
    g.df_dr = r / (qa * sq);
    if (tn >= 0.0f) {
        g.dn_dc = a - q;
        g.dn_dr = -r / (qa * sq);
    }
    return g;
}

// rendering_math.py:91-104 — walk face from precomputed (dp, dpm).
__device__ __forceinline__ void ray_pface_diff(f3 o, f3 d, f3 diff, float pm_diff, float &t, float &dp) {
    dp = dot(d, diff);
    t = (pm_diff - dot(o, diff)) / dp;
}

// rendering_math.py:68-88 — raster face from the two spheres, in two
// parts: what depends on the face and the ray origin only (the plane's
// normal and off - o.n), and the ray's own quotient.  Rays of one camera
// share the first part.
__device__ __forceinline__ float4 pface_plane(f3 o, f3 c, float r, f3 cj, float rj) {
    const f3 n = cj - c;
    const float off = 0.5f * (dot(cj, cj) - dot(c, c) + r * r - rj * rj);
    return make_float4(n.x, n.y, n.z, off - dot(o, n));
}
__device__ __forceinline__ void pface_t(f3 d, float4 pl, float &t, float &dp) {
    dp = dot(d, mk(pl.x, pl.y, pl.z));
    t = pl.w / dp;
}
__device__ __forceinline__ void ray_pface(f3 o, f3 d, f3 c, float r, f3 cj, float rj, float &t, float &dp) {
    pface_t(d, pface_plane(o, c, r, cj, rj), t, dp);
}

// rendering_math.py:107-131.
struct FaceGrad { f3 dc; float dr; f3 dcj; float drj; };
__device__ __forceinline__ FaceGrad ray_pface_bwd(f3 o, f3 d, f3 c, float r, f3 cj, float rj) {
    const f3 n = cj - c;
    const float off = 0.5f * (dot(cj, cj) - dot(c, c) + r * r - rj * rj);
    const float num = off - dot(o, n);
    const float dp = dot(d, n);
    FaceGrad g;
    // gradients carry a tolerance, not bit-exactness: one division
    const float inv = 1.0f / dp;
    const f3 nd = ((num * inv) * inv) * d;
    g.dc = (o - c) * inv + nd;
    g.dcj = (cj - o) * inv - nd;
    g.dr = r * inv;
    g.drj = -rj * inv;
    return g;
}

// rendering_math.py:134-141 — plane n.x = n.p + h.
__device__ __forceinline__ void ray_plane(f3 o, f3 d, f3 p, f3 n, float h, float &t, float &dp) {
    dp = dot(n, d);
    t = (dot(p - o, n) + h) / dp;
}

// rendering_math.py:144-156.
__device__ __forceinline__ void ray_plane_bwd(f3 o, f3 d, f3 p, f3 n, float h, f3 &dtdp, f3 &dtdn, float &dtdh) {
    const float dp = dot(n, d);
    const float inv = 1.0f / dp;
    dtdp = n * inv;
    dtdh = inv;
    dtdn = (p - o) * inv - ((dot(p - o, n) + h) * d) * inv * inv;
}

}  // namespace pf
