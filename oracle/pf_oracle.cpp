// pf_oracle.cpp — CPU restatement of the reference's power-diagram ray
// tracing path.  TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check
// and the bench's CPU baseline as the oracle; nothing in the product path
// links or calls it.
//
// It follows the reference source as text (IsaacSuo/dynamic-powerfoam,
// citations per function) under one arithmetic contract, which the gfx950
// kernels follow independently so that walks agree bit for bit:
//   * -ffp-contract=off; fmaf only in dot products (z*z + (y*y + x*x) order),
//     in multiply-accumulates s += a*b, and in ray points o + t*d;
//   * correctly rounded division and sqrt;  min/max as a<b?a:b / a>b?a:b;
//   * exp through orc_expf (fixed polynomial, below), log through orc_logf.
// The same routines instantiated in double (faces computed exactly from the
// parameters, no fp16) give the FP64 forward used for finite differences.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>
#include <omp.h>

namespace {

// ------------------------------------------------------------ scalar math --
inline float orc_expf(float x) {
    if (!(x > -104.0f)) return (x != x) ? x : 0.0f;
    if (x > 88.72283935546875f) return INFINITY;
    const float k = rintf(x * 1.44269502162933349609375f);
    float r = fmaf(k, -0.693145751953125f, x);
    r = fmaf(k, -1.428606765330187045e-06f, r);
    float p = 1.98412698412698412698e-04f;
    p = fmaf(p, r, 1.38888888888888888889e-03f);
    p = fmaf(p, r, 8.33333333333333333333e-03f);
    p = fmaf(p, r, 4.16666666666666666667e-02f);
    p = fmaf(p, r, 1.66666666666666666667e-01f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)k);
}

inline float orc_logf(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    int e;
    float m = frexpf(x, &e);
    if (m < 0.70710678118654752f) { m = m * 2.0f; e -= 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 0.0909090909090909091f;
    p = fmaf(p, s2, 0.111111111111111111f);
    p = fmaf(p, s2, 0.142857142857142857f);
    p = fmaf(p, s2, 0.2f);
    p = fmaf(p, s2, 0.333333333333333333f);
    p = fmaf(p, s2, 1.0f);
    const float lm = 2.0f * s * p;
    const float ef = (float)e;
    return fmaf(ef, 0.693145751953125f, fmaf(ef, 1.428606765330187045e-06f, lm));
}

inline double xexp(double x) { return std::exp(x); }
inline float xexp(float x) { return orc_expf(x); }
inline double xlog(double x) { return std::log(x); }
inline float xlog(float x) { return orc_logf(x); }
inline double xfma(double a, double b, double c) { return std::fma(a, b, c); }
inline float xfma(float a, float b, float c) { return fmaf(a, b, c); }
inline double xsqrt(double a) { return std::sqrt(a); }
inline float xsqrt(float a) { return sqrtf(a); }

inline float half_to_float(uint16_t h) {
    const uint32_t s = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu;
    const uint32_t m = h & 0x3ffu;
    uint32_t u;
    if (e == 0) {
        if (m == 0) u = s;
        else {  // subnormal half: value m * 2^-24
            float f = (float)m * 5.9604644775390625e-08f;
            uint32_t fu;
            std::memcpy(&fu, &f, 4);
            u = s | fu;
        }
    } else if (e == 31) u = s | 0x7f800000u | (m << 13);
    else u = s | ((e + 112u) << 23) | (m << 13);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// IEEE binary32 -> binary16, round to nearest even (with subnormals).
inline uint16_t float_to_half(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint16_t s = (uint16_t)((u >> 16) & 0x8000u);
    const uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return (uint16_t)(s | 0x7c00u | (a > 0x7f800000u ? 0x200u : 0u));
    if (a >= 0x477ff000u) return (uint16_t)(s | 0x7c00u);  // rounds to >= 65520 -> inf
    if (a < 0x38800000u) {                                  // below smallest normal half
        if (a < 0x33000000u) return s;                      // < 2^-25 rounds to 0
        const uint32_t e = a >> 23;
        const uint32_t m = (a & 0x7fffffu) | 0x800000u;
        const int shift = 126 - (int)e;                      // 14..24 -> bits to drop = shift
        const uint32_t keep = m >> (shift);
        const uint32_t rem = m & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1);
        uint32_t r = keep;
        if (rem > half || (rem == half && (keep & 1u))) r += 1;
        return (uint16_t)(s | r);
    }
    uint32_t r = ((a >> 13) - (112u << 10));
    const uint32_t rem = a & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) r += 1;
    return (uint16_t)(s | r);
}

// Gradient accumulation: plain += (single thread, fixed order) or an atomic
// add when rays run on several threads (the CPU baseline's backward).
thread_local bool g_atomic_acc = false;
template <typename R> inline void acc(R &dst, R v) {
    if (g_atomic_acc) {
#pragma omp atomic
        dst += v;
    } else {
        dst += v;
    }
}

template <typename R> struct V3 { R x, y, z; };
template <typename R> inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <typename R> inline V3<R> sub(V3<R> a, V3<R> b) { return mk<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename R> inline V3<R> add(V3<R> a, V3<R> b) { return mk<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename R> inline V3<R> scl(R s, V3<R> a) { return mk<R>(s * a.x, s * a.y, s * a.z); }
template <typename R> inline V3<R> sclr(V3<R> a, R s) { return mk<R>(a.x * s, a.y * s, a.z * s); }
template <typename R> inline V3<R> dv(V3<R> a, R s) { return mk<R>(a.x / s, a.y / s, a.z / s); }
template <typename R> inline R dot(V3<R> a, V3<R> b) { return xfma(a.z, b.z, xfma(a.y, b.y, a.x * b.x)); }
template <typename R> inline V3<R> at(V3<R> o, R t, V3<R> d) {
    return mk<R>(xfma(t, d.x, o.x), xfma(t, d.y, o.y), xfma(t, d.z, o.z));
}
template <typename R> inline V3<R> macc(V3<R> a, R w, V3<R> v) {
    return mk<R>(xfma(w, v.x, a.x), xfma(w, v.y, a.y), xfma(w, v.z, a.z));
}
template <typename R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return mk<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename R> inline R mn(R a, R b) { return a < b ? a : b; }
template <typename R> inline R mx(R a, R b) { return a > b ? a : b; }
template <typename R> inline R len(V3<R> a) { return xsqrt(dot(a, a)); }

// ---------------------------------------------------------------- camera --
struct CamIn {  // pf_camera layout
    float eye[3], right[3], up[3];
    int32_t start_cell, width, height;
};

template <typename R> struct Cam {
    V3<R> eye, right, up, fwd;
    int W, H, start;
};

template <typename R> Cam<R> make_cam(const CamIn &c) {
    Cam<R> k;
    k.eye = mk<R>(c.eye[0], c.eye[1], c.eye[2]);
    k.right = mk<R>(c.right[0], c.right[1], c.right[2]);
    k.up = mk<R>(c.up[0], c.up[1], c.up[2]);
    k.W = c.width;
    k.H = c.height;
    k.start = c.start_cell;
    const V3<R> f = cross(k.up, k.right);  // camera.py:21
    k.fwd = dv(f, len(f));
    return k;
}

// camera.py:19-24, normalised again at raytrace.py:153-154
template <typename R> V3<R> ray_dir(const Cam<R> &c, int i, int j) {
    const R x = R(2) * (R)j / ((R)c.W - R(1)) - R(1);
    const R y = R(1) - R(2) * (R)i / ((R)c.H - R(1));
    V3<R> d;
    d.x = xfma(y, c.up.x, x * c.right.x) + c.fwd.x;
    d.y = xfma(y, c.up.y, x * c.right.y) + c.fwd.y;
    d.z = xfma(y, c.up.z, x * c.right.z) + c.fwd.z;
    return dv(d, len(d));
}

// -------------------------------------------------------- intersections --
// rendering_math.py:4-27
template <typename R> bool sphere_hit(V3<R> o, V3<R> d, V3<R> c, R r, R &tn, R &tf) {
    const V3<R> oc = sub(o, c);
    const R qa = dot(d, d);
    const R b = dot(oc, d) / qa;
    const V3<R> perp = sub(oc, scl(b, d));
    const R h = (r * r - dot(perp, perp)) / qa;
    tn = tf = R(0);
    if (h < R(0)) return false;
    const R sq = xsqrt(h);
    const R far_ = -b + sq, near_ = -b - sq;
    if (near_ < R(0) && far_ < R(0)) return false;
    if (near_ < R(0)) { tn = R(0); tf = far_; return true; }
    tn = near_;
    tf = far_;
    return true;
}

// rendering_math.py:30-65
template <typename R>
void sphere_bwd(V3<R> o, V3<R> d, V3<R> c, R r, V3<R> &dn_dc, V3<R> &df_dc, R &dn_dr, R &df_dr) {
    dn_dc = df_dc = mk<R>(0, 0, 0);
    dn_dr = df_dr = R(0);
    const V3<R> oc = sub(o, c);
    const R qa = dot(d, d);
    const R b = dot(oc, d) / qa;
    const V3<R> perp = sub(oc, scl(b, d));
    const R h = (r * r - dot(perp, perp)) / qa;
    if (h < R(0)) return;
    const R sq = xsqrt(h);
    const R far_ = -b + sq, near_ = -b - sq;
    if (near_ < R(0) && far_ < R(0)) return;
    const V3<R> a = dv(d, qa);
    const V3<R> q = dv(perp, qa * sq);
    df_dc = add(a, q);
    df_dr = r / (qa * sq);
    if (!(near_ < R(0))) {
        dn_dc = sub(a, q);
        dn_dr = -r / (qa * sq);
    }
}

// rendering_math.py:107-131
template <typename R>
void pface_bwd(V3<R> o, V3<R> d, V3<R> c, R r, V3<R> cj, R rj, V3<R> &dc, R &dr, V3<R> &dcj, R &drj) {
    const V3<R> n = sub(cj, c);
    const R off = R(0.5) * (dot(cj, cj) - dot(c, c) + r * r - rj * rj);
    const R num = off - dot(o, n);
    const R dp = dot(d, n);
    const V3<R> nd = dv(dv(scl(num, d), dp), dp);
    dc = add(dv(sub(o, c), dp), nd);
    dcj = sub(dv(sub(cj, o), dp), nd);
    dr = r / dp;
    drj = -rj / dp;
}

// rendering_math.py:134-141
template <typename R> void plane_hit(V3<R> o, V3<R> d, V3<R> p, V3<R> n, R h, R &t, R &dp) {
    dp = dot(n, d);
    t = (dot(sub(p, o), n) + h) / dp;
}

// rendering_math.py:144-156
template <typename R> void plane_bwd(V3<R> o, V3<R> d, V3<R> p, V3<R> n, R h, V3<R> &dtdp, V3<R> &dtdn, R &dtdh) {
    const R dp = dot(n, d);
    const R inv = R(1) / dp;
    dtdp = sclr(n, inv);
    dtdh = inv;
    dtdn = sub(sclr(sub(p, o), inv), sclr(sclr(sclr(d, (dot(sub(p, o), n) + h)), inv), inv));
}

// ------------------------------------------------------------- scene view --
struct SceneIn {  // host mirror of pf_scene (attributes already f32)
    const float *spheres, *nsigmas, *texel_sites, *texel_rgb, *texel_height;
    const int32_t *adjacency, *offsets;
    const void *faces;
    int64_t num_cells, num_edges;
    int32_t K, face_dtype;  // 0 f32, 1 f16 (uint16 bits)
};

template <typename R> struct Tex {
    R t0, dp0, height, wsum0, t_surf, dp, wsum;
    V3<R> rgb;
};

// raytrace.py:42-113
template <typename R>
Tex<R> texel_fwd(V3<R> o, V3<R> d, R t_near, V3<R> c, V3<R> n, R r, const R *sites, const R *rgbs, const R *heights,
                 int K) {
    Tex<R> t;
    plane_hit(o, d, c, n, R(0), t.t0, t.dp0);
    const R tq0 = t.dp0 >= R(0) ? t_near : mx(t_near, t.t0);
    const V3<R> q0 = at(o, tq0, d);
    const R inv_r2 = R(1) / (r * r);
    R hs = 0, ws = 0;
    for (int i = 0; i < K; ++i) {
        const V3<R> dvv = sub(q0, mk(sites[3 * i], sites[3 * i + 1], sites[3 * i + 2]));
        const R w = xexp(-R(10) * (dot(dvv, dvv) * inv_r2));
        hs = xfma(w, heights[i], hs);
        ws = ws + w;
    }
    ws = mx(ws, (R)1e-20f);
    t.wsum0 = ws;
    t.height = hs / ws;
    plane_hit(o, d, c, n, t.height, t.t_surf, t.dp);
    const R tq = t.dp >= R(0) ? t_near : mx(t_near, t.t_surf);
    const V3<R> q = at(o, tq, d);
    V3<R> acc = mk<R>(0, 0, 0);
    R ws2 = 0;
    for (int i = 0; i < K; ++i) {
        const V3<R> dvv = sub(q, mk(sites[3 * i], sites[3 * i + 1], sites[3 * i + 2]));
        const R w = xexp(-R(10) * (dot(dvv, dvv) * inv_r2));
        acc = macc(acc, w, mk(rgbs[3 * i], rgbs[3 * i + 1], rgbs[3 * i + 2]));
        ws2 = ws2 + w;
    }
    ws2 = mx(ws2, (R)1e-20f);
    t.wsum = ws2;
    t.rgb = dv(acc, ws2);
    return t;
}

// texture.py:74-186 semantics with the walk's weights
template <typename R>
void texel_bwd(V3<R> o, V3<R> d, R tn_in, V3<R> c, V3<R> n, R r, const R *sites, const R *rgbs, const R *heights,
               int K, const Tex<R> &fw, V3<R> g_rgb_in, R dts_in, R *g_sites, R *g_rgb, R *g_height, V3<R> &dc,
               V3<R> &dn, R &dtn, bool exact, R &dr) {
    dc = dn = mk<R>(0, 0, 0);
    R dh = 0;
    const R inv_r2 = R(1) / (r * r);
    const R r2 = r * r;
    const R tq = fw.dp >= R(0) ? tn_in : mx(tn_in, fw.t_surf);
    const V3<R> q = at(o, tq, d);
    V3<R> dq = mk<R>(0, 0, 0);
    const R gdp_in = dot(g_rgb_in, fw.rgb);
    for (int i = 0; i < K; ++i) {
        const V3<R> dvv = sub(q, mk(sites[3 * i], sites[3 * i + 1], sites[3 * i + 2]));
        const R w = xexp(-R(10) * (dot(dvv, dvv) * inv_r2));
        if (!exact && w / fw.wsum < (R)1e-3f) continue;
        const V3<R> dwds = dv(scl((R(2) * R(10)) * w, dvv), r2);
        const V3<R> grgb = dv(scl(w, g_rgb_in), fw.wsum);
        acc(g_rgb[3 * i], grgb.x);
        acc(g_rgb[3 * i + 1], grgb.y);
        acc(g_rgb[3 * i + 2], grgb.z);
        const R gdp = dot(g_rgb_in, mk(rgbs[3 * i], rgbs[3 * i + 1], rgbs[3 * i + 2]));
        const V3<R> gs = dv(sclr(dwds, gdp - gdp_in), fw.wsum);
        acc(g_sites[3 * i], gs.x);
        acc(g_sites[3 * i + 1], gs.y);
        acc(g_sites[3 * i + 2], gs.z);
        dq = add(dq, scl(R(-1), gs));
        if (exact) dr += ((gdp - gdp_in) / fw.wsum) * ((((R(2) * R(10)) * w) * dot(dvv, dvv)) / (r2 * r));
    }
    const R dtq = dot(dq, d);
    dtn = (fw.dp >= R(0) || tn_in >= fw.t_surf) ? dtq : R(0);
    R dts = (fw.dp < R(0) && tn_in < fw.t_surf) ? dtq : R(0);
    dts += dts_in;
    if (std::fabs(dts) > R(0)) {
        V3<R> dtdp, dtdn;
        R dtdh;
        plane_bwd(o, d, c, n, fw.height, dtdp, dtdn, dtdh);
        dc = add(dc, sclr(dtdp, dts));
        dn = add(dn, sclr(dtdn, dts));
        dh += dtdh * dts;
    }
    const R tq0 = fw.dp0 >= R(0) ? tn_in : mx(tn_in, fw.t0);
    const V3<R> q0 = at(o, tq0, d);
    V3<R> dq0 = mk<R>(0, 0, 0);
    const R hdp_in = dh * fw.height;
    for (int i = 0; i < K; ++i) {
        const V3<R> dvv = sub(q0, mk(sites[3 * i], sites[3 * i + 1], sites[3 * i + 2]));
        const R w = xexp(-R(10) * (dot(dvv, dvv) * inv_r2));
        if (!exact && w / fw.wsum0 < (R)1e-3f) continue;
        const V3<R> dwds = dv(scl((R(2) * R(10)) * w, dvv), r2);
        acc(g_height[i], (w * dh) / fw.wsum0);
        const R hdp = dh * heights[i];
        const V3<R> gs = dv(sclr(dwds, hdp - hdp_in), fw.wsum0);
        acc(g_sites[3 * i], gs.x);
        acc(g_sites[3 * i + 1], gs.y);
        acc(g_sites[3 * i + 2], gs.z);
        dq0 = add(dq0, scl(R(-1), gs));
        if (exact) dr += ((hdp - hdp_in) / fw.wsum0) * ((((R(2) * R(10)) * w) * dot(dvv, dvv)) / (r2 * r));
    }
    const R dtq0 = dot(dq0, d);
    dtn += (fw.dp >= R(0) || tn_in >= fw.t0) ? dtq0 : R(0);
    const R dts0 = (fw.dp < R(0) && tn_in < fw.t0) ? dtq0 : R(0);
    if (std::fabs(dts0) > R(0)) {
        V3<R> dtdp, dtdn;
        R dtdh;
        plane_bwd(o, d, c, n, R(0), dtdp, dtdn, dtdh);
        dc = add(dc, sclr(dtdp, dts0));
        dn = add(dn, sclr(dtdn, dts0));
    }
}

// Scene accessor with faces already in R (double path: exact faces).
template <typename R> struct SceneR {
    const R *spheres, *nsigmas, *sites, *rgb, *height, *faces;
    const int32_t *adj, *off;
    int64_t N;
    int K;
};

constexpr int kNoCell = 0x7FFFFFFF;
constexpr uint64_t kFnvOffset = 14695981039346656037ull;
constexpr uint64_t kFnvPrime = 1099511628211ull;

template <typename R> struct RayOut {
    V3<R> rgb;
    R log_t, depth;
    int visits;
    uint64_t hash;
    int status;
};

// Stop rule for walks that would never end.  The next cell is a function of
// the cell alone and termination a function of log T, so a return to an
// earlier (cell, log T) state repeats forever (raytrace.py:163 has no exit
// for it).  The stop happens at the step Brent's method finds the repeat:
// the state is saved at steps 0, 1, 3, 7, ... (after 2^k - 1 moves) and the
// walk stops when the current state equals the saved one.  Status bit 4.
template <typename R> struct Brent {
    int cell = -1;
    R log_t = 0;
    uint32_t power = 1, lam = 1;
    bool repeats(int prim, R lt) {
        if (prim == cell && lt == log_t) return true;
        if (lam == power) { cell = prim; log_t = lt; power <<= 1; lam = 0; }
        ++lam;
        return false;
    }
};

// Side outputs of the tile rasterizer's forward (rasterize.py:1088-1129)
// as one weighted sum, evaluated along a walk: normal distance
// sum_{n.d > 0} (n.d)^P alpha T, normal sum alpha T n, per-cell contrib
// alpha T / (H W), and the depth where T first drops below each quantile
// (quantiles descending).  Along a ray the walk visits cells in the order
// of their power distance from the eye, which is the order of the raster's
// tile lists, so the FP64 walk evaluates the raster's forward.
template <typename R> struct RasterTerms {
    R g_nd = 0;
    int nd_pow = 2;
    V3<R> g_n{0, 0, 0};
    const float *q = nullptr;
    const double *g_q = nullptr;
    int Q = 0, qi = 0;
    const double *g_contrib = nullptr;
    R inv_hw = 0;
    R value = 0;
    void segment(int prim, R tn, R alpha, R trans, R sigma, V3<R> n, V3<R> d, R log_t_after) {
        const R w = alpha * trans;
        const R ndv = dot(n, d);
        if (ndv > R(0)) value += g_nd * (nd_pow == 2 ? ndv * ndv : ndv) * w;
        value += dot(g_n, n) * w;
        if (g_contrib) value += R(g_contrib[prim]) * w * inv_hw;
        const R next = xexp(log_t_after);
        while (qi < Q && next < R(q[qi])) {
            value += R(g_q[qi]) * (tn + xlog(trans / R(q[qi])) / sigma);
            ++qi;
        }
    }
};

// The walk: raytrace.py:115-241 (loop :163-236, face loop :183-200,
// texel :213-226, compositing :230-236); depth as rasterize.py:871-877.
// terms: NULL, or raster side outputs to accumulate per composited segment.
template <typename R, typename FaceLoad>
RayOut<R> walk_ray(const SceneR<R> &s, FaceLoad face, V3<R> o, V3<R> d, int start, R thr, R depth_q, int max_steps,
                   int32_t *seq, int seq_cap, RasterTerms<R> *terms = nullptr) {
    RayOut<R> out{mk<R>(0, 0, 0), R(0), R(0), 0, kFnvOffset, 0};
    R log_t = 0;
    bool found = false;
    int prim = start;
    Brent<R> guard;
    while (true) {
        const R trans = xexp(log_t);
        if (trans < thr || prim == kNoCell) break;
        if (prim < 0 || prim >= s.N || out.visits >= max_steps) { out.status |= 1; break; }
        if (guard.repeats(prim, log_t)) { out.status |= 4; break; }
        if (seq && out.visits < seq_cap) seq[out.visits] = prim;
        out.visits++;
        out.hash = (out.hash ^ (uint64_t)(uint32_t)prim) * kFnvPrime;
        const R *sp = s.spheres + 4 * (int64_t)prim;
        const V3<R> c = mk(sp[0], sp[1], sp[2]);
        const R r = sp[3];
        R tn, tf;
        bool hit = sphere_hit(o, d, c, r, tn, tf);
        if (len(sub(c, o)) < R(4) * r) hit = false;
        const int e0 = s.off[prim], e1 = s.off[prim + 1];
        int next = kNoCell;
        R pt_far = (R)1e10f;
        for (int e = e0; e < e1; ++e) {
            V3<R> fn;
            R fw;
            face(e, fn, fw);
            const R dp = dot(d, fn);
            const R t = (fw - dot(o, fn)) / dp;
            if (dp >= R(0) && t < pt_far) { next = s.adj[e]; pt_far = t; }
            tf = dp >= R(0) ? mn(t, tf) : tf;
            tn = dp < R(0) ? mx(t, tn) : tn;
        }
        const R *ns = s.nsigmas + 4 * (int64_t)prim;
        const V3<R> nrm = mk(ns[0], ns[1], ns[2]);
        const R sigma = ns[3];
        if (!hit || tn > tf || sigma < (R)1e-3f) { prim = next; continue; }
        const int64_t tb = (int64_t)prim * s.K;
        const Tex<R> tx = texel_fwd(o, d, tn, c, nrm, r, s.sites + 3 * tb, s.rgb + 3 * tb, s.height + tb, s.K);
        tf = tx.dp >= R(0) ? mn(tx.t_surf, tf) : tf;
        tn = tx.dp < R(0) ? mx(tx.t_surf, tn) : tn;
        const int prim_here = prim;
        prim = next;
        const R dt = tf - tn;
        if (dt > R(0)) {
            const R dlt = (-sigma) * dt;
            const R alpha = R(1) - xexp(dlt);
            out.rgb.x = xfma(tx.rgb.x * alpha, trans, out.rgb.x);
            out.rgb.y = xfma(tx.rgb.y * alpha, trans, out.rgb.y);
            out.rgb.z = xfma(tx.rgb.z * alpha, trans, out.rgb.z);
            log_t += dlt;
            if (!found && xexp(log_t) < depth_q) {
                out.depth = tn + xlog(trans / depth_q) / sigma;
                found = true;
            }
            if (terms) terms->segment(prim_here, tn, alpha, trans, sigma, nrm, d, log_t);
        }
    }
    out.log_t = log_t;
    return out;
}

// Checkpoint-and-replay backward of one ray (see pf_walk_backward); the
// per-segment routing mirrors rasterize.py:1296-1516.
template <typename R, typename FaceLoad>
int walk_ray_bwd(const SceneR<R> &s, FaceLoad face, V3<R> o, V3<R> d, int start, R thr, int max_steps, V3<R> Cfin,
                 V3<R> g, R gL, R *gsph, R *gns, R *gsite, R *grgb, R *gh, bool exact) {
    R log_t = 0;
    V3<R> rgb = mk<R>(0, 0, 0);
    int prim = start, visits = 0;
    Brent<R> guard;
    while (true) {
        const R trans = xexp(log_t);
        if (trans < thr || prim == kNoCell) break;
        if (prim < 0 || prim >= s.N || visits >= max_steps) return 1;
        if (guard.repeats(prim, log_t)) return 4;
        visits++;
        const R *sp = s.spheres + 4 * (int64_t)prim;
        const V3<R> c = mk(sp[0], sp[1], sp[2]);
        const R r = sp[3];
        R tn, tf;
        bool hit = sphere_hit(o, d, c, r, tn, tf);
        if (len(sub(c, o)) < R(4) * r) hit = false;
        int tn_id = tn < (R)1e-8f ? -2 : -1;
        int tf_id = tf < (R)1e-8f ? -2 : -1;
        const int e0 = s.off[prim], e1 = s.off[prim + 1];
        const int n_adj = e1 - e0;
        int next = kNoCell;
        R pt_far = (R)1e10f;
        for (int e = e0; e < e1; ++e) {
            V3<R> fn;
            R fw;
            face(e, fn, fw);
            const R dp = dot(d, fn);
            const R t = (fw - dot(o, fn)) / dp;
            if (dp >= R(0) && t < pt_far) { next = s.adj[e]; pt_far = t; }
            if (dp >= R(0) && t < tf) tf_id = e - e0;
            tf = dp >= R(0) ? mn(t, tf) : tf;
            if (dp < R(0) && t > tn) tn_id = e - e0;
            tn = dp < R(0) ? mx(t, tn) : tn;
        }
        const R *ns = s.nsigmas + 4 * (int64_t)prim;
        const V3<R> nrm = mk(ns[0], ns[1], ns[2]);
        const R sigma = ns[3];
        if (!hit || tn > tf || sigma < (R)1e-3f) { prim = next; continue; }
        const int64_t tb = (int64_t)prim * s.K;
        const R tn0 = tn;
        const int tn0_id = tn_id, tf0_id = tf_id;
        const Tex<R> tx = texel_fwd(o, d, tn, c, nrm, r, s.sites + 3 * tb, s.rgb + 3 * tb, s.height + tb, s.K);
        if (tx.dp >= R(0) && tx.t_surf < tf) tf_id = n_adj;
        tf = tx.dp >= R(0) ? mn(tx.t_surf, tf) : tf;
        if (tx.dp < R(0) && tx.t_surf > tn) tn_id = n_adj;
        tn = tx.dp < R(0) ? mx(tx.t_surf, tn) : tn;
        const int cell = prim;
        prim = next;
        const R dt = tf - tn;
        if (!(dt > R(0))) continue;
        const R dlt = (-sigma) * dt;
        const R e_dlt = xexp(dlt);
        const R alpha = R(1) - e_dlt;
        rgb.x = xfma(tx.rgb.x * alpha, trans, rgb.x);
        rgb.y = xfma(tx.rgb.y * alpha, trans, rgb.y);
        rgb.z = xfma(tx.rgb.z * alpha, trans, rgb.z);
        log_t += dlt;
        const R dLnext = dot(g, sub(Cfin, rgb)) + gL;
        const R dalpha = dot(g, tx.rgb) * trans;
        const V3<R> dcolor = sclr(sclr(g, trans), alpha);
        const R ddelta = dLnext + (-dalpha * e_dlt);
        const R dsigma = ddelta * (-dt);
        const R dtn = ddelta * sigma;
        const R dtf = ddelta * (-sigma);
        R dts = 0;
        if (tn_id == n_adj) dts = dtn;
        else if (tf_id == n_adj) dts = dtf;
        V3<R> dc, dn;
        R dtn0, dr_tex = 0;
        texel_bwd(o, d, tn0, c, nrm, r, s.sites + 3 * tb, s.rgb + 3 * tb, s.height + tb, s.K, tx, dcolor, dts,
                  gsite + 3 * tb, grgb + 3 * tb, gh + tb, dc, dn, dtn0, exact, dr_tex);
        if (tn_id != n_adj) dtn0 += dtn;
        R dr = dr_tex;
        for (int side = 0; side < 2; ++side) {
            const int tid = side == 0 ? tn0_id : tf0_id;
            const R dLdt = side == 0 ? dtn0 : (tf_id != n_adj ? dtf : R(0));
            if (tid == -1) {
                V3<R> ndc, fdc;
                R ndr, fdr;
                sphere_bwd(o, d, c, r, ndc, fdc, ndr, fdr);
                if (side == 0) { dc = add(dc, sclr(ndc, dLdt)); dr += dLdt * ndr; }
                else { dc = add(dc, sclr(fdc, dLdt)); dr += dLdt * fdr; }
            } else if (tid >= 0 && tid < n_adj) {
                const int a = s.adj[e0 + tid];
                const R *as = s.spheres + 4 * (int64_t)a;
                V3<R> fdc, fdcj;
                R fdr, fdrj;
                pface_bwd(o, d, c, r, mk(as[0], as[1], as[2]), as[3], fdc, fdr, fdcj, fdrj);
                dc = add(dc, sclr(fdc, dLdt));
                dr += dLdt * fdr;
                R *ga = gsph + 4 * (int64_t)a;
                acc(ga[0], dLdt * fdcj.x);
                acc(ga[1], dLdt * fdcj.y);
                acc(ga[2], dLdt * fdcj.z);
                acc(ga[3], dLdt * fdrj);
            }
        }
        R *gs = gsph + 4 * (int64_t)cell;
        acc(gs[0], dc.x);
        acc(gs[1], dc.y);
        acc(gs[2], dc.z);
        acc(gs[3], dr);
        R *gn = gns + 4 * (int64_t)cell;
        acc(gn[0], dn.x);
        acc(gn[1], dn.y);
        acc(gn[2], dn.z);
        acc(gn[3], dsigma);
    }
    return 0;
}

struct FaceLoadF32 {
    const SceneIn *s;
    void operator()(int64_t e, V3<float> &n, float &w) const {
        if (s->face_dtype == 1) {
            const uint16_t *f = (const uint16_t *)s->faces + 4 * e;
            n = mk(half_to_float(f[0]), half_to_float(f[1]), half_to_float(f[2]));
            w = half_to_float(f[3]);
        } else {
            const float *f = (const float *)s->faces + 4 * e;
            n = mk(f[0], f[1], f[2]);
            w = f[3];
        }
    }
};

template <typename R> struct FaceLoadR {
    const R *f;
    void operator()(int64_t e, V3<R> &n, R &w) const {
        n = mk(f[4 * e], f[4 * e + 1], f[4 * e + 2]);
        w = f[4 * e + 3];
    }
};

SceneR<float> view_f32(const SceneIn &s) {
    return SceneR<float>{s.spheres, s.nsigmas, s.texel_sites, s.texel_rgb, s.texel_height, nullptr,
                         s.adjacency, s.offsets, s.num_cells, s.K};
}

}  // namespace

extern "C" {

int orc_num_threads(void) { return omp_get_max_threads(); }

void orc_f32_to_f16(const float *x, int64_t n, uint16_t *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = float_to_half(x[i]);
}
void orc_f16_to_f32(const uint16_t *x, int64_t n, float *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = half_to_float(x[i]);
}
void orc_expf_vec(const float *x, int64_t n, float *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = orc_expf(x[i]);
}
void orc_logf_vec(const float *x, int64_t n, float *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = orc_logf(x[i]);
}

// camera.py:19-24 for every pixel: out (H, W, 3)
void orc_ray_dirs(const CamIn *cam, float *out) {
    const Cam<float> c = make_cam<float>(*cam);
    for (int i = 0; i < c.H; ++i)
        for (int j = 0; j < c.W; ++j) {
            const V3<float> d = ray_dir(c, i, j);
            float *o = out + 3 * ((int64_t)i * c.W + j);
            o[0] = d.x;
            o[1] = d.y;
            o[2] = d.z;
        }
}

// benchmark.py:295-302
void orc_walk_faces(const float *points, const float *radii, int64_t N, const int32_t *adj, const int32_t *off,
                    void *faces, int32_t face_dtype) {
    for (int64_t i = 0; i < N; ++i) {
        const V3<float> pi = mk(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
        const float ni = sqrtf(dot(pi, pi));
        const float pmi = 0.5f * (ni * ni - radii[i] * radii[i]);
        for (int e = off[i]; e < off[i + 1]; ++e) {
            const int64_t j = adj[e];
            const V3<float> pj = mk(points[3 * j], points[3 * j + 1], points[3 * j + 2]);
            const float nj = sqrtf(dot(pj, pj));
            const float pmj = 0.5f * (nj * nj - radii[j] * radii[j]);
            const float v[4] = {pj.x - pi.x, pj.y - pi.y, pj.z - pi.z, pmj - pmi};
            for (int k = 0; k < 4; ++k) {
                if (face_dtype == 1) ((uint16_t *)faces)[4 * (int64_t)e + k] = float_to_half(v[k]);
                else ((float *)faces)[4 * (int64_t)e + k] = v[k];
            }
        }
    }
}

// benchmark.py:320-324 — first index of the minimum
int32_t orc_start_cell(const float *points, const float *radii, int64_t N, const float *eye) {
    const V3<float> e = mk(eye[0], eye[1], eye[2]);
    float best = INFINITY;
    int32_t bi = 0;
    for (int64_t i = 0; i < N; ++i) {
        const V3<float> v = sub(mk(points[3 * i], points[3 * i + 1], points[3 * i + 2]), e);
        const float dist = sqrtf(dot(v, v));
        float k = dist * dist - radii[i] * radii[i];
        if (k != k) k = INFINITY;
        if (k < best) { best = k; bi = (int32_t)i; }
    }
    return bi;
}

// Forward walk over B cameras (f32 contract).  Outputs as pf_walk_out.
// rays: optional list of ray ids (b*H*W + i*W + j) to trace; NULL = all.
int orc_walk_forward(const SceneIn *scene, const CamIn *cams, int32_t B, float thr, float depth_q, int32_t max_steps,
                     const int64_t *rays, int64_t n_rays, float *color, float *log_t, float *depth, int32_t *visits,
                     uint64_t *hash, int32_t *seq, int32_t seq_cap, int32_t nthreads, int32_t *ray_status) {
    const SceneR<float> s = view_f32(*scene);
    const FaceLoadF32 fl{scene};
    const int W = cams[0].width, H = cams[0].height;
    const int64_t total = rays ? n_rays : (int64_t)B * H * W;
    int status = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(| : status)
    for (int64_t k = 0; k < total; ++k) {
        const int64_t ray = rays ? rays[k] : k;
        const int b = (int)(ray / ((int64_t)H * W));
        const int i = (int)((ray / W) % H);
        const int j = (int)(ray % W);
        const Cam<float> c = make_cam<float>(cams[b]);
        const V3<float> d = ray_dir(c, i, j);
        const int64_t o = rays ? k : ray;
        const RayOut<float> r = walk_ray(s, fl, c.eye, d, c.start, thr, depth_q, max_steps,
                                         seq ? seq + o * seq_cap : nullptr, seq_cap);
        color[3 * o] = r.rgb.x;
        color[3 * o + 1] = r.rgb.y;
        color[3 * o + 2] = r.rgb.z;
        if (log_t) log_t[o] = r.log_t;
        if (depth) depth[o] = r.depth;
        if (visits) visits[o] = r.visits;
        if (hash) hash[o] = r.hash;
        if (ray_status) ray_status[o] = r.status;
        status |= r.status;
    }
    return status;
}

// Backward walk (f32 contract), single-threaded so the gradient sums have a
// fixed order.
int orc_walk_backward(const SceneIn *scene, const CamIn *cams, int32_t B, float thr, int32_t max_steps,
                      const float *color, const float *grad_color, const float *grad_log_t, float *g_spheres,
                      float *g_nsigmas, float *g_sites, float *g_rgb, float *g_height, int32_t flags,
                      int32_t nthreads) {
    const SceneR<float> s = view_f32(*scene);
    const FaceLoadF32 fl{scene};
    const int W = cams[0].width, H = cams[0].height;
    int status = 0;
    const bool par = nthreads > 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(par ? nthreads : 1) reduction(| : status)
    for (int64_t ray = 0; ray < (int64_t)B * H * W; ++ray) {
        g_atomic_acc = par;
        const int b = (int)(ray / ((int64_t)H * W));
        const int i = (int)((ray / W) % H);
        const int j = (int)(ray % W);
        const V3<float> g = grad_color ? mk(grad_color[3 * ray], grad_color[3 * ray + 1], grad_color[3 * ray + 2])
                                       : mk<float>(0, 0, 0);
        const float gL = grad_log_t ? grad_log_t[ray] : 0.0f;
        if (g.x == 0.0f && g.y == 0.0f && g.z == 0.0f && gL == 0.0f) continue;
        const Cam<float> c = make_cam<float>(cams[b]);
        const V3<float> d = ray_dir(c, i, j);
        status |= walk_ray_bwd(s, fl, c.eye, d, c.start, thr, max_steps,
                               mk(color[3 * ray], color[3 * ray + 1], color[3 * ray + 2]), g, gL, g_spheres, g_nsigmas,
                               g_sites, g_rgb, g_height, (flags & 1) != 0);
    }
    return status;
}

// FP64 forward for finite differences: parameters in double, walk faces
// computed exactly from points/radii (benchmark.py:297-301 without the fp16
// cast).  Sum over rays of the per-ray loss
//   sum_c (C_c - target_c)^2 + w_alpha * (1 - exp(log_t)) [+ raster terms]
// restricted to the given rays (NULL = all).  rt: NULL, or the weights of
// the raster side outputs (one camera; per-pixel arrays (H,W[,3|,Q])).
struct RasterWeights {
    const double *g_nd;      // (H,W) or NULL
    int32_t nd_pow;          // 2 forward definition, 1 the reference backward's
    const double *g_normal;  // (H,W,3) or NULL
    const float *quantiles;  // (H,W,Q) descending, or NULL
    const double *g_q;       // (H,W,Q)
    int32_t Q;
    const double *g_contrib; // (N,) or NULL
};

// rgb_cam_stride: camera b's texel colours start at rgb + b * rgb_cam_stride
// (0: one set for every camera).
static double loss_f64(const double *points, const double *radii, const double *nsigmas, const double *sites,
                       const double *rgb, const double *height, int32_t K, int64_t N, const int32_t *adj,
                       const int32_t *off, int64_t E, const CamIn *cams, int32_t B, double thr, int32_t max_steps,
                       const int64_t *rays, int64_t n_rays, const float *target, double w_alpha,
                       const RasterWeights *rw, int32_t nthreads, int64_t rgb_cam_stride = 0) {
    std::vector<double> sph(4 * N), faces(4 * E), pm(N);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < N; ++i) {
        sph[4 * i] = points[3 * i];
        sph[4 * i + 1] = points[3 * i + 1];
        sph[4 * i + 2] = points[3 * i + 2];
        sph[4 * i + 3] = radii[i];
        const V3<double> p = mk(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
        pm[i] = 0.5 * (dot(p, p) - radii[i] * radii[i]);
    }
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < N; ++i)
        for (int e = off[i]; e < off[i + 1]; ++e) {
            const int64_t j = adj[e];
            faces[4 * e] = points[3 * j] - points[3 * i];
            faces[4 * e + 1] = points[3 * j + 1] - points[3 * i + 1];
            faces[4 * e + 2] = points[3 * j + 2] - points[3 * i + 2];
            faces[4 * e + 3] = pm[j] - pm[i];
        }
    const SceneR<double> s{sph.data(), nsigmas, sites, rgb, height, faces.data(), adj, off, N, K};
    const FaceLoadR<double> fl{faces.data()};
    const int W = cams[0].width, H = cams[0].height;
    const int64_t total = rays ? n_rays : (int64_t)B * H * W;
    double loss = 0.0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(+ : loss)
    for (int64_t k = 0; k < total; ++k) {
        const int64_t ray = rays ? rays[k] : k;
        const int b = (int)(ray / ((int64_t)H * W));
        const int i = (int)((ray / W) % H);
        const int j = (int)(ray % W);
        const Cam<double> c = make_cam<double>(cams[b]);
        const V3<double> d = ray_dir(c, i, j);
        RasterTerms<double> tm;
        if (rw) {
            const int64_t px = (int64_t)i * W + j;
            tm.g_nd = rw->g_nd ? rw->g_nd[px] : 0.0;
            tm.nd_pow = rw->nd_pow;
            if (rw->g_normal) tm.g_n = mk(rw->g_normal[3 * px], rw->g_normal[3 * px + 1], rw->g_normal[3 * px + 2]);
            if (rw->quantiles) {
                tm.q = rw->quantiles + px * rw->Q;
                tm.g_q = rw->g_q + px * rw->Q;
                tm.Q = rw->Q;
            }
            tm.g_contrib = rw->g_contrib;
            tm.inv_hw = 1.0 / ((double)H * W);
        }
        SceneR<double> sb = s;
        sb.rgb = rgb + b * rgb_cam_stride;
        const RayOut<double> r = walk_ray(sb, fl, c.eye, d, c.start, thr, 0.5, max_steps, (int32_t *)nullptr, 0,
                                          rw ? &tm : nullptr);
        const double e0 = r.rgb.x - target[3 * ray], e1 = r.rgb.y - target[3 * ray + 1],
                     e2 = r.rgb.z - target[3 * ray + 2];
        loss += e0 * e0 + e1 * e1 + e2 * e2 + w_alpha * (1.0 - std::exp(r.log_t)) + tm.value;
    }
    return loss;
}

double orc_walk_loss_f64(const double *points, const double *radii, const double *nsigmas, const double *sites,
                         const double *rgb, const double *height, int32_t K, int64_t N, const int32_t *adj,
                         const int32_t *off, int64_t E, const CamIn *cams, int32_t B, double thr, int32_t max_steps,
                         const int64_t *rays, int64_t n_rays, const float *target, double w_alpha, int32_t nthreads,
                         int64_t rgb_cam_stride) {
    return loss_f64(points, radii, nsigmas, sites, rgb, height, K, N, adj, off, E, cams, B, thr, max_steps, rays,
                    n_rays, target, w_alpha, nullptr, nthreads, rgb_cam_stride);
}

// The same loss plus the raster side outputs weighted by rw (one camera):
// the functional whose gradient the raster backward computes from incoming
// gradients (color, log_t, normal_distance, normal, quantile depths,
// contrib; rasterize.py:1371-1426).
double orc_raster_loss_f64(const double *points, const double *radii, const double *nsigmas, const double *sites,
                           const double *rgb, const double *height, int32_t K, int64_t N, const int32_t *adj,
                           const int32_t *off, int64_t E, const CamIn *cam, double thr, const float *target,
                           double w_alpha, const RasterWeights *rw, int32_t nthreads) {
    return loss_f64(points, radii, nsigmas, sites, rgb, height, K, N, adj, off, E, cam, 1, thr, 1 << 16, nullptr, 0,
                    target, w_alpha, rw, nthreads);
}

}  // extern "C"

// ===================================================================== //
//                       pinhole tile rasterizer                          //
// ===================================================================== //
namespace {

inline int f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483647.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int)x;
}

struct Obb { float uc, vc, r1, r2, vx, vy; };

// camera.py:79-128
Obb proj_obb(const Cam<float> &cam, V3<float> center, float radius) {
    const V3<float> C = sub(center, cam.eye);
    const float C2 = dot(C, C);
    const float R2 = radius * radius;
    const V3<float> X = cam.right, Y = cam.up;
    const V3<float> zc = cross(Y, X);
    const V3<float> Z = dv(zc, len(zc));
    const float cx = dot(C, X), cy = dot(C, Y), cz = dot(C, Z);
    const float a = cx * cx - (C2 - R2) * dot(X, X);
    const float b = cy * cy - (C2 - R2) * dot(Y, Y);
    const float f = cz * cz - (C2 - R2) * dot(Z, Z);
    const float c = 2.0f * cx * cy;
    const float d = 2.0f * cx * cz;
    const float e = 2.0f * cy * cz;
    const float denom = 4.0f * a * b - c * c;
    if (std::fabs(denom) < 1e-8f) return Obb{0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f};
    const float uc = (c * e - 2.0f * b * d) / denom;
    const float vc = (c * d - 2.0f * a * e) / denom;
    const float F0 = a * uc * uc + b * vc * vc + c * uc * vc + d * uc + e * vc + f;
    const float term = sqrtf((a - b) * (a - b) + c * c);
    const float l1 = (a + b + term) / 2.0f;
    const float l2 = (a + b - term) / 2.0f;
    const float r1 = sqrtf(-F0 / l1);
    const float r2 = sqrtf(-F0 / l2);
    float vx = 0.5f * c, vy = l1 - a;
    const float nrm = sqrtf(vx * vx + vy * vy);
    if (nrm < 1e-6f) { vx = 1.0f; vy = 0.0f; }
    else { vx = vx / nrm; vy = vy / nrm; }
    return Obb{uc, vc, r1, r2, vx, vy};
}

// camera.py:131-178
bool tile_obb(int ti, int tj, const Cam<float> &cam, const Obb &o) {
    const float u_min = 2.0f * (float)(tj * 8) / (float)(cam.W - 1) - 1.0f;
    const float u_max = 2.0f * (float)((tj + 1) * 8) / (float)(cam.W - 1) - 1.0f;
    const float v_max = 1.0f - 2.0f * (float)(ti * 8) / (float)(cam.H - 1);
    const float v_min = 1.0f - 2.0f * (float)((ti + 1) * 8) / (float)(cam.H - 1);
    const float ucen = (u_min + u_max) * 0.5f, vcen = (v_min + v_max) * 0.5f;
    const float eu = (u_max - u_min) * 0.5f, ev = (v_max - v_min) * 0.5f;
    const float du = ucen - o.uc, dvv = vcen - o.vc;
    const float p1 = du * o.vx + dvv * o.vy;
    const float t1 = eu * std::fabs(o.vx) + ev * std::fabs(o.vy);
    if (std::fabs(p1) > o.r1 + t1) return false;
    const float p2 = du * -o.vy + dvv * o.vx;
    const float t2 = eu * std::fabs(-o.vy) + ev * std::fabs(o.vx);
    if (std::fabs(p2) > o.r2 + t2) return false;
    return true;
}

// rasterize.py:153-199 / :309-363: visible tiles of one cell
template <typename Fn> void for_tiles(const Cam<float> &cam, const float *sp, Fn fn, float &pow_dist) {
    const V3<float> c = mk(sp[0], sp[1], sp[2]);
    const float r = sp[3];
    const V3<float> v = sub(c, cam.eye);
    const V3<float> fwd = cross(cam.up, cam.right);
    pow_dist = dot(v, v) - r * r;
    if (len(v) < 4.0f * r || dot(v, fwd) < 0.1f) return;
    const Obb o = proj_obb(cam, c, r);
    const float ext_u = std::fabs(o.vx * o.r1) + std::fabs(o.vy * o.r2);
    const float ext_v = std::fabs(o.vy * o.r1) + std::fabs(o.vx * o.r2);
    const float umin = o.uc - ext_u, umax = o.uc + ext_u, vmin = o.vc - ext_v, vmax = o.vc + ext_v;
    const float jmin_f = (umin + 1.0f) * 0.5f * (float)(cam.W - 1);
    const float jmax_f = (umax + 1.0f) * 0.5f * (float)(cam.W - 1);
    const float imin_f = (1.0f - vmax) * 0.5f * (float)(cam.H - 1);
    const float imax_f = (1.0f - vmin) * 0.5f * (float)(cam.H - 1);
    int imin = f2i(floorf(imin_f)); imin = imin > 0 ? imin : 0;
    int imax = f2i(ceilf(imax_f)); imax = imax < cam.H - 1 ? imax : cam.H - 1;
    int jmin = f2i(floorf(jmin_f)); jmin = jmin > 0 ? jmin : 0;
    int jmax = f2i(ceilf(jmax_f)); jmax = jmax < cam.W - 1 ? jmax : cam.W - 1;
    const int tiles_h = (cam.H + 7) / 8;
    for (int ti = imin / 8; ti <= imax / 8; ++ti)
        for (int tj = jmin / 8; tj <= jmax / 8; ++tj)
            if (tile_obb(ti, tj, cam, o)) fn(ti + tj * tiles_h);
}

struct RTex { float t0, dp0, height, wsum0, t_surf, dp, wsum; V3<float> rgb; };

// texture.py:10-71 (raster: weights divide by r*r; zero when not hit).
// tex: [sites 3K | rgb 3K | height K]
RTex rtex_fwd(V3<float> o, V3<float> d, float tn, V3<float> c, V3<float> n, float r, bool hit, const float *tex,
              int K) {
    RTex t{0, 0, 0, 0, 0, 0, 0, mk<float>(0, 0, 0)};
    if (!hit) return t;
    plane_hit(o, d, c, n, 0.0f, t.t0, t.dp0);
    const float tq0 = t.dp0 >= 0.0f ? tn : mx(tn, t.t0);
    const V3<float> q0 = at(o, tq0, d);
    const float r2 = r * r;
    float hs = 0, ws = 0;
    for (int i = 0; i < K; ++i) {
        const V3<float> v = sub(q0, mk(tex[3 * i], tex[3 * i + 1], tex[3 * i + 2]));
        const float w = orc_expf(-10.0f * (dot(v, v) / r2));
        hs = fmaf(w, tex[6 * K + i], hs);
        ws = ws + w;
    }
    ws = mx(ws, 1e-20f);
    t.wsum0 = ws;
    t.height = hs / ws;
    plane_hit(o, d, c, n, t.height, t.t_surf, t.dp);
    const float tq = t.dp >= 0.0f ? tn : mx(tn, t.t_surf);
    const V3<float> q = at(o, tq, d);
    V3<float> acc = mk<float>(0, 0, 0);
    float ws2 = 0;
    for (int i = 0; i < K; ++i) {
        const V3<float> v = sub(q, mk(tex[3 * i], tex[3 * i + 1], tex[3 * i + 2]));
        const float w = orc_expf(-10.0f * (dot(v, v) / r2));
        acc = macc(acc, w, mk(tex[3 * K + 3 * i], tex[3 * K + 3 * i + 1], tex[3 * K + 3 * i + 2]));
        ws2 = ws2 + w;
    }
    ws2 = mx(ws2, 1e-20f);
    t.wsum = ws2;
    t.rgb = dv(acc, ws2);
    return t;
}

// texture.py:74-186 (raster weights); texel grads into red (layout as tex).
// exact == false reproduces the reference's two approximations (texels under
// 1e-3 of the weight sum skipped, texture.py:118-119, :158-159; no radius
// term through the weights); exact == true keeps every texel and adds the
// radius term d w / d r = w * 20 |q - s|^2 / r^3 into dr.
void rtex_bwd(V3<float> o, V3<float> d, float tn_in, V3<float> c, V3<float> n, float r, bool hit, const float *tex,
              int K, const RTex &fw, V3<float> g_rgb, float dts_in, float *red, V3<float> &dc, V3<float> &dn,
              float &dtn, bool exact, float &dr) {
    dc = dn = mk<float>(0, 0, 0);
    dtn = 0.0f;
    if (!hit) return;
    float dh = 0.0f;
    const float r2 = r * r;
    const float tq = fw.dp >= 0.0f ? tn_in : mx(tn_in, fw.t_surf);
    const V3<float> q = at(o, tq, d);
    V3<float> dq = mk<float>(0, 0, 0);
    const float gdp_in = dot(g_rgb, fw.rgb);
    for (int i = 0; i < K; ++i) {
        const V3<float> v = sub(q, mk(tex[3 * i], tex[3 * i + 1], tex[3 * i + 2]));
        const float w = orc_expf(-10.0f * (dot(v, v) / r2));
        if (!exact && w / fw.wsum < 1e-3f) continue;
        const V3<float> dwds = dv(scl((2.0f * 10.0f) * w, v), r2);
        const V3<float> grgb = dv(scl(w, g_rgb), fw.wsum);
        red[3 * K + 3 * i] += grgb.x;
        red[3 * K + 3 * i + 1] += grgb.y;
        red[3 * K + 3 * i + 2] += grgb.z;
        const float gdp = dot(g_rgb, mk(tex[3 * K + 3 * i], tex[3 * K + 3 * i + 1], tex[3 * K + 3 * i + 2]));
        const V3<float> gs = dv(sclr(dwds, gdp - gdp_in), fw.wsum);
        red[3 * i] += gs.x;
        red[3 * i + 1] += gs.y;
        red[3 * i + 2] += gs.z;
        dq = add(dq, scl(-1.0f, gs));
        if (exact) dr += ((gdp - gdp_in) / fw.wsum) * (((20.0f * w) * dot(v, v)) / (r2 * r));
    }
    const float dtq = dot(dq, d);
    dtn = (fw.dp >= 0.0f || tn_in >= fw.t_surf) ? dtq : 0.0f;
    float dts = (fw.dp < 0.0f && tn_in < fw.t_surf) ? dtq : 0.0f;
    dts += dts_in;
    if (std::fabs(dts) > 0.0f) {
        V3<float> a, b;
        float h;
        plane_bwd(o, d, c, n, fw.height, a, b, h);
        dc = add(dc, sclr(a, dts));
        dn = add(dn, sclr(b, dts));
        dh += h * dts;
    }
    const float tq0 = fw.dp0 >= 0.0f ? tn_in : mx(tn_in, fw.t0);
    const V3<float> q0 = at(o, tq0, d);
    V3<float> dq0 = mk<float>(0, 0, 0);
    const float hdp_in = dh * fw.height;
    for (int i = 0; i < K; ++i) {
        const V3<float> v = sub(q0, mk(tex[3 * i], tex[3 * i + 1], tex[3 * i + 2]));
        const float w = orc_expf(-10.0f * (dot(v, v) / r2));
        if (!exact && w / fw.wsum0 < 1e-3f) continue;
        const V3<float> dwds = dv(scl((2.0f * 10.0f) * w, v), r2);
        red[6 * K + i] += (w * dh) / fw.wsum0;
        const float hdp = dh * tex[6 * K + i];
        const V3<float> gs = dv(sclr(dwds, hdp - hdp_in), fw.wsum0);
        red[3 * i] += gs.x;
        red[3 * i + 1] += gs.y;
        red[3 * i + 2] += gs.z;
        dq0 = add(dq0, scl(-1.0f, gs));
        if (exact) dr += ((hdp - hdp_in) / fw.wsum0) * (((20.0f * w) * dot(v, v)) / (r2 * r));
    }
    const float dtq0 = dot(dq0, d);
    dtn += (fw.dp >= 0.0f || tn_in >= fw.t0) ? dtq0 : 0.0f;
    const float dts0 = (fw.dp < 0.0f && tn_in < fw.t0) ? dtq0 : 0.0f;
    if (std::fabs(dts0) > 0.0f) {
        V3<float> a, b;
        float h;
        plane_bwd(o, d, c, n, 0.0f, a, b, h);
        dc = add(dc, sclr(a, dts0));
        dn = add(dn, sclr(b, dts0));
    }
}

struct RScene {  // f32 attributes, faces f16 bits or f32
    const float *spheres, *nsigmas, *sites, *rgb, *height;
    const int32_t *adj, *off;
    const void *faces;
    int64_t N;
    int K, face_f16;
    void face(int64_t e, V3<float> &a, float &b) const {
        if (face_f16) {
            const uint16_t *f = (const uint16_t *)faces + 4 * e;
            a = mk(half_to_float(f[0]), half_to_float(f[1]), half_to_float(f[2]));
            b = half_to_float(f[3]);
        } else {
            const float *f = (const float *)faces + 4 * e;
            a = mk(f[0], f[1], f[2]);
            b = f[3];
        }
    }
    void texels(int prim, float *tex) const {
        const int64_t b = (int64_t)prim * K;
        for (int q = 0; q < 3 * K; ++q) {
            tex[q] = sites[3 * b + q];
            tex[3 * K + q] = rgb[3 * b + q];
        }
        for (int q = 0; q < K; ++q) tex[6 * K + q] = height[b + q];
    }
};

struct Lane {
    int pi, pj;
    bool oob;
    int64_t pix;
    V3<float> d;
};

void tile_lanes(const Cam<float> &cam, int tile, Lane *L) {
    const int tiles_h = (cam.H + 7) / 8;
    const int ti = tile % tiles_h, tj = tile / tiles_h;
    for (int l = 0; l < 64; ++l) {
        int pi = ti * 8 + (l & 7), pj = tj * 8 + (l >> 3);
        L[l].oob = pi >= cam.H || pj >= cam.W;
        L[l].pi = pi < cam.H - 1 ? pi : cam.H - 1;
        L[l].pj = pj < cam.W - 1 ? pj : cam.W - 1;
        L[l].pix = (int64_t)L[l].pi * cam.W + L[l].pj;
        L[l].d = ray_dir(cam, L[l].pi, L[l].pj);
    }
}

inline void atomic_addf(float *p, float v) {
#pragma omp atomic
    *p += v;
}

}  // namespace

extern "C" {

// rasterize.py:580-606
void orc_raster_faces(const float *spheres, int64_t N, const int32_t *adj, const int32_t *off, void *faces,
                      int32_t f16) {
    for (int64_t i = 0; i < N; ++i)
        for (int e = off[i]; e < off[i + 1]; ++e) {
            const float *a = spheres + 4 * (int64_t)adj[e];
            const float v[4] = {a[0] - spheres[4 * i], a[1] - spheres[4 * i + 1], a[2] - spheres[4 * i + 2], a[3]};
            for (int k = 0; k < 4; ++k) {
                if (f16) ((uint16_t *)faces)[4 * (int64_t)e + k] = float_to_half(v[k]);
                else ((float *)faces)[4 * (int64_t)e + k] = v[k];
            }
        }
}

// Binning: counts -> offsets (ntiles+1, int64) -> per-tile lists sorted by
// (power-distance bits, cell).  Returns the number of intersections; the
// lists are written only if it fits in cap.
int64_t orc_raster_bin(const float *spheres, int64_t N, const CamIn *camin, int64_t *offsets, int32_t *prims,
                       int64_t cap) {
    const Cam<float> cam = make_cam<float>(*camin);
    const int ntiles = ((cam.H + 7) / 8) * ((cam.W + 7) / 8);
    std::vector<std::vector<uint64_t>> lists(ntiles);
    for (int64_t i = 0; i < N; ++i) {
        float pd;
        uint32_t bits = 0;
        for_tiles(cam, spheres + 4 * i, [&](int t) {
            std::memcpy(&bits, &pd, 4);
            lists[t].push_back(((uint64_t)bits << 32) | (uint64_t)(uint32_t)i);
        }, pd);
    }
    offsets[0] = 0;
    for (int t = 0; t < ntiles; ++t) offsets[t + 1] = offsets[t] + (int64_t)lists[t].size();
    const int64_t total = offsets[ntiles];
    if (total <= cap && prims) {
        for (int t = 0; t < ntiles; ++t) {
            std::sort(lists[t].begin(), lists[t].end());
            for (size_t k = 0; k < lists[t].size(); ++k) prims[offsets[t] + k] = (int32_t)(lists[t][k] & 0xffffffffu);
        }
    }
    return total;
}

// forward_kernel (rasterize.py:892-1146), tile by tile with 64 lanes.
int orc_raster_forward(const RScene *s, const CamIn *camin, const int64_t *offsets, const int32_t *prims, float thr,
                       const float *depth_q, int32_t Q, const float *ray_gt, int32_t point_err, int32_t surface,
                       float *color, float *log_t_out, float *ndist_out, float *normal_out, float *qdepth,
                       float *err_out, float *contrib, float *perr, int32_t *early_stop, int32_t nthreads) {
    const Cam<float> cam = make_cam<float>(*camin);
    const int ntiles = ((cam.H + 7) / 8) * ((cam.W + 7) / 8);
    const float hw = (float)(cam.H * cam.W);
    const int K = s->K;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int tile = 0; tile < ntiles; ++tile) {
        Lane L[64];
        tile_lanes(cam, tile, L);
        float log_t[64] = {0}, nd[64] = {0}, err[64] = {0}, qv[64];
        int qi[64] = {0};
        V3<float> rgb[64], nrm[64], gt[64];
        for (int l = 0; l < 64; ++l) {
            rgb[l] = nrm[l] = mk<float>(0, 0, 0);
            qv[l] = depth_q ? depth_q[L[l].pix * Q] : 0.0f;
            gt[l] = ray_gt ? mk(ray_gt[3 * L[l].pix], ray_gt[3 * L[l].pix + 1], ray_gt[3 * L[l].pix + 2])
                           : mk<float>(0, 0, 0);
        }
        std::vector<float> tex(7 * K);
        const int64_t start = offsets[tile];
        const int total = (int)(offsets[tile + 1] - start);
        bool early = false;
        for (int i = 0; i < total; ++i) {
            bool all_below = true;
            for (int l = 0; l < 64; ++l) all_below = all_below && (orc_expf(log_t[l]) < thr);
            if (all_below) {
                early = true;
                early_stop[tile] = i;
                break;
            }
            const int prim = prims[start + i];
            const float *sp = s->spheres + 4 * (int64_t)prim;
            const float *ns = s->nsigmas + 4 * (int64_t)prim;
            const V3<float> c = mk(sp[0], sp[1], sp[2]), n = mk(ns[0], ns[1], ns[2]);
            const float r = sp[3], sigma = ns[3];
            bool hit[64];
            float tn[64], tf[64];
            bool any = false;
            for (int l = 0; l < 64; ++l) {
                hit[l] = sphere_hit(cam.eye, L[l].d, c, r, tn[l], tf[l]);
                any = any || hit[l];
            }
            if (!any) continue;
            s->texels(prim, tex.data());
            float csum = 0.0f, psum = 0.0f;
            for (int l = 0; l < 64; ++l) {
                const V3<float> o = cam.eye, d = L[l].d;
                if (hit[l]) {
                    for (int e = s->off[prim]; e < s->off[prim + 1]; ++e) {
                        V3<float> dp3;
                        float rj;
                        s->face(e, dp3, rj);
                        const V3<float> cj = add(c, dp3);
                        const V3<float> fn = sub(cj, c);
                        const float offv = 0.5f * (dot(cj, cj) - dot(c, c) + r * r - rj * rj);
                        const float dp = dot(d, fn);
                        const float t = (offv - dot(o, fn)) / dp;
                        tf[l] = dp >= 0.0f ? mn(t, tf[l]) : tf[l];
                        tn[l] = dp < 0.0f ? mx(t, tn[l]) : tn[l];
                    }
                }
                const RTex tx = rtex_fwd(o, d, tn[l], c, n, r, !L[l].oob && hit[l], tex.data(), K);
                tf[l] = tx.dp >= 0.0f ? mn(tx.t_surf, tf[l]) : tf[l];
                tn[l] = tx.dp < 0.0f ? mx(tx.t_surf, tn[l]) : tn[l];
                const float dt = tf[l] - tn[l];
                if (hit[l] && dt > 0.0f) {
                    const float dlt = (-sigma) * dt;
                    const float alpha = 1.0f - orc_expf(dlt);
                    const float tr = orc_expf(log_t[l]);
                    if (surface) {
                        const V3<float> ev = sub(tx.rgb, gt[l]);
                        err[l] = fmaf(dot(ev, ev) * alpha, tr, err[l]);
                    }
                    rgb[l].x = fmaf(tx.rgb.x * alpha, tr, rgb[l].x);
                    rgb[l].y = fmaf(tx.rgb.y * alpha, tr, rgb[l].y);
                    rgb[l].z = fmaf(tx.rgb.z * alpha, tr, rgb[l].z);
                    log_t[l] += dlt;
                    csum += (alpha * tr) / hw;
                    if (point_err) {
                        const V3<float> ev = sub(tx.rgb, gt[l]);
                        psum += ((alpha * tr) * (std::fabs(ev.x) + std::fabs(ev.y) + std::fabs(ev.z))) / hw;
                    }
                    const float ndv = dot(n, d);
                    if (ndv > 0.0f) nd[l] = fmaf((ndv * ndv) * alpha, tr, nd[l]);
                    nrm[l] = macc(nrm[l], alpha * tr, n);
                    const float next_t = orc_expf(log_t[l]);
                    while (qi[l] < Q && next_t < qv[l]) {
                        qdepth[L[l].pix * Q + qi[l]] = tn[l] + orc_logf(tr / qv[l]) / sigma;
                        ++qi[l];
                        if (qi[l] < Q) qv[l] = depth_q[L[l].pix * Q + qi[l]];
                    }
                }
            }
            if (csum != 0.0f) atomic_addf(contrib + prim, csum);
            if (point_err && psum != 0.0f) atomic_addf(perr + prim, psum);
        }
        if (!early) early_stop[tile] = total;
        for (int l = 0; l < 64; ++l) {
            if (L[l].oob) continue;
            const int64_t px = L[l].pix;
            color[3 * px] = rgb[l].x;
            color[3 * px + 1] = rgb[l].y;
            color[3 * px + 2] = rgb[l].z;
            log_t_out[px] = log_t[l];
            ndist_out[px] = nd[l];
            normal_out[3 * px] = nrm[l].x;
            normal_out[3 * px + 1] = nrm[l].y;
            normal_out[3 * px + 2] = nrm[l].z;
            if (surface) {
                const V3<float> ev = sub(mk<float>(0, 0, 0), gt[l]);
                err_out[px] = fmaf(dot(ev, ev), orc_expf(log_t[l]), err[l]);
            }
        }
    }
    return 0;
}

// backward_kernel (rasterize.py:1150-1516)
int orc_raster_backward(const RScene *s, const CamIn *camin, const int64_t *offsets, const int32_t *prims,
                        const int32_t *early_stop, const float *depth_q, int32_t Q, const float *log_t_in,
                        const float *g_color, const float *g_log_t, const float *g_ndist, const float *g_normal,
                        const float *g_qdepth, const float *g_contrib, float *gs_sph, float *gs_ns, float *gs_sites,
                        float *gs_rgb, float *gs_h, int32_t flags, int32_t nthreads) {
    const bool exact = (flags & 1) != 0;
    const Cam<float> cam = make_cam<float>(*camin);
    const int ntiles = ((cam.H + 7) / 8) * ((cam.W + 7) / 8);
    const float hw = (float)(cam.H * cam.W);
    const int K = s->K;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int tile = 0; tile < ntiles; ++tile) {
        Lane L[64];
        tile_lanes(cam, tile, L);
        float log_t[64], dLdlog_t[64], qv[64], g_nd[64];
        int qi[64];
        V3<float> g_rgb[64], g_n[64];
        for (int l = 0; l < 64; ++l) {
            const int64_t px = L[l].pix;
            log_t[l] = log_t_in[px];
            qi[l] = depth_q ? Q - 1 : -1;
            qv[l] = depth_q ? depth_q[px * Q + qi[l]] : 1.0f;
            g_rgb[l] = mk(g_color[3 * px], g_color[3 * px + 1], g_color[3 * px + 2]);
            g_nd[l] = g_ndist ? g_ndist[px] : 0.0f;
            g_n[l] = g_normal ? mk(g_normal[3 * px], g_normal[3 * px + 1], g_normal[3 * px + 2]) : mk<float>(0, 0, 0);
            dLdlog_t[l] = dot(g_rgb[l], mk<float>(0, 0, 0)) * orc_expf(log_t[l]);
            dLdlog_t[l] += g_log_t ? g_log_t[px] : 0.0f;
        }
        std::vector<float> tex(7 * K), red(7 * K);
        const int64_t start = offsets[tile];
        for (int i = early_stop[tile] - 1; i >= 0; --i) {
            const int prim = prims[start + i];
            const float *sp = s->spheres + 4 * (int64_t)prim;
            const float *ns = s->nsigmas + 4 * (int64_t)prim;
            const V3<float> c = mk(sp[0], sp[1], sp[2]), n = mk(ns[0], ns[1], ns[2]);
            const float r = sp[3], sigma = ns[3];
            bool hit[64];
            float tn[64], tf[64];
            bool any = false;
            for (int l = 0; l < 64; ++l) {
                hit[l] = sphere_hit(cam.eye, L[l].d, c, r, tn[l], tf[l]);
                any = any || hit[l];
            }
            if (!any) continue;
            s->texels(prim, tex.data());
            const int e0 = s->off[prim], n_adj = s->off[prim + 1] - e0;
            for (int l = 0; l < 64; ++l) {
                const V3<float> o = cam.eye, d = L[l].d;
                int tn_id = tn[l] < 1e-8f ? -2 : -1;
                int tf_id = tf[l] < 1e-8f ? -2 : -1;
                if (hit[l]) {
                    for (int j = 0; j < n_adj; ++j) {
                        V3<float> dp3;
                        float rj;
                        s->face(e0 + j, dp3, rj);
                        const V3<float> cj = add(c, dp3);
                        const V3<float> fn = sub(cj, c);
                        const float offv = 0.5f * (dot(cj, cj) - dot(c, c) + r * r - rj * rj);
                        const float dp = dot(d, fn);
                        const float t = (offv - dot(o, fn)) / dp;
                        tf_id = (dp >= 0.0f && t < tf[l]) ? j : tf_id;
                        tf[l] = dp >= 0.0f ? mn(t, tf[l]) : tf[l];
                        tn_id = (dp < 0.0f && t > tn[l]) ? j : tn_id;
                        tn[l] = dp < 0.0f ? mx(t, tn[l]) : tn[l];
                    }
                }
                const float tn0 = tn[l];
                const int tn0_id = tn_id, tf0_id = tf_id;
                const RTex tx = rtex_fwd(o, d, tn[l], c, n, r, !L[l].oob && hit[l], tex.data(), K);
                tf_id = (tx.dp >= 0.0f && tx.t_surf < tf[l]) ? n_adj : tf_id;
                tf[l] = tx.dp >= 0.0f ? mn(tx.t_surf, tf[l]) : tf[l];
                tn_id = (tx.dp < 0.0f && tx.t_surf > tn[l]) ? n_adj : tn_id;
                tn[l] = tx.dp < 0.0f ? mx(tx.t_surf, tn[l]) : tn[l];
                const float dt = tf[l] - tn[l];
                float dLdtn = 0, dLdtf = 0, dLdr = 0, dLdsigma = 0;
                V3<float> dLdc = mk<float>(0, 0, 0), dLdn = mk<float>(0, 0, 0), dLdcolor = mk<float>(0, 0, 0);
                const bool act = !L[l].oob && hit[l] && dt > 0.0f;
                if (act) {
                    const int64_t px = L[l].pix;
                    const float dlt = (-sigma) * dt;
                    log_t[l] -= dlt;
                    const float alpha = 1.0f - orc_expf(dlt);
                    const float trans = orc_expf(log_t[l]);
                    const float ndv = dot(n, d);
                    float dLdlog_t_i = dLdlog_t[l];
                    float dLddelta = dLdlog_t[l];
                    const float cg = g_contrib ? g_contrib[prim] : 0.0f;
                    float dLdalpha = dot(g_rgb[l], tx.rgb) * trans;
                    if (ndv > 0.0f) dLdalpha += g_nd[l] * ndv * trans;
                    dLdalpha += cg * trans / hw;
                    dLdalpha += dot(g_n[l], n) * trans;
                    float dLdtrans = dot(g_rgb[l], tx.rgb) * alpha;
                    if (ndv > 0.0f) dLdtrans += g_nd[l] * ndv * alpha;
                    dLdtrans += cg * alpha / hw;
                    dLdtrans += dot(g_n[l], n) * alpha;
                    dLdcolor = sclr(sclr(g_rgb[l], trans), alpha);
                    dLdlog_t_i += dLdtrans * orc_expf(log_t[l]);
                    dLddelta += -dLdalpha * orc_expf(dlt);
                    dLdsigma += dLddelta * -dt;
                    dLdtn += dLddelta * sigma;
                    dLdtf += dLddelta * -sigma;
                    while (qi[l] >= 0 && trans > qv[l]) {
                        const float gq = g_qdepth[px * Q + qi[l]];
                        dLdsigma += -gq * orc_logf(trans / qv[l]) / (sigma * sigma + 1e-6f);
                        dLdtn += gq;
                        dLdlog_t_i += gq / (sigma + 1e-6f);
                        --qi[l];
                        if (qi[l] >= 0) qv[l] = depth_q[px * Q + qi[l]];
                    }
                    dLdn = add(dLdn, scl(alpha * trans, g_n[l]));
                    if (ndv > 0.0f) dLdn = add(dLdn, scl((g_nd[l] * alpha) * trans, d));
                    dLdlog_t[l] = dLdlog_t_i;
                }
                float dLdts = 0.0f;
                if (tn_id == n_adj) dLdts = dLdtn;
                else if (tf_id == n_adj) dLdts = dLdtf;
                std::fill(red.begin(), red.end(), 0.0f);
                V3<float> pdc, pdn;
                float dtn_in, dr_tex = 0.0f;
                rtex_bwd(o, d, tn0, c, n, r, act, tex.data(), K, tx, dLdcolor, dLdts, red.data(), pdc, pdn, dtn_in,
                         exact, dr_tex);
                dLdr += dr_tex;
                dLdc = add(dLdc, pdc);
                dLdn = add(dLdn, pdn);
                if (tn_id != n_adj) dtn_in += dLdtn;
                if (!act) continue;
                for (int side = 0; side < 2; ++side) {
                    const int tid = side == 0 ? tn0_id : tf0_id;
                    const float dLdt = side == 0 ? dtn_in : (tf_id != n_adj ? dLdtf : 0.0f);
                    if (tid == -1) {
                        V3<float> ndc, fdc;
                        float ndr, fdr;
                        sphere_bwd(o, d, c, r, ndc, fdc, ndr, fdr);
                        if (side == 0) { dLdc = add(dLdc, sclr(ndc, dLdt)); dLdr += dLdt * ndr; }
                        else { dLdc = add(dLdc, sclr(fdc, dLdt)); dLdr += dLdt * fdr; }
                    } else if (tid >= 0 && tid < n_adj) {
                        const int a = s->adj[e0 + tid];
                        const float *as = s->spheres + 4 * (int64_t)a;
                        V3<float> fdc, fdcj;
                        float fdr, fdrj;
                        pface_bwd(o, d, c, r, mk(as[0], as[1], as[2]), as[3], fdc, fdr, fdcj, fdrj);
                        dLdc = add(dLdc, sclr(fdc, dLdt));
                        dLdr += dLdt * fdr;
                        atomic_addf(gs_sph + 4 * (int64_t)a, dLdt * fdcj.x);
                        atomic_addf(gs_sph + 4 * (int64_t)a + 1, dLdt * fdcj.y);
                        atomic_addf(gs_sph + 4 * (int64_t)a + 2, dLdt * fdcj.z);
                        atomic_addf(gs_sph + 4 * (int64_t)a + 3, dLdt * fdrj);
                    }
                }
                atomic_addf(gs_sph + 4 * (int64_t)prim, dLdc.x);
                atomic_addf(gs_sph + 4 * (int64_t)prim + 1, dLdc.y);
                atomic_addf(gs_sph + 4 * (int64_t)prim + 2, dLdc.z);
                atomic_addf(gs_sph + 4 * (int64_t)prim + 3, dLdr);
                atomic_addf(gs_ns + 4 * (int64_t)prim, dLdn.x);
                atomic_addf(gs_ns + 4 * (int64_t)prim + 1, dLdn.y);
                atomic_addf(gs_ns + 4 * (int64_t)prim + 2, dLdn.z);
                atomic_addf(gs_ns + 4 * (int64_t)prim + 3, dLdsigma);
                const int64_t tb = (int64_t)prim * K;
                for (int q = 0; q < 3 * K; ++q) {
                    if (red[q] != 0.0f) atomic_addf(gs_sites + 3 * tb + q, red[q]);
                    if (red[3 * K + q] != 0.0f) atomic_addf(gs_rgb + 3 * tb + q, red[3 * K + q]);
                }
                for (int q = 0; q < K; ++q)
                    if (red[6 * K + q] != 0.0f) atomic_addf(gs_h + tb + q, red[6 * K + q]);
            }
        }
    }
    return 0;
}

// benchmark_kernel (vis = 0, walk faces) and visualization_kernel (vis = 1,
// raster faces): per-lane early break, face-loop early-out, walk texels.
int orc_raster_plain(const RScene *s, const CamIn *camin, const int64_t *offsets, const int32_t *prims, float thr,
                     int32_t vis, int32_t max_isect, float depth_qv, const float *bkgd, float *color, float *depth,
                     float *normal, float *alpha_out, int32_t *isect, int32_t nthreads) {
    const Cam<float> cam = make_cam<float>(*camin);
    const int ntiles = ((cam.H + 7) / 8) * ((cam.W + 7) / 8);
    const int K = s->K;
    const SceneR<float> sr{s->spheres, s->nsigmas, s->sites, s->rgb, s->height, nullptr, s->adj, s->off, s->N, K};
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int tile = 0; tile < ntiles; ++tile) {
        Lane L[64];
        tile_lanes(cam, tile, L);
        const int64_t start = offsets[tile];
        const int total = (int)(offsets[tile + 1] - start);
        for (int l = 0; l < 64; ++l) {
            if (L[l].oob) continue;
            const V3<float> o = cam.eye, d = L[l].d;
            V3<float> rgb = mk<float>(0, 0, 0), nrm = mk<float>(0, 0, 0);
            float log_t = 0.0f, dep = 0.0f;
            bool found = false;
            int nhit = 0;
            for (int i = 0; i < total; ++i) {
                if (vis) {
                    ++nhit;
                    if (i >= max_isect) break;
                }
                const int prim = prims[start + i];
                if (orc_expf(log_t) < thr) break;
                const float *sp = s->spheres + 4 * (int64_t)prim;
                const float *ns = s->nsigmas + 4 * (int64_t)prim;
                const V3<float> c = mk(sp[0], sp[1], sp[2]), n = mk(ns[0], ns[1], ns[2]);
                const float r = sp[3], sigma = ns[3];
                float tn, tf;
                if (!sphere_hit(o, d, c, r, tn, tf)) continue;
                for (int e = s->off[prim]; e < s->off[prim + 1]; ++e) {
                    V3<float> fd;
                    float fw;
                    s->face(e, fd, fw);
                    float t, dp;
                    if (vis) {
                        const V3<float> cj = add(c, fd);
                        const V3<float> fn = sub(cj, c);
                        const float offv = 0.5f * (dot(cj, cj) - dot(c, c) + r * r - fw * fw);
                        dp = dot(d, fn);
                        t = (offv - dot(o, fn)) / dp;
                    } else {
                        dp = dot(d, fd);
                        t = (fw - dot(o, fd)) / dp;
                    }
                    tf = dp >= 0.0f ? mn(t, tf) : tf;
                    tn = dp < 0.0f ? mx(t, tn) : tn;
                    if (tn > tf) break;
                }
                if (tn > tf) continue;
                const int64_t tb = (int64_t)prim * K;
                const Tex<float> tx = texel_fwd(o, d, tn, c, n, r, sr.sites + 3 * tb, sr.rgb + 3 * tb, sr.height + tb, K);
                tf = tx.dp >= 0.0f ? mn(tx.t_surf, tf) : tf;
                tn = tx.dp < 0.0f ? mx(tx.t_surf, tn) : tn;
                const float dt = tf - tn;
                if (dt > 0.0f) {
                    const float dlt = (-sigma) * dt;
                    const float alpha = 1.0f - orc_expf(dlt);
                    const float trn = orc_expf(log_t);
                    rgb.x = fmaf(tx.rgb.x * alpha, trn, rgb.x);
                    rgb.y = fmaf(tx.rgb.y * alpha, trn, rgb.y);
                    rgb.z = fmaf(tx.rgb.z * alpha, trn, rgb.z);
                    if (vis) {
                        nrm.x = fmaf(n.x * alpha, trn, nrm.x);
                        nrm.y = fmaf(n.y * alpha, trn, nrm.y);
                        nrm.z = fmaf(n.z * alpha, trn, nrm.z);
                    }
                    log_t += dlt;
                    if (vis && !found && orc_expf(log_t) < depth_qv) {
                        dep = tn + orc_logf(trn / depth_qv) / sigma;
                        found = true;
                    }
                }
            }
            const float T = orc_expf(log_t);
            const V3<float> bk = vis ? mk(bkgd[0], bkgd[1], bkgd[2]) : mk<float>(0, 0, 0);
            rgb = macc(rgb, T, bk);
            const int64_t px = L[l].pix;
            color[3 * px] = rgb.x;
            color[3 * px + 1] = rgb.y;
            color[3 * px + 2] = rgb.z;
            if (vis) {
                depth[px] = dep;
                normal[3 * px] = nrm.x;
                normal[3 * px + 1] = nrm.y;
                normal[3 * px + 2] = nrm.z;
                alpha_out[px] = 1.0f - T;
                isect[px] = nhit;
            }
        }
    }
    return 0;
}

}  // extern "C"

// Čech complex by brute force over all pairs (the invariant the reference
// itself checks, experiments/affine_probe.py:127-134), with the AABB
// round trip of bvh.py:77-82, :17-18 and the strict test of :31-33.  Rows
// ascending.  Returns the edge count; writes rows only if it fits in cap.
extern "C" int64_t orc_cech(const float *centers, const float *radii, int64_t N, int32_t *offsets, int32_t *adj,
                            int64_t cap) {
    std::vector<V3<float>> c(N);
    std::vector<float> r(N);
    for (int64_t i = 0; i < N; ++i) {
        const float ri = radii[i];
        const V3<float> p = mk(centers[3 * i], centers[3 * i + 1], centers[3 * i + 2]);
        const V3<float> lo = mk(p.x - ri, p.y - ri, p.z - ri), hi = mk(p.x + ri, p.y + ri, p.z + ri);
        c[i] = scl(0.5f, add(lo, hi));
        r[i] = 0.5f * (hi.x - lo.x);
    }
    std::vector<std::vector<int32_t>> rows(N);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < N; ++i)
        for (int64_t j = 0; j < N; ++j)
            if (j != i && len(sub(c[i], c[j])) < r[i] + r[j]) rows[i].push_back((int32_t)j);
    offsets[0] = 0;
    for (int64_t i = 0; i < N; ++i) offsets[i + 1] = offsets[i] + (int32_t)rows[i].size();
    const int64_t total = offsets[N];
    if (adj && total <= cap)
        for (int64_t i = 0; i < N; ++i) std::copy(rows[i].begin(), rows[i].end(), adj + offsets[i]);
    return total;
}

/* ---- spherical-Voronoi colour pre-pass ----------------------------------
 * Restates spherical_voronoi_fwd_kernel / _bwd_kernel
 * (powerfoam/color_fn.py:54-101, 105-184): per texel site, lobe-major lobe
 * arrays (dof, n, 3) / (dof, n).  R = float follows the device arithmetic
 * (fma dot products, correctly rounded sqrt/div, the shared exp); R = double
 * is the exact form the finite-difference check uses.  Attribute storage is
 * f32 here; the f16 variant converts on the caller's side.                */
namespace {
template <typename R> V3<R> sv_dir(const float *s, const float *eye) {
    V3<R> v = mk<R>((R)s[0] - (R)eye[0], (R)s[1] - (R)eye[1], (R)s[2] - (R)eye[2]);
    R l = len(v);
    return l > (R)0 ? dv(v, l) : mk<R>(0, 0, 0);
}
template <typename R> bool sv_in_view(V3<R> d, const float *fwd, float cut) {
    return !(dot(d, mk<R>((R)fwd[0], (R)fwd[1], (R)fwd[2])) < (R)cut);
}
template <typename R> V3<R> ld3r(const float *p) { return mk<R>((R)p[0], (R)p[1], (R)p[2]); }

template <typename R>
void sv_fwd(const float *sites, int64_t n, const float *eye, const float *fwd, float cut, const float *axes,
            const float *vals, const float *temps, int dof, R *value, R *wsum) {
    for (int64_t t = 0; t < n; ++t) {
        V3<R> d = sv_dir<R>(sites + 3 * t, eye);
        if (!sv_in_view(d, fwd, cut)) {
            value[3 * t] = value[3 * t + 1] = value[3 * t + 2] = (R)0.5;
            wsum[t] = 0;
            continue;
        }
        R ws = 0;
        V3<R> vs = mk<R>(0, 0, 0);
        for (int i = 0; i < dof; ++i) {
            const int64_t q = (int64_t)i * n + t;
            R w = xexp(-(R)temps[q] * len(sub(d, ld3r<R>(axes + 3 * q))));
            ws = ws + w;
            vs = macc(vs, w, ld3r<R>(vals + 3 * q));
        }
        wsum[t] = ws;
        V3<R> v = add(dv(vs, ws), mk<R>(0.5, 0.5, 0.5));
        value[3 * t] = mx(v.x, (R)0);
        value[3 * t + 1] = mx(v.y, (R)0);
        value[3 * t + 2] = mx(v.z, (R)0);
    }
}

template <typename R>
void sv_bwd(const float *sites, int64_t n, const float *eye, const float *fwd, float cut, const float *axes,
            const float *vals, const float *temps, int dof, const R *wsum, const R *value, const float *grad,
            R *g_axes, R *g_vals, R *g_temps) {
    for (int64_t t = 0; t < n; ++t) {
        V3<R> d = sv_dir<R>(sites + 3 * t, eye);
        if (!sv_in_view(d, fwd, cut)) continue;
        const R ws = wsum[t];
        V3<R> val = mk<R>(value[3 * t], value[3 * t + 1], value[3 * t + 2]);
        V3<R> g = ld3r<R>(grad + 3 * t);
        if (val.x <= 0) g.x = 0;
        if (val.y <= 0) g.y = 0;
        if (val.z <= 0) g.z = 0;
        val = sub(val, mk<R>(0.5, 0.5, 0.5));
        const R dp_out = dot(g, val);
        for (int i = 0; i < dof; ++i) {
            const int64_t q = (int64_t)i * n + t;
            const V3<R> ax = ld3r<R>(axes + 3 * q), v = ld3r<R>(vals + 3 * q);
            const R temp = (R)temps[q];
            const V3<R> dd = sub(d, ax);
            const R dist = len(dd);
            const R w = xexp(-temp * dist);
            if (w / ws < (R)1e-3f) continue;
            const V3<R> gv = dv(scl(w, g), ws);
            g_vals[3 * q] = gv.x;
            g_vals[3 * q + 1] = gv.y;
            g_vals[3 * q + 2] = gv.z;
            const V3<R> dwds = dv(scl(temp * w, dd), dist);
            const R dwdt = -dist * w;
            const R dp = dot(g, v);
            const V3<R> ga = dv(sclr(dwds, dp - dp_out), ws);
            g_axes[3 * q] = ga.x;
            g_axes[3 * q + 1] = ga.y;
            g_axes[3 * q + 2] = ga.z;
            g_temps[q] = (dwdt * (dp - dp_out)) / ws;
        }
    }
}
}  // namespace

extern "C" void orc_sv_forward(const float *sites, int64_t n, const float *eye, const float *fwd, float cut,
                               const float *axes, const float *vals, const float *temps, int32_t dof, float *value,
                               float *wsum) {
    sv_fwd<float>(sites, n, eye, fwd, cut, axes, vals, temps, dof, value, wsum);
}
extern "C" void orc_sv_forward_f64(const float *sites, int64_t n, const float *eye, const float *fwd, float cut,
                                   const float *axes, const float *vals, const float *temps, int32_t dof,
                                   double *value, double *wsum) {
    sv_fwd<double>(sites, n, eye, fwd, cut, axes, vals, temps, dof, value, wsum);
}
extern "C" void orc_sv_backward(const float *sites, int64_t n, const float *eye, const float *fwd, float cut,
                                const float *axes, const float *vals, const float *temps, int32_t dof,
                                const float *wsum, const float *value, const float *grad, float *g_axes,
                                float *g_vals, float *g_temps) {
    sv_bwd<float>(sites, n, eye, fwd, cut, axes, vals, temps, dof, wsum, value, grad, g_axes, g_vals, g_temps);
}
extern "C" void orc_sv_backward_f64(const float *sites, int64_t n, const float *eye, const float *fwd, float cut,
                                    const float *axes, const float *vals, const float *temps, int32_t dof,
                                    const double *wsum, const double *value, const float *grad, double *g_axes,
                                    double *g_vals, double *g_temps) {
    sv_bwd<double>(sites, n, eye, fwd, cut, axes, vals, temps, dof, wsum, value, grad, g_axes, g_vals, g_temps);
}
