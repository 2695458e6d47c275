// pf_oracle.cpp — CPU restatement of the reference's power-diagram ray
// tracing path.  TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check
// and the bench's CPU baseline as the checker; nothing in the product path
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

// The walk: raytrace.py:115-241 (loop :163-236, face loop :183-200,
// texel :213-226, compositing :230-236); depth as rasterize.py:871-877.
template <typename R, typename FaceLoad>
RayOut<R> walk_ray(const SceneR<R> &s, FaceLoad face, V3<R> o, V3<R> d, int start, R thr, R depth_q, int max_steps,
                   int32_t *seq, int seq_cap) {
    RayOut<R> out{mk<R>(0, 0, 0), R(0), R(0), 0, kFnvOffset, 0};
    R log_t = 0;
    bool found = false;
    int prim = start;
    while (true) {
        const R trans = xexp(log_t);
        if (trans < thr || prim == kNoCell) break;
        if (prim < 0 || prim >= s.N || out.visits >= max_steps) { out.status |= 1; break; }
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
    while (true) {
        const R trans = xexp(log_t);
        if (trans < thr || prim == kNoCell) break;
        if (prim < 0 || prim >= s.N || visits >= max_steps) return 1;
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
                     uint64_t *hash, int32_t *seq, int32_t seq_cap, int32_t nthreads) {
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
// cast).  Returns sum over rays of the per-ray loss
//   sum_c w_c (C_c - target_c)^2 + w_alpha * (1 - exp(log_t))
// restricted to the given rays (NULL = all), and writes per-ray colour.
double orc_walk_loss_f64(const double *points, const double *radii, const double *nsigmas, const double *sites,
                         const double *rgb, const double *height, int32_t K, int64_t N, const int32_t *adj,
                         const int32_t *off, int64_t E, const CamIn *cams, int32_t B, double thr, int32_t max_steps,
                         const int64_t *rays, int64_t n_rays, const float *target, double w_alpha, int32_t nthreads) {
    std::vector<double> sph(4 * N), faces(4 * E), pm(N);
    for (int64_t i = 0; i < N; ++i) {
        sph[4 * i] = points[3 * i];
        sph[4 * i + 1] = points[3 * i + 1];
        sph[4 * i + 2] = points[3 * i + 2];
        sph[4 * i + 3] = radii[i];
        const V3<double> p = mk(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
        pm[i] = 0.5 * (dot(p, p) - radii[i] * radii[i]);
    }
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
        const RayOut<double> r = walk_ray(s, fl, c.eye, d, c.start, thr, 0.5, max_steps, (int32_t *)nullptr, 0);
        const double e0 = r.rgb.x - target[3 * ray], e1 = r.rgb.y - target[3 * ray + 1],
                     e2 = r.rgb.z - target[3 * ray + 2];
        loss += e0 * e0 + e1 * e1 + e2 * e2 + w_alpha * (1.0 - std::exp(r.log_t));
    }
    return loss;
}

}  // extern "C"
