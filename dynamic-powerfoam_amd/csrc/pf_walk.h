// pf_walk.h — the power-cell walk of one ray (forward, and backward by
// replay) and its texel lookup, as device functions shared by the plain
// walk kernels (walk.hip), the rasterizer's benchmark / visualize kernels
// (raster.hip) and the serving kernel (serve.hip).
//
// Forward semantics follow RayTracer.benchmark_kernel
// (powerfoam/raytrace.py:115-241) and plane_intersection_fwd_local
// (raytrace.py:42-113).  A texel's colour comes from a colour source: stored
// texel_rgb (the reference's input), or spherical harmonics evaluated at the
// texel site's view direction (identical bits to the SH pre-pass).
#pragma once
#include <type_traits>

#include "pf_device.h"
#include "pf_hip.h"
#include "pf_sh.h"

namespace pf {

constexpr int kNoCell = 0x7FFFFFFF;
constexpr uint64_t kFnvOffset = 14695981039346656037ull;
constexpr uint64_t kFnvPrime = 1099511628211ull;
constexpr float kTemp = 10.0f;  // texel weight temperature (raytrace.py:39)

struct alignas(8) h4 { _Float16 x, y, z, w; };

template <typename F> struct FaceT;
template <> struct FaceT<_Float16> {
    using Raw = h4;  // 8 bytes, kept packed until used
    __device__ static __forceinline__ Raw raw(const void *p, int64_t e) { return reinterpret_cast<const h4 *>(p)[e]; }
    __device__ static __forceinline__ void decode(const Raw &v, f3 &n, float &w) {
        n = mk((float)v.x, (float)v.y, (float)v.z);
        w = (float)v.w;
    }
    __device__ static __forceinline__ void load(const void *p, int64_t e, f3 &n, float &w) { decode(raw(p, e), n, w); }
};
template <> struct FaceT<float> {
    using Raw = float4;
    __device__ static __forceinline__ Raw raw(const void *p, int64_t e) {
        return reinterpret_cast<const float4 *>(p)[e];
    }
    __device__ static __forceinline__ void decode(const Raw &v, f3 &n, float &w) {
        n = mk(v.x, v.y, v.z);
        w = v.w;
    }
    __device__ static __forceinline__ void load(const void *p, int64_t e, f3 &n, float &w) { decode(raw(p, e), n, w); }
};

// Packed walk faces (PF_F16_PACKED): the fp16 face and the neighbour index
// of an edge in one 12-byte record.  A lane's face fetch is then one load
// instead of two; the walk's lanes gather from different cells, so the
// number of load instructions sets the L1 line traffic.
struct PackedF16 {};
struct PackedFace {
    _Float16 x, y, z, w;
    int32_t nb;
};
static_assert(sizeof(PackedFace) == 12, "packed face record");
template <> struct FaceT<PackedF16> {
    using Raw = PackedFace;
    __device__ static __forceinline__ Raw raw(const void *p, int64_t e) {
        return reinterpret_cast<const PackedFace *>(p)[e];
    }
    __device__ static __forceinline__ void decode(const Raw &v, f3 &n, float &w) {
        n = mk((float)v.x, (float)v.y, (float)v.z);
        w = (float)v.w;
    }
    __device__ static __forceinline__ void load(const void *p, int64_t e, f3 &n, float &w) { decode(raw(p, e), n, w); }
};
// Visits the faces e0 <= e < e1 of a CSR row in order, calling
// fn(e, normal, weight).  The neighbour across a face is not fetched here:
// the walk needs it for the exit face only (face_neighbour after the scan),
// and every load saved counts, since a wave's lanes gather from 64
// different cells.  Faces are fetched B at a time before any is used.  The
// tail of the last batch re-reads the row's last face and passes it to fn
// again, with its own index: every update the walk makes from a face is a
// strict comparison or a min / max, so a repeated face changes nothing,
// and the batch needs no branch per face.
#ifndef PF_FACE_BATCH
#define PF_FACE_BATCH 4
#endif
constexpr int kFaceBatch = PF_FACE_BATCH;
template <typename F, int B = kFaceBatch, typename Fn>
__device__ __forceinline__ void for_faces(const void *faces, int64_t, int e0, int e1, Fn &&fn) {
    for (int e = e0; e < e1; e += B) {
        typename FaceT<F>::Raw raw[B];
#pragma unroll
        for (int k = 0; k < B; ++k) raw[k] = FaceT<F>::raw(faces, e + k < e1 ? e + k : e1 - 1);
#pragma unroll
        for (int k = 0; k < B; ++k) {
            f3 n;
            float w;
            FaceT<F>::decode(raw[k], n, w);
            fn(e + k < e1 ? e + k : e1 - 1, n, w);
        }
    }
}

// The neighbour across face e (the adjacency entry, or the packed record's).
template <typename F> __device__ __forceinline__ int face_neighbour(const void *, const int *adj, int64_t e) {
    return adj[e];
}
template <> __device__ __forceinline__ int face_neighbour<PackedF16>(const void *faces, const int *, int64_t e) {
    return reinterpret_cast<const PackedFace *>(faces)[e].nb;
}

template <typename F, int B = kFaceBatch, typename Fn>
__device__ __forceinline__ void scan_row(const void *faces, int64_t E, int e0, int e1, Fn &&fn) {
    for_faces<F, B>(faces, E, e0, e1, fn);
}

// fp16 faces (E, 4) in pairs: one 16-byte buffer load brings two faces,
// which halves the loads a row costs.  The walk's forward is bound by the
// vector memory pipeline's requests (each lane of a gather touches its own
// cache line), not by bytes: 12.65 instead of 13.45 ms at 1080p.  Pairs
// start at the row's first face; a pair that runs past the row's end
// brings the next row's first face, passed to fn as not valid.  A batch is
// two pairs, and the next batch is in flight while the current one is
// evaluated (two register buffers, used alternately), so that a row waits
// for about one load latency instead of one per batch.  A pair past the
// row's end is not read: its offset lies past the array, where a buffer
// load returns zeros without a memory request.  The array must be at most
// 0xFFFFFF00 bytes; RowReader falls back to for_faces beyond that.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t f16_face_rsrc(const void *faces, int64_t E) {
    // raw buffer (stride 0), num_records in bytes; word 3 is the gfx9 value
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(faces), 0, (int)(uint32_t)(E * 8), 0x00020000);
}
constexpr uint32_t kNoPair = 0xFFFFFF00u;  // a buffer offset past every paired face array
template <typename Fn>
__device__ __forceinline__ void for_f16_face_pairs(__amdgpu_buffer_rsrc_t rs, int e0, int e1, Fn &&fn) {
    uint4 a[2], b[2];
    auto fetch = [&](uint4(&buf)[2], int f) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint32_t off = f + 2 * j < e1 ? (uint32_t)(f + 2 * j) * 8u : kNoPair;
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
            buf[j] = make_uint4(v[0], v[1], v[2], v[3]);
        }
        // issued here: not sunk past the loop's exit test to their first use
        __asm__ volatile("" ::: "memory");
    };
    auto eval = [&](const uint4(&buf)[2], int f) {
        const int rem = e1 - f;  // faces of the row from f on
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint4 q = buf[k >> 1];
            const uint32_t lo = (k & 1) ? q.z : q.x, hi = (k & 1) ? q.w : q.y;
            const f3 n = mk((float)__builtin_bit_cast(_Float16, (uint16_t)(lo & 0xffffu)),
                            (float)__builtin_bit_cast(_Float16, (uint16_t)(lo >> 16)),
                            (float)__builtin_bit_cast(_Float16, (uint16_t)(hi & 0xffffu)));
            const float w = (float)__builtin_bit_cast(_Float16, (uint16_t)(hi >> 16));
            fn(f + k, n, w, k < rem);
        }
    };
    if (e0 >= e1) return;
    fetch(a, e0);
    for (int f = e0;; f += 8) {
        fetch(b, f + 4);
        eval(a, f);
        if (f + 4 >= e1) break;
        fetch(a, f + 8);
        eval(b, f + 4);
        if (f + 8 >= e1) break;
    }
}

// How a kernel reads the faces of a row: fp16 faces in pairs where the
// array allows it (one test per kernel, the same for every wave), else one
// face per load (for_faces, whose repeated tail faces are passed as valid:
// every use of a face is a strict comparison or a min / max).
// fn(e, normal, weight, valid).
template <typename F> struct RowReader {
    const void *faces;
    int64_t E;
    __amdgpu_buffer_rsrc_t rs;
    bool paired;
    __device__ __forceinline__ RowReader(const void *f, int64_t e) : faces(f), E(e) {
        paired = std::is_same_v<F, _Float16> && e * 8 <= (int64_t)kNoPair;
        rs = f16_face_rsrc(f, paired ? e : 0);
    }
    template <int B = kFaceBatch, typename Fn> __device__ __forceinline__ void scan(int e0, int e1, Fn &&fn) const {
        if constexpr (std::is_same_v<F, _Float16>) {
            if (paired) {
                for_f16_face_pairs(rs, e0, e1, fn);
                return;
            }
        }
        for_faces<F, B>(faces, E, e0, e1, [&](int e, f3 n, float w) { fn(e, n, w, true); });
    }
};

// The folds the walk makes over a cell's faces (raytrace.py's kernel: the
// exit face from 1e10, t_far from the sphere's far bound, t_near from its
// near bound, each with strict comparisons, the near / far face ids at the
// first face that moves them), kept as one strict-min and one strict-max
// fold from -inf / +inf with the index of the first face at each extremum.
// A strict fold from a start value x0 ends at x0 when no face beats it, and
// otherwise at the first face at the extremum, so each of the walk's folds
// is this fold compared once with its start value (finish below): three
// compares and four selects a face instead of five and five.
struct FaceFold {
    float lo = __builtin_inff(), hi = -__builtin_inff();  // least t leaving (dp >= 0), greatest entering
    int lo_e = -1, hi_e = -1;
    __device__ __forceinline__ void add(int e, f3 o, f3 d, f3 fn, float fw, bool valid) {
        float t, dp;
        ray_pface_diff(o, d, fn, fw, t, dp);
        const bool leave = dp >= 0.0f;
        if (valid && leave && t < lo) {
            lo = t;
            lo_e = e;
        }
        if (valid && !leave && t > hi) {
            hi = t;
            hi_e = e;
        }
    }
    // exit face (-1: none below 1e10) and its t (1e10 when none)
    __device__ __forceinline__ int exit_face() const { return lo < 1e10f ? lo_e : -1; }
    __device__ __forceinline__ float exit_t() const { return lo < 1e10f ? lo : 1e10f; }
    __device__ __forceinline__ float t_far(float sphere_tf) const { return fmin_(lo, sphere_tf); }
    __device__ __forceinline__ float t_near(float sphere_tn) const { return fmax_(hi, sphere_tn); }
    // face ids relative to the row start e0, or the sphere's id when it bounds
    __device__ __forceinline__ int far_id(float sphere_tf, int e0, int id0) const {
        return lo < sphere_tf ? lo_e - e0 : id0;
    }
    __device__ __forceinline__ int near_id(float sphere_tn, int e0, int id0) const {
        return hi > sphere_tn ? hi_e - e0 : id0;
    }
};

template <typename A> __device__ __forceinline__ f3 ld3(const A *p) {
    return mk((float)p[0], (float)p[1], (float)p[2]);
}
template <typename A> __device__ __forceinline__ void ld4(const A *p, f3 &n, float &s) {
    n = mk((float)p[0], (float)p[1], (float)p[2]);
    s = (float)p[3];
}

__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }

// LDS access through an explicitly LDS-qualified pointer, so that it
// compiles to ds_* instructions rather than flat ones.
typedef __attribute__((address_space(3))) float lds_float;
typedef __attribute__((address_space(3))) int lds_int;
__device__ __forceinline__ lds_float *as_lds(float *p) { return (lds_float *)p; }
__device__ __forceinline__ void lds_add(float *p, float v) {
    __hip_atomic_fetch_add(as_lds(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---------------------------------------------------------- colour sources --
template <typename A> struct MemRGBCell {
    const A *p;
    __device__ __forceinline__ f3 operator()(int i, f3) const { return ld3(p + 3 * i); }
};
template <typename A> struct MemRGB {
    const A *base;
    __device__ __forceinline__ MemRGBCell<A> cell(int64_t cell, int K) const {
        return MemRGBCell<A>{base + 3 * cell * K};
    }
};
template <int DEG> struct SHRGBCell {
    const float *k;
    f3 eye;
    __device__ __forceinline__ f3 operator()(int, f3 site) const {
        float Y[16];
        return sh_color(sh_value<DEG>(k, site, eye, Y));
    }
};
template <int DEG> struct SHRGB {
    const float *coeffs;
    f3 eye;
    __device__ __forceinline__ SHRGBCell<DEG> cell(int64_t cell, int) const {
        return SHRGBCell<DEG>{coeffs + cell * (DEG + 1) * (DEG + 1) * 3, eye};
    }
};

// ------------------------------------------------------------ texel lookup --
// Two passes over the K texel sites: height at the base-plane query point,
// then colour at the displaced plane's query point; weights use the walk's
// r^-2 multiply (raytrace.py:61, :68).
template <int KT> struct TexFwdT {
    float t0, dp0, height, wsum0, t_surf, dp, wsum;
    f3 rgb;
    // K fixed: the texel weights of the height pass (w0) and the colour
    // pass (w1), which the backward reuses instead of recomputing
    float w0[KT > 0 ? KT : 1], w1[KT > 0 ? KT : 1];
};
using TexFwd = TexFwdT<0>;

template <typename A, int KT, typename RGBCell>
__device__ __forceinline__ TexFwdT<KT> texel_fwd(f3 o, f3 d, float t_near, f3 c, f3 n, float r, const A *sites,
                                                 const RGBCell &rgbs, const A *heights, int K) {
    TexFwdT<KT> out;
    if constexpr (KT > 0) K = KT;
    // With K fixed, all sites and heights are fetched before the first use
    // (one memory round trip instead of one per texel) and stay in
    // registers for both passes.
    constexpr int kHeld = KT > 0 ? KT : 1;
    f3 held_site[kHeld];
    float held_h[kHeld];
    if constexpr (KT > 0) {
#pragma unroll
        for (int i = 0; i < KT; ++i) {
            held_site[i] = ld3(sites + 3 * i);
            held_h[i] = (float)heights[i];
        }
    }
    ray_plane(o, d, c, n, 0.0f, out.t0, out.dp0);
    const float tq0 = out.dp0 >= 0.0f ? t_near : fmax_(t_near, out.t0);
    const f3 q0 = ray_at(o, tq0, d);
    const float inv_r2 = 1.0f / (r * r);
    float hs = 0.0f, ws = 0.0f;
#pragma unroll
    for (int i = 0; i < (KT > 0 ? KT : 16); ++i) {
        if (KT == 0 && i >= K) break;
        const f3 s = KT > 0 ? held_site[KT > 0 ? i : 0] : ld3(sites + 3 * i);
        const f3 dv = q0 - s;
        const float dsq = dot(dv, dv) * inv_r2;
        const float w = pf_expf(-kTemp * dsq);
        if constexpr (KT > 0) out.w0[i] = w;
        hs = __builtin_fmaf(w, KT > 0 ? held_h[KT > 0 ? i : 0] : (float)heights[i], hs);
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
        const f3 s = KT > 0 ? held_site[KT > 0 ? i : 0] : ld3(sites + 3 * i);
        const f3 dv = q - s;
        const float dsq = dot(dv, dv) * inv_r2;
        const float w = pf_expf(-kTemp * dsq);
        if constexpr (KT > 0) out.w1[i] = w;
        acc = macc(acc, w, rgbs(i, s));
        ws2 = ws2 + w;
    }
    ws2 = fmax_(ws2, 1e-20f);
    out.wsum = ws2;
    out.rgb = divs(acc, ws2);
    return out;
}

template <typename A, int KT>
__device__ __forceinline__ TexFwdT<KT> texel_fwd(f3 o, f3 d, float t_near, f3 c, f3 n, float r, const A *sites,
                                                 const A *rgbs, const A *heights, int K) {
    return texel_fwd<A, KT>(o, d, t_near, c, n, r, sites, MemRGBCell<A>{rgbs}, heights, K);
}

// Texel backward: the semantics of plane_intersection_bwd
// (powerfoam/texture.py:74-186) with the walk's weights (r^-2 multiply).
// Returns d/d(plane origin), d/d(normal) and d/d(t_near before the plane).
// exact == false reproduces the reference's two approximations: texels
// whose weight is below 1e-3 of the weight sum are skipped
// (texture.py:118-119, :158-159) and the radius gets no gradient through
// the texel weights.  exact == true keeps every texel and adds that radius
// term, d w_k / d r = 2 temp w_k |q - s_k|^2 / r^3, into dL_dr.
template <typename A, int KT, typename RGBCell, typename Sink>
__device__ __forceinline__ void texel_bwd(f3 o, f3 d, float t_near_in, f3 c, f3 n, float r, const A *sites,
                                          const RGBCell &rgbs, const A *heights, int K, const TexFwdT<KT> &fw,
                                          f3 dL_drgb, float dL_dtsurf_in, const Sink &sink, f3 &dL_dc, f3 &dL_dn,
                                          float &dL_dtnear, bool exact, float &dL_dr) {
    if constexpr (KT > 0) K = KT;
    dL_dc = mk(0, 0, 0);
    dL_dn = mk(0, 0, 0);
    float dL_dheight = 0.0f;
    const float inv_r2 = 1.0f / (r * r);

    const float tq = fw.dp >= 0.0f ? t_near_in : fmax_(t_near_in, fw.t_surf);
    const f3 q = ray_at(o, tq, d);
    f3 dL_dq = mk(0, 0, 0);
    const float gdp_in = dot(dL_drgb, fw.rgb);
    // Gradient arithmetic multiplies by reciprocals (one division each per
    // segment); the 1e-3 skip keeps its exact division, so the set of
    // contributing texels is the forward's.
    const float k2 = (2.0f * kTemp) * inv_r2;  // d w / d s = k2 w (q - s)
    const float inv_ws = 1.0f / fw.wsum;
#pragma unroll
    for (int i = 0; i < (KT > 0 ? KT : 16); ++i) {
        if (KT == 0 && i >= K) break;
        const f3 s = ld3(sites + 3 * i);
        const f3 dv = q - s;
        float w;
        if constexpr (KT > 0) w = fw.w1[i];
        else w = pf_expf(-kTemp * (dot(dv, dv) * inv_r2));
        const bool keep = exact || !(w / fw.wsum < 1e-3f);
        if constexpr (!Sink::kUniform) {
            if (!keep) continue;
        }
        const float kw = k2 * w;
        f3 g_rgb = (w * inv_ws) * dL_drgb;
        const float gdp = dot(dL_drgb, rgbs(i, s));
        const float coef = (gdp - gdp_in) * inv_ws;
        f3 gs = (kw * coef) * dv;
        if constexpr (Sink::kUniform) {
            // every lane makes the sink calls; a skipped texel adds zeros
            g_rgb = keep ? g_rgb : mk(0, 0, 0);
            gs = keep ? gs : mk(0, 0, 0);
        }
        sink.rgb(i, g_rgb);
        sink.site(i, gs);
        dL_dq = dL_dq + (-1.0f * gs);
        if (exact) dL_dr += coef * (kw * dot(dv, dv) / r);
    }
    const float dL_dtq = dot(dL_dq, d);
    dL_dtnear = (fw.dp >= 0.0f || t_near_in >= fw.t_surf) ? dL_dtq : 0.0f;
    float dL_dts = (fw.dp < 0.0f && t_near_in < fw.t_surf) ? dL_dtq : 0.0f;
    dL_dts += dL_dtsurf_in;
    if (__builtin_fabsf(dL_dts) > 0.0f) {
        f3 dtdp, dtdn;
        float dtdh;
        ray_plane_bwd(o, d, c, n, fw.height, dtdp, dtdn, dtdh);
        dL_dc = dL_dc + dtdp * dL_dts;
        dL_dn = dL_dn + dtdn * dL_dts;
        dL_dheight += dtdh * dL_dts;
    }

    const float tq0 = fw.dp0 >= 0.0f ? t_near_in : fmax_(t_near_in, fw.t0);
    const f3 q0 = ray_at(o, tq0, d);
    f3 dL_dq0 = mk(0, 0, 0);
    const float hdp_in = dL_dheight * fw.height;
    // A sink whose calls every lane makes together (kUniformHeight) runs
    // this pass on every lane: lanes without a height gradient, and skipped
    // texels, add zeros.
    if (Sink::kUniformHeight || dL_dheight != 0.0f) {
        const float inv_ws0 = 1.0f / fw.wsum0;
#pragma unroll
        for (int i = 0; i < (KT > 0 ? KT : 16); ++i) {
            if (KT == 0 && i >= K) break;
            const f3 s = ld3(sites + 3 * i);
            const f3 dv = q0 - s;
            float w;
            if constexpr (KT > 0) w = fw.w0[i];
            else w = pf_expf(-kTemp * (dot(dv, dv) * inv_r2));
            const bool keep = exact || !(w / fw.wsum0 < 1e-3f);
            if constexpr (!Sink::kUniformHeight) {
                if (!keep) continue;
            }
            const float kw = k2 * w;
            float hg = (w * dL_dheight) * inv_ws0;
            const float hdp = dL_dheight * (float)heights[i];
            const float coef = (hdp - hdp_in) * inv_ws0;
            f3 gs = (kw * coef) * dv;
            if constexpr (Sink::kUniformHeight) {
                hg = keep ? hg : 0.0f;
                gs = keep ? gs : mk(0, 0, 0);
            }
            sink.height(i, hg);
            sink.site_div(i, gs);
            dL_dq0 = dL_dq0 + (-1.0f * gs);
            if (exact) dL_dr += coef * (kw * dot(dv, dv) / r);
        }
    }
    const float dL_dtq0 = dot(dL_dq0, d);
    dL_dtnear += (fw.dp >= 0.0f || t_near_in >= fw.t0) ? dL_dtq0 : 0.0f;
    const float dL_dts0 = (fw.dp < 0.0f && t_near_in < fw.t0) ? dL_dtq0 : 0.0f;
    if (__builtin_fabsf(dL_dts0) > 0.0f) {
        f3 dtdp, dtdn;
        float dtdh;
        ray_plane_bwd(o, d, c, n, 0.0f, dtdp, dtdn, dtdh);
        dL_dc = dL_dc + dtdp * dL_dts0;
        dL_dn = dL_dn + dtdn * dL_dts0;
    }
}

// fl(w / wsum) < 1e-3f, decided from products except within 2^-20 of the
// boundary, where the division is done: c * wsum is within 2^-23 of the
// exact product, so outside that band the product comparison agrees with
// the correctly rounded quotient's (texture.py:118-119, :158-159 test the
// quotient).
__device__ __forceinline__ bool below_1e3(float w, float wsum) {
    const float t = 1e-3f * wsum;
    if (w < t * (1.0f - 0x1p-20f)) return true;
    if (w > t * (1.0f + 0x1p-20f)) return false;
    return w / wsum < 1e-3f;
}

// texel_bwd for a sink whose calls all 64 lanes make together
// (Sink::kUniformHeight): both passes run on every lane (zeros where a
// texel is skipped or no height gradient flows), and each texel site's two
// terms (colour pass, height pass) are summed before one sink call.  The
// arithmetic of each term is texel_bwd's.
template <typename A, int KT, typename RGBCell, typename Sink>
__device__ __forceinline__ void texel_bwd_uniform(f3 o, f3 d, float t_near_in, f3 c, f3 n, float r, const A *sites,
                                                  const RGBCell &rgbs, const A *heights, const TexFwdT<KT> &fw,
                                                  f3 dL_drgb, float dL_dtsurf_in, const Sink &sink, f3 &dL_dc,
                                                  f3 &dL_dn, float &dL_dtnear, bool exact, float &dL_dr) {
    static_assert(KT > 0, "fixed K");
    dL_dc = mk(0, 0, 0);
    dL_dn = mk(0, 0, 0);
    float dL_dheight = 0.0f;
    const float inv_r2 = 1.0f / (r * r);
    const float tq = fw.dp >= 0.0f ? t_near_in : fmax_(t_near_in, fw.t_surf);
    const f3 q = ray_at(o, tq, d);
    f3 dL_dq = mk(0, 0, 0);
    const float gdp_in = dot(dL_drgb, fw.rgb);
    const float k2 = (2.0f * kTemp) * inv_r2;
    const float inv_ws = 1.0f / fw.wsum;
    float a[KT];  // colour-pass site coefficient: site term a_i (q - s_i)
#pragma unroll
    for (int i = 0; i < KT; ++i) {
        const f3 s = ld3(sites + 3 * i);
        const f3 dv = q - s;
        const float w = fw.w1[i];
        const bool keep = exact || !below_1e3(w, fw.wsum);
        const float kw = k2 * w;
        f3 g_rgb = (w * inv_ws) * dL_drgb;
        const float gdp = dot(dL_drgb, rgbs(i, s));
        const float coef = (gdp - gdp_in) * inv_ws;
        const float ai = kw * coef;
        const f3 gs = ai * dv;
        g_rgb = keep ? g_rgb : mk(0, 0, 0);
        a[i] = keep ? ai : 0.0f;
        sink.rgb(i, g_rgb);
        if (keep) dL_dq = dL_dq + (-1.0f * gs);
        if (exact) dL_dr += coef * (kw * dot(dv, dv) / r);
    }
    const float dL_dtq = dot(dL_dq, d);
    dL_dtnear = (fw.dp >= 0.0f || t_near_in >= fw.t_surf) ? dL_dtq : 0.0f;
    float dL_dts = (fw.dp < 0.0f && t_near_in < fw.t_surf) ? dL_dtq : 0.0f;
    dL_dts += dL_dtsurf_in;
    if (__builtin_fabsf(dL_dts) > 0.0f) {
        f3 dtdp, dtdn;
        float dtdh;
        ray_plane_bwd(o, d, c, n, fw.height, dtdp, dtdn, dtdh);
        dL_dc = dL_dc + dtdp * dL_dts;
        dL_dn = dL_dn + dtdn * dL_dts;
        dL_dheight += dtdh * dL_dts;
    }
    const float tq0 = fw.dp0 >= 0.0f ? t_near_in : fmax_(t_near_in, fw.t0);
    const f3 q0 = ray_at(o, tq0, d);
    f3 dL_dq0 = mk(0, 0, 0);
    const float hdp_in = dL_dheight * fw.height;
    const float inv_ws0 = 1.0f / fw.wsum0;
#pragma unroll
    for (int i = 0; i < KT; ++i) {
        const f3 s = ld3(sites + 3 * i);
        const f3 dv0 = q0 - s;
        const float w = fw.w0[i];
        const bool keep = exact || !below_1e3(w, fw.wsum0);
        const float kw = k2 * w;
        float hg = (w * dL_dheight) * inv_ws0;
        const float hdp = dL_dheight * (float)heights[i];
        const float coef = (hdp - hdp_in) * inv_ws0;
        f3 gs0 = (kw * coef) * dv0;
        hg = keep ? hg : 0.0f;
        gs0 = keep ? gs0 : mk(0, 0, 0);
        sink.height(i, hg);
        sink.site(i, a[i] * (q - s) + gs0);
        dL_dq0 = dL_dq0 + (-1.0f * gs0);
        if (exact) dL_dr += coef * (kw * dot(dv0, dv0) / r);
    }
    const float dL_dtq0 = dot(dL_dq0, d);
    dL_dtnear += (fw.dp >= 0.0f || t_near_in >= fw.t0) ? dL_dtq0 : 0.0f;
    const float dL_dts0 = (fw.dp < 0.0f && t_near_in < fw.t0) ? dL_dtq0 : 0.0f;
    if (__builtin_fabsf(dL_dts0) > 0.0f) {
        f3 dtdp, dtdn;
        float dtdh;
        ray_plane_bwd(o, d, c, n, 0.0f, dtdp, dtdn, dtdh);
        dL_dc = dL_dc + dtdp * dL_dts0;
        dL_dn = dL_dn + dtdn * dL_dts0;
    }
}

// --------------------------------------------------------------- the walk --
// Algorithmic bytes (what a visit, a face or a shaded segment must read at
// least once): sphere 16 B + CSR row bounds 8 B + [n, sigma]; per face the
// neighbour index and the face; per shaded segment K texels of
// (site, colour, height).
template <typename A, typename F> struct WalkBytes {
    static constexpr uint64_t visit = 16 + 8 + 4 * sizeof(A);
    static constexpr uint64_t face = 4 + 4 * sizeof(F);
    static constexpr uint64_t texel = 7 * sizeof(A);
};
template <typename A> struct WalkBytes<A, PackedF16> : WalkBytes<A, _Float16> {};

template <typename A> struct SceneDev {
    const float4 *spheres;
    const A *nsig, *sites, *height;
    const int *adj, *off;
    const void *faces;
    int N, K;
    int64_t E;
};

template <typename A> __device__ __forceinline__ SceneDev<A> scene_dev(const pf_scene &s) {
    return SceneDev<A>{reinterpret_cast<const float4 *>(s.spheres), reinterpret_cast<const A *>(s.nsigmas),
                       reinterpret_cast<const A *>(s.texel_sites), reinterpret_cast<const A *>(s.texel_height),
                       s.adjacency, s.adjacency_offsets, s.faces, (int)s.num_cells, s.num_texel_sites,
                       s.num_edges};
}

// Brent's cycle detection on the walk state.  The next cell is a function of
// the current cell alone (the exit face of a fixed ray), and termination a
// function of log T, so a walk that comes back to a (cell, log T) pair it
// held before repeats that loop forever; the reference's loop
// (raytrace.py:163) never ends on such a ray.  Three fp16 face planes that
// no longer meet in a common line make such loops.  The walk stops on the
// repeat and reports PF_STATUS_CYCLE; a walk whose state never repeats takes
// exactly the steps it took before, and is found within twice the length of
// its tail plus loop.
struct CycleGuard {
    int cell = kNoCell;
    float log_t = 0.0f;
    uint32_t power = 1, lam = 1;
    __device__ __forceinline__ bool repeats(int prim, float lt) {
        if (prim == cell && lt == log_t) return true;
        if (lam == power) {
            cell = prim;
            log_t = lt;
            power <<= 1;
            lam = 0;
        }
        ++lam;
        return false;
    }
};

struct RayFwd {
    f3 rgb;
    float log_t, depth;
    int visits, status;
    uint64_t hash, bytes;
    int segs;  // contributing segments (recorded or not)
};

// A ray's record of contributing segments: segment k of the ray in slot
// `slot` of a record over S slots is entry rec_index(k, slot, S), as
// {cell, face-bounded t_near bits, face-bounded t_far bits,
// near face id (low 16 bits) | far face id (high 16 bits)}; ids index the
// cell's CSR row, -1 for the sphere, -2 for a sphere bound behind the
// origin.  Segments past cap are counted but not stored.  Eight
// consecutive segments of one ray share a 128-byte line (the forward's
// scattered writes fill lines), and a tile's 64 rays read 64 lines that
// serve their next seven segments too.  cap is a multiple of 8.
constexpr int kRecChunk = 8;
__device__ __forceinline__ int64_t rec_index(int k, int64_t slot, int64_t S) {
    return ((int64_t)(k / kRecChunk) * S + slot) * kRecChunk + (k % kRecChunk);
}
struct SegRecord {
    int4 *rec;  // the record's first entry
    int64_t slot, S;
    int cap;
    // optional, same indexing as rec: {transmittance before the segment,
    // the ray's colour after it}, which makes a segment's gradients
    // independent of the other segments of its ray
    float4 *part = nullptr;
};
constexpr uint64_t kSegPartBytes = 16;
constexpr uint64_t kSegRecordBytes = 16;

__device__ __forceinline__ int4 pack_segment(int cell, float tn, float tf, int tn_id, int tf_id) {
    return make_int4(cell, __float_as_int(tn), __float_as_int(tf),
                     (int)(((uint32_t)tn_id & 0xffffu) | ((uint32_t)tf_id << 16)));
}

// raytrace.py:157-241; depth as rasterize.py:871-877.  REC: also write
// the record of contributing segments the recorded backward reads.
template <typename A, typename F, int KT, typename RGB, bool REC = false>
__device__ __forceinline__ RayFwd walk_ray_fwd(const SceneDev<A> &s, const RGB &colors, f3 o, f3 d, int start,
                                               float thr, float depth_q, int max_steps, int *seq, int seq_cap,
                                               SegRecord rec = SegRecord{nullptr, 0, 0, 0}) {
    constexpr uint64_t kVisitBytes = WalkBytes<A, F>::visit, kFaceBytes = WalkBytes<A, F>::face,
                       kTexelBytes = WalkBytes<A, F>::texel;
    const int K = KT > 0 ? KT : s.K;
    const RowReader<F> rows(s.faces, s.E);
    RayFwd out{mk(0, 0, 0), 0.0f, 0.0f, 0, 0, kFnvOffset, 0, 0};
    bool depth_found = false;
    int prim = start;
    CycleGuard guard;
    while (true) {
        const float trans = pf_expf(out.log_t);
        if (trans < thr || prim == kNoCell) break;
        if ((unsigned)prim >= (unsigned)s.N || out.visits >= max_steps) { out.status |= PF_STATUS_STEP_CAP; break; }
        if (guard.repeats(prim, out.log_t)) { out.status |= PF_STATUS_CYCLE; break; }
        if (seq && out.visits < seq_cap) seq[out.visits] = prim;
        ++out.visits;
        out.hash = (out.hash ^ (uint64_t)(uint32_t)prim) * kFnvPrime;

        const float4 sp = s.spheres[prim];
        f3 nrm;
        float sigma;
        ld4(s.nsig + 4 * (int64_t)prim, nrm, sigma);  // fetched with the sphere, used after the faces
        const f3 c = mk(sp.x, sp.y, sp.z);
        const float r = sp.w;
        SphereHit sh = ray_sphere(o, d, c, r);
        if (length(c - o) < 4.0f * r) sh.hit = false;
        float tn = sh.tn, tf = sh.tf;
        // the bounding faces: the near one tracked as walk_ray_bwd's replay
        // tracks it, the far one from the exit face after the scan
        int tn_id = 0, tf_id = 0;
        if constexpr (REC) {
            tn_id = tn < 1e-8f ? -2 : -1;
            tf_id = tf < 1e-8f ? -2 : -1;
        }

        const int e0 = s.off[prim];
        const int e1 = s.off[prim + 1];
        out.bytes += kVisitBytes + (uint64_t)(e1 - e0) * kFaceBytes;
        FaceFold ff;
        rows.scan(e0, e1, [&](int e, f3 fn, float fw, bool valid) { ff.add(e, o, d, fn, fw, valid); });
        const int exit_e = ff.exit_face();
        const float pt_far = ff.exit_t();
        if constexpr (REC) tn_id = ff.near_id(tn, e0, tn_id);
        tf = ff.t_far(tf);
        tn = ff.t_near(tn);
        const int next = exit_e >= 0 ? face_neighbour<F>(s.faces, s.adj, exit_e) : kNoCell;
        if (!sh.hit || tn > tf || sigma < 1e-3f) { prim = next; continue; }
        if constexpr (REC) {
            // the face bounding t_far is the exit face when the exit comes
            // before the sphere's far bound (both are the first face at the
            // minimum t of the faces ahead); the exit search starts at 1e10,
            // so this holds for sphere bounds up to 1e10 (larger ones send
            // the ray to the replay, below)
            if (exit_e >= 0 && pt_far < sh.tf) tf_id = exit_e - e0;
        }

        const int64_t tbase = (int64_t)prim * K;
        out.bytes += (uint64_t)K * kTexelBytes;
        const TexFwdT<KT> tx = texel_fwd<A, KT>(o, d, tn, c, nrm, r, s.sites + 3 * tbase, colors.cell(prim, K),
                                          s.height + tbase, K);
        const float tn_face = tn, tf_face = tf;
        tf = tx.dp >= 0.0f ? fmin_(tx.t_surf, tf) : tf;
        tn = tx.dp < 0.0f ? fmax_(tx.t_surf, tn) : tn;
        const int cell = prim;
        prim = next;
        const float dt = tf - tn;
        if (dt > 0.0f) {
            int64_t rec_at = -1;
            if constexpr (REC) {
                if (!(sh.tf <= 1e10f)) out.segs = out.segs > rec.cap ? out.segs : rec.cap;
                if (out.segs < rec.cap) {
                    rec_at = rec_index(out.segs, rec.slot, rec.S);
                    rec.rec[rec_at] = pack_segment(cell, tn_face, tf_face, tn_id, tf_id);
                    out.bytes += kSegRecordBytes;
                }
                ++out.segs;
            }
            const float dlt = (-sigma) * dt;
            const float alpha = 1.0f - pf_expf(dlt);
            out.rgb.x = __builtin_fmaf(tx.rgb.x * alpha, trans, out.rgb.x);
            out.rgb.y = __builtin_fmaf(tx.rgb.y * alpha, trans, out.rgb.y);
            out.rgb.z = __builtin_fmaf(tx.rgb.z * alpha, trans, out.rgb.z);
            out.log_t += dlt;
            if constexpr (REC) {
                if (rec.part && rec_at >= 0) {
                    rec.part[rec_at] = make_float4(trans, out.rgb.x, out.rgb.y, out.rgb.z);
                    out.bytes += kSegPartBytes;
                }
            }
            if (!depth_found) {
                const float next_trans = pf_expf(out.log_t);
                if (next_trans < depth_q) {
                    out.depth = tn + pf_logf(trans / depth_q) / sigma;
                    depth_found = true;
                }
            }
        }
    }
    return out;
}

struct GradDev {
    float *spheres, *nsigmas, *sites, *rgb, *height;
};

// The value of v in lane l ^ 1 (quad_perm [1, 0, 3, 2]).
__device__ __forceinline__ float lane_xor1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ int lane_xor1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false); }
// The value of v in lane l ^ 8 (row_ror:8 within each row of 16 lanes).
__device__ __forceinline__ float lane_xor8(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xF, 0xF, false));
}
__device__ __forceinline__ int lane_xor8(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false); }

// Lane pairing of a wave whose 64 lanes all make every paired sink call:
// lanes l and l ^ 1 (vertically adjacent pixels of the tile) that add into
// the same cache row sum their terms first, the even lane keeping the sum
// and the odd one zero; then lanes l and l ^ 8 (horizontally adjacent) do
// the same, so a 2x2 quad of rays on one cell makes one LDS atomic instead
// of four.  Each exchange moves a term to a lane of the same row, so every
// row receives the same total.  A lane without a segment adds zeros.
struct LanePair {
    bool active;  // this lane has a segment
    bool match1;  // lane l ^ 1 adds into the same row
    bool match8;  // lane l ^ 8 adds into the same row
    bool lead1;   // l & 1 == 0
    bool lead8;   // l & 8 == 0
};

// Destination of one shaded segment's gradients for its own cell: the
// sphere [c, r], [n, sigma] and the K texels' sites, colours and heights.
// lds == nullptr adds straight into the global arrays; otherwise into the
// cell's row of a wave's LDS accumulator (CellCache below), in the layout
// sphere 0-3, nsigma 4-7, sites 8.., colours 8+3K.., heights 8+6K...
// Zero contributions never reach global memory.  PAIR: the calls of
// sphere, nsig, rgb and site are made by all 64 lanes together and use
// LanePair; site_div and height may be called from divergent code.
template <bool PAIR> struct CellSink {
    static constexpr bool kUniform = PAIR;
    static constexpr bool kUniformHeight = false;
    float *lds;
    float *gs, *gn, *gsite, *grgb, *gh;
    int K;
    LanePair pr;
    __device__ __forceinline__ void add_div(int off, float *g, float v) const {
        // zero contributions are skipped: an LDS float atomic costs per
        // active lane, and many of a segment's terms are exactly zero
        if (v == 0.0f) return;
        if (lds) lds_add(lds + off, v);
        else atomic_add(g, v);
    }
    __device__ __forceinline__ void add(int off, float *g, float v) const {
        if constexpr (PAIR) {
            v = pr.active ? v : 0.0f;
            const float o1 = lane_xor1(v);
            v = pr.match1 ? (pr.lead1 ? v + o1 : 0.0f) : v;
            const float o8 = lane_xor8(v);
            v = pr.match8 ? (pr.lead8 ? v + o8 : 0.0f) : v;
        }
        add_div(off, g, v);
    }
    __device__ __forceinline__ void sphere(f3 dc, float dr) const {
        add(0, gs + 0, dc.x);
        add(1, gs + 1, dc.y);
        add(2, gs + 2, dc.z);
        add(3, gs + 3, dr);
    }
    __device__ __forceinline__ void nsig(f3 dn, float dsigma) const {
        add(4, gn + 0, dn.x);
        add(5, gn + 1, dn.y);
        add(6, gn + 2, dn.z);
        add(7, gn + 3, dsigma);
    }
    __device__ __forceinline__ void site(int i, f3 v) const {
        add(8 + 3 * i, gsite + 3 * i, v.x);
        add(9 + 3 * i, gsite + 3 * i + 1, v.y);
        add(10 + 3 * i, gsite + 3 * i + 2, v.z);
    }
    __device__ __forceinline__ void rgb(int i, f3 v) const {
        add(8 + 3 * K + 3 * i, grgb + 3 * i, v.x);
        add(9 + 3 * K + 3 * i, grgb + 3 * i + 1, v.y);
        add(10 + 3 * K + 3 * i, grgb + 3 * i + 2, v.z);
    }
    __device__ __forceinline__ void site_div(int i, f3 v) const {
        add_div(8 + 3 * i, gsite + 3 * i, v.x);
        add_div(9 + 3 * i, gsite + 3 * i + 1, v.y);
        add_div(10 + 3 * i, gsite + 3 * i + 2, v.z);
    }
    __device__ __forceinline__ void height(int i, float v) const { add_div(8 + 6 * K + i, gh + i, v); }
};

// Per-lane policy: every lane sends its gradients straight to global memory
// with atomics, and the walk loop may diverge freely.
struct LaneCache {
    static constexpr bool kWave = false;
    __device__ __forceinline__ int assign(int) { return -1; }
    __device__ __forceinline__ float *row(int) const { return nullptr; }
    __device__ __forceinline__ void flush_all() {}
};

// Wave policy: a software cache of per-cell gradient rows in LDS shared by
// the 64 lanes (one 8x8 tile) of a wave.  Neighbouring rays cross the same
// cells, so a tile's shaded segments fall on far fewer cells than segments;
// the rows absorb them with LDS atomics and each cell reaches global memory
// with one 64-lane atomic per residency instead of one per segment.  Slots
// are fully associative: lane j holds the id of the cell in slot j; each
// step assigns slots to the distinct cells the lanes shade (one ballot
// round per distinct cell), evicting first-in-first-out among slots not
// used this step.  Lanes whose cell finds no slot (more distinct cells in a
// step than slots) use global atomics for that segment.  Every lane of the
// wave must call assign() and flush_all() together (uniform control flow).
template <int S, int KT> struct CellCache {
    static constexpr bool kWave = true;
    static constexpr int kRow = 8 + 7 * KT;    // floats per cell row
    static constexpr int kPitch = kRow + 1;    // odd pitch: rows start in distinct banks
    static_assert(S <= 64 && kRow <= 64, "one slot per lane and one row element per lane");
    float *acc;   // S * kPitch floats of LDS, zeroed
    int owner;    // lane j < S: cell in slot j, or kNoCell
    uint32_t rr;  // next slot to evict
    GradDev G;

    __device__ __forceinline__ void init(float *lds, const GradDev &g) {
        acc = lds;
        G = g;
        const int lane = __lane_id();
        owner = kNoCell;
        rr = 0;
        for (int k = lane; k < S * kPitch; k += 64) as_lds(acc)[k] = 0.0f;
        __builtin_amdgcn_wave_barrier();
    }
    __device__ __forceinline__ float *row(int slot) const { return slot >= 0 ? acc + slot * kPitch : nullptr; }

    // Adds slot s's row to cell's global gradients and clears it: lane k
    // moves element k.
    __device__ __forceinline__ void flush(int s, int cell) {
        const int k = __lane_id();
        if (k < kRow) {
            lds_float *p = as_lds(acc) + s * kPitch + k;
            const float v = *p;
            *p = 0.0f;
            if (v != 0.0f) {
                const int64_t c = cell;
                float *dst;
                if (k < 4) dst = G.spheres + 4 * c + k;
                else if (k < 8) dst = G.nsigmas + 4 * c + (k - 4);
                else if (k < 8 + 3 * KT) dst = G.sites + 3 * KT * c + (k - 8);
                else if (k < 8 + 6 * KT) dst = G.rgb + 3 * KT * c + (k - 8 - 3 * KT);
                else dst = G.height + KT * c + (k - 8 - 6 * KT);
                atomic_add(dst, v);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }

    // cell: this lane's shaded cell this step, or kNoCell.  Returns the slot
    // holding the cell's row, or -1 (use global atomics).
    __device__ __forceinline__ int assign(int cell) {
        constexpr uint64_t kMask = S == 64 ? ~0ull : ((1ull << S) - 1);
        const int lane = __lane_id();
        uint64_t pending = __ballot(cell != kNoCell);
        uint64_t used = 0;
        int mine = -1;
        while (pending) {
            const int lead = __ffsll((unsigned long long)pending) - 1;
            const int lc = __builtin_amdgcn_readlane(cell, lead);
            const uint64_t members = __ballot(cell == lc);
            pending &= ~members;
            const uint64_t hit = __ballot(owner == lc) & kMask;
            int slot = -1;
            if (hit) {
                slot = __ffsll((unsigned long long)hit) - 1;
            } else {
                const uint64_t free = ~used & kMask;
                if (free) {
                    const uint64_t after = free & (~0ull << rr);
                    slot = __ffsll((unsigned long long)(after ? after : free)) - 1;
                    rr = slot + 1 == S ? 0 : slot + 1;
                    const int old = __builtin_amdgcn_readlane(owner, slot);
                    if (old != kNoCell) flush(slot, old);
                    if (lane == slot) owner = lc;
                }
            }
            if (slot >= 0) {
                used |= 1ull << slot;
                if ((members >> lane) & 1) mine = slot;
            }
        }
        __builtin_amdgcn_wave_barrier();
        return mine;
    }

    __device__ __forceinline__ void flush_all() {
        uint64_t m = __ballot(owner != kNoCell) & (S == 64 ? ~0ull : ((1ull << S) - 1));
        while (m) {
            const int slot = __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
            flush(slot, __builtin_amdgcn_readlane(owner, slot));
        }
    }
};

// One contributing segment of a replayed ray: what its gradients need
// besides the ray (o, d), its incoming gradients and the texel pass.
struct Seg {
    int cell;
    f3 c, nrm;
    float r, sigma, tn0, dt, e_dlt, alpha, trans;
    f3 rgb;  // the ray's colour after this segment
    int tn_id, tf_id, tn0_id, tf0_id, e0, n_adj;
    int64_t tbase;
};

// Algorithmic bytes of one segment's gradients: the read-modify-write of
// its cell's row, and 32 per bounding neighbour face.
__device__ __forceinline__ uint64_t segment_grad_bytes(int K, int tn0_id, int tf0_id, int n_adj) {
    return 2ull * (32ull + (uint64_t)K * 28ull) + (tn0_id >= 0 && tn0_id < n_adj ? 32ull : 0ull) +
           (tf0_id >= 0 && tf0_id < n_adj ? 32ull : 0ull);
}

// The gradients a segment sends to the neighbours across its bounding
// faces (side 0: near, 1: far): the neighbour's sphere [c, r] terms, or
// nb = -1 when that bound is not a face.
struct FaceTerms {
    int nb[2];
    f3 dc[2];
    float dr[2];
};

// Gradients of one segment (rasterize.py:1371-1516, texture.py:74-186):
//   dL/d(delta_s) = g.(C - C_{<=s}) + g_logT - (g.c_s) T_s exp(delta_s),
// routed to the sphere, a power face or the texel plane.  The cell's own
// gradients go to sink (sphere, [n, sigma], texel sites, colours,
// heights); neighbour spheres get global atomics, or, with ft, are handed
// back for the caller to sum across lanes first.
// The segment's incoming terms: Tg = T_s g (the ray's colour gradient
// scaled by the transmittance before the segment) and Lnext =
// g.(C - C_{<=s}) + g_logT, the gradient of the log transmittance after it.
__device__ __forceinline__ f3 seg_Tg(f3 g, float trans) { return g * trans; }
__device__ __forceinline__ float seg_Lnext(f3 g, float gL, f3 Cfin, f3 Cs) { return dot(g, Cfin - Cs) + gL; }

template <typename A, typename F, int KT, typename RGB, typename Sink>
__device__ __forceinline__ void segment_grads_to(const SceneDev<A> &s, const RGB &colors, f3 o, f3 d, const Seg &q,
                                                 const TexFwdT<KT> &tx, f3 Tg, float Lnext, const GradDev &G,
                                                 bool exact, const Sink &sink, FaceTerms *ft = nullptr) {
    if (ft) ft->nb[0] = ft->nb[1] = -1;
    const int K = KT > 0 ? KT : s.K;
    const int cell = q.cell, n_adj = q.n_adj;
    const f3 c = q.c;
    const float r = q.r, sigma = q.sigma, dt = q.dt;
    const float dL_dLnext = Lnext;
    const float dL_dalpha = dot(Tg, tx.rgb);
    const f3 dL_dcolor = Tg * q.alpha;
    const float dL_ddelta = dL_dLnext + (-dL_dalpha * q.e_dlt);
    const float dL_dsigma = dL_ddelta * (-dt);
    const float dL_dtn = dL_ddelta * sigma;
    const float dL_dtf = dL_ddelta * (-sigma);

    float dL_dts = 0.0f;
    if (q.tn_id == n_adj) dL_dts = dL_dtn;
    else if (q.tf_id == n_adj) dL_dts = dL_dtf;

    const int64_t tbase = q.tbase;
    const auto rgbs = colors.cell(cell, K);
    f3 dL_dc, dL_dn;
    float dL_dtn0, dL_dr = 0.0f;
    if constexpr (Sink::kUniformHeight)
        texel_bwd_uniform<A, KT>(o, d, q.tn0, c, q.nrm, r, s.sites + 3 * tbase, rgbs, s.height + tbase, tx,
                                 dL_dcolor, dL_dts, sink, dL_dc, dL_dn, dL_dtn0, exact, dL_dr);
    else
        texel_bwd<A, KT>(o, d, q.tn0, c, q.nrm, r, s.sites + 3 * tbase, rgbs, s.height + tbase, K, tx, dL_dcolor,
                         dL_dts, sink, dL_dc, dL_dn, dL_dtn0, exact, dL_dr);
    if (q.tn_id != n_adj) dL_dtn0 += dL_dtn;
    for (int side = 0; side < 2; ++side) {
        const int tid = side == 0 ? q.tn0_id : q.tf0_id;
        const float dLdt = side == 0 ? dL_dtn0 : (q.tf_id != n_adj ? dL_dtf : 0.0f);
        if (tid == -1) {
            const SphereGrad sg = ray_sphere_bwd(o, d, c, r);
            if (side == 0) {
                dL_dc = dL_dc + sg.dn_dc * dLdt;
                dL_dr += dLdt * sg.dn_dr;
            } else {
                dL_dc = dL_dc + sg.df_dc * dLdt;
                dL_dr += dLdt * sg.df_dr;
            }
        } else if (tid >= 0 && tid < n_adj) {
            const int adjc = s.adj[q.e0 + tid];
            const float4 asp = s.spheres[adjc];
            const FaceGrad fg = ray_pface_bwd(o, d, c, r, mk(asp.x, asp.y, asp.z), asp.w);
            dL_dc = dL_dc + fg.dc * dLdt;
            dL_dr += dLdt * fg.dr;
            if (ft) {
                ft->nb[side] = adjc;
                ft->dc[side] = fg.dcj * dLdt;
                ft->dr[side] = dLdt * fg.drj;
            } else {
                float *ga = G.spheres + 4 * (int64_t)adjc;
                atomic_add(ga + 0, dLdt * fg.dcj.x);
                atomic_add(ga + 1, dLdt * fg.dcj.y);
                atomic_add(ga + 2, dLdt * fg.dcj.z);
                atomic_add(ga + 3, dLdt * fg.drj);
            }
        }
    }
    sink.sphere(dL_dc, dL_dr);
    sink.nsig(dL_dn, dL_dsigma);
}

// segment_grads_to with the cell's own gradients going to row (an LDS row
// of the wave's cache) or, row == nullptr, to global memory.
template <typename A, typename F, int KT, typename RGB, bool PAIR = false>
__device__ __forceinline__ void segment_grads(const SceneDev<A> &s, const RGB &colors, f3 o, f3 d, const Seg &q,
                                              const TexFwdT<KT> &tx, f3 g, float gL, f3 Cfin, const GradDev &G,
                                              bool exact, float *row,
                                              LanePair pr = LanePair{true, false, false, false, false}) {
    const int K = KT > 0 ? KT : s.K;
    const int64_t tbase = q.tbase;
    const CellSink<PAIR> sink{row,
                              G.spheres + 4 * (int64_t)q.cell,
                              G.nsigmas + 4 * (int64_t)q.cell,
                              G.sites + 3 * tbase,
                              G.rgb + 3 * tbase,
                              G.height + tbase,
                              K,
                              pr};
    segment_grads_to<A, F, KT>(s, colors, o, d, q, tx, seg_Tg(g, q.trans), seg_Lnext(g, gL, Cfin, q.rgb), G, exact,
                               sink);
}

// Backward by checkpoint and replay.  The forward's per-ray colour C and
// log transmittance are the checkpoint.  The ray replays its walk in
// forward order with the forward's arithmetic (so the cell sequence and
// every segment are bit-identical), and at each shaded segment s uses
//   dL/d(delta_s) = g.(C - C_{<=s}) + g_logT - (g.c_s) T_s exp(delta_s)
// which equals the reverse-order accumulation of the rasterizer backward
// (rasterize.py:1371-1426).  The per-segment routing of dL/dt to the
// sphere, a power face, or the texel plane mirrors rasterize.py:1296-1516.
// live == false: the lane has no ray (or a zero incoming gradient) but
// still takes part in a wave policy's collective steps.  Returns the
// status bits; adds the ray's algorithmic bytes to bytes.
template <typename A, typename F, int KT, typename RGB, typename Cache = LaneCache>
__device__ __forceinline__ int walk_ray_bwd(const SceneDev<A> &s, const RGB &colors, f3 o, f3 d, int start, float thr,
                                            int max_steps, f3 Cfin, f3 g, float gL, const GradDev &G, bool exact,
                                            uint64_t &bytes, bool live = true, Cache cache = Cache{}) {
    constexpr uint64_t kVisitBytes = WalkBytes<A, F>::visit, kFaceBytes = WalkBytes<A, F>::face,
                       kTexelBytes = WalkBytes<A, F>::texel;
    const int K = KT > 0 ? KT : s.K;
    const RowReader<F> rows(s.faces, s.E);
    f3 rgb = mk(0, 0, 0);
    float log_t = 0.0f;
    int prim = start;
    int visits = 0;
    int status = 0;
    CycleGuard guard;
    while (true) {
        float trans = 0.0f;
        if (live) {
            trans = pf_expf(log_t);
            if (trans < thr || prim == kNoCell) {
                live = false;
            } else if ((unsigned)prim >= (unsigned)s.N || visits >= max_steps) {
                status = PF_STATUS_STEP_CAP;
                live = false;
            } else if (guard.repeats(prim, log_t)) {
                status = PF_STATUS_CYCLE;
                live = false;
            }
        }
        if constexpr (Cache::kWave) {
            if (!__any(live)) break;
        } else {
            if (!live) break;
        }

        // ---- replay one step; shade = this segment contributes
        int cell = kNoCell;
        f3 c = mk(0, 0, 0), nrm = mk(0, 0, 0);
        float r = 1.0f, sigma = 0.0f, tn = 0.0f, tf = 0.0f, tn0 = 0.0f, dt = 0.0f, e_dlt = 1.0f, alpha = 0.0f;
        int tn_id = -1, tf_id = -1, tn0_id = -1, tf0_id = -1, e0 = 0, n_adj = 0;
        int64_t tbase = 0;
        TexFwdT<KT> tx{};
        if (live) {
            ++visits;
            const float4 sp = s.spheres[prim];
            ld4(s.nsig + 4 * (int64_t)prim, nrm, sigma);  // fetched with the sphere, used after the faces
            c = mk(sp.x, sp.y, sp.z);
            r = sp.w;
            SphereHit sh = ray_sphere(o, d, c, r);
            if (length(c - o) < 4.0f * r) sh.hit = false;
            tn = sh.tn;
            tf = sh.tf;
            tn_id = tn < 1e-8f ? -2 : -1;
            tf_id = tf < 1e-8f ? -2 : -1;

            e0 = s.off[prim];
            const int e1 = s.off[prim + 1];
            bytes += kVisitBytes + (uint64_t)(e1 - e0) * kFaceBytes;
            n_adj = e1 - e0;
            // unpaired faces: batches of 8 measured faster than 4 here (and
            // slower in the forward)
            FaceFold ff;
            rows.template scan<8>(e0, e1, [&](int e, f3 fn, float fw, bool valid) { ff.add(e, o, d, fn, fw, valid); });
            const int exit_e = ff.exit_face();
            tf_id = ff.far_id(tf, e0, tf_id);
            tn_id = ff.near_id(tn, e0, tn_id);
            tf = ff.t_far(tf);
            tn = ff.t_near(tn);
            const int next = exit_e >= 0 ? face_neighbour<F>(s.faces, s.adj, exit_e) : kNoCell;
            if (sh.hit && !(tn > tf) && !(sigma < 1e-3f)) {
                tbase = (int64_t)prim * K;
                bytes += (uint64_t)K * kTexelBytes;
                tn0 = tn;
                tn0_id = tn_id;
                tf0_id = tf_id;
                tx = texel_fwd<A, KT>(o, d, tn, c, nrm, r, s.sites + 3 * tbase, colors.cell(prim, K),
                                      s.height + tbase, K);
                if (tx.dp >= 0.0f && tx.t_surf < tf) tf_id = n_adj;
                tf = tx.dp >= 0.0f ? fmin_(tx.t_surf, tf) : tf;
                if (tx.dp < 0.0f && tx.t_surf > tn) tn_id = n_adj;
                tn = tx.dp < 0.0f ? fmax_(tx.t_surf, tn) : tn;
                dt = tf - tn;
                if (dt > 0.0f) {
                    cell = prim;
                    const float dlt = (-sigma) * dt;
                    e_dlt = pf_expf(dlt);
                    alpha = 1.0f - e_dlt;
                    rgb.x = __builtin_fmaf(tx.rgb.x * alpha, trans, rgb.x);
                    rgb.y = __builtin_fmaf(tx.rgb.y * alpha, trans, rgb.y);
                    rgb.z = __builtin_fmaf(tx.rgb.z * alpha, trans, rgb.z);
                    log_t += dlt;
                }
            }
            prim = next;
        }

        // ---- gradients of the shaded segment
        const int slot = cache.assign(cell);
        if (cell == kNoCell) continue;
        bytes += segment_grad_bytes(K, tn0_id, tf0_id, n_adj);
        segment_grads<A, F, KT>(s, colors, o, d, Seg{cell, c, nrm, r, sigma, tn0, dt, e_dlt, alpha, trans, rgb, tn_id,
                                                     tf_id, tn0_id, tf0_id, e0, n_adj, tbase},
                                tx, g, gL, Cfin, G, exact, cache.row(slot));
    }
    cache.flush_all();
    return status;
}

// Backward from the forward's record of contributing segments (SegRecord):
// the same per-segment arithmetic as walk_ray_bwd's replay, from the same
// face-bounded t values, so every segment, the running colour and the
// transmittance are bit-identical; no face row is read.  count is the
// ray's number of contributing segments, all of them in the record (a ray
// whose count exceeds the record's cap goes through walk_ray_bwd instead).
// A recorded cell outside the scene (a record of another walk) is skipped
// and reported as PF_STATUS_BAD_RECORD, which the function returns.
template <typename A, typename F, int KT, typename RGB, typename Cache = LaneCache>
__device__ __forceinline__ int walk_rec_bwd(const SceneDev<A> &s, const RGB &colors, f3 o, f3 d, const int4 *rec,
                                             int64_t slot, int64_t S, int count, f3 Cfin, f3 g, float gL, const GradDev &G,
                                             bool exact, uint64_t &bytes, bool live, Cache cache = Cache{}) {
    constexpr uint64_t kVisitBytes = WalkBytes<A, F>::visit, kTexelBytes = WalkBytes<A, F>::texel;
    const int K = KT > 0 ? KT : s.K;
    f3 rgb = mk(0, 0, 0);
    float log_t = 0.0f;
    int status = 0;
    for (int k = 0;; ++k) {
        bool act = live && k < count;
        int4 q = make_int4(0, 0, 0, 0);
        if (act) {
            q = rec[rec_index(k, slot, S)];
            if ((unsigned)q.x >= (unsigned)s.N) {
                status = PF_STATUS_BAD_RECORD;
                live = act = false;
            }
        }
        if constexpr (Cache::kWave) {
            if (!__any(act)) break;
        } else {
            if (!act) break;
        }
        int cell = kNoCell;
        f3 c = mk(0, 0, 0), nrm = mk(0, 0, 0);
        float r = 1.0f, sigma = 0.0f, tn0 = 0.0f, dt = 0.0f, e_dlt = 1.0f, alpha = 0.0f, trans = 0.0f;
        int tn_id = -1, tf_id = -1, tn0_id = -1, tf0_id = -1, e0 = 0, n_adj = 0;
        int64_t tbase = 0;
        TexFwdT<KT> tx{};
        if (act) {
            cell = q.x;
            float tn = __int_as_float(q.y), tf = __int_as_float(q.z);
            tn_id = (int)(int16_t)(q.w & 0xffff);
            tf_id = q.w >> 16;
            trans = pf_expf(log_t);
            const float4 sp = s.spheres[cell];
            ld4(s.nsig + 4 * (int64_t)cell, nrm, sigma);
            c = mk(sp.x, sp.y, sp.z);
            r = sp.w;
            e0 = s.off[cell];
            n_adj = s.off[cell + 1] - e0;
            tbase = (int64_t)cell * K;
            bytes += kSegRecordBytes + kVisitBytes + (uint64_t)K * kTexelBytes;
            tn0 = tn;
            tn0_id = tn_id;
            tf0_id = tf_id;
            tx = texel_fwd<A, KT>(o, d, tn, c, nrm, r, s.sites + 3 * tbase, colors.cell(cell, K), s.height + tbase,
                                  K);
            if (tx.dp >= 0.0f && tx.t_surf < tf) tf_id = n_adj;
            tf = tx.dp >= 0.0f ? fmin_(tx.t_surf, tf) : tf;
            if (tx.dp < 0.0f && tx.t_surf > tn) tn_id = n_adj;
            tn = tx.dp < 0.0f ? fmax_(tx.t_surf, tn) : tn;
            dt = tf - tn;  // > 0: only contributing segments are recorded
            const float dlt = (-sigma) * dt;
            e_dlt = pf_expf(dlt);
            alpha = 1.0f - e_dlt;
            rgb.x = __builtin_fmaf(tx.rgb.x * alpha, trans, rgb.x);
            rgb.y = __builtin_fmaf(tx.rgb.y * alpha, trans, rgb.y);
            rgb.z = __builtin_fmaf(tx.rgb.z * alpha, trans, rgb.z);
            log_t += dlt;
        }
        const int slot = cache.assign(cell);
        if constexpr (Cache::kWave) {
            // all 64 lanes make the gradient calls, so that pairs of lanes
            // on one cell can sum their terms before the LDS atomics; a
            // lane without a segment computes on cell 0 and adds zeros
            const int s1 = lane_xor1(slot), s8 = lane_xor8(slot);
            const int lane = __lane_id();
            const LanePair pr{cell != kNoCell, slot >= 0 && s1 == slot, slot >= 0 && s8 == slot, (lane & 1) == 0,
                              (lane & 8) == 0};
            if (cell != kNoCell) bytes += segment_grad_bytes(K, tn0_id, tf0_id, n_adj);
            else tn0_id = tf0_id = -1;  // the sphere bounds, so no neighbour is touched
            segment_grads<A, F, KT, RGB, true>(s, colors, o, d,
                                               Seg{cell != kNoCell ? cell : 0, c, nrm, r, sigma, tn0, dt, e_dlt,
                                                   alpha, trans, rgb, tn_id, tf_id, tn0_id, tf0_id, e0, n_adj, tbase},
                                               tx, g, gL, Cfin, G, exact, cache.row(slot), pr);
        } else {
            if (cell == kNoCell) continue;
            bytes += segment_grad_bytes(K, tn0_id, tf0_id, n_adj);
            segment_grads<A, F, KT>(s, colors, o, d,
                                    Seg{cell, c, nrm, r, sigma, tn0, dt, e_dlt, alpha, trans, rgb, tn_id, tf_id,
                                        tn0_id, tf0_id, e0, n_adj, tbase},
                                    tx, g, gL, Cfin, G, exact, cache.row(slot));
        }
    }
    cache.flush_all();
    return status;
}

}  // namespace pf
