// walk.hip — the power-cell walk (forward and checkpoint-and-replay
// backward), walk face precompute and start-cell lookup for gfx950.
//
// Forward semantics follow RayTracer.benchmark_kernel
// (powerfoam/raytrace.py:115-241) and its texel lookup
// plane_intersection_fwd_local (raytrace.py:42-113).  One ray per lane, one
// 8x8 tile per 64-wide wavefront, tiles numbered column-major as in the
// reference (raytrace.py:130-151), so lane l of a wave shades pixel
// (tile_i*8 + l%8, tile_j*8 + l/8).
#include "pf_device.h"
#include "pf_hip.h"
#include "pf_launch.h"
#include "pf_texel.h"

namespace pf {

constexpr int kNoCell = 0x7FFFFFFF;
constexpr uint64_t kFnvOffset = 14695981039346656037ull;
constexpr uint64_t kFnvPrime = 1099511628211ull;

struct alignas(8) h4 { _Float16 x, y, z, w; };

template <typename F> struct FaceT;
template <> struct FaceT<_Float16> {
    __device__ static __forceinline__ void load(const void *p, int64_t e, f3 &n, float &w) {
        const h4 v = reinterpret_cast<const h4 *>(p)[e];
        n = mk((float)v.x, (float)v.y, (float)v.z);
        w = (float)v.w;
    }
};
template <> struct FaceT<float> {
    __device__ static __forceinline__ void load(const void *p, int64_t e, f3 &n, float &w) {
        const float4 v = reinterpret_cast<const float4 *>(p)[e];
        n = mk(v.x, v.y, v.z);
        w = v.w;
    }
};

template <typename A> __device__ __forceinline__ void ld4(const A *p, f3 &n, float &s) {
    n = mk((float)p[0], (float)p[1], (float)p[2]);
    s = (float)p[3];
}

__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }

// Texel backward: the semantics of plane_intersection_bwd
// (powerfoam/texture.py:74-186) with the walk's weights (r^-2 multiply).
// Returns d/d(plane origin), d/d(normal) and d/d(t_near before the plane).
// exact == false reproduces the reference's two approximations: texels
// whose weight is below 1e-3 of the weight sum are skipped
// (texture.py:118-119, :158-159) and the radius gets no gradient through
// the texel weights.  exact == true keeps every texel and adds that radius
// term, d w_k / d r = 2 temp w_k |q - s_k|^2 / r^3, into dL_dr.
template <typename A, int KT>
__device__ __forceinline__ void texel_bwd(f3 o, f3 d, float t_near_in, f3 c, f3 n, float r,
                                          const A *sites, const A *rgbs, const A *heights, int K,
                                          const TexFwd &fw, f3 dL_drgb, float dL_dtsurf_in,
                                          float *g_sites, float *g_rgb, float *g_height,
                                          f3 &dL_dc, f3 &dL_dn, float &dL_dtnear, bool exact,
                                          float &dL_dr) {
    if constexpr (KT > 0) K = KT;
    dL_dc = mk(0, 0, 0);
    dL_dn = mk(0, 0, 0);
    float dL_dheight = 0.0f;
    const float inv_r2 = 1.0f / (r * r);
    const float r2 = r * r;

    const float tq = fw.dp >= 0.0f ? t_near_in : fmax_(t_near_in, fw.t_surf);
    const f3 q = ray_at(o, tq, d);
    f3 dL_dq = mk(0, 0, 0);
    const float gdp_in = dot(dL_drgb, fw.rgb);
#pragma unroll
    for (int i = 0; i < (KT > 0 ? KT : 16); ++i) {
        if (KT == 0 && i >= K) break;
        const f3 s = ld3(sites + 3 * i);
        const f3 dv = q - s;
        const float dsq = dot(dv, dv) * inv_r2;
        const float w = pf_expf(-kTemp * dsq);
        if (!exact && w / fw.wsum < 1e-3f) continue;
        const f3 dw_ds = divs(((2.0f * kTemp) * w) * dv, r2);
        const f3 grgb = divs(w * dL_drgb, fw.wsum);
        atomic_add(g_rgb + 3 * i + 0, grgb.x);
        atomic_add(g_rgb + 3 * i + 1, grgb.y);
        atomic_add(g_rgb + 3 * i + 2, grgb.z);
        const float gdp = dot(dL_drgb, ld3(rgbs + 3 * i));
        const f3 gs = divs(dw_ds * (gdp - gdp_in), fw.wsum);
        atomic_add(g_sites + 3 * i + 0, gs.x);
        atomic_add(g_sites + 3 * i + 1, gs.y);
        atomic_add(g_sites + 3 * i + 2, gs.z);
        dL_dq = dL_dq + (-1.0f * gs);
        if (exact) dL_dr += ((gdp - gdp_in) / fw.wsum) * ((((2.0f * kTemp) * w) * dot(dv, dv)) / (r2 * r));
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
#pragma unroll
    for (int i = 0; i < (KT > 0 ? KT : 16); ++i) {
        if (KT == 0 && i >= K) break;
        const f3 s = ld3(sites + 3 * i);
        const f3 dv = q0 - s;
        const float dsq = dot(dv, dv) * inv_r2;
        const float w = pf_expf(-kTemp * dsq);
        if (!exact && w / fw.wsum0 < 1e-3f) continue;
        const f3 dw_ds = divs(((2.0f * kTemp) * w) * dv, r2);
        atomic_add(g_height + i, (w * dL_dheight) / fw.wsum0);
        const float hdp = dL_dheight * (float)heights[i];
        const f3 gs = divs(dw_ds * (hdp - hdp_in), fw.wsum0);
        atomic_add(g_sites + 3 * i + 0, gs.x);
        atomic_add(g_sites + 3 * i + 1, gs.y);
        atomic_add(g_sites + 3 * i + 2, gs.z);
        dL_dq0 = dL_dq0 + (-1.0f * gs);
        if (exact) dL_dr += ((hdp - hdp_in) / fw.wsum0) * ((((2.0f * kTemp) * w) * dot(dv, dv)) / (r2 * r));
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

struct WalkParams {
    pf_scene s;
    const float *cams;  // (B, 12) pf_camera
    int B, W, H, tiles_h, tiles_per_cam;
    int64_t total_tiles;
    float thr, depth_q;
    int max_steps;
    pf_walk_out out;
    // backward only
    const float *ck_color, *ck_log_t, *g_color, *g_log_t;
    pf_grads grads;
    int *status;
    bool exact_texel;
    uint64_t *alg_bytes;  // backward per-ray byte counter (may be null)
};

struct TileLane {
    int b, pix_i, pix_j;
    bool valid;
    int64_t ray;
};

__device__ __forceinline__ TileLane tile_lane(const WalkParams &p, int waves_per_block) {
    TileLane tl;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * waves_per_block + wave;
    tl.valid = t < p.total_tiles;
    const int64_t tt = tl.valid ? t : 0;
    tl.b = (int)(tt / p.tiles_per_cam);
    const int tile = (int)(tt % p.tiles_per_cam);
    const int tile_i = tile % p.tiles_h;
    const int tile_j = tile / p.tiles_h;
    tl.pix_i = tile_i * 8 + (lane & 7);
    tl.pix_j = tile_j * 8 + (lane >> 3);
    tl.valid = tl.valid && tl.pix_i < p.H && tl.pix_j < p.W;
    tl.ray = ((int64_t)tl.b * p.H + tl.pix_i) * p.W + tl.pix_j;
    return tl;
}

// Algorithmic bytes of the walk (what a visit, a face or a shaded segment
// must read at least once): sphere 16 B + CSR row bounds 8 B + [n, sigma];
// per face the neighbour index and the face; per shaded segment K texels
// of (site, rgb, height).
template <typename A, typename F> struct WalkBytes {
    static constexpr uint64_t visit = 16 + 8 + 4 * sizeof(A);
    static constexpr uint64_t face = 4 + 4 * sizeof(F);
    static constexpr uint64_t texel = 7 * sizeof(A);
};

template <typename A, typename F, int KT>
__global__ __launch_bounds__(256) void walk_fwd_kernel(WalkParams p) {
    constexpr uint64_t kVisitBytes = WalkBytes<A, F>::visit, kFaceBytes = WalkBytes<A, F>::face,
                       kTexelBytes = WalkBytes<A, F>::texel;
    const TileLane tl = tile_lane(p, 4);
    if (!tl.valid) return;
    const Cam cam = load_cam(p.cams + 12 * tl.b);
    const f3 o = cam.eye;
    const f3 d = ray_dir(cam, tl.pix_i, tl.pix_j);

    const float4 *spheres = reinterpret_cast<const float4 *>(p.s.spheres);
    const A *nsig = reinterpret_cast<const A *>(p.s.nsigmas);
    const A *tsites = reinterpret_cast<const A *>(p.s.texel_sites);
    const A *trgb = reinterpret_cast<const A *>(p.s.texel_rgb) + (int64_t)tl.b * p.s.texel_rgb_cam_stride;
    const A *theight = reinterpret_cast<const A *>(p.s.texel_height);
    const int K = p.s.num_texel_sites;
    const int N = (int)p.s.num_cells;

    f3 rgb = mk(0, 0, 0);
    float log_t = 0.0f;
    float depth = 0.0f;
    bool depth_found = false;
    int prim = cam.start;
    int visits = 0;
    uint64_t hash = kFnvOffset;
    uint64_t bytes = 0;
    int st = 0;

    while (true) {
        const float trans = pf_expf(log_t);
        if (trans < p.thr || prim == kNoCell) break;
        if ((unsigned)prim >= (unsigned)N) { st |= PF_STATUS_STEP_CAP; break; }
        if (visits >= p.max_steps) { st |= PF_STATUS_STEP_CAP; break; }
        if (p.out.seq && visits < p.out.seq_cap) p.out.seq[tl.ray * p.out.seq_cap + visits] = prim;
        ++visits;
        hash = (hash ^ (uint64_t)(uint32_t)prim) * kFnvPrime;

        const float4 sp = spheres[prim];
        const f3 c = mk(sp.x, sp.y, sp.z);
        const float r = sp.w;
        SphereHit sh = ray_sphere(o, d, c, r);
        const f3 v = c - o;
        if (length(v) < 4.0f * r) sh.hit = false;
        float tn = sh.tn, tf = sh.tf;

        const int e0 = p.s.adjacency_offsets[prim];
        const int e1 = p.s.adjacency_offsets[prim + 1];
        bytes += kVisitBytes + (uint64_t)(e1 - e0) * kFaceBytes;
        int next = kNoCell;
        float pt_far = 1e10f;
        for (int e = e0; e < e1; ++e) {
            const int nb = p.s.adjacency[e];
            f3 fn;
            float fw;
            FaceT<F>::load(p.s.faces, e, fn, fw);
            float t, dp;
            ray_pface_diff(o, d, fn, fw, t, dp);
            if (dp >= 0.0f && t < pt_far) { next = nb; pt_far = t; }
            tf = dp >= 0.0f ? fmin_(t, tf) : tf;
            tn = dp < 0.0f ? fmax_(t, tn) : tn;
        }
        f3 nrm;
        float sigma;
        ld4(nsig + 4 * (int64_t)prim, nrm, sigma);
        if (!sh.hit || tn > tf || sigma < 1e-3f) { prim = next; continue; }

        const int64_t tbase = (int64_t)prim * K;
        bytes += (uint64_t)K * kTexelBytes;
        const TexFwd tx = texel_fwd<A, KT>(o, d, tn, c, nrm, r, tsites + 3 * tbase, trgb + 3 * tbase,
                                          theight + tbase, K);
        tf = tx.dp >= 0.0f ? fmin_(tx.t_surf, tf) : tf;
        tn = tx.dp < 0.0f ? fmax_(tx.t_surf, tn) : tn;
        prim = next;
        const float dt = tf - tn;
        if (dt > 0.0f) {
            const float dlt = (-sigma) * dt;
            const float alpha = 1.0f - pf_expf(dlt);
            rgb.x = __builtin_fmaf(tx.rgb.x * alpha, trans, rgb.x);
            rgb.y = __builtin_fmaf(tx.rgb.y * alpha, trans, rgb.y);
            rgb.z = __builtin_fmaf(tx.rgb.z * alpha, trans, rgb.z);
            log_t += dlt;
            if (!depth_found) {
                const float next_trans = pf_expf(log_t);
                if (next_trans < p.depth_q) {
                    depth = tn + pf_logf(trans / p.depth_q) / sigma;
                    depth_found = true;
                }
            }
        }
    }
    float *co = p.out.color + 3 * tl.ray;
    co[0] = rgb.x;
    co[1] = rgb.y;
    co[2] = rgb.z;
    if (p.out.log_t) p.out.log_t[tl.ray] = log_t;
    if (p.out.depth) p.out.depth[tl.ray] = depth;
    if (p.out.visits) p.out.visits[tl.ray] = visits;
    if (p.out.seq_hash) p.out.seq_hash[tl.ray] = hash;
    if (p.out.alg_bytes) p.out.alg_bytes[tl.ray] = bytes + 12 + (p.out.log_t ? 4 : 0) + (p.out.depth ? 4 : 0);
    if (st && p.out.status) atomicOr(p.out.status, st);
}

// Backward by checkpoint and replay.  The forward's per-ray colour C and
// log transmittance are the checkpoint.  Each ray replays its walk in
// forward order with the forward's arithmetic (so the cell sequence and
// every segment are bit-identical), and at each shaded segment s uses
//   dL/d(delta_s) = g.(C - C_{<=s}) + g_logT - (g.c_s) T_s exp(delta_s)
// which equals the reverse-order accumulation of the rasterizer backward
// (rasterize.py:1371-1426).  The per-segment routing of dL/dt to the
// sphere, a power face, or the texel plane mirrors rasterize.py:1296-1516.
template <typename A, typename F, int KT>
__global__ __launch_bounds__(256) void walk_bwd_kernel(WalkParams p) {
    constexpr uint64_t kVisitBytes = WalkBytes<A, F>::visit, kFaceBytes = WalkBytes<A, F>::face,
                       kTexelBytes = WalkBytes<A, F>::texel;
    const TileLane tl = tile_lane(p, 4);
    if (!tl.valid) return;
    const Cam cam = load_cam(p.cams + 12 * tl.b);
    const f3 o = cam.eye;
    const f3 d = ray_dir(cam, tl.pix_i, tl.pix_j);

    const float4 *spheres = reinterpret_cast<const float4 *>(p.s.spheres);
    const A *nsig = reinterpret_cast<const A *>(p.s.nsigmas);
    const A *tsites = reinterpret_cast<const A *>(p.s.texel_sites);
    const A *trgb = reinterpret_cast<const A *>(p.s.texel_rgb) + (int64_t)tl.b * p.s.texel_rgb_cam_stride;
    const A *theight = reinterpret_cast<const A *>(p.s.texel_height);
    const int K = p.s.num_texel_sites;
    const int N = (int)p.s.num_cells;

    const f3 Cfin = mk(p.ck_color[3 * tl.ray], p.ck_color[3 * tl.ray + 1], p.ck_color[3 * tl.ray + 2]);
    const f3 g = p.g_color ? mk(p.g_color[3 * tl.ray], p.g_color[3 * tl.ray + 1], p.g_color[3 * tl.ray + 2])
                           : mk(0, 0, 0);
    const float gL = p.g_log_t ? p.g_log_t[tl.ray] : 0.0f;
    if (g.x == 0.0f && g.y == 0.0f && g.z == 0.0f && gL == 0.0f) {
        if (p.alg_bytes) p.alg_bytes[tl.ray] = 12 + (p.g_log_t ? 4 : 0);
        return;
    }

    f3 rgb = mk(0, 0, 0);
    float log_t = 0.0f;
    int prim = cam.start;
    int visits = 0;
    int st = 0;
    uint64_t bytes = 12 + 12 + (p.g_log_t ? 4 : 0);  // checkpoint colour + incoming gradients

    while (true) {
        const float trans = pf_expf(log_t);
        if (trans < p.thr || prim == kNoCell) break;
        if ((unsigned)prim >= (unsigned)N || visits >= p.max_steps) { st |= PF_STATUS_STEP_CAP; break; }
        ++visits;

        const float4 sp = spheres[prim];
        const f3 c = mk(sp.x, sp.y, sp.z);
        const float r = sp.w;
        SphereHit sh = ray_sphere(o, d, c, r);
        const f3 v = c - o;
        if (length(v) < 4.0f * r) sh.hit = false;
        float tn = sh.tn, tf = sh.tf;
        int tn_id = tn < 1e-8f ? -2 : -1;
        int tf_id = tf < 1e-8f ? -2 : -1;

        const int e0 = p.s.adjacency_offsets[prim];
        const int e1 = p.s.adjacency_offsets[prim + 1];
        bytes += kVisitBytes + (uint64_t)(e1 - e0) * kFaceBytes;
        const int n_adj = e1 - e0;
        int next = kNoCell;
        float pt_far = 1e10f;
        for (int e = e0; e < e1; ++e) {
            const int nb = p.s.adjacency[e];
            f3 fn;
            float fw;
            FaceT<F>::load(p.s.faces, e, fn, fw);
            float t, dp;
            ray_pface_diff(o, d, fn, fw, t, dp);
            if (dp >= 0.0f && t < pt_far) { next = nb; pt_far = t; }
            if (dp >= 0.0f && t < tf) tf_id = e - e0;
            tf = dp >= 0.0f ? fmin_(t, tf) : tf;
            if (dp < 0.0f && t > tn) tn_id = e - e0;
            tn = dp < 0.0f ? fmax_(t, tn) : tn;
        }
        f3 nrm;
        float sigma;
        ld4(nsig + 4 * (int64_t)prim, nrm, sigma);
        if (!sh.hit || tn > tf || sigma < 1e-3f) { prim = next; continue; }

        const int64_t tbase = (int64_t)prim * K;
        bytes += (uint64_t)K * kTexelBytes;
        const float tn0 = tn;
        const int tn0_id = tn_id, tf0_id = tf_id;
        const TexFwd tx = texel_fwd<A, KT>(o, d, tn, c, nrm, r, tsites + 3 * tbase, trgb + 3 * tbase,
                                          theight + tbase, K);
        if (tx.dp >= 0.0f && tx.t_surf < tf) tf_id = n_adj;
        tf = tx.dp >= 0.0f ? fmin_(tx.t_surf, tf) : tf;
        if (tx.dp < 0.0f && tx.t_surf > tn) tn_id = n_adj;
        tn = tx.dp < 0.0f ? fmax_(tx.t_surf, tn) : tn;
        const int cell = prim;
        prim = next;
        const float dt = tf - tn;
        if (!(dt > 0.0f)) continue;

        const float dlt = (-sigma) * dt;
        const float e_dlt = pf_expf(dlt);
        const float alpha = 1.0f - e_dlt;
        rgb.x = __builtin_fmaf(tx.rgb.x * alpha, trans, rgb.x);
        rgb.y = __builtin_fmaf(tx.rgb.y * alpha, trans, rgb.y);
        rgb.z = __builtin_fmaf(tx.rgb.z * alpha, trans, rgb.z);
        log_t += dlt;

        bytes += 2ull * (32ull + (uint64_t)K * 28ull);  // read-modify-write of the cell's gradients
        const float dL_dLnext = dot(g, Cfin - rgb) + gL;
        const float dL_dalpha = dot(g, tx.rgb) * trans;
        const f3 dL_dcolor = (g * trans) * alpha;
        const float dL_ddelta = dL_dLnext + (-dL_dalpha * e_dlt);
        const float dL_dsigma = dL_ddelta * (-dt);
        const float dL_dtn = dL_ddelta * sigma;
        const float dL_dtf = dL_ddelta * (-sigma);

        float dL_dts = 0.0f;
        if (tn_id == n_adj) dL_dts = dL_dtn;
        else if (tf_id == n_adj) dL_dts = dL_dtf;

        f3 dL_dc, dL_dn;
        float dL_dtn0, dL_dr_tex = 0.0f;
        texel_bwd<A, KT>(o, d, tn0, c, nrm, r, tsites + 3 * tbase, trgb + 3 * tbase, theight + tbase, K, tx,
                         dL_dcolor, dL_dts, p.grads.texel_sites + 3 * tbase,
                         p.grads.texel_rgb + (int64_t)tl.b * p.grads.texel_rgb_cam_stride + 3 * tbase,
                         p.grads.texel_height + tbase, dL_dc, dL_dn, dL_dtn0, p.exact_texel, dL_dr_tex);
        if (tn_id != n_adj) dL_dtn0 += dL_dtn;
        float dL_dr = dL_dr_tex;
        for (int side = 0; side < 2; ++side) {
            const int tid = side == 0 ? tn0_id : tf0_id;
            const float dLdt = side == 0 ? dL_dtn0 : (tf_id != n_adj ? dL_dtf : 0.0f);
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
                const int adjc = p.s.adjacency[e0 + tid];
                const float4 asp = spheres[adjc];
                const FaceGrad fg = ray_pface_bwd(o, d, c, r, mk(asp.x, asp.y, asp.z), asp.w);
                dL_dc = dL_dc + fg.dc * dLdt;
                dL_dr += dLdt * fg.dr;
                float *ga = p.grads.spheres + 4 * (int64_t)adjc;
                atomic_add(ga + 0, dLdt * fg.dcj.x);
                atomic_add(ga + 1, dLdt * fg.dcj.y);
                atomic_add(ga + 2, dLdt * fg.dcj.z);
                atomic_add(ga + 3, dLdt * fg.drj);
                bytes += 32;  // neighbour sphere gradient read-modify-write
            }
        }
        float *gs = p.grads.spheres + 4 * (int64_t)cell;
        atomic_add(gs + 0, dL_dc.x);
        atomic_add(gs + 1, dL_dc.y);
        atomic_add(gs + 2, dL_dc.z);
        atomic_add(gs + 3, dL_dr);
        float *gn = p.grads.nsigmas + 4 * (int64_t)cell;
        atomic_add(gn + 0, dL_dn.x);
        atomic_add(gn + 1, dL_dn.y);
        atomic_add(gn + 2, dL_dn.z);
        atomic_add(gn + 3, dL_dsigma);
    }
    if (p.alg_bytes) p.alg_bytes[tl.ray] = bytes;
    if (st && p.status) atomicOr(p.status, st);
}

// Walk faces: benchmark.py:295-302, one thread per cell over its CSR row.
template <typename F>
__global__ void walk_faces_kernel(const float *points, const float *radii, int64_t N, const int32_t *adj,
                                  const int32_t *off, F *faces) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const f3 pi = mk(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
    const float ri = radii[i];
    const float ni = __builtin_sqrtf(dot(pi, pi));
    const float pmi = 0.5f * (ni * ni - ri * ri);
    for (int e = off[i]; e < off[i + 1]; ++e) {
        const int j = adj[e];
        const f3 pj = mk(points[3 * (int64_t)j], points[3 * (int64_t)j + 1], points[3 * (int64_t)j + 2]);
        const float rj = radii[j];
        const float nj = __builtin_sqrtf(dot(pj, pj));
        const float pmj = 0.5f * (nj * nj - rj * rj);
        faces[4 * (int64_t)e + 0] = (F)(pj.x - pi.x);
        faces[4 * (int64_t)e + 1] = (F)(pj.y - pi.y);
        faces[4 * (int64_t)e + 2] = (F)(pj.z - pi.z);
        faces[4 * (int64_t)e + 3] = (F)(pmj - pmi);
    }
}

// Start cell: argmin |p - eye|^2 - r^2, first index on ties
// (benchmark.py:320-324).  Keys are packed as (orderable f32 bits, index)
// in a u64 so that one atomicMin per block gives the first minimum.
__device__ __forceinline__ uint32_t order_f32(float x) {
    const uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void start_init_kernel(unsigned long long *ws, int B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) ws[b] = ~0ull;
}

__global__ __launch_bounds__(256) void start_argmin_kernel(const float *points, const float *radii, int64_t N,
                                                           const float *cams, int B, unsigned long long *ws) {
    const int b = blockIdx.y;
    const f3 eye = mk(cams[12 * b], cams[12 * b + 1], cams[12 * b + 2]);
    unsigned long long best = ~0ull;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const f3 v = mk(points[3 * i], points[3 * i + 1], points[3 * i + 2]) - eye;
        const float dist = __builtin_sqrtf(dot(v, v));
        const float r = radii[i];
        float key = dist * dist - r * r;
        if (key != key) key = __builtin_inff();
        const unsigned long long k = ((unsigned long long)order_f32(key) << 32) | (unsigned long long)(uint32_t)i;
        best = k < best ? k : best;
    }
    // wave then block reduction
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_down(best, off, 64);
        best = o < best ? o : best;
    }
    __shared__ unsigned long long red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) best = red[w] < best ? red[w] : best;
        atomicMin(ws + b, best);
    }
}

__global__ void start_write_kernel(const unsigned long long *ws, float *cams, int B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) reinterpret_cast<int *>(cams)[12 * b + 9] = (int)(uint32_t)(ws[b] & 0xffffffffull);
}

}  // namespace pf

using namespace pf;

template <typename A, typename F>
static void launch_walk(const WalkParams &p, bool bwd, int K, hipStream_t st) {
    const int64_t blocks = (p.total_tiles + 3) / 4;
    dim3 grid((unsigned)blocks), block(256);
    if (K == 8) {
        if (bwd) hipLaunchKernelGGL((walk_bwd_kernel<A, F, 8>), grid, block, 0, st, p);
        else hipLaunchKernelGGL((walk_fwd_kernel<A, F, 8>), grid, block, 0, st, p);
    } else {
        if (bwd) hipLaunchKernelGGL((walk_bwd_kernel<A, F, 0>), grid, block, 0, st, p);
        else hipLaunchKernelGGL((walk_fwd_kernel<A, F, 0>), grid, block, 0, st, p);
    }
}

static int dispatch_walk(const WalkParams &p, bool bwd, hipStream_t st) {
    const int K = p.s.num_texel_sites;
    if (p.s.attr_dtype == PF_F32 && p.s.face_dtype == PF_F16) launch_walk<float, _Float16>(p, bwd, K, st);
    else if (p.s.attr_dtype == PF_F16 && p.s.face_dtype == PF_F16) launch_walk<_Float16, _Float16>(p, bwd, K, st);
    else if (p.s.attr_dtype == PF_F32 && p.s.face_dtype == PF_F32) launch_walk<float, float>(p, bwd, K, st);
    else if (p.s.attr_dtype == PF_F16 && p.s.face_dtype == PF_F32) launch_walk<_Float16, float>(p, bwd, K, st);
    else return PF_ERR_DTYPE;
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}

static int check_scene(const pf_scene *s) {
    if (!s || !s->spheres || !s->nsigmas || !s->texel_sites || !s->texel_rgb || !s->texel_height || !s->adjacency_offsets)
        return PF_ERR_ARG;
    if (s->num_edges > 0 && (!s->adjacency || !s->faces)) return PF_ERR_ARG;
    if (s->num_cells <= 0 || s->num_cells >= 0x7FFFFFFF) return PF_ERR_ARG;
    if (s->num_texel_sites < 1 || s->num_texel_sites > 16) return PF_ERR_ARG;
    if ((s->attr_dtype != PF_F32 && s->attr_dtype != PF_F16) || (s->face_dtype != PF_F32 && s->face_dtype != PF_F16))
        return PF_ERR_DTYPE;
    return PF_OK;
}

static WalkParams make_params(const pf_scene *scene, const pf_camera *cams, int32_t B, int32_t W, int32_t H,
                              float thr, int32_t max_steps) {
    WalkParams p{};
    p.s = *scene;
    p.cams = reinterpret_cast<const float *>(cams);
    p.B = B;
    p.W = W;
    p.H = H;
    p.tiles_h = (H + 7) / 8;
    p.tiles_per_cam = p.tiles_h * ((W + 7) / 8);
    p.total_tiles = (int64_t)p.tiles_per_cam * B;
    p.thr = thr;
    p.max_steps = max_steps;
    return p;
}

extern "C" int pf_walk_forward(const pf_scene *scene, const pf_camera *cams, int32_t num_cams, int32_t width,
                               int32_t height, float transmittance_threshold, float depth_quantile, int32_t max_steps,
                               const pf_walk_out *out, void *stream) {
    int rc = check_scene(scene);
    if (rc) return rc;
    if (!cams || num_cams <= 0 || width <= 0 || height <= 0 || !out || !out->color || max_steps <= 0) return PF_ERR_ARG;
    if (out->seq && out->seq_cap <= 0) return PF_ERR_ARG;
    WalkParams p = make_params(scene, cams, num_cams, width, height, transmittance_threshold, max_steps);
    p.depth_q = depth_quantile;
    p.out = *out;
    return dispatch_walk(p, false, (hipStream_t)stream);
}

extern "C" int pf_walk_backward(const pf_scene *scene, const pf_camera *cams, int32_t num_cams, int32_t width,
                                int32_t height, float transmittance_threshold, int32_t max_steps, const float *color,
                                const float *log_t, const float *grad_color, const float *grad_log_t,
                                const pf_grads *grads, int32_t flags, int32_t *status, uint64_t *alg_bytes,
                                void *stream) {
    int rc = check_scene(scene);
    if (rc) return rc;
    if (!cams || num_cams <= 0 || width <= 0 || height <= 0 || !color || !grads || max_steps <= 0) return PF_ERR_ARG;
    if (!grads->spheres || !grads->nsigmas || !grads->texel_sites || !grads->texel_rgb || !grads->texel_height)
        return PF_ERR_ARG;
    (void)log_t;
    WalkParams p = make_params(scene, cams, num_cams, width, height, transmittance_threshold, max_steps);
    p.ck_color = color;
    p.ck_log_t = log_t;
    p.g_color = grad_color;
    p.g_log_t = grad_log_t;
    p.grads = *grads;
    p.status = status;
    p.exact_texel = (flags & PF_BWD_EXACT_TEXEL) != 0;
    p.alg_bytes = alg_bytes;
    return dispatch_walk(p, true, (hipStream_t)stream);
}

extern "C" int pf_walk_faces(const float *points, const float *radii, int64_t num_cells, const int32_t *adjacency,
                             const int32_t *adjacency_offsets, int64_t num_edges, void *faces, int32_t face_dtype,
                             void *stream) {
    if (!points || !radii || num_cells <= 0 || !adjacency_offsets || (num_edges > 0 && (!adjacency || !faces)))
        return PF_ERR_ARG;
    const int blocks = (int)((num_cells + 255) / 256);
    hipStream_t st = (hipStream_t)stream;
    if (face_dtype == PF_F16)
        hipLaunchKernelGGL((walk_faces_kernel<_Float16>), dim3(blocks), dim3(256), 0, st, points, radii, num_cells,
                           adjacency, adjacency_offsets, (_Float16 *)faces);
    else if (face_dtype == PF_F32)
        hipLaunchKernelGGL((walk_faces_kernel<float>), dim3(blocks), dim3(256), 0, st, points, radii, num_cells,
                           adjacency, adjacency_offsets, (float *)faces);
    else return PF_ERR_DTYPE;
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}

extern "C" int pf_start_cells(const float *points, const float *radii, int64_t num_cells, pf_camera *cams,
                              int32_t num_cams, void *workspace, void *stream) {
    if (!points || !radii || num_cells <= 0 || !cams || num_cams <= 0 || !workspace) return PF_ERR_ARG;
    unsigned long long *ws = (unsigned long long *)workspace;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(start_init_kernel, dim3((num_cams + 255) / 256), dim3(256), 0, st, ws, num_cams);
    int blocks = (int)((num_cells + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(start_argmin_kernel, dim3(blocks, num_cams), dim3(256), 0, st, points, radii, num_cells,
                       (const float *)cams, num_cams, ws);
    hipLaunchKernelGGL(start_write_kernel, dim3((num_cams + 255) / 256), dim3(256), 0, st, ws, (float *)cams,
                       num_cams);
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}
