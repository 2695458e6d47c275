// raster.hip — pinhole tile rasterizer over power cells for gfx950.
//
// Binning follows count_visible_pinhole_kernel / write_visible_pinhole_kernel
// (powerfoam/rasterize.py:140-200, :293-363) with the sphere->OBB projection
// and tile SAT of powerfoam/camera.py:79-178; the per-tile lists are ordered
// by power distance from the eye (rasterize.py:360-363, 1771) with the cell
// index as a deterministic tie-break.  Compositing follows forward_kernel
// (rasterize.py:892-1146), backward_kernel (:1150-1516), benchmark_kernel
// (:614-734) and visualization_kernel (:738-888): one 8x8 tile per
// 64-lane workgroup (one wavefront), so the reference's block-wide
// tile_max / tile_sum are wavefront ballots here, and the cooperative loads
// of 64 cells / 64 faces go through LDS.
#include "pf_device.h"
#include "pf_hip.h"
#include "pf_launch.h"
#include "pf_walk.h"

namespace pf {

struct alignas(8) rh4 { _Float16 x, y, z, w; };

template <typename F> struct RFace;
template <> struct RFace<_Float16> {
    using T = rh4;
    __device__ static __forceinline__ void get(const T &v, f3 &a, float &b) {
        a = mk((float)v.x, (float)v.y, (float)v.z);
        b = (float)v.w;
    }
};
template <> struct RFace<float> {
    using T = float4;
    __device__ static __forceinline__ void get(const T &v, f3 &a, float &b) {
        a = mk(v.x, v.y, v.z);
        b = v.w;
    }
};

// ------------------------------------------------------------- binning --
// Float -> int as the GPU converts: NaN -> 0, saturating.
__device__ __forceinline__ int f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483647.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int)x;
}

struct Obb { float uc, vc, r1, r2, vx, vy; };

// camera.py:79-128
__device__ __forceinline__ Obb proj_sphere_to_obb(const Cam &cam, f3 center, float radius) {
    const f3 C = center - cam.eye;
    const float C2 = dot(C, C);
    const float R2 = radius * radius;
    const f3 X = cam.right, Y = cam.up;
    const f3 zc = cross(Y, X);
    const f3 Z = divs(zc, length(zc));
    const float cx = dot(C, X), cy = dot(C, Y), cz = dot(C, Z);
    const float a = cx * cx - (C2 - R2) * dot(X, X);
    const float b = cy * cy - (C2 - R2) * dot(Y, Y);
    const float f = cz * cz - (C2 - R2) * dot(Z, Z);
    const float c = 2.0f * cx * cy;
    const float d = 2.0f * cx * cz;
    const float e = 2.0f * cy * cz;
    const float denom = 4.0f * a * b - c * c;
    if (__builtin_fabsf(denom) < 1e-8f) return Obb{0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f};
    const float uc = (c * e - 2.0f * b * d) / denom;
    const float vc = (c * d - 2.0f * a * e) / denom;
    const float F0 = a * uc * uc + b * vc * vc + c * uc * vc + d * uc + e * vc + f;
    const float term = __builtin_sqrtf((a - b) * (a - b) + c * c);
    const float l1 = (a + b + term) / 2.0f;
    const float l2 = (a + b - term) / 2.0f;
    const float r1 = __builtin_sqrtf(-F0 / l1);
    const float r2 = __builtin_sqrtf(-F0 / l2);
    float vx = 0.5f * c, vy = l1 - a;
    const float nrm = __builtin_sqrtf(vx * vx + vy * vy);
    if (nrm < 1e-6f) { vx = 1.0f; vy = 0.0f; }
    else { vx = vx / nrm; vy = vy / nrm; }
    return Obb{uc, vc, r1, r2, vx, vy};
}

// camera.py:131-178
__device__ __forceinline__ bool tile_obb(int ti, int tj, const Cam &cam, const Obb &o) {
    const float u_min = 2.0f * (float)(tj * 8) / (float)(cam.W - 1) - 1.0f;
    const float u_max = 2.0f * (float)((tj + 1) * 8) / (float)(cam.W - 1) - 1.0f;
    const float v_max = 1.0f - 2.0f * (float)(ti * 8) / (float)(cam.H - 1);
    const float v_min = 1.0f - 2.0f * (float)((ti + 1) * 8) / (float)(cam.H - 1);
    const float ucen = (u_min + u_max) * 0.5f, vcen = (v_min + v_max) * 0.5f;
    const float eu = (u_max - u_min) * 0.5f, ev = (v_max - v_min) * 0.5f;
    const float du = ucen - o.uc, dv = vcen - o.vc;
    const float p1 = du * o.vx + dv * o.vy;
    const float t1 = eu * __builtin_fabsf(o.vx) + ev * __builtin_fabsf(o.vy);
    if (__builtin_fabsf(p1) > o.r1 + t1) return false;
    const float p2 = du * -o.vy + dv * o.vx;
    const float t2 = eu * __builtin_fabsf(-o.vy) + ev * __builtin_fabsf(o.vx);
    if (__builtin_fabsf(p2) > o.r2 + t2) return false;
    return true;
}

struct TileRange { int ti0, ti1, tj0, tj1; bool visible; Obb o; float pow_dist; };

// rasterize.py:153-190: cull, OBB, pixel then tile bounds
__device__ __forceinline__ TileRange tile_range(const Cam &cam, float4 sp) {
    TileRange tr;
    const f3 c = mk(sp.x, sp.y, sp.z);
    const float r = sp.w;
    const f3 v = c - cam.eye;
    const f3 fwd = cross(cam.up, cam.right);
    tr.visible = !(length(v) < 4.0f * r || dot(v, fwd) < 0.1f);
    tr.pow_dist = dot(v, v) - r * r;
    if (!tr.visible) return tr;
    tr.o = proj_sphere_to_obb(cam, c, r);
    const Obb &o = tr.o;
    const float ext_u = __builtin_fabsf(o.vx * o.r1) + __builtin_fabsf(o.vy * o.r2);
    const float ext_v = __builtin_fabsf(o.vy * o.r1) + __builtin_fabsf(o.vx * o.r2);
    const float umin = o.uc - ext_u, umax = o.uc + ext_u, vmin = o.vc - ext_v, vmax = o.vc + ext_v;
    const float jmin_f = (umin + 1.0f) * 0.5f * (float)(cam.W - 1);
    const float jmax_f = (umax + 1.0f) * 0.5f * (float)(cam.W - 1);
    const float imin_f = (1.0f - vmax) * 0.5f * (float)(cam.H - 1);
    const float imax_f = (1.0f - vmin) * 0.5f * (float)(cam.H - 1);
    int imin = f2i(__builtin_floorf(imin_f)); imin = imin > 0 ? imin : 0;
    int imax = f2i(__builtin_ceilf(imax_f)); imax = imax < cam.H - 1 ? imax : cam.H - 1;
    int jmin = f2i(__builtin_floorf(jmin_f)); jmin = jmin > 0 ? jmin : 0;
    int jmax = f2i(__builtin_ceilf(jmax_f)); jmax = jmax < cam.W - 1 ? jmax : cam.W - 1;
    tr.ti0 = imin / 8; tr.ti1 = imax / 8; tr.tj0 = jmin / 8; tr.tj1 = jmax / 8;
    return tr;
}

__global__ __launch_bounds__(256) void bin_count_kernel(const float4 *spheres, int64_t N, const float *camw,
                                                        int *counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const Cam cam = load_cam(camw);
    const TileRange tr = tile_range(cam, spheres[i]);
    if (!tr.visible) return;
    const int tiles_h = (cam.H + 7) / 8;
    for (int ti = tr.ti0; ti <= tr.ti1; ++ti)
        for (int tj = tr.tj0; tj <= tr.tj1; ++tj)
            if (tile_obb(ti, tj, cam, tr.o)) atomicAdd(counts + ti + tj * tiles_h + 1, 1);
}

__global__ __launch_bounds__(256) void bin_write_kernel(const float4 *spheres, int64_t N, const float *camw,
                                                        const int64_t *offsets, int *write_counter,
                                                        unsigned long long *keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const Cam cam = load_cam(camw);
    const TileRange tr = tile_range(cam, spheres[i]);
    if (!tr.visible) return;
    const int tiles_h = (cam.H + 7) / 8;
    const unsigned long long key =
        ((unsigned long long)__float_as_uint(tr.pow_dist) << 32) | (unsigned long long)(uint32_t)i;
    for (int ti = tr.ti0; ti <= tr.ti1; ++ti)
        for (int tj = tr.tj0; tj <= tr.tj1; ++tj) {
            if (!tile_obb(ti, tj, cam, tr.o)) continue;
            const int t = ti + tj * tiles_h;
            const int slot = atomicAdd(write_counter + t, 1);
            const int64_t idx = offsets[t] + slot;
            if (idx < offsets[t + 1]) keys[idx] = key;
        }
}

// Inclusive scan of n int32 counts into int64 offsets, one workgroup of
// 1024 threads (tile counts: tens of thousands of entries).
// cap >= 0: offsets are clamped to cap (a caller-sized list buffer); a
// total above cap sets PF_STATUS_LIST_CAP in *status, and every later
// kernel reads only the clamped ranges.
__global__ __launch_bounds__(1024) void scan_kernel(const int *in, int64_t n, int64_t *out, int64_t cap,
                                                    int *status) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t b = t * per, e = b + per < n ? b + per : n;
    int64_t s = 0;
    for (int64_t i = b; i < e; ++i) s += in[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t > 0 ? part[t - 1] : 0;
    for (int64_t i = b; i < e; ++i) {
        run += in[i];
        out[i] = cap >= 0 && run > cap ? cap : run;
    }
    if (cap >= 0 && t == 1023 && part[1023] > cap && status) atomicOr(status, PF_STATUS_LIST_CAP);
}

// Per-tile sort of the 64-bit keys (power-distance bits, cell index):
// bitonic sort in LDS for lists up to kSortLds, otherwise LDS-sorted runs
// merged by rank through the scratch buffer.  Emits the cell indices.
constexpr int kSortLds = 4096;

__device__ __forceinline__ void bitonic_lds(unsigned long long *s, int n2) {
    for (int k = 2; k <= n2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const unsigned long long a = s[i], b = s[l];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) { s[i] = b; s[l] = a; }
                }
            }
            __syncthreads();
        }
}

__device__ __forceinline__ int64_t count_less(const unsigned long long *a, int64_t n, unsigned long long k) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t m = (lo + hi) >> 1;
        if (a[m] < k) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__global__ __launch_bounds__(256) void tile_sort_kernel(const int64_t *offsets, int ntiles,
                                                        unsigned long long *keys, unsigned long long *scratch,
                                                        int *prims) {
    __shared__ unsigned long long s[kSortLds];
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const int64_t b = offsets[t], n = offsets[t + 1] - b;
    if (n <= 0) return;
    unsigned long long *k = keys + b;
    for (int64_t c0 = 0; c0 < n; c0 += kSortLds) {
        const int m = (int)(n - c0 < kSortLds ? n - c0 : kSortLds);
        int n2 = 1;
        while (n2 < m) n2 <<= 1;
        for (int i = threadIdx.x; i < n2; i += blockDim.x) s[i] = i < m ? k[c0 + i] : ~0ull;
        __syncthreads();
        bitonic_lds(s, n2);
        for (int i = threadIdx.x; i < m; i += blockDim.x) k[c0 + i] = s[i];
        __syncthreads();
    }
    unsigned long long *src = k, *dst = scratch + b;
    for (int64_t w = kSortLds; w < n; w <<= 1) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            const int64_t na = (n - lo < w ? n - lo : w);
            const int64_t nb = (n - lo - na < w ? n - lo - na : w);
            const unsigned long long *A = src + lo, *B = src + lo + na;
            for (int64_t i = threadIdx.x; i < na + nb; i += blockDim.x) {
                if (i < na) dst[lo + i + count_less(B, nb, A[i])] = A[i];
                else dst[lo + (i - na) + count_less(A, na, B[i - na])] = B[i - na];
            }
        }
        __syncthreads();
        unsigned long long *tmp = src; src = dst; dst = tmp;
    }
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) prims[b + i] = (int)(uint32_t)(src[i] & 0xffffffffull);
}

// Marks listed cells; reads the list length from the (possibly clamped)
// offsets on the device, so n may be the buffer capacity.
__global__ void visible_mask_kernel(const int *prims, int64_t n, const int64_t *offsets, int ntiles,
                                    unsigned char *mask) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && i < offsets[ntiles]) mask[prims[i]] = 1;
}

// Raster faces: rasterize.py:580-606 — fp16 (p_j - p_i, r_j) per CSR edge.
template <typename F>
__global__ void raster_faces_kernel(const float4 *spheres, int64_t N, const int *adj, const int *off, F *faces) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float4 s = spheres[i];
    for (int e = off[i]; e < off[i + 1]; ++e) {
        const float4 a = spheres[adj[e]];
        faces[4 * (int64_t)e + 0] = (F)(a.x - s.x);
        faces[4 * (int64_t)e + 1] = (F)(a.y - s.y);
        faces[4 * (int64_t)e + 2] = (F)(a.z - s.z);
        faces[4 * (int64_t)e + 3] = (F)(a.w);
    }
}

// ----------------------------------------------------------- texel fwd --
// w0 / w1: the texel weights at the plane entry q0 and at the clipped entry
// q, kept for the backward (at most 8 texels on the raster path).
struct RTex {
    float t0, dp0, height, wsum0, t_surf, dp, wsum;
    f3 rgb;
    float w0[8], w1[8];
};

// powerfoam/texture.py:10-71 (raster variant: weights divide by r*r; if the
// ray misses, everything is zero).  tex: [sites 3K | rgb 3K | height K].
// KEEP: also keep the weights (K <= 8), for the backward.
template <bool KEEP>
__device__ __forceinline__ RTex texel_fwd_raster(f3 o, f3 d, float t_near, f3 c, f3 n, float r, bool hit,
                                                 const float *tex, int K) {
    RTex out{0, 0, 0, 0, 0, 0, 0, mk(0, 0, 0), {}, {}};
    if (!hit) return out;
    ray_plane(o, d, c, n, 0.0f, out.t0, out.dp0);
    const float tq0 = out.dp0 >= 0.0f ? t_near : fmax_(t_near, out.t0);
    const f3 q0 = ray_at(o, tq0, d);
    const float r2 = r * r;
    // exp(-kTemp |q - s|^2 / r^2) as the hardware base-2 exponential of one
    // product per texel (the scale folded once per cell), where a correctly
    // rounded division and a software exp took about thirty VALU
    // instructions; weights agree to a few ulp and the rasterizer carries a
    // tolerance, not bit-exactness (the walk keeps the exact form)
    const float ke = (-kTemp * 1.44269504088896341f) / r2;
    float hs = 0.0f, ws = 0.0f;
    auto height_term = [&](int i) {
        const f3 dv = q0 - mk(tex[3 * i], tex[3 * i + 1], tex[3 * i + 2]);
        const float w = __builtin_amdgcn_exp2f(ke * dot(dv, dv));
        hs = __builtin_fmaf(w, tex[6 * K + i], hs);
        ws = ws + w;
        return w;
    };
    if constexpr (KEEP) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (i >= K) break;
            out.w0[i] = height_term(i);
        }
    } else {
        for (int i = 0; i < K; ++i) height_term(i);
    }
    ws = fmax_(ws, 1e-20f);
    out.wsum0 = ws;
    out.height = hs / ws;
    ray_plane(o, d, c, n, out.height, out.t_surf, out.dp);
    const float tq = out.dp >= 0.0f ? t_near : fmax_(t_near, out.t_surf);
    const f3 q = ray_at(o, tq, d);
    f3 acc = mk(0, 0, 0);
    float ws2 = 0.0f;
    auto colour_term = [&](int i) {
        const f3 dv = q - mk(tex[3 * i], tex[3 * i + 1], tex[3 * i + 2]);
        const float w = __builtin_amdgcn_exp2f(ke * dot(dv, dv));
        acc = macc(acc, w, mk(tex[3 * K + 3 * i], tex[3 * K + 3 * i + 1], tex[3 * K + 3 * i + 2]));
        ws2 = ws2 + w;
        return w;
    };
    if constexpr (KEEP) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (i >= K) break;
            out.w1[i] = colour_term(i);
        }
    } else {
        for (int i = 0; i < K; ++i) colour_term(i);
    }
    ws2 = fmax_(ws2, 1e-20f);
    out.wsum = ws2;
    out.rgb = divs(acc, ws2);
    return out;
}

// ---------------------------------------------------------- parameters --
struct RasterParams {
    const float4 *spheres;
    const float4 *nsigmas;  // f32 [n, sigma] (forward/backward take f32, rasterize.py:1620-1621)
    const float *tsites, *trgb, *theight;
    const void *nsig_any;   // benchmark/visualize: attribute dtype
    int attr_f16;
    const int *adj, *off;
    const void *faces;
    int K;
    int64_t N;
    const float *camw;
    int W, H, tiles_h, ntiles;
    const int64_t *offsets;
    const int *prims;
    int *early_stop;
    float thr;
    // forward extras
    const float *depth_q;  // (H,W,Q) or null
    int Q;
    const float *ray_gt;   // (H,W,3) or null
    int point_err, surface;
    float *color, *log_t, *ndist, *normal, *qdepth, *err, *contrib, *perr;
    // backward
    const float *g_color, *g_log_t, *g_ndist, *g_normal, *g_qdepth, *g_contrib;
    float *gs_sph, *gs_ns, *gs_sites, *gs_rgb, *gs_h;
    // visualize
    int max_isect;
    float vis_depth_q;
    float bkgd[3];
    float *depth_out, *alpha_out;
    int *isect_out;
};

struct TileCtx {
    int tile, lane, pix_i, pix_j;
    bool oob;
    f3 o, d;
    Cam cam;
};

__device__ __forceinline__ TileCtx tile_ctx(const RasterParams &p) {
    TileCtx t;
    t.tile = xcd_block(blockIdx.x, gridDim.x);
    t.lane = threadIdx.x;
    const int ti = t.tile % p.tiles_h, tj = t.tile / p.tiles_h;
    int pi = ti * 8 + (t.lane & 7), pj = tj * 8 + (t.lane >> 3);
    t.oob = pi >= p.H || pj >= p.W;
    t.pix_i = pi < p.H - 1 ? pi : p.H - 1;
    t.pix_j = pj < p.W - 1 ? pj : p.W - 1;
    t.cam = load_cam(p.camw);
    t.o = t.cam.eye;
    t.d = ray_dir(t.cam, t.pix_i, t.pix_j);
    return t;
}

// One chunk of up to 64 list entries staged in LDS (the reference's
// cooperative prim load, rasterize.py:1001-1026).
struct PrimChunk {
    float4 sph[64];
    float4 ns[64];
    int prim[64];
    int e0[64], e1[64];
};

__device__ __forceinline__ void stage_prims(PrimChunk &pc, const RasterParams &p, int64_t first, int64_t lo,
                                            int64_t hi, bool reverse) {
    __syncthreads();
    const int64_t k = reverse ? first - threadIdx.x : first + threadIdx.x;
    if (k >= lo && k < hi) {
        const int prim = p.prims[k];
        pc.prim[threadIdx.x] = prim;
        pc.sph[threadIdx.x] = p.spheres[prim];
        pc.ns[threadIdx.x] = p.nsigmas[prim];
        pc.e0[threadIdx.x] = p.off[prim];
        pc.e1[threadIdx.x] = p.off[prim + 1];
    }
    __syncthreads();
}

__device__ __forceinline__ void stage_texels(float *tex, const RasterParams &p, int prim) {
    __syncthreads();
    const int K = p.K, l = threadIdx.x;
    const int64_t b = (int64_t)prim * K;
    if (l < 3 * K) {
        tex[l] = p.tsites[3 * b + l];
        tex[3 * K + l] = p.trgb[3 * b + l];
    }
    if (l < K) tex[6 * K + l] = p.theight[b + l];
    __syncthreads();
}

// DPP reads within rows of 16 lanes (a source off in EXEC reads 0).
template <int CTRL, int ROWS = 0xF> __device__ __forceinline__ float dpp_row(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROWS, 0xF, false));
}

// Sum of x over the wavefront, in lane 63: row shifts by 1, 2, 4, 8, then
// row_bcast:15 and row_bcast:31 (six DPP adds, where a butterfly of lane
// shuffles takes six LDS-unit round trips).  Every lane must take part.
__device__ __forceinline__ float wave_sum_last(float x) {
    x += dpp_row<0x111>(x);
    x += dpp_row<0x112>(x);
    x += dpp_row<0x114>(x);
    x += dpp_row<0x118>(x);
    x += dpp_row<0x142, 0xA>(x);
    x += dpp_row<0x143, 0xC>(x);
    return x;
}

// The planes of faces [e, e + min(n, 64)) of a row, one face per lane:
// normal and off - o.n (pface_plane), which every ray of the tile shares.
template <typename F>
__device__ __forceinline__ void stage_planes_t(float4 *splane, const RasterParams &p, f3 o, f3 c, float r, int e,
                                               int n) {
    if ((int)threadIdx.x < n) {
        f3 dp3;
        float rj;
        RFace<F>::get(reinterpret_cast<const typename RFace<F>::T *>(p.faces)[e + threadIdx.x], dp3, rj);
        splane[threadIdx.x] = pface_plane(o, c, r, c + dp3, rj);
    } else if ((int)threadIdx.x < ((n + 3) & ~3)) {
        // neutral planes up to a multiple of four: a NaN crossing, which the
        // strict compares of every fold pass over (its zero normal's dot
        // product may be -0, so an infinite offset could read as -inf)
        splane[threadIdx.x] = make_float4(0.0f, 0.0f, 0.0f, __builtin_nanf(""));
    }
}

// The face planes splane[0, m) in order, read four at a time so that four
// LDS reads are in flight where one at a time each face waits out its own
// (the row is padded with neutral planes to a multiple of four).
template <typename Fn> __device__ __forceinline__ void for_planes(const float4 *splane, int m, Fn &&fn) {
    for (int j = 0; j < m; j += 4) {
        const float4 a0 = splane[j], a1 = splane[j + 1], a2 = splane[j + 2], a3 = splane[j + 3];
        fn(j, a0);
        fn(j + 1, a1);
        fn(j + 2, a2);
        fn(j + 3, a3);
    }
}

// The raster's face test: the plane's offset times the hardware reciprocal
// of the ray's dot product with its normal (v_rcp_f32, 1 ulp), two VALU
// instructions where a correctly rounded division takes about ten.  The
// rasterizer carries a tolerance, not bit-exactness; forward and backward
// both use this form, so they agree on every segment.  A zero dot product
// gives an infinite (or, over a zero offset, NaN) crossing, as division does.
__device__ __forceinline__ void pface_t_rcp(f3 d, float4 pl, float &t, float &dp) {
    dp = dot(d, mk(pl.x, pl.y, pl.z));
    t = pl.w * __builtin_amdgcn_rcpf(dp);
}

// ------------------------------------------------------------- forward --
template <typename F>
__global__ __launch_bounds__(64) void raster_fwd_kernel(RasterParams p) {
    __shared__ PrimChunk pc;
    __shared__ float4 splane[64];
    __shared__ float tex[7 * 16];
    const TileCtx tc = tile_ctx(p);
    if (tc.tile >= p.ntiles) return;
    const f3 o = tc.o, d = tc.d;
    const int64_t pix = (int64_t)tc.pix_i * p.W + tc.pix_j;
    const float hw = (float)(p.H * p.W);

    int qi = 0;
    float qv = 0.0f;
    if (p.depth_q) qv = p.depth_q[pix * p.Q];
    f3 gt = mk(0, 0, 0);
    if (p.ray_gt) gt = mk(p.ray_gt[3 * pix], p.ray_gt[3 * pix + 1], p.ray_gt[3 * pix + 2]);
    f3 rgb = mk(0, 0, 0), normal = mk(0, 0, 0);
    float log_t = 0.0f, ndist = 0.0f, err = 0.0f;
    bool early = false;
    const int64_t start = p.offsets[tc.tile], end = p.offsets[tc.tile + 1];
    const int total = (int)(end - start);

    for (int i = 0; i < total; ++i) {
        const float trans = pf_expf(log_t);
        if (__all(trans < p.thr)) {  // tile_max(T) < threshold (rasterize.py:985-991)
            early = true;
            p.early_stop[tc.tile] = i;
            break;
        }
        const int sub = i & 63;
        if (sub == 0) stage_prims(pc, p, start + i, start, end, false);
        const int prim = pc.prim[sub];
        const float4 sp = pc.sph[sub];
        const float4 ns = pc.ns[sub];
        const f3 c = mk(sp.x, sp.y, sp.z), nrm = mk(ns.x, ns.y, ns.z);
        const float r = sp.w, sigma = ns.w;
        SphereHit sh = ray_sphere(o, d, c, r);
        if (!__any(sh.hit)) continue;  // tile_sum(hit) == 0 (:1029-1032)
        float tn = sh.tn, tf = sh.tf;
        const int e0 = __builtin_amdgcn_readfirstlane(pc.e0[sub]);
        const int n_adj = __builtin_amdgcn_readfirstlane(pc.e1[sub]) - e0;
        for (int fb = 0; fb < n_adj; fb += 64) {
            __syncthreads();
            stage_planes_t<F>(splane, p, o, c, r, e0 + fb, n_adj - fb);
            __syncthreads();
            const int m = n_adj - fb < 64 ? n_adj - fb : 64;
            if (sh.hit) {
                for_planes(splane, m, [&](int, float4 pl) {
                    float t, dp;
                    pface_t_rcp(d, pl, t, dp);
                    tf = dp >= 0.0f ? fmin_(t, tf) : tf;
                    tn = dp < 0.0f ? fmax_(t, tn) : tn;
                });
            }
        }
        stage_texels(tex, p, prim);
        const RTex tx = texel_fwd_raster<false>(o, d, tn, c, nrm, r, !tc.oob && sh.hit, tex, p.K);
        tf = tx.dp >= 0.0f ? fmin_(tx.t_surf, tf) : tf;
        tn = tx.dp < 0.0f ? fmax_(tx.t_surf, tn) : tn;
        const float dt = tf - tn;
        float contrib = 0.0f, perr = 0.0f;
        if (sh.hit && dt > 0.0f) {
            const float dlt = (-sigma) * dt;
            const float alpha = 1.0f - pf_expf(dlt);
            const float tr = pf_expf(log_t);
            if (p.surface) {
                const f3 e = tx.rgb - gt;
                err = __builtin_fmaf(dot(e, e) * alpha, tr, err);
            }
            rgb.x = __builtin_fmaf(tx.rgb.x * alpha, tr, rgb.x);
            rgb.y = __builtin_fmaf(tx.rgb.y * alpha, tr, rgb.y);
            rgb.z = __builtin_fmaf(tx.rgb.z * alpha, tr, rgb.z);
            log_t += dlt;
            contrib = (alpha * tr) / hw;
            if (p.point_err) {
                const f3 e = tx.rgb - gt;
                perr = ((alpha * tr) * (__builtin_fabsf(e.x) + __builtin_fabsf(e.y) + __builtin_fabsf(e.z))) / hw;
            }
            const float ndv = dot(nrm, d);
            if (ndv > 0.0f) ndist = __builtin_fmaf((ndv * ndv) * alpha, tr, ndist);
            normal = macc(normal, alpha * tr, nrm);
            const float next_t = pf_expf(log_t);
            while (qi < p.Q && next_t < qv) {
                p.qdepth[pix * p.Q + qi] = tn + pf_logf(tr / qv) / sigma;
                ++qi;
                if (qi < p.Q) qv = p.depth_q[pix * p.Q + qi];
            }
        }
        // per-cell sums over the tile's lanes, one atomic per wavefront
        if (__any(contrib != 0.0f)) {
            const float s = wave_sum_last(contrib);
            if (threadIdx.x == 63) unsafeAtomicAdd(p.contrib + prim, s);
        }
        if (p.point_err && __any(perr != 0.0f)) {
            const float s = wave_sum_last(perr);
            if (threadIdx.x == 63) unsafeAtomicAdd(p.perr + prim, s);
        }
    }
    if (!tc.oob) {
        if (!early) p.early_stop[tc.tile] = total;
        p.color[3 * pix] = rgb.x;
        p.color[3 * pix + 1] = rgb.y;
        p.color[3 * pix + 2] = rgb.z;
        p.log_t[pix] = log_t;
        p.ndist[pix] = ndist;
        p.normal[3 * pix] = normal.x;
        p.normal[3 * pix + 1] = normal.y;
        p.normal[3 * pix + 2] = normal.z;
        if (p.surface) {
            const f3 e = mk(0, 0, 0) - gt;
            p.err[pix] = __builtin_fmaf(dot(e, e), pf_expf(log_t), err);
        }
    }
}

// ----------------------------------------------------------- texel bwd --
// powerfoam/texture.py:74-186 (raster weights).  Texel gradients go to the
// per-lane row red[0..7K) as [sites 3K | rgb 3K | height K].
template <int KT>
__device__ __forceinline__ void texel_bwd_raster(f3 o, f3 d, float tn_in, f3 c, f3 n, float r, bool hit,
                                                 const float *tex, int K, const RTex &fw, f3 g_rgb, float dts_in,
                                                 float *red, f3 &dc, f3 &dn, float &dtn) {
    if constexpr (KT > 0) K = KT;
    dc = mk(0, 0, 0);
    dn = mk(0, 0, 0);
    dtn = 0.0f;
    if (!hit) return;
    float dh = 0.0f;
    const float r2 = r * r;
    const float tq = fw.dp >= 0.0f ? tn_in : fmax_(tn_in, fw.t_surf);
    const f3 q = ray_at(o, tq, d);
    f3 dq = mk(0, 0, 0);
    const float gdp_in = dot(g_rgb, fw.rgb);
    // The forward's weights (no second exp), and reciprocals in place of the
    // per-texel divisions.  The skip of texels under 1e-3 of the weight sum
    // (texture.py:118, :158 divide) compares w with 1e-3 * sum, which differs
    // from the quotient's test only where the quotient rounds across 1e-3.
    const float k2 = (2.0f * kTemp) * (1.0f / r2);  // d w / d s = k2 w (q - s)
    const float inv_ws = 1.0f / fw.wsum;
#pragma unroll
    for (int i = 0; i < (KT > 0 ? KT : 8); ++i) {
        if (KT == 0 && i >= K) break;
        const f3 dv = q - mk(tex[3 * i], tex[3 * i + 1], tex[3 * i + 2]);
        const float w = fw.w1[i];
        if (w < 1e-3f * fw.wsum) continue;
        const f3 dwds = (k2 * w) * dv;
        const f3 grgb = (w * inv_ws) * g_rgb;
        red[3 * K + 3 * i] += grgb.x;
        red[3 * K + 3 * i + 1] += grgb.y;
        red[3 * K + 3 * i + 2] += grgb.z;
        const float gdp = dot(g_rgb, mk(tex[3 * K + 3 * i], tex[3 * K + 3 * i + 1], tex[3 * K + 3 * i + 2]));
        const f3 gs = dwds * ((gdp - gdp_in) * inv_ws);
        red[3 * i] += gs.x;
        red[3 * i + 1] += gs.y;
        red[3 * i + 2] += gs.z;
        dq = dq + (-1.0f * gs);
    }
    const float dtq = dot(dq, d);
    dtn = (fw.dp >= 0.0f || tn_in >= fw.t_surf) ? dtq : 0.0f;
    float dts = (fw.dp < 0.0f && tn_in < fw.t_surf) ? dtq : 0.0f;
    dts += dts_in;
    if (__builtin_fabsf(dts) > 0.0f) {
        f3 a, b;
        float h;
        ray_plane_bwd(o, d, c, n, fw.height, a, b, h);
        dc = dc + a * dts;
        dn = dn + b * dts;
        dh += h * dts;
    }
    const float tq0 = fw.dp0 >= 0.0f ? tn_in : fmax_(tn_in, fw.t0);
    const f3 q0 = ray_at(o, tq0, d);
    f3 dq0 = mk(0, 0, 0);
    const float hdp_in = dh * fw.height;
    const float inv_ws0 = 1.0f / fw.wsum0;
#pragma unroll
    for (int i = 0; i < (KT > 0 ? KT : 8); ++i) {
        if (KT == 0 && i >= K) break;
        const f3 dv = q0 - mk(tex[3 * i], tex[3 * i + 1], tex[3 * i + 2]);
        const float w = fw.w0[i];
        if (w < 1e-3f * fw.wsum0) continue;
        const f3 dwds = (k2 * w) * dv;
        red[6 * K + i] += (w * dh) * inv_ws0;
        const float hdp = dh * tex[6 * K + i];
        const f3 gs = dwds * ((hdp - hdp_in) * inv_ws0);
        red[3 * i] += gs.x;
        red[3 * i + 1] += gs.y;
        red[3 * i + 2] += gs.z;
        dq0 = dq0 + (-1.0f * gs);
    }
    const float dtq0 = dot(dq0, d);
    dtn += (fw.dp >= 0.0f || tn_in >= fw.t0) ? dtq0 : 0.0f;
    const float dts0 = (fw.dp < 0.0f && tn_in < fw.t0) ? dtq0 : 0.0f;
    if (__builtin_fabsf(dts0) > 0.0f) {
        f3 a, b;
        float h;
        ray_plane_bwd(o, d, c, n, 0.0f, a, b, h);
        dc = dc + a * dts0;
        dn = dn + b * dts0;
    }
}

// ------------------------------------------------------------ backward --
// Sums each of 64 per-lane values over the wavefront and leaves column c's
// total in lane c, by recursive halving: at each step a lane and its
// partner keep complementary halves of the columns still held, each sends
// the other half and adds what it receives.  Partners: lane ^ 32 and
// lane ^ 16 through v_permlane32_swap / v_permlane16_swap (which also do the
// sending), then DPP row_ror:8 (lane ^ 8 in a row), row_half_mirror (the
// lanes of an 8-lane group in reverse order: partners differ in bit 2),
// quad_perm xor 2 and xor 1.  141 VALU instructions and no LDS, for what a
// transpose through LDS did with 64 writes, 128 reads and four barriers.
// Every lane of the wavefront must take part.
__device__ __forceinline__ float wave_column_sums(float (&v)[64]) {
    const int lane = __lane_id();
#pragma unroll
    for (int i = 0; i < 32; ++i) {  // v[i] -> column (lane & 32) + i
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[i]), __float_as_uint(v[32 + i]), false,
                                                        false);
        v[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {  // v[i] -> column (lane & 48) + i
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[i]), __float_as_uint(v[16 + i]), false,
                                                        false);
        v[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    auto half = [&](auto dpp, int n, bool upper) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (i >= n) break;
            const float send = upper ? v[i] : v[n + i];
            const float keep = upper ? v[n + i] : v[i];
            v[i] = keep + dpp(send);
        }
    };
    half([](float x) { return dpp_row<0x128>(x); }, 8, lane & 8);  // row_ror:8
    half([](float x) { return dpp_row<0x141>(x); }, 4, lane & 4);  // row_half_mirror
    half([](float x) { return dpp_row<0x4E>(x); }, 2, lane & 2);   // quad_perm [2,3,0,1]
    half([](float x) { return dpp_row<0xB1>(x); }, 1, lane & 1);   // quad_perm [1,0,3,2]
    return v[0];
}

// Gradient row per lane: [sphere 4 | nsigma 4 | texels 7K]; summed over the
// wavefront by wave_column_sums (lane v gets column v), one atomic per value.
constexpr int kRedCols = 8 + 7 * 8;  // K <= 8 on this path
static_assert(kRedCols == 64, "one column per lane");

// Three waves per SIMD: the uncapped kernel takes 206 VGPRs (two waves);
// capped at 168 it measured 10% faster (1M-cell scene at 1920x1080:
// forward+backward 50.9 against 56.1 ms).  Four waves (128 VGPRs, 46
// spilled) measured 46.3 against 45.7 ms.
template <typename F, int KT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3))) void raster_bwd_kernel(RasterParams p) {
    __shared__ PrimChunk pc;
    __shared__ float4 splane[64];
    __shared__ float tex[7 * 8];
    __shared__ float sadj[64 * 4];
    const TileCtx tc = tile_ctx(p);
    if (tc.tile >= p.ntiles) return;
    const f3 o = tc.o, d = tc.d;
    const int64_t pix = (int64_t)tc.pix_i * p.W + tc.pix_j;
    const float hw = (float)(p.H * p.W);
    const int K = KT > 0 ? KT : p.K;

    float log_t = p.log_t[pix];
    int qi = p.depth_q ? p.Q - 1 : -1;
    float qv = p.depth_q ? p.depth_q[pix * p.Q + qi] : 1.0f;
    const f3 g_rgb = mk(p.g_color[3 * pix], p.g_color[3 * pix + 1], p.g_color[3 * pix + 2]);
    const float g_lt = p.g_log_t ? p.g_log_t[pix] : 0.0f;
    const float g_nd = p.g_ndist ? p.g_ndist[pix] : 0.0f;
    const f3 g_n = p.g_normal ? mk(p.g_normal[3 * pix], p.g_normal[3 * pix + 1], p.g_normal[3 * pix + 2]) : mk(0, 0, 0);
    const f3 bk = mk(0, 0, 0);
    float dLdlog_t = dot(g_rgb, bk) * pf_expf(log_t);
    dLdlog_t += g_lt;

    const int64_t start = p.offsets[tc.tile];
    const int stop = p.early_stop[tc.tile];
    float red[kRedCols];
    for (int i = stop - 1; i >= 0; --i) {
        const int sub = (stop - 1 - i) & 63;
        if (sub == 0) stage_prims(pc, p, start + i, start, start + stop, true);
        const int prim = pc.prim[sub];
        const float4 sp = pc.sph[sub];
        const float4 ns = pc.ns[sub];
        const f3 c = mk(sp.x, sp.y, sp.z), nrm = mk(ns.x, ns.y, ns.z);
        const float r = sp.w, sigma = ns.w;
        SphereHit sh = ray_sphere(o, d, c, r);
        if (!__any(sh.hit)) continue;
        float tn = sh.tn, tf = sh.tf;
        int tn_id = tn < 1e-8f ? -2 : -1;
        int tf_id = tf < 1e-8f ? -2 : -1;
        const int e0 = __builtin_amdgcn_readfirstlane(pc.e0[sub]);
        const int n_adj = __builtin_amdgcn_readfirstlane(pc.e1[sub]) - e0;
        for (int fb = 0; fb < n_adj; fb += 64) {
            __syncthreads();
            stage_planes_t<F>(splane, p, o, c, r, e0 + fb, n_adj - fb);
            __syncthreads();
            const int m = n_adj - fb < 64 ? n_adj - fb : 64;
            if (sh.hit) {
                for_planes(splane, m, [&](int j, float4 pl) {
                    float t, dp;
                    pface_t_rcp(d, pl, t, dp);
                    tf_id = (dp >= 0.0f && t < tf) ? fb + j : tf_id;
                    tf = dp >= 0.0f ? fmin_(t, tf) : tf;
                    tn_id = (dp < 0.0f && t > tn) ? fb + j : tn_id;
                    tn = dp < 0.0f ? fmax_(t, tn) : tn;
                });
            }
        }
        const float tn0 = tn;
        const int tn0_id = tn_id, tf0_id = tf_id;
        stage_texels(tex, p, prim);
        const RTex tx = texel_fwd_raster<true>(o, d, tn, c, nrm, r, !tc.oob && sh.hit, tex, K);
        tf_id = (tx.dp >= 0.0f && tx.t_surf < tf) ? n_adj : tf_id;
        tf = tx.dp >= 0.0f ? fmin_(tx.t_surf, tf) : tf;
        tn_id = (tx.dp < 0.0f && tx.t_surf > tn) ? n_adj : tn_id;
        tn = tx.dp < 0.0f ? fmax_(tx.t_surf, tn) : tn;
        const float dt = tf - tn;
        float dLdtn = 0.0f, dLdtf = 0.0f, dLdr = 0.0f, dLdsigma = 0.0f;
        f3 dLdc = mk(0, 0, 0), dLdn = mk(0, 0, 0), dLdcolor = mk(0, 0, 0);
        const bool act = !tc.oob && sh.hit && dt > 0.0f;
        if (act) {
            const float dlt = (-sigma) * dt;
            log_t -= dlt;
            const float alpha = 1.0f - pf_expf(dlt);
            const float trans = pf_expf(log_t);
            const float ndv = dot(nrm, d);
            float dLdlog_t_i = dLdlog_t;
            float dLddelta = dLdlog_t;
            const float cg = p.g_contrib ? p.g_contrib[prim] : 0.0f;
            float dLdalpha = dot(g_rgb, tx.rgb) * trans;
            if (ndv > 0.0f) dLdalpha += g_nd * ndv * trans;
            dLdalpha += cg * trans / hw;
            dLdalpha += dot(g_n, nrm) * trans;
            float dLdtrans = dot(g_rgb, tx.rgb) * alpha;
            if (ndv > 0.0f) dLdtrans += g_nd * ndv * alpha;
            dLdtrans += cg * alpha / hw;
            dLdtrans += dot(g_n, nrm) * alpha;
            dLdcolor = (g_rgb * trans) * alpha;
            dLdlog_t_i += dLdtrans * pf_expf(log_t);
            dLddelta += -dLdalpha * pf_expf(dlt);
            dLdsigma += dLddelta * -dt;
            dLdtn += dLddelta * sigma;
            dLdtf += dLddelta * -sigma;
            while (qi >= 0 && trans > qv) {
                const float gq = p.g_qdepth[pix * p.Q + qi];
                dLdsigma += -gq * pf_logf(trans / qv) / (sigma * sigma + 1e-6f);
                dLdtn += gq;
                dLdlog_t_i += gq / (sigma + 1e-6f);
                --qi;
                if (qi >= 0) qv = p.depth_q[pix * p.Q + qi];
            }
            dLdn = dLdn + (alpha * trans) * g_n;
            if (ndv > 0.0f) dLdn = dLdn + ((g_nd * alpha) * trans) * d;
            dLdlog_t = dLdlog_t_i;
        }
        float dLdts = 0.0f;
        if (tn_id == n_adj) dLdts = dLdtn;
        else if (tf_id == n_adj) dLdts = dLdtf;
#pragma unroll
        for (int v = 0; v < kRedCols; ++v) red[v] = 0.0f;
        f3 pc_dc, pc_dn;
        float dtn_in;
        texel_bwd_raster<KT>(o, d, tn0, c, nrm, r, act, tex, K, tx, dLdcolor, dLdts, red + 8, pc_dc, pc_dn, dtn_in);
        dLdc = dLdc + pc_dc;
        dLdn = dLdn + pc_dn;
        if (tn_id != n_adj) dtn_in += dLdtn;
        // neighbour sphere gradients: aggregated per face in LDS
        __syncthreads();
        for (int v = threadIdx.x; v < 64 * 4; v += 64) sadj[v] = 0.0f;
        __syncthreads();
        if (act) {
            for (int side = 0; side < 2; ++side) {
                const int tid = side == 0 ? tn0_id : tf0_id;
                const float dLdt = side == 0 ? dtn_in : (tf_id != n_adj ? dLdtf : 0.0f);
                if (tid == -1) {
                    const SphereGrad sg = ray_sphere_bwd(o, d, c, r);
                    if (side == 0) { dLdc = dLdc + sg.dn_dc * dLdt; dLdr += dLdt * sg.dn_dr; }
                    else { dLdc = dLdc + sg.df_dc * dLdt; dLdr += dLdt * sg.df_dr; }
                } else if (tid >= 0 && tid < n_adj) {
                    const int adjc = p.adj[e0 + tid];
                    const float4 as = p.spheres[adjc];
                    const FaceGrad fg = ray_pface_bwd(o, d, c, r, mk(as.x, as.y, as.z), as.w);
                    dLdc = dLdc + fg.dc * dLdt;
                    dLdr += dLdt * fg.dr;
                    const float ga[4] = {dLdt * fg.dcj.x, dLdt * fg.dcj.y, dLdt * fg.dcj.z, dLdt * fg.drj};
                    if (tid < 64) {
                        for (int q = 0; q < 4; ++q) atomicAdd(&sadj[4 * tid + q], ga[q]);
                    } else {
                        for (int q = 0; q < 4; ++q) unsafeAtomicAdd(p.gs_sph + 4 * (int64_t)adjc + q, ga[q]);
                    }
                }
            }
            red[0] = dLdc.x; red[1] = dLdc.y; red[2] = dLdc.z; red[3] = dLdr;
            red[4] = dLdn.x; red[5] = dLdn.y; red[6] = dLdn.z; red[7] = dLdsigma;
        }
        __syncthreads();
        if (threadIdx.x < n_adj && threadIdx.x < 64) {
            const int adjc = p.adj[e0 + threadIdx.x];
            for (int q = 0; q < 4; ++q) {
                const float v = sadj[4 * threadIdx.x + q];
                if (v != 0.0f) unsafeAtomicAdd(p.gs_sph + 4 * (int64_t)adjc + q, v);
            }
        }
        if (__any(act)) {
            // columns past 8 + 7K hold zeros (red is cleared per entry)
            const int ncols = 8 + 7 * K;
            const float s = wave_column_sums(red);
            const int col = threadIdx.x;
            if (col < ncols && s != 0.0f) {
                float *dst;
                if (col < 4) dst = p.gs_sph + 4 * (int64_t)prim + col;
                else if (col < 8) dst = p.gs_ns + 4 * (int64_t)prim + (col - 4);
                else if (col < 8 + 3 * K) dst = p.gs_sites + 3 * (int64_t)prim * K + (col - 8);
                else if (col < 8 + 6 * K) dst = p.gs_rgb + 3 * (int64_t)prim * K + (col - 8 - 3 * K);
                else dst = p.gs_h + (int64_t)prim * K + (col - 8 - 6 * K);
                unsafeAtomicAdd(dst, s);
            }
        }
    }
}

// ----------------------------------------------- benchmark / visualize --
template <typename A> __device__ __forceinline__ float4 ld_attr4(const void *p, int64_t i) {
    const A *a = reinterpret_cast<const A *>(p) + 4 * i;
    return make_float4((float)a[0], (float)a[1], (float)a[2], (float)a[3]);
}

// rasterize.py:614-734: per-lane early break, walk faces (dp, dpm) with the
// face-loop early-out, walk-variant texel lookup; colour only.
// rasterize.py:738-888 (vis != 0): raster faces (dp, r_j), colour, depth at a
// transmittance quantile, normal, alpha, intersection count, background.
template <typename A, bool VIS>
__global__ __launch_bounds__(64) void raster_plain_kernel(RasterParams p) {
    const TileCtx tc = tile_ctx(p);
    if (tc.tile >= p.ntiles) return;
    const f3 o = tc.o, d = tc.d;
    const int64_t pix = (int64_t)tc.pix_i * p.W + tc.pix_j;
    const int K = p.K;
    const A *ts = reinterpret_cast<const A *>(p.tsites);
    const A *tr_ = reinterpret_cast<const A *>(p.trgb);
    const A *th = reinterpret_cast<const A *>(p.theight);
    f3 rgb = mk(0, 0, 0), normal = mk(0, 0, 0);
    float log_t = 0.0f, depth = 0.0f;
    bool found = false;
    int nhit = 0;
    const int64_t start = p.offsets[tc.tile], end = p.offsets[tc.tile + 1];
    const int total = (int)(end - start);
    for (int i = 0; i < total; ++i) {
        if (VIS) {
            ++nhit;
            if (i >= p.max_isect) break;
        }
        const int prim = p.prims[start + i];
        const float trans = pf_expf(log_t);
        if (trans < p.thr) break;
        const float4 sp = p.spheres[prim];
        const float4 ns = ld_attr4<A>(p.nsig_any, prim);
        const f3 c = mk(sp.x, sp.y, sp.z), nrm = mk(ns.x, ns.y, ns.z);
        const float r = sp.w, sigma = ns.w;
        SphereHit sh = ray_sphere(o, d, c, r);
        if (!sh.hit) continue;
        float tn = sh.tn, tf = sh.tf;
        const int e0 = p.off[prim], e1 = p.off[prim + 1];
        for (int e = e0; e < e1; ++e) {
            const rh4 fv = reinterpret_cast<const rh4 *>(p.faces)[e];
            const f3 fd = mk((float)fv.x, (float)fv.y, (float)fv.z);
            float t, dp;
            if (VIS) ray_pface(o, d, c, r, c + fd, (float)fv.w, t, dp);
            else ray_pface_diff(o, d, fd, (float)fv.w, t, dp);
            tf = dp >= 0.0f ? fmin_(t, tf) : tf;
            tn = dp < 0.0f ? fmax_(t, tn) : tn;
            if (tn > tf) break;
        }
        if (tn > tf) continue;
        const int64_t b = (int64_t)prim * K;
        const TexFwd tx = texel_fwd<A, 0>(o, d, tn, c, nrm, r, ts + 3 * b, tr_ + 3 * b, th + b, K);
        tf = tx.dp >= 0.0f ? fmin_(tx.t_surf, tf) : tf;
        tn = tx.dp < 0.0f ? fmax_(tx.t_surf, tn) : tn;
        const float dt = tf - tn;
        if (dt > 0.0f) {
            const float dlt = (-sigma) * dt;
            const float alpha = 1.0f - pf_expf(dlt);
            const float trn = pf_expf(log_t);
            rgb.x = __builtin_fmaf(tx.rgb.x * alpha, trn, rgb.x);
            rgb.y = __builtin_fmaf(tx.rgb.y * alpha, trn, rgb.y);
            rgb.z = __builtin_fmaf(tx.rgb.z * alpha, trn, rgb.z);
            if (VIS) {
                normal.x = __builtin_fmaf(nrm.x * alpha, trn, normal.x);
                normal.y = __builtin_fmaf(nrm.y * alpha, trn, normal.y);
                normal.z = __builtin_fmaf(nrm.z * alpha, trn, normal.z);
            }
            log_t += dlt;
            if (VIS && !found) {
                const float nt = pf_expf(log_t);
                if (nt < p.vis_depth_q) {
                    depth = tn + pf_logf(trn / p.vis_depth_q) / sigma;
                    found = true;
                }
            }
        }
    }
    if (!tc.oob) {
        const float T = pf_expf(log_t);
        const f3 bk = VIS ? mk(p.bkgd[0], p.bkgd[1], p.bkgd[2]) : mk(0, 0, 0);
        rgb = macc(rgb, T, bk);
        p.color[3 * pix] = rgb.x;
        p.color[3 * pix + 1] = rgb.y;
        p.color[3 * pix + 2] = rgb.z;
        if (VIS) {
            p.depth_out[pix] = depth;
            p.normal[3 * pix] = normal.x;
            p.normal[3 * pix + 1] = normal.y;
            p.normal[3 * pix + 2] = normal.z;
            p.alpha_out[pix] = 1.0f - T;
            p.isect_out[pix] = nhit;
        }
    }
}

}  // namespace pf

using namespace pf;

static int hip_ok() { return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP; }

extern "C" int pf_raster_faces(const float *spheres, int64_t num_cells, const int32_t *adjacency,
                               const int32_t *adjacency_offsets, void *faces, int32_t face_dtype, void *stream) {
    if (!spheres || num_cells <= 0 || !adjacency_offsets) return PF_ERR_ARG;
    const unsigned blocks = (unsigned)((num_cells + 255) / 256);
    hipStream_t st = (hipStream_t)stream;
    if (face_dtype == PF_F16)
        hipLaunchKernelGGL((raster_faces_kernel<_Float16>), dim3(blocks), dim3(256), 0, st, (const float4 *)spheres,
                           num_cells, adjacency, adjacency_offsets, (_Float16 *)faces);
    else if (face_dtype == PF_F32)
        hipLaunchKernelGGL((raster_faces_kernel<float>), dim3(blocks), dim3(256), 0, st, (const float4 *)spheres,
                           num_cells, adjacency, adjacency_offsets, (float *)faces);
    else return PF_ERR_DTYPE;
    return hip_ok();
}

// Fewer than 2^24 8x8 tiles (1.07 G pixels) per image: the per-tile sort
// runs one 256-thread workgroup per tile, under 2^32 work-items per
// dispatch.
constexpr int64_t kMaxRasterTiles = int64_t(1) << 24;

static bool raster_grid_fits(int32_t W, int32_t H) {
    return (int64_t)((H + 7) / 8) * ((W + 7) / 8) < kMaxRasterTiles;
}

extern "C" int pf_raster_bin_count(const float *spheres, int64_t num_cells, const pf_camera *cam,
                                   int32_t *tile_counts, void *stream) {
    if (!spheres || num_cells <= 0 || !cam || !tile_counts) return PF_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(bin_count_kernel, dim3((unsigned)((num_cells + 255) / 256)), dim3(256), 0, st,
                       (const float4 *)spheres, num_cells, (const float *)cam, tile_counts);
    return hip_ok();
}

extern "C" int pf_raster_scan(const int32_t *counts, int64_t n, int64_t *offsets, void *stream) {
    if (!counts || !offsets || n <= 0) return PF_ERR_ARG;
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, counts, n, offsets, (int64_t)-1,
                       (int *)nullptr);
    return hip_ok();
}

extern "C" int pf_raster_scan_bounded(const int32_t *counts, int64_t n, int64_t *offsets, int64_t capacity,
                                      int32_t *status, void *stream) {
    if (!counts || !offsets || n <= 0 || capacity < 0 || !status) return PF_ERR_ARG;
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, counts, n, offsets, capacity,
                       status);
    return hip_ok();
}

extern "C" int pf_raster_bin_write(const float *spheres, int64_t num_cells, const pf_camera *cam, int32_t num_tiles,
                                   const int64_t *tile_offsets, int32_t *write_counter, uint64_t *keys,
                                   uint64_t *scratch, int32_t *tile_prims, uint8_t *visible_mask,
                                   int64_t num_isect, void *stream) {
    if (!spheres || num_cells <= 0 || !cam || !tile_offsets || !write_counter || num_tiles <= 0) return PF_ERR_ARG;
    if (num_tiles >= kMaxRasterTiles) return PF_ERR_ARG;
    if (num_isect > 0 && (!keys || !scratch || !tile_prims)) return PF_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(bin_write_kernel, dim3((unsigned)((num_cells + 255) / 256)), dim3(256), 0, st,
                       (const float4 *)spheres, num_cells, (const float *)cam, tile_offsets, write_counter,
                       (unsigned long long *)keys);
    hipLaunchKernelGGL(tile_sort_kernel, dim3((unsigned)num_tiles), dim3(256), 0, st, tile_offsets, num_tiles,
                       (unsigned long long *)keys, (unsigned long long *)scratch, tile_prims);
    if (visible_mask && num_isect > 0)
        hipLaunchKernelGGL(visible_mask_kernel, dim3((unsigned)((num_isect + 255) / 256)), dim3(256), 0, st,
                           tile_prims, num_isect, tile_offsets, num_tiles, visible_mask);
    return hip_ok();
}

static void fill_common(RasterParams &p, const pf_scene *s, const pf_camera *cam, int W, int H,
                        const int64_t *offsets, const int32_t *prims) {
    p.spheres = (const float4 *)s->spheres;
    p.nsigmas = (const float4 *)s->nsigmas;
    p.nsig_any = s->nsigmas;
    p.attr_f16 = s->attr_dtype == PF_F16;
    p.tsites = (const float *)s->texel_sites;
    p.trgb = (const float *)s->texel_rgb;
    p.theight = (const float *)s->texel_height;
    p.adj = s->adjacency;
    p.off = s->adjacency_offsets;
    p.faces = s->faces;
    p.K = s->num_texel_sites;
    p.N = s->num_cells;
    p.camw = (const float *)cam;
    p.W = W;
    p.H = H;
    p.tiles_h = (H + 7) / 8;
    p.ntiles = p.tiles_h * ((W + 7) / 8);
    p.offsets = offsets;
    p.prims = prims;
}

extern "C" int pf_raster_forward(const pf_scene *scene, const pf_camera *cam, int32_t width, int32_t height,
                                 const int64_t *tile_offsets, const int32_t *tile_prims, int32_t *tile_early_stop,
                                 float transmittance_threshold, const float *depth_quantiles, int32_t num_quantiles,
                                 const float *ray_gt, int32_t return_point_err, int32_t surface_objective,
                                 const pf_raster_out *out, void *stream) {
    if (!scene || !cam || !tile_offsets || !tile_early_stop || !out || width <= 0 || height <= 0) return PF_ERR_ARG;
    if (!raster_grid_fits(width, height)) return PF_ERR_ARG;
    if (scene->attr_dtype != PF_F32) return PF_ERR_DTYPE;
    if (scene->num_texel_sites < 1 || scene->num_texel_sites > 16) return PF_ERR_ARG;
    if (!out->color || !out->log_t || !out->normal_distance || !out->normal || !out->contrib) return PF_ERR_ARG;
    if ((return_point_err || surface_objective) && !ray_gt) return PF_ERR_ARG;
    if (depth_quantiles && (num_quantiles <= 0 || !out->quantile_depths)) return PF_ERR_ARG;
    RasterParams p{};
    fill_common(p, scene, cam, width, height, tile_offsets, tile_prims);
    p.early_stop = tile_early_stop;
    p.thr = transmittance_threshold;
    p.depth_q = depth_quantiles;
    p.Q = depth_quantiles ? num_quantiles : 0;
    p.ray_gt = ray_gt;
    p.point_err = return_point_err;
    p.surface = surface_objective;
    p.color = out->color;
    p.log_t = out->log_t;
    p.ndist = out->normal_distance;
    p.normal = out->normal;
    p.qdepth = out->quantile_depths;
    p.err = out->err;
    p.contrib = out->contrib;
    p.perr = out->point_err;
    if (return_point_err && !p.perr) return PF_ERR_ARG;
    if (surface_objective && !p.err) return PF_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    if (scene->face_dtype == PF_F16)
        hipLaunchKernelGGL((raster_fwd_kernel<_Float16>), dim3(p.ntiles), dim3(64), 0, st, p);
    else
        hipLaunchKernelGGL((raster_fwd_kernel<float>), dim3(p.ntiles), dim3(64), 0, st, p);
    return hip_ok();
}

extern "C" int pf_raster_backward(const pf_scene *scene, const pf_camera *cam, int32_t width, int32_t height,
                                  const int64_t *tile_offsets, const int32_t *tile_prims,
                                  const int32_t *tile_early_stop, const float *depth_quantiles,
                                  int32_t num_quantiles, const float *log_t, const pf_raster_grad_in *gin,
                                  const pf_grads *grads, void *stream) {
    if (!scene || !cam || !tile_offsets || !tile_early_stop || !log_t || !gin || !grads || width <= 0 || height <= 0)
        return PF_ERR_ARG;
    if (!raster_grid_fits(width, height)) return PF_ERR_ARG;
    if (scene->attr_dtype != PF_F32) return PF_ERR_DTYPE;
    if (scene->num_texel_sites < 1 || scene->num_texel_sites > 8) return PF_ERR_ARG;
    if (!gin->color) return PF_ERR_ARG;
    if (depth_quantiles && (num_quantiles <= 0 || !gin->quantile_depths)) return PF_ERR_ARG;
    RasterParams p{};
    fill_common(p, scene, cam, width, height, tile_offsets, tile_prims);
    p.early_stop = const_cast<int32_t *>(tile_early_stop);
    p.depth_q = depth_quantiles;
    p.Q = depth_quantiles ? num_quantiles : 0;
    p.log_t = const_cast<float *>(log_t);
    p.g_color = gin->color;
    p.g_log_t = gin->log_t;
    p.g_ndist = gin->normal_distance;
    p.g_normal = gin->normal;
    p.g_qdepth = gin->quantile_depths;
    p.g_contrib = gin->contrib;
    p.gs_sph = grads->spheres;
    p.gs_ns = grads->nsigmas;
    p.gs_sites = grads->texel_sites;
    p.gs_rgb = grads->texel_rgb;
    p.gs_h = grads->texel_height;
    hipStream_t st = (hipStream_t)stream;
    const bool k8 = scene->num_texel_sites == 8;
    if (scene->face_dtype == PF_F16) {
        if (k8) hipLaunchKernelGGL((raster_bwd_kernel<_Float16, 8>), dim3(p.ntiles), dim3(64), 0, st, p);
        else hipLaunchKernelGGL((raster_bwd_kernel<_Float16, 0>), dim3(p.ntiles), dim3(64), 0, st, p);
    } else {
        if (k8) hipLaunchKernelGGL((raster_bwd_kernel<float, 8>), dim3(p.ntiles), dim3(64), 0, st, p);
        else hipLaunchKernelGGL((raster_bwd_kernel<float, 0>), dim3(p.ntiles), dim3(64), 0, st, p);
    }
    return hip_ok();
}

extern "C" int pf_raster_benchmark(const pf_scene *scene, const pf_camera *cam, int32_t width, int32_t height,
                                   const int64_t *tile_offsets, const int32_t *tile_prims,
                                   float transmittance_threshold, float *color, void *stream) {
    if (!scene || !cam || !tile_offsets || !color || width <= 0 || height <= 0) return PF_ERR_ARG;
    if (!raster_grid_fits(width, height)) return PF_ERR_ARG;
    if (scene->face_dtype != PF_F16) return PF_ERR_DTYPE;
    if (scene->num_texel_sites < 1 || scene->num_texel_sites > 16) return PF_ERR_ARG;
    RasterParams p{};
    fill_common(p, scene, cam, width, height, tile_offsets, tile_prims);
    p.thr = transmittance_threshold;
    p.color = color;
    hipStream_t st = (hipStream_t)stream;
    if (scene->attr_dtype == PF_F16)
        hipLaunchKernelGGL((raster_plain_kernel<_Float16, false>), dim3(p.ntiles), dim3(64), 0, st, p);
    else
        hipLaunchKernelGGL((raster_plain_kernel<float, false>), dim3(p.ntiles), dim3(64), 0, st, p);
    return hip_ok();
}

extern "C" int pf_raster_visualize(const pf_scene *scene, const pf_camera *cam, int32_t width, int32_t height,
                                   const int64_t *tile_offsets, const int32_t *tile_prims,
                                   const pf_vis_options *options, const pf_vis_out *out, void *stream) {
    if (!scene || !cam || !tile_offsets || !options || !out || width <= 0 || height <= 0) return PF_ERR_ARG;
    if (!raster_grid_fits(width, height)) return PF_ERR_ARG;
    if (!out->color || !out->depth || !out->normal || !out->alpha || !out->intersections) return PF_ERR_ARG;
    if (scene->face_dtype != PF_F16) return PF_ERR_DTYPE;
    if (scene->num_texel_sites < 1 || scene->num_texel_sites > 16) return PF_ERR_ARG;
    RasterParams p{};
    fill_common(p, scene, cam, width, height, tile_offsets, tile_prims);
    p.thr = options->transmittance_threshold;
    p.max_isect = options->max_intersections;
    p.vis_depth_q = options->depth_quantile;
    p.bkgd[0] = options->bkgd_color[0];
    p.bkgd[1] = options->bkgd_color[1];
    p.bkgd[2] = options->bkgd_color[2];
    p.color = out->color;
    p.depth_out = out->depth;
    p.normal = out->normal;
    p.alpha_out = out->alpha;
    p.isect_out = out->intersections;
    hipStream_t st = (hipStream_t)stream;
    if (scene->attr_dtype == PF_F16)
        hipLaunchKernelGGL((raster_plain_kernel<_Float16, true>), dim3(p.ntiles), dim3(64), 0, st, p);
    else
        hipLaunchKernelGGL((raster_plain_kernel<float, true>), dim3(p.ntiles), dim3(64), 0, st, p);
    return hip_ok();
}
