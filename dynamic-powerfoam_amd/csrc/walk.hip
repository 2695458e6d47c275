// walk.hip — the power-cell walk kernels (forward and checkpoint-and-replay
// backward), walk face precompute and start-cell lookup for gfx950.
//
// One ray per lane, one 8x8 tile per 64-wide wavefront, tiles numbered
// column-major as in the reference (raytrace.py:130-151), so lane l of a
// wave shades pixel (tile_i*8 + l%8, tile_j*8 + l/8).  The per-ray walk is
// in pf_walk.h.
#include "pf_device.h"
#include "pf_hip.h"
#include "pf_launch.h"
#include "pf_walk.h"
#include "pf_walk_deferred.h"
#include "pf_walk_params.h"

namespace pf {

// K = 8 runs the wave walk with deferred, dense texel shading
// (pf_walk_deferred.h); other K the per-lane walk.  Same results.  Five
// waves per SIMD (at most 96 VGPRs; the recording walk needs 101 unbounded
// and fits without spilling).
template <typename A, typename F, int KT, bool REC> __device__ __forceinline__ void walk_fwd_tile(const WalkParams &p) {
    __shared__ WaveShade shade[KT == 8 ? 4 : 1];
    const TileLane tl = tile_lane(p, 4, xcd_block(blockIdx.x, gridDim.x));
    if (!tl.valid) return;
    const Cam cam = load_cam(p.cams + 12 * tl.b);
    const f3 d = ray_dir(cam, tl.pix_i, tl.pix_j);
    const SceneDev<A> s = scene_dev<A>(p.s);
    const MemRGB<A> colors{reinterpret_cast<const A *>(p.s.texel_rgb) + (int64_t)tl.b * p.s.texel_rgb_cam_stride};
    int *seq = p.out.seq ? p.out.seq + tl.ray * p.out.seq_cap : nullptr;
    const SegRecord rec{p.rec_segs, tl.slot, p.total_tiles * 64, p.rec_cap, p.rec_part};
    RayFwd r;
    if constexpr (KT == 8)
        r = walk_wave_fwd<A, F, MemRGB<A>, REC>(s, colors, cam.eye, d, cam.start, p.thr, p.depth_q, p.max_steps, seq,
                                                p.out.seq_cap, rec, shade[threadIdx.x >> 6]);
    else
        r = walk_ray_fwd<A, F, KT, MemRGB<A>, REC>(s, colors, cam.eye, d, cam.start, p.thr, p.depth_q, p.max_steps,
                                                   seq, p.out.seq_cap, rec);
    if constexpr (REC) p.rec_counts[tl.slot] = r.segs;
    float *co = p.out.color + 3 * tl.ray;
    co[0] = r.rgb.x;
    co[1] = r.rgb.y;
    co[2] = r.rgb.z;
    if (p.out.log_t) p.out.log_t[tl.ray] = r.log_t;
    if (p.out.depth) p.out.depth[tl.ray] = r.depth;
    if (p.out.visits) p.out.visits[tl.ray] = r.visits;
    if (p.out.seq_hash) p.out.seq_hash[tl.ray] = r.hash;
    if (p.out.ray_status) p.out.ray_status[tl.ray] = r.status;
    if (p.out.alg_bytes)
        p.out.alg_bytes[tl.ray] = r.bytes + 12 + (p.out.log_t ? 4 : 0) + (p.out.depth ? 4 : 0) + (REC ? 4 : 0);
    if (r.status && p.out.status) atomicOr(p.out.status, r.status);
}

template <typename A, typename F, int KT, bool REC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void walk_fwd_kernel(WalkParams p) {
    walk_fwd_tile<A, F, KT, REC>(p);
}

// The same recording forward over one band of a banded replay backward, a
// kernel of its own name so that profiles keep its band-sized launches
// apart from the frame's.
template <typename A, typename F, int KT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void walk_band_fwd_kernel(WalkParams p) {
    walk_fwd_tile<A, F, KT, true>(p);
}

// Slots of the per-wave gradient cache (K = 8: 65-float rows, 9.9 KB of
// LDS per wave, 39.5 KB per block): four blocks per CU, which matches the
// kernel's register budget of 128 VGPRs (four waves per SIMD).  48 slots
// catch more reuse but hold three blocks per CU and measured slower.
constexpr int kCacheSlots = 38;

using WaveCache = CellCache<kCacheSlots, 8>;
constexpr int kRowsPerWave = kCacheSlots * WaveCache::kPitch;

// One tile's backward by checkpoint and replay.  only_over: only the rays
// the forward's record could not hold (count above its cap).  rows: the
// block's LDS gradient cache (K = 8), four waves' worth.
template <typename A, typename F, int KT>
__device__ __forceinline__ void replay_tile(const WalkParams &p, const TileLane &tl, bool only_over, float *rows) {
    constexpr bool kWave = KT == 8;
    if (only_over && p.rec_header && p.rec_header[7] != 1) return;  // a record of another walk
    if (!kWave && !tl.valid) return;
    const bool mine = tl.valid && (!only_over || p.rec_counts[tl.slot] > p.rec_cap);
    if (!kWave && !mine) return;
    const int64_t ray = tl.valid ? tl.ray : 0;
    const f3 g = p.g_color && mine ? mk(p.g_color[3 * ray], p.g_color[3 * ray + 1], p.g_color[3 * ray + 2])
                                   : mk(0, 0, 0);
    const float gL = p.g_log_t && mine ? p.g_log_t[ray] : 0.0f;
    uint64_t bytes = 12 + (p.g_log_t ? 4 : 0) + (only_over ? 4 : 0);  // incoming gradients (and count)
    const bool live = mine && (g.x != 0.0f || g.y != 0.0f || g.z != 0.0f || gL != 0.0f);
    if (kWave && !__any(live)) {
        if (mine && p.alg_bytes) p.alg_bytes[tl.ray] = bytes;
        return;
    }
    const Cam cam = load_cam(p.cams + 12 * tl.b);
    const f3 d = ray_dir(cam, tl.pix_i, tl.pix_j);
    const SceneDev<A> s = scene_dev<A>(p.s);
    const MemRGB<A> colors{reinterpret_cast<const A *>(p.s.texel_rgb) + (int64_t)tl.b * p.s.texel_rgb_cam_stride};
    const f3 Cfin = live ? mk(p.ck_color[3 * ray], p.ck_color[3 * ray + 1], p.ck_color[3 * ray + 2]) : mk(0, 0, 0);
    if (live) bytes += 12;  // checkpoint colour
    const GradDev G{p.grads.spheres, p.grads.nsigmas, p.grads.texel_sites,
                    p.grads.texel_rgb + (int64_t)tl.b * p.grads.texel_rgb_cam_stride, p.grads.texel_height};
    int st = 0;
    if constexpr (kWave) {
        WaveCache cache;
        cache.init(rows + (threadIdx.x >> 6) * kRowsPerWave, G);
        st = walk_ray_bwd<A, F, KT, MemRGB<A>, WaveCache>(s, colors, cam.eye, d, cam.start, p.thr, p.max_steps, Cfin,
                                                          g, gL, G, p.exact_texel, bytes, live, cache);
    } else if (live) {
        st = walk_ray_bwd<A, F, KT>(s, colors, cam.eye, d, cam.start, p.thr, p.max_steps, Cfin, g, gL, G,
                                    p.exact_texel, bytes);
    }
    if (!mine) return;
    if (p.alg_bytes) p.alg_bytes[tl.ray] = bytes;
    if (st && p.status) atomicOr(p.status, st);
}

// One tile's backward from the forward's record of contributing segments:
// no replay, one pass over each ray's recorded segments, with the same
// gradient cache.  Rays whose count exceeds the cap are replay_tile's.
template <typename A, typename F, int KT>
__device__ __forceinline__ void record_tile(const WalkParams &p, const TileLane &tl, float *rows) {
    constexpr bool kWave = KT == 8;
    if (p.rec_header && p.rec_header[7] != 1) return;  // a record of another walk
    if (!kWave && !tl.valid) return;
    const int count = tl.valid ? p.rec_counts[tl.slot] : 0;
    const bool mine = tl.valid && count <= p.rec_cap;
    if (!kWave && !mine) return;
    const int64_t ray = tl.valid ? tl.ray : 0;
    const f3 g = p.g_color && mine ? mk(p.g_color[3 * ray], p.g_color[3 * ray + 1], p.g_color[3 * ray + 2])
                                   : mk(0, 0, 0);
    const float gL = p.g_log_t && mine ? p.g_log_t[ray] : 0.0f;
    uint64_t bytes = 12 + (p.g_log_t ? 4 : 0) + 4;  // incoming gradients, segment count
    const bool live = mine && count > 0 && (g.x != 0.0f || g.y != 0.0f || g.z != 0.0f || gL != 0.0f);
    if (kWave && !__any(live)) {
        if (mine && p.alg_bytes) p.alg_bytes[tl.ray] = bytes;
        return;
    }
    const Cam cam = load_cam(p.cams + 12 * tl.b);
    const f3 d = ray_dir(cam, tl.pix_i, tl.pix_j);
    const SceneDev<A> s = scene_dev<A>(p.s);
    const MemRGB<A> colors{reinterpret_cast<const A *>(p.s.texel_rgb) + (int64_t)tl.b * p.s.texel_rgb_cam_stride};
    const f3 Cfin = live ? mk(p.ck_color[3 * ray], p.ck_color[3 * ray + 1], p.ck_color[3 * ray + 2]) : mk(0, 0, 0);
    if (live) bytes += 12;  // checkpoint colour
    const GradDev G{p.grads.spheres, p.grads.nsigmas, p.grads.texel_sites,
                    p.grads.texel_rgb + (int64_t)tl.b * p.grads.texel_rgb_cam_stride, p.grads.texel_height};
    const int64_t S = p.total_tiles * 64;
    int st = 0;
    if constexpr (kWave) {
        WaveCache cache;
        cache.init(rows + (threadIdx.x >> 6) * kRowsPerWave, G);
        st = walk_rec_bwd<A, F, KT, MemRGB<A>, WaveCache>(s, colors, cam.eye, d, p.rec_segs, tl.slot, S, count, Cfin,
                                                          g, gL, G, p.exact_texel, bytes, live, cache);
    } else if (live) {
        st = walk_rec_bwd<A, F, KT>(s, colors, cam.eye, d, p.rec_segs, tl.slot, S, count, Cfin, g, gL, G,
                                    p.exact_texel, bytes, true);
    }
    if (!mine) return;
    if (p.alg_bytes) p.alg_bytes[tl.ray] = bytes;
    if (st && p.status) atomicOr(p.status, st);
}

// Four waves per SIMD (the LDS cache allows four blocks per CU): the
// register budget is capped at 128, which costs a few spilled registers and
// measured faster than three waves with no spills.
template <typename A, typename F, int KT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void walk_bwd_kernel(WalkParams p) {
    __shared__ float rows[KT == 8 ? 4 * kRowsPerWave : 1];
    replay_tile<A, F, KT>(p, tile_lane(p, 4, xcd_block(blockIdx.x, gridDim.x)), false, rows);
}

// The recorded backward: grid of twice the tile blocks.  The first half
// replays the tiles that hold a ray the record could not (most of its
// waves find none and end at once); being dispatched first, their long
// walks overlap the second half, which differentiates every other ray from
// its record.
template <typename A, typename F, int KT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void walk_bwd_rec_kernel(WalkParams p) {
    __shared__ float rows[KT == 8 ? 4 * kRowsPerWave : 1];
    const int nb = (int)(gridDim.x / 2);
    if ((int)blockIdx.x < nb) replay_tile<A, F, KT>(p, tile_lane(p, 4, xcd_block(blockIdx.x, nb)), true, rows);
    else record_tile<A, F, KT>(p, tile_lane(p, 4, xcd_block(blockIdx.x - nb, nb)), rows);
}

// Stores one walk face; the packed layout carries the neighbour too.
template <typename F> __device__ __forceinline__ void store_face(void *faces, int64_t e, f3 dp, float dpm, int) {
    F *f = reinterpret_cast<F *>(faces) + 4 * e;
    f[0] = (F)dp.x;
    f[1] = (F)dp.y;
    f[2] = (F)dp.z;
    f[3] = (F)dpm;
}
template <> __device__ __forceinline__ void store_face<PackedF16>(void *faces, int64_t e, f3 dp, float dpm, int nb) {
    reinterpret_cast<PackedFace *>(faces)[e] = PackedFace{(_Float16)dp.x, (_Float16)dp.y, (_Float16)dp.z,
                                                          (_Float16)dpm, nb};
}

// Reference-layout faces (E,4) f16 plus the adjacency into packed records.
__global__ __launch_bounds__(256) void pack_faces_kernel(const h4 *faces, const int32_t *adj, int64_t E,
                                                         PackedFace *out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const h4 f = faces[e];
    out[e] = PackedFace{f.x, f.y, f.z, f.w, adj[e]};
}

// Walk faces: benchmark.py:295-302, one thread per cell over its CSR row.
template <typename F>
__global__ void walk_faces_kernel(const float *points, const float *radii, int64_t N, const int32_t *adj,
                                  const int32_t *off, void *faces) {
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
        store_face<F>(faces, e, mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), pmj - pmi, j);
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

// Record stamp: the scene's cell and edge counts, the cap, the tile count
// and a hash of the camera words, written by the recording forward and
// compared by each backward that reads the record.  The hash sums a mix of
// each (index, word) pair (the 32-bit "lowbias32" integer hash of
// C. Wellons), so one block computes it in parallel.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
constexpr int kStampMagic = 0x50465752;  // "RWFP"

__global__ __launch_bounds__(256) void record_stamp_kernel(const int *cams, int64_t words, int64_t N, int64_t E,
                                                           int cap, int64_t tiles, int *header, int check,
                                                           int *status) {
    uint32_t h = 0;
    for (int64_t i = threadIdx.x; i < words; i += 256) h += mix32((uint32_t)cams[i] ^ mix32((uint32_t)i * 0x9E3779B9u));
    for (int off = 32; off > 0; off >>= 1) h += (uint32_t)__shfl_down((int)h, off, 64);
    __shared__ uint32_t part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = h;
    __syncthreads();
    if (threadIdx.x != 0) return;
    h = part[0] + part[1] + part[2] + part[3];
    const int want[7] = {kStampMagic, (int)N, (int)(E & 0x7fffffff), (int)(E >> 31), cap, (int)tiles, (int)h};
    if (!check) {
        for (int k = 0; k < 7; ++k) header[k] = want[k];
        header[7] = 1;
        return;
    }
    bool ok = true;
    for (int k = 0; k < 7; ++k) ok = ok && header[k] == want[k];
    header[7] = ok ? 1 : 0;
    if (!ok && status) atomicOr(status, PF_STATUS_BAD_RECORD);
}

}  // namespace pf

using namespace pf;

// Stamps (check == 0) or checks (check == 1) a record's header; no-op
// without one.
static int record_stamp(const pf_walk_record *r, const pf_scene *s, const pf_camera *cams, int32_t num_cams,
                        int64_t total_tiles, int check, int32_t *status, hipStream_t st) {
    if (!r->header) return PF_OK;
    hipLaunchKernelGGL(record_stamp_kernel, dim3(1), dim3(256), 0, st, reinterpret_cast<const int *>(cams),
                       (int64_t)num_cams * 12, s->num_cells, s->num_edges, r->cap, total_tiles, r->header, check,
                       status);
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}

int pf_walk_check_stamp(const pf_walk_record *record, const pf_scene *scene, const pf_camera *cams,
                        int32_t num_cams, int64_t total_tiles, int32_t *status, hipStream_t st) {
    return record_stamp(record, scene, cams, num_cams, total_tiles, 1, status, st);
}

enum WalkMode { kFwd, kFwdRecord, kFwdBand, kBwd, kBwdRecord };

template <typename A, typename F, int KT> static void launch_walk_k(const WalkParams &p, WalkMode m, hipStream_t st) {
    const int64_t blocks = (p.total_tiles + 3) / 4;
    dim3 grid((unsigned)blocks), block(256);
    switch (m) {
    case kFwd: hipLaunchKernelGGL((walk_fwd_kernel<A, F, KT, false>), grid, block, 0, st, p); break;
    case kFwdRecord: hipLaunchKernelGGL((walk_fwd_kernel<A, F, KT, true>), grid, block, 0, st, p); break;
    case kFwdBand: hipLaunchKernelGGL((walk_band_fwd_kernel<A, F, KT>), grid, block, 0, st, p); break;
    case kBwd: hipLaunchKernelGGL((walk_bwd_kernel<A, F, KT>), grid, block, 0, st, p); break;
    case kBwdRecord: hipLaunchKernelGGL((walk_bwd_rec_kernel<A, F, KT>), dim3(2 * (unsigned)blocks), block, 0, st, p);
        break;
    }
}

template <typename A, typename F> static void launch_walk(const WalkParams &p, WalkMode m, int K, hipStream_t st) {
    if (K == 8) launch_walk_k<A, F, 8>(p, m, st);
    else launch_walk_k<A, F, 0>(p, m, st);
}

static int dispatch_walk(const WalkParams &p, WalkMode m, hipStream_t st) {
    const int K = p.s.num_texel_sites;
    if (p.s.attr_dtype == PF_F32 && p.s.face_dtype == PF_F16_PACKED) launch_walk<float, PackedF16>(p, m, K, st);
    else if (p.s.attr_dtype == PF_F16 && p.s.face_dtype == PF_F16_PACKED)
        launch_walk<_Float16, PackedF16>(p, m, K, st);
    else if (p.s.attr_dtype == PF_F32 && p.s.face_dtype == PF_F16) launch_walk<float, _Float16>(p, m, K, st);
    else if (p.s.attr_dtype == PF_F16 && p.s.face_dtype == PF_F16) launch_walk<_Float16, _Float16>(p, m, K, st);
    else if (p.s.attr_dtype == PF_F32 && p.s.face_dtype == PF_F32) launch_walk<float, float>(p, m, K, st);
    else if (p.s.attr_dtype == PF_F16 && p.s.face_dtype == PF_F32) launch_walk<_Float16, float>(p, m, K, st);
    else return PF_ERR_DTYPE;
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}

int pf::walk_record_launch(const WalkParams &p, hipStream_t st) { return dispatch_walk(p, kFwdBand, st); }

static int check_scene(const pf_scene *s) {
    if (!s || !s->spheres || !s->nsigmas || !s->texel_sites || !s->texel_rgb || !s->texel_height || !s->adjacency_offsets)
        return PF_ERR_ARG;
    if (s->num_edges > 0 && (!s->adjacency || !s->faces)) return PF_ERR_ARG;
    if (s->num_cells <= 0 || s->num_cells >= 0x7FFFFFFF) return PF_ERR_ARG;
    if (s->num_texel_sites < 1 || s->num_texel_sites > 16) return PF_ERR_ARG;
    if ((s->attr_dtype != PF_F32 && s->attr_dtype != PF_F16) ||
        (s->face_dtype != PF_F32 && s->face_dtype != PF_F16 && s->face_dtype != PF_F16_PACKED))
        return PF_ERR_DTYPE;
    return PF_OK;
}

// The recorded backward launches two 256-thread workgroups per four tiles
// (128 work-items per tile), and a dispatch holds fewer than 2^32
// work-items: fewer than 2^25 tiles of 8x8 rays (2.1 G rays) per call.
constexpr int64_t kMaxWalkTiles = int64_t(1) << 25;

static bool grid_fits(int32_t B, int32_t W, int32_t H) {
    return (int64_t)B * ((H + 7) / 8) * ((W + 7) / 8) < kMaxWalkTiles;
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
    if (!grid_fits(num_cams, width, height)) return PF_ERR_ARG;
    if (out->seq && out->seq_cap <= 0) return PF_ERR_ARG;
    WalkParams p = make_params(scene, cams, num_cams, width, height, transmittance_threshold, max_steps);
    p.depth_q = depth_quantile;
    p.out = *out;
    if (out->record) {
        const pf_walk_record *r = out->record;
        if (!r->segments || !r->counts || r->cap <= 0 || r->cap % kRecChunk) return PF_ERR_ARG;
        p.rec_segs = reinterpret_cast<int4 *>(r->segments);
        p.rec_counts = r->counts;
        p.rec_cap = r->cap;
        p.rec_part = reinterpret_cast<float4 *>(r->partials);
        rc = record_stamp(r, scene, cams, num_cams, p.total_tiles, 0, nullptr, (hipStream_t)stream);
        if (rc) return rc;
        return dispatch_walk(p, kFwdRecord, (hipStream_t)stream);
    }
    return dispatch_walk(p, kFwd, (hipStream_t)stream);
}

extern "C" int pf_walk_backward(const pf_scene *scene, const pf_camera *cams, int32_t num_cams, int32_t width,
                                int32_t height, float transmittance_threshold, int32_t max_steps, const float *color,
                                const float *log_t, const float *grad_color, const float *grad_log_t,
                                const pf_grads *grads, int32_t flags, int32_t *status, uint64_t *alg_bytes,
                                void *stream) {
    int rc = check_scene(scene);
    if (rc) return rc;
    if (!cams || num_cams <= 0 || width <= 0 || height <= 0 || !color || !grads || max_steps <= 0) return PF_ERR_ARG;
    if (!grid_fits(num_cams, width, height)) return PF_ERR_ARG;
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
    return dispatch_walk(p, kBwd, (hipStream_t)stream);
}

extern "C" int64_t pf_walk_record_slots(int32_t num_cams, int32_t width, int32_t height) {
    if (num_cams <= 0 || width <= 0 || height <= 0) return 0;
    return (int64_t)num_cams * ((height + 7) / 8) * ((width + 7) / 8) * 64;
}

extern "C" int pf_walk_backward_recorded(const pf_scene *scene, const pf_camera *cams, int32_t num_cams,
                                         int32_t width, int32_t height, float transmittance_threshold,
                                         int32_t max_steps, const float *color, const float *log_t,
                                         const float *grad_color, const float *grad_log_t,
                                         const pf_walk_record *record, const pf_grads *grads, int32_t flags,
                                         int32_t *status, uint64_t *alg_bytes, void *stream) {
    int rc = check_scene(scene);
    if (rc) return rc;
    if (!cams || num_cams <= 0 || width <= 0 || height <= 0 || !color || !grads || max_steps <= 0) return PF_ERR_ARG;
    if (!grid_fits(num_cams, width, height)) return PF_ERR_ARG;
    if (!grads->spheres || !grads->nsigmas || !grads->texel_sites || !grads->texel_rgb || !grads->texel_height)
        return PF_ERR_ARG;
    if (!record || !record->segments || !record->counts || record->cap <= 0 || record->cap % kRecChunk)
        return PF_ERR_ARG;
    WalkParams p = make_params(scene, cams, num_cams, width, height, transmittance_threshold, max_steps);
    p.ck_color = color;
    p.ck_log_t = log_t;
    p.g_color = grad_color;
    p.g_log_t = grad_log_t;
    p.grads = *grads;
    p.status = status;
    p.exact_texel = (flags & PF_BWD_EXACT_TEXEL) != 0;
    p.alg_bytes = alg_bytes;
    p.rec_segs = reinterpret_cast<int4 *>(record->segments);
    p.rec_counts = record->counts;
    p.rec_cap = record->cap;
    p.rec_header = record->header;
    rc = record_stamp(record, scene, cams, num_cams, p.total_tiles, 1, status, (hipStream_t)stream);
    if (rc) return rc;
    return dispatch_walk(p, kBwdRecord, (hipStream_t)stream);
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
                           adjacency, adjacency_offsets, faces);
    else if (face_dtype == PF_F32)
        hipLaunchKernelGGL((walk_faces_kernel<float>), dim3(blocks), dim3(256), 0, st, points, radii, num_cells,
                           adjacency, adjacency_offsets, faces);
    else if (face_dtype == PF_F16_PACKED)
        hipLaunchKernelGGL((walk_faces_kernel<PackedF16>), dim3(blocks), dim3(256), 0, st, points, radii, num_cells,
                           adjacency, adjacency_offsets, faces);
    else return PF_ERR_DTYPE;
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}

extern "C" int pf_walk_pack_faces(const void *faces_f16, const int32_t *adjacency, int64_t num_edges, void *packed,
                                  void *stream) {
    if (num_edges < 0 || (num_edges > 0 && (!faces_f16 || !adjacency || !packed))) return PF_ERR_ARG;
    if (num_edges == 0) return PF_OK;
    hipLaunchKernelGGL(pack_faces_kernel, dim3((unsigned)((num_edges + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, (const h4 *)faces_f16, adjacency, num_edges, (PackedFace *)packed);
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
