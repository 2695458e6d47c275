// serve.hip — one serving step: a fixed-size launch that walks the tiles a
// step covers, drawn from many requests (different scenes, cameras and
// resolutions) listed in a device request table.
//
// The launch shape is fixed (grid sized for the step capacity) and every
// per-step quantity is read from device memory, so the host can capture the
// launch once in a HIP graph and replay it every step after refreshing the
// table.  Each wavefront takes one 8x8 tile; the request owning a tile is
// found by binary search over the table's step_tile0 prefix.  Forward
// requests write colour and log transmittance; forward+backward requests
// also compute the gradient of sum_pixels |C - target|^2 (checkpoint: the
// colour just computed, kept in registers) through the wave's LDS gradient
// cache, from the segments their forward recorded (or, without a record
// workspace or past its cap, by replaying the walk).  The two
// kinds run in two launches of one step, so forward tiles are not held to
// the occupancy of the backward's register footprint.
#include "pf_device.h"
#include "pf_hip.h"
#include "pf_launch.h"
#include "pf_walk.h"

namespace pf {

__device__ __forceinline__ int find_request(const pf_serve_request *reqs, int n, int64_t g) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (reqs[mid].step_tile0 <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One ray of a forward request.
template <typename RGB>
__device__ __forceinline__ int serve_ray(const SceneDev<float> &s, const RGB &colors, const pf_serve_request &rq,
                                         const Cam &cam, f3 d, int64_t pix, float thr, int max_steps) {
    const RayFwd r = walk_ray_fwd<float, PackedF16, 8>(s, colors, cam.eye, d, cam.start, thr, 0.5f, max_steps,
                                                       nullptr, 0);
    rq.color[3 * pix] = r.rgb.x;
    rq.color[3 * pix + 1] = r.rgb.y;
    rq.color[3 * pix + 2] = r.rgb.z;
    if (rq.log_t) rq.log_t[pix] = r.log_t;
    return r.status;
}

// Gradient-cache slots of the forward+backward launch (as the walk
// backward's: 9.9 KB of LDS per wave).
constexpr int kServeSlots = 38;
using ServeCache = CellCache<kServeSlots, 8>;

// One ray of a forward+backward request: the forward, then the backward of
// sum |C - target|^2 (the checkpoint is the colour just computed).  A wave
// serves one tile of one request, so its gradients go through the wave's
// LDS cache (rows) into that request's arrays; every lane takes part
// (valid == false: a pixel outside the image).  With a record (rec.cap >
// 0) the forward records the ray's contributing segments in its slot and
// the backward reads them back (this lane's own writes); rays past the cap
// are replayed.
template <typename RGB>
__device__ __forceinline__ int serve_ray_bwd(const SceneDev<float> &s, const RGB &colors,
                                             const pf_serve_request &rq, const Cam &cam, f3 d, int64_t pix,
                                             float thr, int max_steps, float *rows, bool valid, SegRecord rec) {
    RayFwd r{};
    f3 gc = mk(0, 0, 0);
    if (valid) {
        if (rec.cap > 0)
            r = walk_ray_fwd<float, PackedF16, 8, RGB, true>(s, colors, cam.eye, d, cam.start, thr, 0.5f, max_steps,
                                                             nullptr, 0, rec);
        else
            r = walk_ray_fwd<float, PackedF16, 8>(s, colors, cam.eye, d, cam.start, thr, 0.5f, max_steps, nullptr,
                                                  0);
        rq.color[3 * pix] = r.rgb.x;
        rq.color[3 * pix + 1] = r.rgb.y;
        rq.color[3 * pix + 2] = r.rgb.z;
        if (rq.log_t) rq.log_t[pix] = r.log_t;
        const f3 tg = mk(rq.target[3 * pix], rq.target[3 * pix + 1], rq.target[3 * pix + 2]);
        gc = 2.0f * (r.rgb - tg);
    }
    const GradDev G{rq.grads.spheres, rq.grads.nsigmas, rq.grads.texel_sites, rq.grads.texel_rgb,
                    rq.grads.texel_height};
    ServeCache cache;
    cache.init(rows, G);
    uint64_t bytes = 0;
    const bool live = valid && (gc.x != 0.0f || gc.y != 0.0f || gc.z != 0.0f);
    int st = r.status;
    bool replay = live;
    if (rec.cap > 0) {
        const bool held = r.segs <= rec.cap;
        st |= walk_rec_bwd<float, PackedF16, 8, RGB, ServeCache>(s, colors, cam.eye, d, rec.rec, rec.slot, rec.S,
                                                                 r.segs, r.rgb, gc, 0.0f, G, false, bytes,
                                                                 live && held, cache);
        replay = live && !held;
    }
    return st | walk_ray_bwd<float, PackedF16, 8, RGB, ServeCache>(s, colors, cam.eye, d, cam.start, thr, max_steps,
                                                                   r.rgb, gc, 0.0f, G, false, bytes, replay, cache);
}

// Served scenes use f32 attributes, K = 8 and packed walk faces
// (PF_F16_PACKED).  The table lists forward entries first: tiles [0, desc[2]) are forward
// only and [desc[2], desc[1]) forward+backward.  BWD selects the range, so
// forward-only tiles run in a kernel without the backward's registers.
// rec_segs / rec_cap: the forward+backward tiles' record workspace (one
// slot per lane of the launch, 64 * 4 * gridDim.x slots of rec_cap
// segments), or rec_cap == 0 to replay.
template <int DEG, bool BWD, bool SH>
__device__ __forceinline__ void serve_tile(const pf_serve_scene *scenes, const pf_serve_request *reqs,
                                           const int64_t *desc, float thr, int max_steps, int *status,
                                           int4 *rec_segs, int rec_cap) {
    const int nreq = (int)desc[0];
    const int64_t begin = BWD ? desc[2] : 0, end = BWD ? desc[1] : desc[2];
    // The grid covers the capacity; the XCD-aware numbering is taken over
    // this step's blocks only, so that a small step spreads over every XCD
    // (over the whole grid it would put a short prefix of tiles on one).
    const int nact = (int)((end - begin + 3) / 4);
    if ((int)blockIdx.x >= nact || nreq <= 0) return;
    const int64_t g = begin + (int64_t)xcd_block(blockIdx.x, nact) * 4 + (threadIdx.x >> 6);
    if (g >= end) return;
    const int lane = threadIdx.x & 63;
    const int ri = __builtin_amdgcn_readfirstlane(find_request(reqs, nreq, g));
    const pf_serve_request &rq = reqs[ri];
    const int W = rq.cam.width, H = rq.cam.height;
    const int tile = rq.tile_begin + (int)(g - rq.step_tile0);
    const int tiles_h = (H + 7) / 8;
    const int pi = (tile % tiles_h) * 8 + (lane & 7);
    const int pj = (tile / tiles_h) * 8 + (lane >> 3);
    const bool valid = pi < H && pj < W;
    if (!BWD && !valid) return;
    const int64_t pix = valid ? (int64_t)pi * W + pj : 0;
    const pf_serve_scene &sc = scenes[rq.scene];
    Cam cam = load_cam(reinterpret_cast<const float *>(&rq.cam));
    if (rq.start_cell) cam.start = __builtin_amdgcn_readfirstlane(*rq.start_cell);
    const f3 d = ray_dir(cam, pi, pj);
    const SceneDev<float> s = scene_dev<float>(sc.scene);
    const float *rgb = rq.texel_rgb ? rq.texel_rgb : reinterpret_cast<const float *>(sc.scene.texel_rgb);
    const bool sh_in_walk = SH && !rq.texel_rgb && sc.sh;
    int st;
    if constexpr (BWD) {
        __shared__ float rows[4][kServeSlots * ServeCache::kPitch];
        float *wave_rows = rows[threadIdx.x >> 6];
        const SegRecord rec{rec_segs, (g - begin) * 64 + lane, (int64_t)gridDim.x * 4 * 64, rec_cap};
        if (sh_in_walk)
            st = serve_ray_bwd(s, SHRGB<DEG>{sc.sh, cam.eye}, rq, cam, d, pix, thr, max_steps, wave_rows, valid, rec);
        else
            st = serve_ray_bwd(s, MemRGB<float>{rgb}, rq, cam, d, pix, thr, max_steps, wave_rows, valid, rec);
    } else {
        if (sh_in_walk) st = serve_ray(s, SHRGB<DEG>{sc.sh, cam.eye}, rq, cam, d, pix, thr, max_steps);
        else st = serve_ray(s, MemRGB<float>{rgb}, rq, cam, d, pix, thr, max_steps);
    }
    if (st && status) atomicOr(status, st);
}

template <int DEG, bool SH>
__global__ __launch_bounds__(256) void serve_fwd_kernel(const pf_serve_scene *scenes, const pf_serve_request *reqs,
                                                        const int64_t *desc, float thr, int max_steps, int *status) {
    serve_tile<DEG, false, SH>(scenes, reqs, desc, thr, max_steps, status, nullptr, 0);
}

// Colours from memory: four waves per SIMD, as the walk backward (the LDS
// gradient cache allows four blocks per CU; the register budget is capped
// to match).  SH evaluated in the walk needs more registers than that cap
// and runs uncapped.
template <int DEG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void serve_bwd_kernel(
    const pf_serve_scene *scenes, const pf_serve_request *reqs, const int64_t *desc, float thr, int max_steps,
    int *status, int4 *rec_segs, int rec_cap) {
    serve_tile<DEG, true, false>(scenes, reqs, desc, thr, max_steps, status, rec_segs, rec_cap);
}
template <int DEG>
__global__ __launch_bounds__(256) void serve_bwd_sh_kernel(const pf_serve_scene *scenes,
                                                           const pf_serve_request *reqs, const int64_t *desc,
                                                           float thr, int max_steps, int *status, int4 *rec_segs,
                                                           int rec_cap) {
    serve_tile<DEG, true, true>(scenes, reqs, desc, thr, max_steps, status, rec_segs, rec_cap);
}

}  // namespace pf

using namespace pf;

template <bool SH>
static void launch_step(const pf_serve_scene *scenes, const pf_serve_request *requests, const int64_t *desc,
                        unsigned blocks, float thr, int32_t max_steps, int32_t *status, int4 *rec_segs, int rec_cap,
                        hipStream_t st) {
    hipLaunchKernelGGL((serve_fwd_kernel<3, SH>), dim3(blocks), dim3(256), 0, st, scenes, requests, desc, thr,
                       max_steps, status);
    if constexpr (SH)
        hipLaunchKernelGGL((serve_bwd_sh_kernel<3>), dim3(blocks), dim3(256), 0, st, scenes, requests, desc, thr,
                           max_steps, status, rec_segs, rec_cap);
    else
        hipLaunchKernelGGL((serve_bwd_kernel<3>), dim3(blocks), dim3(256), 0, st, scenes, requests, desc, thr,
                           max_steps, status, rec_segs, rec_cap);
}

extern "C" int64_t pf_serve_record_slots(int64_t tile_capacity) {
    return tile_capacity > 0 ? (tile_capacity + 3) / 4 * 4 * 64 : 0;
}

extern "C" int pf_serve_step(const pf_serve_scene *scenes, const pf_serve_request *requests, const int64_t *desc,
                             int64_t tile_capacity, float transmittance_threshold, int32_t max_steps, int32_t flags,
                             int32_t *status, const pf_walk_record *record, void *stream) {
    if (!scenes || !requests || !desc || tile_capacity <= 0 || max_steps <= 0) return PF_ERR_ARG;
    // one 256-thread workgroup per four tiles (64 work-items per tile), under
    // 2^32 work-items: fewer than 2^26 tiles
    if (tile_capacity >= (int64_t(1) << 26)) return PF_ERR_ARG;
    if (flags & ~PF_SERVE_SH_IN_WALK) return PF_ERR_ARG;
    if (record && (!record->segments || record->cap <= 0 || record->cap % kRecChunk)) return PF_ERR_ARG;
    const unsigned blocks = (unsigned)((tile_capacity + 3) / 4);
    int4 *segs = record ? reinterpret_cast<int4 *>(record->segments) : nullptr;
    const int cap = record ? record->cap : 0;
    if (flags & PF_SERVE_SH_IN_WALK)
        launch_step<true>(scenes, requests, desc, blocks, transmittance_threshold, max_steps, status, segs, cap,
                          (hipStream_t)stream);
    else
        launch_step<false>(scenes, requests, desc, blocks, transmittance_threshold, max_steps, status, segs, cap,
                           (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}
