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
//
// pf_serve_step_grouped differentiates the forward+backward tiles grouped
// by cell instead (pf_grouped.h): their recording forward keeps each
// contributing segment with its partials, and the step's segments of all
// forward+backward requests are sorted by (request, cell) and
// differentiated in that order, a request's cells numbered from its
// bucket0.  Every launch of the step keeps a fixed shape.
#include "pf_device.h"
#include "pf_grouped.h"
#include "pf_hip.h"
#include "pf_launch.h"
#include "pf_walk.h"
#include "pf_walk_deferred.h"

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
    const RayFwd r = walk_ray_fwd<float, _Float16, 8>(s, colors, cam.eye, d, cam.start, thr, 0.5f, max_steps,
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
            r = walk_ray_fwd<float, _Float16, 8, RGB, true>(s, colors, cam.eye, d, cam.start, thr, 0.5f, max_steps,
                                                             nullptr, 0, rec);
        else
            r = walk_ray_fwd<float, _Float16, 8>(s, colors, cam.eye, d, cam.start, thr, 0.5f, max_steps, nullptr,
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
        st |= walk_rec_bwd<float, _Float16, 8, RGB, ServeCache>(s, colors, cam.eye, d, rec.rec, rec.slot, rec.S,
                                                                 r.segs, r.rgb, gc, 0.0f, G, false, bytes,
                                                                 live && held, cache);
        replay = live && !held;
    }
    return st | walk_ray_bwd<float, _Float16, 8, RGB, ServeCache>(s, colors, cam.eye, d, cam.start, thr, max_steps,
                                                                   r.rgb, gc, 0.0f, G, false, bytes, replay, cache);
}

// Served scenes use f32 attributes, K = 8 and (E, 4) fp16 walk faces
// (PF_F16), read in pairs.  The table lists forward entries first: tiles [0, desc[2]) are forward
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

// ------------------------------------------------ grouped serving step --
// The step's forward+backward tiles [desc[2], desc[1]) hold slots
// (g - desc[2]) * 64 + lane of a record of S slots (S / 64 tiles at most).
struct ServeGroup {
    const pf_serve_scene *scenes;
    const pf_serve_request *reqs;
    const int64_t *desc;  // {requests, step tiles, forward tiles, buckets}
    int4 *segs;
    float4 *part;
    int *counts;
    int64_t S;
    int cap;
    float thr;
    int max_steps;
    int *status;
    bool exact_texel;
    BucketScan scan;  // its count is desc[3]
    int *over_count, *over_list;  // rays replayed (S)
    int *ent_used;    // entries reserved by the count pass
    int64_t max_entries;
    GroupEntry *ent;
    int64_t max_buckets;
    int grouped_blocks;
};

// The range of forward+backward tiles, or an empty one (and
// PF_STATUS_LIST_CAP) when the table exceeds the record or the buckets.
__device__ __forceinline__ bool grouped_range_of(const ServeGroup &q, int64_t &begin, int64_t &end) {
    begin = q.desc[2];
    end = q.desc[1];
    if (end - begin > q.S / 64 || q.desc[3] > q.max_buckets) {
        if (q.status && blockIdx.x == 0 && threadIdx.x == 0) atomicOr(q.status, PF_STATUS_LIST_CAP);
        return false;
    }
    return (int)q.desc[0] > 0 && end > begin;
}

// The tile of a slot's wave: request, pixel, ray direction (per lane).
struct ServeLane {
    int ri;
    int64_t pix;
    bool valid;
    int pi, pj;
};
__device__ __forceinline__ ServeLane serve_lane(const ServeGroup &q, int64_t g) {
    ServeLane l;
    const int lane = __lane_id();
    l.ri = __builtin_amdgcn_readfirstlane(find_request(q.reqs, (int)q.desc[0], g));
    const pf_serve_request &rq = q.reqs[l.ri];
    const int W = rq.cam.width, H = rq.cam.height;
    const int tile = rq.tile_begin + (int)(g - rq.step_tile0);
    const int tiles_h = (H + 7) / 8;
    l.pi = (tile % tiles_h) * 8 + (lane & 7);
    l.pj = (tile / tiles_h) * 8 + (lane >> 3);
    l.valid = l.pi < H && l.pj < W;
    l.pix = l.valid ? (int64_t)l.pi * W + l.pj : 0;
    return l;
}

// The forward of the forward+backward tiles, recording each ray's
// contributing segments with their partials (the deferred wave walk).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void serve_recfwd_kernel(
    ServeGroup q) {
    __shared__ WaveShade shade[4];
    int64_t begin, end;
    if (!grouped_range_of(q, begin, end)) return;
    const int nact = (int)((end - begin + 3) / 4);
    if ((int)blockIdx.x >= nact) return;
    const int64_t g = begin + (int64_t)xcd_block(blockIdx.x, nact) * 4 + (threadIdx.x >> 6);
    if (g >= end) return;
    const ServeLane l = serve_lane(q, g);
    const int64_t slot = (g - begin) * 64 + __lane_id();
    if (!l.valid) {
        q.counts[slot] = 0;
        return;
    }
    const pf_serve_request &rq = q.reqs[l.ri];
    const pf_serve_scene &sc = q.scenes[rq.scene];
    Cam cam = load_cam(reinterpret_cast<const float *>(&rq.cam));
    if (rq.start_cell) cam.start = __builtin_amdgcn_readfirstlane(*rq.start_cell);
    const f3 d = ray_dir(cam, l.pi, l.pj);
    const SceneDev<float> s = scene_dev<float>(sc.scene);
    const float *rgb = rq.texel_rgb ? rq.texel_rgb : reinterpret_cast<const float *>(sc.scene.texel_rgb);
    const RayFwd r = walk_wave_fwd<float, _Float16, MemRGB<float>, true>(
        s, MemRGB<float>{rgb}, cam.eye, d, cam.start, q.thr, 0.5f, q.max_steps, nullptr, 0,
        SegRecord{q.segs, slot, q.S, q.cap, q.part}, shade[threadIdx.x >> 6]);
    q.counts[slot] = r.segs;
    rq.color[3 * l.pix] = r.rgb.x;
    rq.color[3 * l.pix + 1] = r.rgb.y;
    rq.color[3 * l.pix + 2] = r.rgb.z;
    if (rq.log_t) rq.log_t[l.pix] = r.log_t;
    if (r.status && q.status) atomicOr(q.status, r.status);
}

// The incoming gradient of sum |C - target|^2 at a pixel of request rq.
__device__ __forceinline__ f3 l2_grad(const pf_serve_request &rq, int64_t pix, f3 &C) {
    C = mk(rq.color[3 * pix], rq.color[3 * pix + 1], rq.color[3 * pix + 2]);
    const f3 t = mk(rq.target[3 * pix], rq.target[3 * pix + 1], rq.target[3 * pix + 2]);
    return 2.0f * (C - t);
}

// One thread per slot (one wave per tile).  Count pass: rays with more
// segments than the cap, or past the entry budget, go to the replay list
// (their count is set past the cap so that the fill pass skips them); the
// others count their segments per bucket.  Fill pass: each segment is
// written as a self-contained entry at its bucket's next position.
template <bool FILL> __global__ __launch_bounds__(256) void serve_bucket_kernel(ServeGroup q) {
    int64_t begin, end;
    if (!grouped_range_of(q, begin, end)) return;
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t g = begin + (slot >> 6);
    if (g >= end) return;  // whole waves
    const ServeLane l = serve_lane(q, g);
    const pf_serve_request &rq = q.reqs[l.ri];
    const int N = (int)q.scenes[rq.scene].scene.num_cells;
    const int64_t b0 = rq.bucket0, nb = q.desc[3];
    f3 C = mk(0, 0, 0), gc = mk(0, 0, 0);
    int count = 0;
    if (l.valid) {
        gc = l2_grad(rq, l.pix, C);
        if (gc.x != 0.0f || gc.y != 0.0f || gc.z != 0.0f) count = q.counts[slot];
    }
    if (!FILL) {
        const bool over = count > q.cap;
        // the entry budget: a wave reserves its rays' segments with one atomic
        int want = over ? 0 : count;
        int incl = want;
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(incl, off, 64);
            if (__lane_id() >= off) incl += o;
        }
        const int total = __shfl(incl, 63, 64);
        int base = 0;
        if (__lane_id() == 63 && total > 0) base = atomicAdd(q.ent_used, total);
        base = __shfl(base, 63, 64);
        const bool past = want > 0 && (int64_t)base + incl > q.max_entries;
        if (over || past) {
            q.over_list[atomicAdd(q.over_count, 1)] = (int)slot;
            if (past) q.counts[slot] = q.cap + 1;
            count = 0;
        }
    } else if (count > q.cap) {
        count = 0;
    }
    const f3 dir = FILL && count > 0 ? ray_dir(load_cam(reinterpret_cast<const float *>(&rq.cam)), l.pi, l.pj)
                                     : mk(0, 0, 0);
    for (int c = 0; __any(kRecChunk * c < count); ++c) {
        int4 rc[kRecChunk];
        float4 pc[kRecChunk];
        if (kRecChunk * c < count) {
            const int64_t e0 = rec_index(kRecChunk * c, slot, q.S);
#pragma unroll
            for (int j = 0; j < kRecChunk; ++j) rc[j] = q.segs[e0 + j];
            if (FILL) {
#pragma unroll
                for (int j = 0; j < kRecChunk; ++j) pc[j] = q.part[e0 + j];
            }
        }
#pragma unroll
        for (int j = 0; j < kRecChunk; ++j) {
            const int k = kRecChunk * c + j;
            int bk = -1;
            int4 sg = make_int4(0, 0, 0, 0);
            if (k < count) {
                sg = rc[j];
                const int64_t b = b0 + sg.x;
                if ((unsigned)sg.x < (unsigned)N && b < nb) bk = (int)b;
                else if (!FILL && q.status) atomicOr(q.status, PF_STATUS_BAD_RECORD);
            }
            const BucketLead bl = bucket_lead(bk);
            const bool lead_here = bk >= 0 && bl.leader == __lane_id();
            if (FILL) {
                int base = 0;
                if (lead_here) base = atomicAdd(q.scan.cursor + bk, bl.n);
                base = __shfl(base, bl.leader, 64);
                if (bk >= 0) {
                    const float4 pt = pc[j];
                    const f3 Tg = seg_Tg(gc, pt.x);
                    const float Ln = seg_Lnext(gc, 0.0f, C, mk(pt.y, pt.z, pt.w));
                    GroupEntry *en = q.ent + (base + bl.rank);
                    en->seg = sg;
                    en->coef = make_float4(Tg.x, Tg.y, Tg.z, Ln);
                    en->dirb = make_float4(dir.x, dir.y, dir.z, __int_as_float(l.ri));
                    en->aux = make_int4((int)l.pix, (int)slot, 0, 0);
                }
            } else if (lead_here) {
                atomicAdd(q.scan.offsets + bk, bl.n);
            }
        }
    }
}

// The serving step's view of the sorted entries: an entry's group is its
// request (its scene, camera, colours and gradient arrays); a chunk holds
// one request's entries, so all of these stay in scalar registers.
struct ServeGroupView {
    SceneDev<float> s;
    f3 o;
    const float *rgb;
    GradDev G;
    int bucket0;
    __device__ __forceinline__ f3 eye(int) const { return o; }
    __device__ __forceinline__ MemRGB<float> colors(int) const { return MemRGB<float>{rgb}; }
    __device__ __forceinline__ GradDev grads(int) const { return G; }
    __device__ __forceinline__ int key(int cell, int) const { return bucket0 + cell; }
};

__device__ __forceinline__ GradDev request_grads(const pf_serve_request &rq) {
    return GradDev{rq.grads.spheres, rq.grads.nsigmas, rq.grads.texel_sites, rq.grads.texel_rgb,
                   rq.grads.texel_height};
}

struct ServeGroupCtx {
    static constexpr bool kOneGroup = true;
    bool exact;
    const pf_serve_scene *scenes;
    const pf_serve_request *reqs;
    __device__ __forceinline__ ServeGroupView view(int ri) const {
        const pf_serve_request &rq = reqs[ri];
        const pf_serve_scene &sc = scenes[rq.scene];
        const float *c = reinterpret_cast<const float *>(&rq.cam);
        return ServeGroupView{scene_dev<float>(sc.scene), mk(c[0], c[1], c[2]),
                              rq.texel_rgb ? rq.texel_rgb : reinterpret_cast<const float *>(sc.scene.texel_rgb),
                              request_grads(rq), (int)rq.bucket0};
    }
    __device__ __forceinline__ void flush(float *row, int cell, int ri) const {
        flush_row(row, request_grads(reqs[ri]), cell, 0, 0);
    }
    __device__ __forceinline__ void count_bytes(const GroupEntry *, int, int, int) const {}
};

// Replays the listed rays (past the cap or the entry budget), 64 per wave,
// one per lane; the lanes of one request go together so that its scene
// stays in scalar registers (per-segment global atomics: these are few).
__device__ __forceinline__ void serve_replay(const ServeGroup &q, int64_t begin, int64_t first_wave,
                                             int64_t waves) {
    const int n = *q.over_count;
    const int lane = __lane_id();
    for (int64_t i0 = first_wave * 64; i0 < n; i0 += waves * 64) {
        const int64_t i = i0 + lane;
        const bool has = i < n;
        const int64_t slot = has ? q.over_list[i] : 0;
        const int64_t g = begin + (slot >> 6);
        const int ri = has ? find_request(q.reqs, (int)q.desc[0], g) : -1;
        uint64_t pend = __ballot(has);
        while (pend) {
            const int r0 = __builtin_amdgcn_readlane(ri, __ffsll((unsigned long long)pend) - 1);
            const bool m = has && ri == r0;
            pend &= ~__ballot(m);
            if (!m) continue;
            const pf_serve_request &rq = q.reqs[r0];
            const pf_serve_scene &sc = q.scenes[rq.scene];
            const int W = rq.cam.width, H = rq.cam.height;
            const int tile = rq.tile_begin + (int)(g - rq.step_tile0);
            const int tiles_h = (H + 7) / 8;
            const int pi = (tile % tiles_h) * 8 + (int)(slot & 7), pj = (tile / tiles_h) * 8 + (int)((slot & 63) >> 3);
            const int64_t pix = (int64_t)pi * W + pj;
            Cam cam = load_cam(reinterpret_cast<const float *>(&rq.cam));
            if (rq.start_cell) cam.start = *rq.start_cell;
            const f3 d = ray_dir(cam, pi, pj);
            f3 C;
            const f3 gc = l2_grad(rq, pix, C);
            const SceneDev<float> s = scene_dev<float>(sc.scene);
            const float *rgb = rq.texel_rgb ? rq.texel_rgb : reinterpret_cast<const float *>(sc.scene.texel_rgb);
            uint64_t bytes = 0;
            const int st = walk_ray_bwd<float, _Float16, 8>(s, MemRGB<float>{rgb}, cam.eye, d, cam.start, q.thr,
                                                             q.max_steps, C, gc, 0.0f, request_grads(rq),
                                                             q.exact_texel, bytes);
            if (st && q.status) atomicOr(q.status, st);
        }
    }
}

// Blocks [0, kServeReplayBlocks) replay the listed rays; the others split
// the sorted entries into equal contiguous ranges, one per wave, numbered
// XCD-aware.
constexpr int kServeReplayBlocks = 64;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void serve_grad_kernel(ServeGroup q) {
    __shared__ float rows[4 * kGroupRows * kGroupPitch];
    int64_t begin, end;
    if (!grouped_range_of(q, begin, end)) return;
    if ((int)blockIdx.x < kServeReplayBlocks) {
        serve_replay(q, begin, (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6), (int64_t)kServeReplayBlocks * 4);
        return;
    }
    const int gb = xcd_block((int)blockIdx.x - kServeReplayBlocks, q.grouped_blocks);
    const int64_t waves = (int64_t)q.grouped_blocks * 4;
    const int64_t w = (int64_t)gb * 4 + (threadIdx.x >> 6);
    const int64_t M = q.scan.offsets[q.desc[3]];
    const int64_t per = (M + waves - 1) / waves;
    const int64_t lo = w * per;
    const int64_t hi = lo + per < M ? lo + per : M;
    if (lo >= hi) return;
    grouped_range<float, _Float16>(ServeGroupCtx{q.exact_texel, q.scenes, q.reqs}, q.ent, lo, hi,
                                    rows + (threadIdx.x >> 6) * kGroupRows * kGroupPitch);
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

// ------------------------------------------------ grouped serving step --
// Workspace of pf_serve_step_grouped, 256-byte aligned parts: bucket counts
// and offsets (max_buckets + 1), scatter cursors (max_buckets), scan chunk
// totals, the two counters (replayed rays, reserved entries), the replay
// list (S slots), then the sorted entries, as many as the rest holds (at
// most S * cap): a ray whose segments would pass that budget is replayed
// instead.  The default size holds S * min(cap, 48) entries.
constexpr int64_t kServeEntriesPerSlot = 48;

struct ServeGroupLayout {
    int64_t S, chunks;
    int64_t off_bucket, off_cursor, off_blk, off_counters, off_over_list, off_ent;
};

static inline int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

static bool serve_group_layout(int64_t bwd_tiles, int32_t cap, int64_t max_buckets, ServeGroupLayout &L) {
    if (bwd_tiles <= 0 || bwd_tiles >= (int64_t(1) << 24) || cap <= 0 || cap % kRecChunk || max_buckets <= 0 ||
        max_buckets >= 0x7FFFFFFF)
        return false;
    L.S = bwd_tiles * 64;
    L.chunks = (max_buckets + 1 + kScanChunk - 1) / kScanChunk;
    int64_t o = 0;
    L.off_bucket = o;
    o = align256(o + 4 * (max_buckets + 1));
    L.off_cursor = o;
    o = align256(o + 4 * max_buckets);
    L.off_blk = o;
    o = align256(o + 4 * (L.chunks + 1));
    L.off_counters = o;
    o = align256(o + 8);
    L.off_over_list = o;
    o = align256(o + 4 * L.S);
    L.off_ent = o;
    return true;
}

// Entries a workspace of the given size holds (entry positions are int32).
static int64_t serve_group_entries(const ServeGroupLayout &L, int32_t cap, int64_t workspace_bytes) {
    int64_t n = (workspace_bytes - L.off_ent) / (int64_t)sizeof(GroupEntry);
    if (n > L.S * cap) n = L.S * cap;
    if (n > 0x7FFFFFFF - 1) n = 0x7FFFFFFF - 1;
    return n;
}

extern "C" int64_t pf_serve_grouped_workspace_bytes(int64_t bwd_tile_capacity, int32_t record_cap,
                                                    int64_t max_buckets) {
    ServeGroupLayout L;
    if (!serve_group_layout(bwd_tile_capacity, record_cap, max_buckets, L)) return 0;
    const int64_t per = record_cap < kServeEntriesPerSlot ? record_cap : kServeEntriesPerSlot;
    if (L.S * per >= 0x7FFFFFFF) return 0;
    return align256(L.off_ent + (int64_t)sizeof(GroupEntry) * L.S * per);
}

extern "C" int pf_serve_step_grouped(const pf_serve_scene *scenes, const pf_serve_request *requests,
                                     const int64_t *desc, int64_t tile_capacity, int64_t bwd_tile_capacity,
                                     float transmittance_threshold, int32_t max_steps, int32_t *status,
                                     const pf_walk_record *record, int64_t max_buckets, int32_t flags,
                                     void *workspace, int64_t workspace_bytes, void *stream) {
    if (!scenes || !requests || !desc || tile_capacity <= 0 || max_steps <= 0) return PF_ERR_ARG;
    if (flags & ~PF_BWD_EXACT_TEXEL) return PF_ERR_ARG;
    if (tile_capacity >= (int64_t(1) << 26)) return PF_ERR_ARG;
    if (bwd_tile_capacity <= 0 || bwd_tile_capacity > tile_capacity) return PF_ERR_ARG;
    if (!record || !record->segments || !record->partials || !record->counts || record->cap <= 0 ||
        record->cap % kRecChunk)
        return PF_ERR_ARG;
    ServeGroupLayout L;
    if (!serve_group_layout(bwd_tile_capacity, record->cap, max_buckets, L)) return PF_ERR_ARG;
    // the fixed parts and at least one entry
    if (!workspace || workspace_bytes < L.off_ent + (int64_t)sizeof(GroupEntry)) return PF_ERR_OVERFLOW;
    hipStream_t st = (hipStream_t)stream;
    char *ws = reinterpret_cast<char *>(workspace);
    ServeGroup q{};
    q.scenes = scenes;
    q.reqs = requests;
    q.desc = desc;
    q.segs = reinterpret_cast<int4 *>(record->segments);
    q.part = reinterpret_cast<float4 *>(record->partials);
    q.counts = record->counts;
    q.S = L.S;
    q.cap = record->cap;
    q.thr = transmittance_threshold;
    q.max_steps = max_steps;
    q.status = status;
    q.exact_texel = (flags & PF_BWD_EXACT_TEXEL) != 0;
    q.scan = BucketScan{reinterpret_cast<int *>(ws + L.off_bucket), reinterpret_cast<int *>(ws + L.off_cursor),
                        reinterpret_cast<int *>(ws + L.off_blk), desc + 3, 0};
    int *counters = reinterpret_cast<int *>(ws + L.off_counters);
    q.over_count = counters;
    q.ent_used = counters + 1;
    q.over_list = reinterpret_cast<int *>(ws + L.off_over_list);
    q.max_entries = serve_group_entries(L, record->cap, workspace_bytes);
    q.ent = reinterpret_cast<GroupEntry *>(ws + L.off_ent);
    q.max_buckets = max_buckets;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    q.grouped_blocks = cus * 4;

    // forward tiles
    hipLaunchKernelGGL((serve_fwd_kernel<3, false>), dim3((unsigned)((tile_capacity + 3) / 4)), dim3(256), 0, st,
                       scenes, requests, desc, transmittance_threshold, max_steps, status);
    // forward+backward tiles: recording forward, counting sort, gradients
    hipLaunchKernelGGL(serve_recfwd_kernel, dim3((unsigned)((bwd_tile_capacity + 3) / 4)), dim3(256), 0, st, q);
    hipLaunchKernelGGL(bucket_zero_kernel, dim3(1024), dim3(256), 0, st, q.scan, counters, 2);
    const dim3 slot_grid((unsigned)(L.S / 256 + (L.S % 256 ? 1 : 0)));
    hipLaunchKernelGGL(serve_bucket_kernel<false>, slot_grid, dim3(256), 0, st, q);
    bucket_scan(q.scan, (unsigned)L.chunks, st);
    hipLaunchKernelGGL(serve_bucket_kernel<true>, slot_grid, dim3(256), 0, st, q);
    hipLaunchKernelGGL(serve_grad_kernel, dim3((unsigned)(kServeReplayBlocks + q.grouped_blocks)), dim3(256), 0,
                       st, q);
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}
