// walk_grouped.hip — the walk backward from a record, grouped by cell.
//
// The recording forward stores each contributing segment with its
// partials (transmittance before it, the ray's colour after it), so its
// gradients need nothing from the other segments of its ray.  This file
// sorts the segments by cell (per camera when each camera has its own texel
// colours) and differentiates them in cell order with the building blocks
// of pf_grouped.h: a counting sort on the device (count, scan, scatter of
// self-contained 64-byte entries) and a gradient pass over contiguous
// ranges of the sorted entries that sums each term over a cell's run across
// the lanes and adds it to global memory once per run.
//
// Rays whose segment count exceeded the record's cap are replayed (the
// walk backward by checkpoint and replay, one lane per ray) by the first
// blocks of the same launch.
#include "pf_device.h"
#include "pf_grouped.h"
#include "pf_hip.h"
#include "pf_launch.h"
#include "pf_walk.h"
#include "pf_walk_params.h"

#include <mutex>

namespace pf {

#ifndef PF_GROUP_WAVES
#define PF_GROUP_WAVES 4  // waves per SIMD of the grouped gradient kernel
#endif
#ifndef PF_REPLAY_BLOCKS
#define PF_REPLAY_BLOCKS 256
#endif
constexpr int kReplayBlocks = PF_REPLAY_BLOCKS;  // blocks replaying rays past the cap
constexpr int kGroupBlocksPerCU = PF_GROUP_WAVES;

struct GroupWs {
    int *bucket_off;  // nb + 1: counts, then exclusive offsets (bucket_off[nb] = total)
    int *cursor;      // nb: scatter cursors
    int *blk;         // scan block totals (nblk + 1)
    int *over_count;  // rays past the cap
    int *over_list;   // S: their slots
    // the segments in bucket order, each self-contained (S * cap entries):
    GroupEntry *ent;  // the sorted segments
};

struct GroupParams {
    WalkParams w;
    float4 *part;
    int *header;      // stamp of the record (may be null)
    GroupWs ws;
    int64_t S;        // slots
    int nb;           // buckets: N, or B * N with per-camera texel colour gradients
    int nblk;         // scan blocks
    int grouped_blocks;
};

__host__ __device__ inline BucketScan bucket_scan_of(const GroupParams &q) {
    return BucketScan{q.ws.bucket_off, q.ws.cursor, q.ws.blk, nullptr, (int64_t)q.nb};
}

// Slot -> the lane's camera, pixel and ray (tile_lane's map, from the slot).
struct SlotRay {
    int b, pix_i, pix_j;
    bool valid;
    int64_t ray;
};
__device__ __forceinline__ SlotRay slot_ray(const WalkParams &p, int64_t slot) {
    SlotRay r;
    const int64_t t = p.tile0 + (slot >> 6);
    const int lane = (int)(slot & 63);
    r.b = (int)(t / p.tiles_per_cam);
    const int tile = (int)(t % p.tiles_per_cam);
    r.pix_i = (tile % p.tiles_h) * 8 + (lane & 7);
    r.pix_j = (tile / p.tiles_h) * 8 + (lane >> 3);
    r.valid = r.pix_i < p.H && r.pix_j < p.W;
    r.ray = ((int64_t)r.b * p.H + r.pix_i) * p.W + r.pix_j;
    return r;
}

__device__ __forceinline__ bool ray_live(const WalkParams &p, int64_t ray) {
    if (p.g_log_t && p.g_log_t[ray] != 0.0f) return true;
    if (!p.g_color) return false;
    const float *g = p.g_color + 3 * ray;
    return g[0] != 0.0f || g[1] != 0.0f || g[2] != 0.0f;
}

__device__ __forceinline__ bool record_ok(const int *header) { return !header || header[7] == 1; }

__device__ __forceinline__ int bucket_of(const GroupParams &q, int cell, int b) {
    return q.nb == q.w.s.num_cells ? cell : b * (int)q.w.s.num_cells + cell;
}

// One thread per slot: rays past the cap go to the replay list; the others
// count their recorded segments per bucket.  FILL: scatter each segment as
// a self-contained entry to its bucket's next position instead (the
// leader's returned base plus the lane's rank among the bucket's lanes).
template <bool FILL> __global__ __launch_bounds__(256) void grp_bucket_kernel(GroupParams q) {
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = record_ok(q.header);
    const WalkParams &p = q.w;
    const SlotRay sr = slot_ray(p, slot < q.S ? slot : 0);
    int count = 0;
    if (ok && slot < q.S && sr.valid && ray_live(p, sr.ray)) {
        count = p.rec_counts[slot];
        if (count > p.rec_cap) {
            if (!FILL) q.ws.over_list[atomicAdd(q.ws.over_count, 1)] = (int)slot;
            count = 0;
        }
    }
    const int N = (int)p.s.num_cells;
    // the ray's incoming gradient, final colour and direction, once per ray
    f3 g = mk(0, 0, 0), Cfin = mk(0, 0, 0), dir = mk(0, 0, 0);
    float gL = 0.0f;
    if (FILL && count > 0) {
        const int64_t ray = sr.ray;
        dir = ray_dir(load_cam(p.cams + 12 * sr.b), sr.pix_i, sr.pix_j);
        if (p.g_color) g = mk(p.g_color[3 * ray], p.g_color[3 * ray + 1], p.g_color[3 * ray + 2]);
        if (p.g_log_t) gL = p.g_log_t[ray];
        Cfin = mk(p.ck_color[3 * ray], p.ck_color[3 * ray + 1], p.ck_color[3 * ray + 2]);
        // per ray: incoming gradients, checkpoint colour, segment count
        if (p.alg_bytes) atomicAdd((unsigned long long *)(p.alg_bytes + ray), 12ull + 12 + 4 + (p.g_log_t ? 4 : 0));
    }
    // A ray's segments come eight to a 128-byte line (rec_index): each lane
    // reads its line's eight entries (and partials) at once.  Read one at a
    // time, the wave's 64 lines were evicted from L2 between the reads and
    // fetched again for every segment.
    for (int c = 0; __any(kRecChunk * c < count); ++c) {
        int4 rc[kRecChunk];
        float4 pc[kRecChunk];
        if (kRecChunk * c < count) {
            const int64_t e0 = rec_index(kRecChunk * c, slot, q.S);
#pragma unroll
            for (int j = 0; j < kRecChunk; ++j) rc[j] = p.rec_segs[e0 + j];
            if (FILL) {
#pragma unroll
                for (int j = 0; j < kRecChunk; ++j) pc[j] = q.part[e0 + j];
            }
        }
#pragma unroll
        for (int j = 0; j < kRecChunk; ++j) {
            const int k = kRecChunk * c + j;
            int bk = -1;
            int4 rq = make_int4(0, 0, 0, 0);
            if (k < count) {
                rq = rc[j];
                const int cell = rq.x;
                if ((unsigned)cell < (unsigned)N) bk = bucket_of(q, cell, sr.b);
                else if (!FILL && p.status) atomicOr(p.status, PF_STATUS_BAD_RECORD);
            }
            const BucketLead bl = bucket_lead(bk);
            const bool lead_here = bk >= 0 && bl.leader == __lane_id();
            if (FILL) {
                int base = 0;
                if (lead_here) base = atomicAdd(q.ws.cursor + bk, bl.n);
                base = __shfl(base, bl.leader, 64);
                if (bk >= 0) {
                    const float4 pt = pc[j];
                    const f3 Tg = seg_Tg(g, pt.x);
                    const float Ln = seg_Lnext(g, gL, Cfin, mk(pt.y, pt.z, pt.w));
                    GroupEntry *en = q.ws.ent + (base + bl.rank);
                    en->seg = rq;
                    en->coef = make_float4(Tg.x, Tg.y, Tg.z, Ln);
                    en->dirb = make_float4(dir.x, dir.y, dir.z, __int_as_float(sr.b));
                    en->aux = make_int4((int)sr.ray, (int)slot, 0, 0);
                }
            } else if (lead_here) {
                atomicAdd(q.ws.bucket_off + bk, bl.n);
            }
        }
    }
}

// The walk's view of the sorted entries: one scene; an entry's group is its
// camera, which selects the ray origin and, with per-camera texel colours,
// the colours and their gradients.
template <typename A> struct WalkGroupView {
    SceneDev<A> s;
    const float *cams;
    const A *rgb;
    int64_t rgb_stride;
    GradDev G;
    int64_t grad_rgb_stride;
    int N;
    bool per_cam;
    __device__ __forceinline__ f3 eye(int b) const {
        const float *e = cams + 12 * b;
        return mk(e[0], e[1], e[2]);
    }
    __device__ __forceinline__ MemRGB<A> colors(int b) const { return MemRGB<A>{rgb + (int64_t)b * rgb_stride}; }
    __device__ __forceinline__ GradDev grads(int b) const {
        GradDev g = G;
        g.rgb += (int64_t)b * grad_rgb_stride;
        return g;
    }
    __device__ __forceinline__ int key(int cell, int b) const { return per_cam ? b * N + cell : cell; }
};

template <typename A, typename F> struct WalkGroupCtx {
    static constexpr bool kOneGroup = false;
    WalkGroupView<A> v;
    bool exact;
    uint64_t *alg_bytes;
    __device__ __forceinline__ const WalkGroupView<A> &view(int) const { return v; }
    __device__ __forceinline__ void flush(float *row, int cell, int b) const {
        flush_row(row, v.G, cell, b, v.grad_rgb_stride);
    }
    __device__ __forceinline__ void count_bytes(const GroupEntry *en, int tn0_id, int tf0_id, int n_adj) const {
        if (!alg_bytes) return;
        constexpr uint64_t kVisit = WalkBytes<A, F>::visit, kTexel = WalkBytes<A, F>::texel;
        // record entry and partials (read once), the sorted entry (written
        // and read), the cell, texels and gradient rows
        const uint64_t bytes = kSegRecordBytes + kSegPartBytes + 2 * sizeof(GroupEntry) + kVisit + 8 * kTexel +
                               segment_grad_bytes(8, tn0_id, tf0_id, n_adj);
        atomicAdd((unsigned long long *)(alg_bytes + en->aux.x), (unsigned long long)bytes);
    }
};

// Replays the rays past the record's cap, one lane per ray (per-segment
// global atomics: these are a small share of the rays).
template <typename A, typename F>
__device__ __forceinline__ void replay_over(const GroupParams &q, int64_t first, int64_t stride) {
    const WalkParams &p = q.w;
    const int n = *q.ws.over_count;
    const SceneDev<A> s = scene_dev<A>(p.s);
    for (int64_t i = first; i < n; i += stride) {
        const int64_t slot = q.ws.over_list[i];
        const SlotRay sr = slot_ray(p, slot);
        const int64_t ray = sr.ray;
        const Cam cam = load_cam(p.cams + 12 * sr.b);
        const f3 d = ray_dir(cam, sr.pix_i, sr.pix_j);
        const MemRGB<A> colors{reinterpret_cast<const A *>(p.s.texel_rgb) + (int64_t)sr.b * p.s.texel_rgb_cam_stride};
        const f3 g = p.g_color ? mk(p.g_color[3 * ray], p.g_color[3 * ray + 1], p.g_color[3 * ray + 2])
                               : mk(0, 0, 0);
        const float gL = p.g_log_t ? p.g_log_t[ray] : 0.0f;
        const f3 Cfin = mk(p.ck_color[3 * ray], p.ck_color[3 * ray + 1], p.ck_color[3 * ray + 2]);
        const GradDev G{p.grads.spheres, p.grads.nsigmas, p.grads.texel_sites,
                        p.grads.texel_rgb + (int64_t)sr.b * p.grads.texel_rgb_cam_stride, p.grads.texel_height};
        uint64_t bytes = 12 + (p.g_log_t ? 4 : 0) + 4 + 12;
        const int st = walk_ray_bwd<A, F, 8>(s, colors, cam.eye, d, cam.start, p.thr, p.max_steps, Cfin, g, gL, G,
                                             p.exact_texel, bytes);
        if (p.alg_bytes) atomicAdd((unsigned long long *)(p.alg_bytes + ray), (unsigned long long)bytes);
        if (st && p.status) atomicOr(p.status, st);
    }
}

// Blocks [0, kReplayBlocks) replay the rays past the cap, one per lane
// (dispatched first, their long walks overlap the rest; a block with no such
// ray returns at once, so many blocks cost nothing when few rays pass the cap
// and keep a low cap, where a tenth of the rays may pass it, to one pass of
// replays instead of dozens); the others split the sorted segments
// into equal contiguous ranges, one per wave, numbered XCD-aware so that one
// XCD's waves take neighbouring ranges (the same cells' rows).
template <typename A, typename F>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PF_GROUP_WAVES, PF_GROUP_WAVES))) void
grp_grad_kernel(GroupParams q) {
    __shared__ float rows[4 * kGroupRows * kGroupPitch];
    if (!record_ok(q.header)) return;
    if ((int)blockIdx.x < kReplayBlocks) {
        replay_over<A, F>(q, (int64_t)blockIdx.x * 256 + threadIdx.x, (int64_t)kReplayBlocks * 256);
        return;
    }
    const int gb = xcd_block((int)blockIdx.x - kReplayBlocks, q.grouped_blocks);
    const int64_t waves = (int64_t)q.grouped_blocks * 4;
    const int64_t w = (int64_t)gb * 4 + (threadIdx.x >> 6);
    const int64_t M = q.ws.bucket_off[q.nb];
    const int64_t per = (M + waves - 1) / waves;
    const int64_t lo = w * per;
    const int64_t hi = lo + per < M ? lo + per : M;
    if (lo >= hi) return;
    const WalkParams &p = q.w;
    const WalkGroupCtx<A, F> ctx{
        WalkGroupView<A>{scene_dev<A>(p.s), p.cams, reinterpret_cast<const A *>(p.s.texel_rgb),
                         p.s.texel_rgb_cam_stride,
                         GradDev{p.grads.spheres, p.grads.nsigmas, p.grads.texel_sites, p.grads.texel_rgb,
                                 p.grads.texel_height},
                         p.grads.texel_rgb_cam_stride, (int)p.s.num_cells, q.nb != (int)p.s.num_cells},
        p.exact_texel, p.alg_bytes};
    grouped_range<A, F>(ctx, q.ws.ent, lo, hi, rows + (threadIdx.x >> 6) * kGroupRows * kGroupPitch);
}
}  // namespace pf

using namespace pf;

static inline int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

struct GroupLayout {
    int64_t off_bucket, off_cursor, off_blk, off_over_count, off_over_list, off_ent, total;
    int64_t nb, nblk;
};

// S record slots (64 per tile), buckets per cell or per (camera, cell).
static bool group_layout_slots(int64_t num_cells, int64_t S, int32_t num_cams, int32_t cap, bool per_camera,
                               GroupLayout &L) {
    if (num_cells <= 0 || num_cells >= 0x7FFFFFFF || num_cams <= 0 || S <= 0 || cap <= 0) return false;
    const int64_t nb = per_camera ? num_cells * num_cams : num_cells;
    // entry ids and bucket offsets are int32
    if (S * cap >= (int64_t(1) << 31) || nb >= 0x7FFFFFFF) return false;
    L.nb = nb;
    L.nblk = (nb + 1 + kScanChunk - 1) / kScanChunk;
    int64_t o = 0;
    L.off_bucket = o;
    o = align256(o + 4 * (nb + 1));
    L.off_cursor = o;
    o = align256(o + 4 * nb);
    L.off_blk = o;
    o = align256(o + 4 * (L.nblk + 1));
    L.off_over_count = o;
    o = align256(o + 4);
    L.off_over_list = o;
    o = align256(o + 4 * S);
    L.off_ent = o;
    o = align256(o + (int64_t)sizeof(GroupEntry) * S * cap);
    L.total = o;
    return true;
}

static bool group_layout(int64_t num_cells, int32_t num_cams, int32_t width, int32_t height, int32_t cap,
                         bool per_camera, GroupLayout &L) {
    if (num_cams <= 0 || width <= 0 || height <= 0) return false;
    const int64_t S = (int64_t)num_cams * ((height + 7) / 8) * ((width + 7) / 8) * 64;
    return group_layout_slots(num_cells, S, num_cams, cap, per_camera, L);
}

// The grouped pipeline over q's record slots [0, q.S): bucket counts, scan,
// scatter of the sorted entries, gradient pass; its workspace is carved
// from ws by L.
static int run_grouped(GroupParams &q, const GroupLayout &L, char *ws, hipStream_t st) {
    q.ws.bucket_off = reinterpret_cast<int *>(ws + L.off_bucket);
    q.ws.cursor = reinterpret_cast<int *>(ws + L.off_cursor);
    q.ws.blk = reinterpret_cast<int *>(ws + L.off_blk);
    q.ws.over_count = reinterpret_cast<int *>(ws + L.off_over_count);
    q.ws.over_list = reinterpret_cast<int *>(ws + L.off_over_list);
    q.ws.ent = reinterpret_cast<GroupEntry *>(ws + L.off_ent);
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    q.grouped_blocks = cus * kGroupBlocksPerCU;
    const unsigned slot_blocks = (unsigned)((q.S + 255) / 256);
    hipLaunchKernelGGL(bucket_zero_kernel, dim3(1024), dim3(256), 0, st, bucket_scan_of(q), q.ws.over_count, 1);
    hipLaunchKernelGGL(grp_bucket_kernel<false>, dim3(slot_blocks), dim3(256), 0, st, q);
    bucket_scan(bucket_scan_of(q), (unsigned)q.nblk, st);
    hipLaunchKernelGGL(grp_bucket_kernel<true>, dim3(slot_blocks), dim3(256), 0, st, q);
    const dim3 grid((unsigned)(kReplayBlocks + q.grouped_blocks));
    const pf_scene &sc = q.w.s;
    if (sc.attr_dtype == PF_F32) {
        if (sc.face_dtype == PF_F16_PACKED)
            hipLaunchKernelGGL((grp_grad_kernel<float, PackedF16>), grid, dim3(256), 0, st, q);
        else if (sc.face_dtype == PF_F16)
            hipLaunchKernelGGL((grp_grad_kernel<float, _Float16>), grid, dim3(256), 0, st, q);
        else hipLaunchKernelGGL((grp_grad_kernel<float, float>), grid, dim3(256), 0, st, q);
    } else {
        if (sc.face_dtype == PF_F16_PACKED)
            hipLaunchKernelGGL((grp_grad_kernel<_Float16, PackedF16>), grid, dim3(256), 0, st, q);
        else if (sc.face_dtype == PF_F16)
            hipLaunchKernelGGL((grp_grad_kernel<_Float16, _Float16>), grid, dim3(256), 0, st, q);
        else hipLaunchKernelGGL((grp_grad_kernel<_Float16, float>), grid, dim3(256), 0, st, q);
    }
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}

extern "C" int64_t pf_walk_grouped_workspace_bytes(int64_t num_cells, int32_t num_cams, int32_t width,
                                                   int32_t height, int32_t cap) {
    // sized for per-camera buckets, so one workspace serves either layout
    GroupLayout L;
    return group_layout(num_cells, num_cams, width, height, cap, num_cams > 1, L) ? L.total : 0;
}

int pf_walk_check_stamp(const pf_walk_record *record, const pf_scene *scene, const pf_camera *cams,
                        int32_t num_cams, int64_t total_tiles, int32_t *status, hipStream_t st);

// The grouped backward with its workspace; workspace == nullptr: the
// library's static scratch array of this device (the no-allocation entry
// point), whose address is looked up here.
static int grouped_backward(const pf_scene *scene, const pf_camera *cams, int32_t num_cams, int32_t width,
                            int32_t height, float transmittance_threshold, int32_t max_steps, const float *color,
                            const float *log_t, const float *grad_color, const float *grad_log_t,
                            const pf_walk_record *record, const pf_grads *grads, int32_t flags, void *workspace,
                            int64_t workspace_bytes, bool use_scratch, int32_t *status, uint64_t *alg_bytes,
                            void *stream);

extern "C" int pf_walk_backward_grouped(const pf_scene *scene, const pf_camera *cams, int32_t num_cams,
                                        int32_t width, int32_t height, float transmittance_threshold,
                                        int32_t max_steps, const float *color, const float *log_t,
                                        const float *grad_color, const float *grad_log_t,
                                        const pf_walk_record *record, const pf_grads *grads, int32_t flags,
                                        void *workspace, int64_t workspace_bytes, int32_t *status,
                                        uint64_t *alg_bytes, void *stream) {
    return grouped_backward(scene, cams, num_cams, width, height, transmittance_threshold, max_steps, color, log_t,
                            grad_color, grad_log_t, record, grads, flags, workspace, workspace_bytes, false, status,
                            alg_bytes, stream);
}

// ------------------------------------------------ no-allocation variant --
// A file-scope device array serves as the grouped backward's workspace for
// callers that cannot allocate (code under HIP graph capture, serving
// loops, harnesses that forbid allocation).  Each GPU holds its own copy
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns with that device current (the array's name is
// a host-side stand-in, not a device address), looked up once per device.
// Every call on one GPU shares that GPU's array: the kernels clear what
// must start at zero, and calls that overlap on one GPU race.
constexpr int64_t kScratchBytes = int64_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void *g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_for_current_device(void **ptr) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return PF_ERR_HIP;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PF_ERR_HIP;
        g_scratch[dev] = p;
    }
    *ptr = g_scratch[dev];
    return PF_OK;
}

extern "C" int64_t pf_walk_grouped_scratch_bytes(void) { return kScratchBytes; }

extern "C" int pf_walk_backward_grouped_noalloc(const pf_scene *scene, const pf_camera *cams, int32_t num_cams,
                                                int32_t width, int32_t height, float transmittance_threshold,
                                                int32_t max_steps, const float *color, const float *log_t,
                                                const float *grad_color, const float *grad_log_t,
                                                const pf_walk_record *record, const pf_grads *grads, int32_t flags,
                                                int32_t *status, uint64_t *alg_bytes, void *stream) {
    return grouped_backward(scene, cams, num_cams, width, height, transmittance_threshold, max_steps, color, log_t,
                            grad_color, grad_log_t, record, grads, flags, nullptr, 0, true, status, alg_bytes,
                            stream);
}

static int grouped_backward(const pf_scene *scene, const pf_camera *cams, int32_t num_cams, int32_t width,
                            int32_t height, float transmittance_threshold, int32_t max_steps, const float *color,
                            const float *log_t, const float *grad_color, const float *grad_log_t,
                            const pf_walk_record *record, const pf_grads *grads, int32_t flags, void *workspace,
                            int64_t workspace_bytes, bool use_scratch, int32_t *status, uint64_t *alg_bytes,
                            void *stream) {
    if (!scene || !cams || num_cams <= 0 || width <= 0 || height <= 0 || !color || !grads || max_steps <= 0)
        return PF_ERR_ARG;
    if (!scene->spheres || !scene->nsigmas || !scene->texel_sites || !scene->texel_rgb || !scene->texel_height ||
        !scene->adjacency_offsets || scene->num_cells <= 0 || scene->num_cells >= 0x7FFFFFFF)
        return PF_ERR_ARG;
    if (scene->num_edges > 0 && (!scene->adjacency || !scene->faces)) return PF_ERR_ARG;
    if (scene->num_texel_sites != 8) return PF_ERR_ARG;
    if (!grads->spheres || !grads->nsigmas || !grads->texel_sites || !grads->texel_rgb || !grads->texel_height)
        return PF_ERR_ARG;
    if (!record || !record->segments || !record->counts || !record->partials || record->cap <= 0 ||
        record->cap % kRecChunk)
        return PF_ERR_ARG;
    if ((int64_t)num_cams * ((height + 7) / 8) * ((width + 7) / 8) > (int64_t(1) << 25)) return PF_ERR_ARG;
    const bool per_camera = grads->texel_rgb_cam_stride != 0 && num_cams > 1;
    GroupLayout L;
    if (!group_layout(scene->num_cells, num_cams, width, height, record->cap, per_camera, L)) return PF_ERR_ARG;
    if (use_scratch) {
        // sizes past the static array are refused before any device work
        if (L.total > kScratchBytes) return PF_ERR_OVERFLOW;
        const int rc = scratch_for_current_device(&workspace);
        if (rc) return rc;
        workspace_bytes = kScratchBytes;
    }
    if (!workspace || workspace_bytes < L.total) return PF_ERR_OVERFLOW;
    if (scene->attr_dtype != PF_F32 && scene->attr_dtype != PF_F16) return PF_ERR_DTYPE;
    if (scene->face_dtype != PF_F32 && scene->face_dtype != PF_F16 && scene->face_dtype != PF_F16_PACKED)
        return PF_ERR_DTYPE;

    GroupParams q{};
    WalkParams &p = q.w;
    p.s = *scene;
    p.cams = reinterpret_cast<const float *>(cams);
    p.B = num_cams;
    p.W = width;
    p.H = height;
    p.tiles_h = (height + 7) / 8;
    p.tiles_per_cam = p.tiles_h * ((width + 7) / 8);
    p.total_tiles = (int64_t)p.tiles_per_cam * num_cams;
    p.thr = transmittance_threshold;
    p.max_steps = max_steps;
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
    q.part = reinterpret_cast<float4 *>(record->partials);
    q.header = record->header;
    q.S = p.total_tiles * 64;
    q.nb = (int)L.nb;
    q.nblk = (int)L.nblk;
    hipStream_t st = (hipStream_t)stream;
    const int rc = pf_walk_check_stamp(record, scene, cams, num_cams, p.total_tiles, status, st);
    if (rc) return rc;
    return run_grouped(q, L, reinterpret_cast<char *>(workspace), st);
}

// ----------------------------------- checkpoint and replay, grouped --
// The backward with no record from the forward: its outputs (colour and
// log T per ray) are the checkpoint, and the walk is replayed band by band
// (band_tiles tiles of 8x8 rays, in the cameras' tile order) by the
// recording forward into a band-sized record, whose segments the grouped
// pipeline above sorts by cell and differentiates.  Gradients accumulate
// over the bands (every gradient store is an atomic add).  The workspace
// holds the replay's colours, one band's record and its sorted entries, so
// its size follows the band, not the frame.
struct ReplayLayout {
    int64_t off_color, off_counts, off_segs, off_part, off_group, total;
    GroupLayout g;
};

static bool replay_layout(int64_t num_cells, int32_t num_cams, int32_t width, int32_t height, int32_t cap,
                          int64_t band, bool per_camera, ReplayLayout &R) {
    if (num_cams <= 0 || width <= 0 || height <= 0 || cap <= 0 || cap % kRecChunk || band <= 0) return false;
    const int64_t S = band * 64;
    if (!group_layout_slots(num_cells, S, num_cams, cap, per_camera, R.g)) return false;
    int64_t o = 0;
    R.off_color = o;
    o = align256(o + 12 * (int64_t)num_cams * width * height);
    R.off_counts = o;
    o = align256(o + 4 * S);
    R.off_segs = o;
    o = align256(o + 16 * S * cap);
    R.off_part = o;
    o = align256(o + 16 * S * cap);
    R.off_group = o;
    R.total = align256(o + R.g.total);
    return true;
}

static int64_t replay_band(int32_t num_cams, int32_t width, int32_t height, int32_t band_tiles) {
    const int64_t total = (int64_t)num_cams * ((height + 7) / 8) * ((width + 7) / 8);
    return band_tiles < total ? band_tiles : total;
}

extern "C" int64_t pf_walk_replay_grouped_workspace_bytes(int64_t num_cells, int32_t num_cams, int32_t width,
                                                          int32_t height, int32_t cap, int32_t band_tiles) {
    ReplayLayout R;
    if (num_cams <= 0 || width <= 0 || height <= 0 || band_tiles <= 0) return 0;
    // sized for per-camera buckets, so one workspace serves either layout
    return replay_layout(num_cells, num_cams, width, height, cap, replay_band(num_cams, width, height, band_tiles),
                         num_cams > 1, R)
               ? R.total
               : 0;
}

extern "C" int pf_walk_backward_replay_grouped(const pf_scene *scene, const pf_camera *cams, int32_t num_cams,
                                               int32_t width, int32_t height, float transmittance_threshold,
                                               int32_t max_steps, const float *color, const float *log_t,
                                               const float *grad_color, const float *grad_log_t,
                                               const pf_grads *grads, int32_t flags, int32_t record_cap,
                                               int32_t band_tiles, void *workspace, int64_t workspace_bytes,
                                               int32_t *status, void *stream) {
    if (!scene || !cams || num_cams <= 0 || width <= 0 || height <= 0 || !color || !grads || max_steps <= 0 ||
        band_tiles <= 0)
        return PF_ERR_ARG;
    if (!scene->spheres || !scene->nsigmas || !scene->texel_sites || !scene->texel_rgb || !scene->texel_height ||
        !scene->adjacency_offsets || scene->num_cells <= 0 || scene->num_cells >= 0x7FFFFFFF)
        return PF_ERR_ARG;
    if (scene->num_edges > 0 && (!scene->adjacency || !scene->faces)) return PF_ERR_ARG;
    if (scene->num_texel_sites != 8) return PF_ERR_ARG;
    if (!grads->spheres || !grads->nsigmas || !grads->texel_sites || !grads->texel_rgb || !grads->texel_height)
        return PF_ERR_ARG;
    // the walk launches' limit: fewer than 2^25 tiles per call
    if ((int64_t)num_cams * ((height + 7) / 8) * ((width + 7) / 8) >= (int64_t(1) << 25)) return PF_ERR_ARG;
    if (scene->attr_dtype != PF_F32 && scene->attr_dtype != PF_F16) return PF_ERR_DTYPE;
    if (scene->face_dtype != PF_F32 && scene->face_dtype != PF_F16 && scene->face_dtype != PF_F16_PACKED)
        return PF_ERR_DTYPE;
    const bool per_camera = grads->texel_rgb_cam_stride != 0 && num_cams > 1;
    const int64_t band = replay_band(num_cams, width, height, band_tiles);
    ReplayLayout R;
    if (!replay_layout(scene->num_cells, num_cams, width, height, record_cap, band, per_camera, R)) return PF_ERR_ARG;
    if (!workspace || workspace_bytes < R.total) return PF_ERR_OVERFLOW;
    char *ws = reinterpret_cast<char *>(workspace);

    GroupParams q{};
    WalkParams &p = q.w;
    p.s = *scene;
    p.cams = reinterpret_cast<const float *>(cams);
    p.B = num_cams;
    p.W = width;
    p.H = height;
    p.tiles_h = (height + 7) / 8;
    p.tiles_per_cam = p.tiles_h * ((width + 7) / 8);
    p.thr = transmittance_threshold;
    p.max_steps = max_steps;
    p.status = status;
    p.rec_segs = reinterpret_cast<int4 *>(ws + R.off_segs);
    p.rec_counts = reinterpret_cast<int *>(ws + R.off_counts);
    p.rec_cap = record_cap;
    p.rec_part = reinterpret_cast<float4 *>(ws + R.off_part);
    WalkParams f = p;  // the replay: the recording forward, colours to the workspace
    f.out.color = reinterpret_cast<float *>(ws + R.off_color);
    f.out.status = status;
    p.ck_color = color;
    p.ck_log_t = log_t;
    p.g_color = grad_color;
    p.g_log_t = grad_log_t;
    p.grads = *grads;
    p.exact_texel = (flags & PF_BWD_EXACT_TEXEL) != 0;
    q.part = p.rec_part;
    q.header = nullptr;
    q.nb = (int)R.g.nb;
    q.nblk = (int)R.g.nblk;
    hipStream_t st = (hipStream_t)stream;
    const int64_t total = (int64_t)p.tiles_per_cam * num_cams;
    for (int64_t t0 = 0; t0 < total; t0 += band) {
        const int64_t n = total - t0 < band ? total - t0 : band;
        f.tile0 = p.tile0 = t0;
        f.total_tiles = p.total_tiles = n;
        q.S = n * 64;
        int rc = walk_record_launch(f, st);
        if (rc) return rc;
        rc = run_grouped(q, R.g, ws + R.off_group, st);
        if (rc) return rc;
    }
    return PF_OK;
}
