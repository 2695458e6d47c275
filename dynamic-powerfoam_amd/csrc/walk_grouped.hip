// walk_grouped.hip — the walk backward from a record, grouped by cell.
//
// The recording forward stores, with each contributing segment, the
// transmittance before it and the ray's colour after it (the record's
// partials).  With those, a segment's gradients no longer depend on the
// other segments of its ray:
//   dL/d(delta_s) = g.(C - C_{<=s}) + g_logT - (g.c_s) T_s exp(delta_s)
// needs only the ray's incoming gradient g, g_logT and final colour C.  So
// the segments can be differentiated in any order.  This file sorts them by
// cell (a counting sort on the device: count, scan, scatter; the scatter
// writes each segment as a self-contained payload, its record entry, the
// ray's slot and the two scalars of the ray it needs, T_s g and
// g.(C - C_{<=s}) + g_logT, so the gradient pass reads the sorted segments
// in order and gathers nothing per ray) and differentiates them in cell
// order: each wave takes a contiguous range of
// the sorted segments, 64 at a time with every lane busy, so runs of one
// cell sit in neighbouring lanes.  A segmented scan across the lanes (DPP
// row shifts and row broadcasts) sums each gradient term over its run, and
// the run's last lane adds the sum into the run's LDS row; a row reaches
// global memory once per run, as one 64-lane atomic add.  A run that
// continues into the wave's next 64 segments keeps its row.  No LDS float
// atomic is issued per segment.
//
// Rays whose segment count exceeded the record's cap are replayed (the
// walk backward by checkpoint and replay, one lane per ray) by the first
// blocks of the same launch.
#include "pf_device.h"
#include "pf_hip.h"
#include "pf_launch.h"
#include "pf_walk.h"
#include "pf_walk_params.h"

#include <mutex>

namespace pf {

#ifndef PF_GROUP_WAVES
#define PF_GROUP_WAVES 4  // waves per SIMD of the grouped gradient kernel
#endif
constexpr int kGroupRows = 16;                  // LDS rows per wave (runs held at once)
constexpr int kGroupPitch = 8 + 7 * 8 + 1;      // K = 8: 64 terms, odd pitch
constexpr int kScanChunk = 4096;                // entries per scan block (256 x 16)
#ifndef PF_REPLAY_BLOCKS
#define PF_REPLAY_BLOCKS 256
#endif
constexpr int kReplayBlocks = PF_REPLAY_BLOCKS;  // blocks replaying rays past the cap
constexpr int kGroupBlocksPerCU = PF_GROUP_WAVES;

// One sorted segment, self-contained: its record entry, the two scalars of
// its ray that its gradients need, the ray's direction and camera, its ray
// and slot.  64 bytes, 64-byte aligned: the scatter writes whole 64-byte
// sectors (a partial sector costs HBM a read-modify-write).
struct alignas(64) GroupEntry {
    int4 seg;     // {cell, t_near bits, t_far bits, face ids}
    float4 coef;  // {T_s g, g.(C - C_<=s) + g_logT}
    float4 dirb;  // {ray direction, camera (int bits)}
    int4 aux;     // {ray, slot, 0, 0}
};
static_assert(sizeof(GroupEntry) == 64, "sorted segment entry");

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

// Slot -> the lane's camera, pixel and ray (tile_lane's map, from the slot).
struct SlotRay {
    int b, pix_i, pix_j;
    bool valid;
    int64_t ray;
};
__device__ __forceinline__ SlotRay slot_ray(const WalkParams &p, int64_t slot) {
    SlotRay r;
    const int64_t t = slot >> 6;
    const int lane = (int)(slot & 63);
    r.b = (int)(t / p.tiles_per_cam);
    const int tile = (int)(t % p.tiles_per_cam);
    r.pix_i = (tile % p.tiles_h) * 8 + (lane & 7);
    r.pix_j = (tile / p.tiles_h) * 8 + (lane >> 3);
    r.valid = r.pix_i < p.H && r.pix_j < p.W;
    r.ray = ((int64_t)r.b * p.H + r.pix_i) * p.W + r.pix_j;
    return r;
}

// The same map in 32-bit arithmetic, for slots below 2^31 (the grouped
// backward's records hold fewer than 2^31 entries, so fewer slots).
__device__ __forceinline__ SlotRay slot_ray32(const WalkParams &p, uint32_t slot) {
    SlotRay r;
    const uint32_t t = slot >> 6;
    const uint32_t lane = slot & 63u;
    const uint32_t tpc = (uint32_t)p.tiles_per_cam, th = (uint32_t)p.tiles_h;
    const uint32_t b = t / tpc;
    const uint32_t tile = t - b * tpc;
    const uint32_t tj = tile / th;
    const uint32_t ti = tile - tj * th;
    r.b = (int)b;
    r.pix_i = (int)(ti * 8 + (lane & 7));
    r.pix_j = (int)(tj * 8 + (lane >> 3));
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

__global__ void grp_zero_kernel(GroupParams q) {
    const int64_t n = (int64_t)q.nb + 1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        q.ws.bucket_off[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *q.ws.over_count = 0;
}

// One thread per slot: rays past the cap go to the replay list; the others
// count their recorded segments per bucket.  FILL: scatter each segment's
// record entry id to its bucket's next position instead.  Neighbouring
// rays (lanes of a tile) record the same cells at the same step, so the
// lanes of one bucket first agree on a leader (one ballot round per
// distinct bucket), and each leader makes one atomic for all of them;
// with FILL the leader's returned base plus the lane's rank among them is
// its position.
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
    const int lane = __lane_id();
    const uint64_t lt = (1ull << lane) - 1;
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
            uint64_t pend = __ballot(bk >= 0);
            int leader = 0, rank = 0, n = 0;
            while (pend) {
                const int lead = __ffsll((unsigned long long)pend) - 1;
                const int b = __builtin_amdgcn_readlane(bk, lead);
                const uint64_t mem = __ballot(bk == b);
                pend &= ~mem;
                if (bk == b) {
                    leader = lead;
                    rank = __builtin_popcountll(mem & lt);
                    n = __builtin_popcountll(mem);
                }
            }
            const bool lead_here = bk >= 0 && leader == lane;
            if (FILL) {
                int base = 0;
                if (lead_here) base = atomicAdd(q.ws.cursor + bk, n);
                base = __shfl(base, leader, 64);
                if (bk >= 0) {
                    const float4 pt = pc[j];
                    const f3 Tg = seg_Tg(g, pt.x);
                    const float Ln = seg_Lnext(g, gL, Cfin, mk(pt.y, pt.z, pt.w));
                    GroupEntry *en = q.ws.ent + (base + rank);
                    en->seg = rq;
                    en->coef = make_float4(Tg.x, Tg.y, Tg.z, Ln);
                    en->dirb = make_float4(dir.x, dir.y, dir.z, __int_as_float(sr.b));
                    en->aux = make_int4((int)sr.ray, (int)slot, 0, 0);
                }
            } else if (lead_here) {
                atomicAdd(q.ws.bucket_off + bk, n);
            }
        }
    }
}

// Exclusive scan of bucket_off[0 .. nb] in place, in chunks of kScanChunk:
// chunk totals, a one-block scan of the totals, then each chunk from its
// prefix (also copied into the scatter cursors).
__device__ __forceinline__ int block_incl_scan(int v, int *sh) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    if (lane == 63) sh[wave] = v;
    __syncthreads();
    int add = 0;
    for (int w = 0; w < wave; ++w) add += sh[w];
    __syncthreads();
    return v + add;
}

__global__ __launch_bounds__(256) void grp_scan_totals_kernel(GroupParams q) {
    __shared__ int sh[4];
    const int64_t n = (int64_t)q.nb + 1;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    int s = 0;
    for (int k = threadIdx.x; k < kScanChunk; k += 256) {
        const int64_t i = base + k;
        s += i < n ? q.ws.bucket_off[i] : 0;
    }
    s = block_incl_scan(s, sh);
    if (threadIdx.x == 255) q.ws.blk[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void grp_scan_blocks_kernel(GroupParams q) {
    __shared__ int sh[4];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < q.nblk; base += 256) {
        const int i = base + (int)threadIdx.x;
        const int v = i < q.nblk ? q.ws.blk[i] : 0;
        const int incl = block_incl_scan(v, sh);
        const int c = carry;
        if (i < q.nblk) q.ws.blk[i] = c + incl - v;
        __syncthreads();
        if (threadIdx.x == 255) carry = c + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) q.ws.blk[q.nblk] = carry;
}

__global__ __launch_bounds__(256) void grp_scan_chunks_kernel(GroupParams q) {
    __shared__ int sh[4];
    const int64_t n = (int64_t)q.nb + 1;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    // 16 consecutive entries per thread
    int v[16];
    int s = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int64_t i = base + threadIdx.x * 16 + k;
        v[k] = i < n ? q.ws.bucket_off[i] : 0;
        s += v[k];
    }
    const int incl = block_incl_scan(s, sh);
    int run = q.ws.blk[blockIdx.x] + incl - s;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int64_t i = base + threadIdx.x * 16 + k;
        if (i < n) {
            q.ws.bucket_off[i] = run;
            if (i < q.nb) q.ws.cursor[i] = run;
        }
        run += v[k];
    }
}

// ------------------------------------------------- segmented sum over runs --
template <int CTRL, int ROWS> __device__ __forceinline__ float dpp_f(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROWS, 0xF, false));
}

// Inclusive scan of a value over each run of lanes (runs: maximal ranges of
// lanes with one bucket).  m_k: all ones when lane - k still lies in the
// lane's run, else zero.  Row shifts by 1, 2, 4, 8 scan within each row of
// 16 lanes; lane 15's value then goes to row 1 and lane 47's to row 3
// (row_bcast:15), and lane 31's to rows 2 and 3 (row_bcast:31).  The run's
// last lane ends with its total.  All 64 lanes must call it together: a DPP
// read from a lane that is off in EXEC returns zero, so every shift runs on
// every lane and the run masks are applied to its result with an AND.
struct RunScan {
    int m1, m2, m4, m8, m15, m31;
    template <int CTRL, int ROWS> static __device__ __forceinline__ float step(float x, int m) {
        return x + __int_as_float(__float_as_int(dpp_f<CTRL, ROWS>(x)) & m);
    }
    __device__ __forceinline__ float operator()(float x) const {
        x = step<0x111, 0xF>(x, m1);
        x = step<0x112, 0xF>(x, m2);
        x = step<0x114, 0xF>(x, m4);
        x = step<0x118, 0xF>(x, m8);
        x = step<0x142, 0xA>(x, m15);
        x = step<0x143, 0xC>(x, m31);
        return x;
    }
};

// Sum of x over all 64 lanes, in lane 63 (the in-row scan by row shifts,
// then row_bcast:15 and row_bcast:31).  All 64 lanes must call it together.
__device__ __forceinline__ float wave_total(float x) {
    x += dpp_f<0x111, 0xF>(x);
    x += dpp_f<0x112, 0xF>(x);
    x += dpp_f<0x114, 0xF>(x);
    x += dpp_f<0x118, 0xF>(x);
    x += dpp_f<0x142, 0xA>(x);
    x += dpp_f<0x143, 0xC>(x);
    return x;
}

// Adds the lanes' neighbour-face terms (one side) to the neighbours'
// sphere gradients: one round per distinct neighbour among the lanes, whose
// terms are summed across the wave and added by one lane.  Neighbouring
// rays through a cell leave it by few faces, so a wave has few distinct
// neighbours per side and the atomics no longer pile up on one address.
__device__ __forceinline__ void add_face_terms(int nb, f3 dc, float dr, float *spheres) {
    uint64_t pend = __ballot(nb >= 0);
    while (pend) {
        const int lead = __ffsll((unsigned long long)pend) - 1;
        const int tgt = __builtin_amdgcn_readlane(nb, lead);
        const bool m = nb == tgt;
        pend &= ~__ballot(m);
        const float sx = wave_total(m ? dc.x : 0.0f);
        const float sy = wave_total(m ? dc.y : 0.0f);
        const float sz = wave_total(m ? dc.z : 0.0f);
        const float sr = wave_total(m ? dr : 0.0f);
        if (__lane_id() == 63) {
            float *ga = spheres + 4 * (int64_t)tgt;
            if (sx != 0.0f) atomic_add(ga + 0, sx);
            if (sy != 0.0f) atomic_add(ga + 1, sy);
            if (sz != 0.0f) atomic_add(ga + 2, sz);
            if (sr != 0.0f) atomic_add(ga + 3, sr);
        }
    }
}

// Sink of the grouped pass: every term is summed over its run; the run's
// last lane adds the total into the run's LDS row, or straight into global
// memory when the wave holds more runs than rows.
struct GroupSink {
    static constexpr bool kUniform = true;
    static constexpr bool kUniformHeight = true;
    RunScan scan;
    bool active;  // the lane holds a segment
    bool write;   // the lane ends its run
    float *row;   // the run's LDS row, or nullptr
    float *gs, *gn, *gsite, *grgb, *gh;
    __device__ __forceinline__ void add(int off, float *g, float v) const {
        const float t = scan(active ? v : 0.0f);
        if (write && t != 0.0f) {
            if (row) lds_add(row + off, t);
            else atomic_add(g, t);
        }
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
        add(32 + 3 * i, grgb + 3 * i, v.x);
        add(33 + 3 * i, grgb + 3 * i + 1, v.y);
        add(34 + 3 * i, grgb + 3 * i + 2, v.z);
    }
    __device__ __forceinline__ void site_div(int i, f3 v) const { site(i, v); }
    __device__ __forceinline__ void height(int i, float v) const { add(56 + i, gh + i, v); }
};

// Adds a run's row to its cell's gradients and clears it: lane k moves term
// k (sphere 0-3, [n, sigma] 4-7, sites 8-31, colours 32-55, heights 56-63).
__device__ __forceinline__ void flush_row(float *row, const GradDev &G, int cell, int b, int64_t rgb_stride) {
    const int k = __lane_id();
    lds_float *p = as_lds(row) + k;
    const float v = *p;
    *p = 0.0f;
    if (v != 0.0f) {
        const int64_t c = cell;
        float *dst;
        if (k < 4) dst = G.spheres + 4 * c + k;
        else if (k < 8) dst = G.nsigmas + 4 * c + (k - 4);
        else if (k < 32) dst = G.sites + 24 * c + (k - 8);
        else if (k < 56) dst = G.rgb + (int64_t)b * rgb_stride + 24 * c + (k - 32);
        else dst = G.height + 8 * c + (k - 56);
        atomic_add(dst, v);
    }
    __builtin_amdgcn_wave_barrier();
}

// A contiguous range [lo, hi) of the sorted segments, 64 at a time.
template <typename A, typename F>
__device__ __forceinline__ void grouped_range(const GroupParams &q, int64_t lo, int64_t hi, float *rows) {
    const WalkParams &p = q.w;
    const int lane = __lane_id();
    const int N = (int)p.s.num_cells;
    const SceneDev<A> s = scene_dev<A>(p.s);
    const GradDev G{p.grads.spheres, p.grads.nsigmas, p.grads.texel_sites, p.grads.texel_rgb, p.grads.texel_height};
    const int64_t rgb_stride = p.grads.texel_rgb_cam_stride;
    for (int k = lane; k < kGroupRows * kGroupPitch; k += 64) as_lds(rows)[k] = 0.0f;
    __builtin_amdgcn_wave_barrier();
    int carry = -1;  // bucket of the run held in row 0
    int carry_cell = 0, carry_b = 0;
    for (int64_t i0 = lo; i0 < hi; i0 += 64) {
        const int64_t i = i0 + lane;
        const bool act = i < hi;
        const GroupEntry *en = q.ws.ent + (act ? i : 0);
        const int4 rq = act ? en->seg : make_int4(0, 0, 0, 0);
        const float4 cf = act ? en->coef : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 db = act ? en->dirb : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        const int ray_b = __float_as_int(db.w);
        const int cell = act ? rq.x : 0;
        const int key = act ? bucket_of(q, cell, ray_b) : -2;
        // runs of equal buckets; lane 0 continues the carried run if equal
        const int prev = __shfl_up(key, 1, 64);
        const bool head = lane == 0 ? key != carry : key != prev;
        const uint64_t heads = __ballot(head);
        const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1);
        const uint64_t hle = heads & le;
        const int rs = hle ? 63 - __builtin_clzll(hle) : 0;
        const bool cont = !(heads & 1ull);
        const int run = __builtin_popcountll(hle) - (cont ? 0 : 1);
        const bool end = act && (lane == 63 || ((heads >> (lane + 1)) & 1ull));
        if (!cont && carry >= 0) {
            flush_row(rows, G, carry_cell, carry_b, rgb_stride);
            carry = -1;
        }
        const RunScan scan{-(int)(lane - 1 >= rs), -(int)(lane - 2 >= rs), -(int)(lane - 4 >= rs),
                           -(int)(lane - 8 >= rs), -(int)(lane >= 16 && (lane & ~15) - 1 >= rs),
                           -(int)(lane >= 32 && 31 >= rs)};

        // ---- the segment: recorded t values, partials, texel pass
        const float *eye = p.cams + 12 * ray_b;
        const f3 d = mk(db.x, db.y, db.z);
        const f3 o = mk(eye[0], eye[1], eye[2]);
        const MemRGB<A> colors{reinterpret_cast<const A *>(p.s.texel_rgb) + (int64_t)ray_b * p.s.texel_rgb_cam_stride};
        float tn = __int_as_float(rq.y), tf = __int_as_float(rq.z);
        int tn_id = (int)(int16_t)(rq.w & 0xffff);
        int tf_id = rq.w >> 16;
        const float4 sp = s.spheres[cell];
        f3 nrm;
        float sigma;
        ld4(s.nsig + 4 * (int64_t)cell, nrm, sigma);
        const f3 c = mk(sp.x, sp.y, sp.z);
        const float r = sp.w;
        const int e0 = s.off[cell];
        const int n_adj = s.off[cell + 1] - e0;
        const int64_t tbase = (int64_t)cell * 8;
        const float tn0 = tn;
        int tn0_id = act ? tn_id : -1, tf0_id = act ? tf_id : -1;
        const TexFwdT<8> tx = texel_fwd<A, 8>(o, d, tn, c, nrm, r, s.sites + 3 * tbase, colors.cell(cell, 8),
                                              s.height + tbase, 8);
        if (tx.dp >= 0.0f && tx.t_surf < tf) tf_id = n_adj;
        tf = tx.dp >= 0.0f ? fmin_(tx.t_surf, tf) : tf;
        if (tx.dp < 0.0f && tx.t_surf > tn) tn_id = n_adj;
        tn = tx.dp < 0.0f ? fmax_(tx.t_surf, tn) : tn;
        const float dt = tf - tn;
        const float dlt = (-sigma) * dt;
        const float e_dlt = pf_expf(dlt);
        const float alpha = 1.0f - e_dlt;
        const GroupSink sink{scan, act, end, run < kGroupRows ? rows + run * kGroupPitch : nullptr,
                             G.spheres + 4 * (int64_t)cell, G.nsigmas + 4 * (int64_t)cell, G.sites + 3 * tbase,
                             G.rgb + (int64_t)ray_b * rgb_stride + 3 * tbase, G.height + tbase};
        FaceTerms ft;
        segment_grads_to<A, F, 8>(s, colors, o, d,
                                  Seg{cell, c, nrm, r, sigma, tn0, dt, e_dlt, alpha, 0.0f, mk(0, 0, 0), tn_id,
                                      tf_id, tn0_id, tf0_id, e0, n_adj, tbase},
                                  tx, mk(cf.x, cf.y, cf.z), cf.w, G, p.exact_texel, sink, &ft);
        add_face_terms(act ? ft.nb[0] : -1, ft.dc[0], ft.dr[0], G.spheres);
        add_face_terms(act ? ft.nb[1] : -1, ft.dc[1], ft.dr[1], G.spheres);
        if (act && p.alg_bytes) {
            constexpr uint64_t kVisit = WalkBytes<A, F>::visit, kTexel = WalkBytes<A, F>::texel;
            // record entry and partials (read once), the sorted entry
            // (written and read), the cell, texels and gradient rows
            const uint64_t bytes = kSegRecordBytes + kSegPartBytes + 2 * sizeof(GroupEntry) + kVisit +
                                   8 * kTexel + segment_grad_bytes(8, tn0_id, tf0_id, n_adj);
            atomicAdd((unsigned long long *)(p.alg_bytes + en->aux.x), (unsigned long long)bytes);
        }
        __builtin_amdgcn_wave_barrier();

        // ---- complete runs reach global memory; the last one is carried
        const uint64_t actm = __ballot(act);
        uint64_t hm = heads & actm;
        const int nr = __builtin_popcountll(hm) + (cont ? 1 : 0);
        int r_idx = 0;
        if (cont) {
            if (nr > 1) {
                flush_row(rows, G, carry_cell, carry_b, rgb_stride);
                carry = -1;
            }
            r_idx = 1;
        }
        while (hm) {
            const int lead = __ffsll((unsigned long long)hm) - 1;
            hm &= hm - 1;
            const int rc = __builtin_amdgcn_readlane(cell, lead);
            const int rb = __builtin_amdgcn_readlane(ray_b, lead);
            const int rk = __builtin_amdgcn_readlane(key, lead);
            if (hm) {
                if (r_idx < kGroupRows) flush_row(rows + r_idx * kGroupPitch, G, rc, rb, rgb_stride);
            } else if (r_idx < kGroupRows) {
                // the chunk's last run: keep its row for the next chunk
                if (r_idx != 0) {
                    const int k = lane;
                    lds_float *src = as_lds(rows) + r_idx * kGroupPitch + k;
                    as_lds(rows)[k] = *src;
                    *src = 0.0f;
                    __builtin_amdgcn_wave_barrier();
                }
                carry = rk;
                carry_cell = rc;
                carry_b = rb;
            } else {
                carry = -1;  // its terms went to global memory directly
            }
            ++r_idx;
        }
    }
    if (carry >= 0) flush_row(rows, G, carry_cell, carry_b, rgb_stride);
}

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
    grouped_range<A, F>(q, lo, hi, rows + (threadIdx.x >> 6) * kGroupRows * kGroupPitch);
}

}  // namespace pf

using namespace pf;

static inline int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

struct GroupLayout {
    int64_t off_bucket, off_cursor, off_blk, off_over_count, off_over_list, off_ent, total;
    int64_t nb, nblk;
};

static bool group_layout(int64_t num_cells, int32_t num_cams, int32_t width, int32_t height, int32_t cap,
                         bool per_camera, GroupLayout &L) {
    if (num_cells <= 0 || num_cells >= 0x7FFFFFFF || num_cams <= 0 || width <= 0 || height <= 0 || cap <= 0)
        return false;
    const int64_t S = (int64_t)num_cams * ((height + 7) / 8) * ((width + 7) / 8) * 64;
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
    char *ws = reinterpret_cast<char *>(workspace);
    q.ws.bucket_off = reinterpret_cast<int *>(ws + L.off_bucket);
    q.ws.cursor = reinterpret_cast<int *>(ws + L.off_cursor);
    q.ws.blk = reinterpret_cast<int *>(ws + L.off_blk);
    q.ws.over_count = reinterpret_cast<int *>(ws + L.off_over_count);
    q.ws.over_list = reinterpret_cast<int *>(ws + L.off_over_list);
    q.ws.ent = reinterpret_cast<GroupEntry *>(ws + L.off_ent);
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    q.grouped_blocks = cus * kGroupBlocksPerCU;

    hipStream_t st = (hipStream_t)stream;
    int rc = pf_walk_check_stamp(record, scene, cams, num_cams, p.total_tiles, status, st);
    if (rc) return rc;
    const unsigned slot_blocks = (unsigned)((q.S + 255) / 256);
    hipLaunchKernelGGL(grp_zero_kernel, dim3(1024), dim3(256), 0, st, q);
    hipLaunchKernelGGL(grp_bucket_kernel<false>, dim3(slot_blocks), dim3(256), 0, st, q);
    hipLaunchKernelGGL(grp_scan_totals_kernel, dim3((unsigned)q.nblk), dim3(256), 0, st, q);
    hipLaunchKernelGGL(grp_scan_blocks_kernel, dim3(1), dim3(256), 0, st, q);
    hipLaunchKernelGGL(grp_scan_chunks_kernel, dim3((unsigned)q.nblk), dim3(256), 0, st, q);
    hipLaunchKernelGGL(grp_bucket_kernel<true>, dim3(slot_blocks), dim3(256), 0, st, q);
    const dim3 grid((unsigned)(kReplayBlocks + q.grouped_blocks));
    if (scene->attr_dtype == PF_F32) {
        if (scene->face_dtype == PF_F16_PACKED)
            hipLaunchKernelGGL((grp_grad_kernel<float, PackedF16>), grid, dim3(256), 0, st, q);
        else if (scene->face_dtype == PF_F16)
            hipLaunchKernelGGL((grp_grad_kernel<float, _Float16>), grid, dim3(256), 0, st, q);
        else hipLaunchKernelGGL((grp_grad_kernel<float, float>), grid, dim3(256), 0, st, q);
    } else {
        if (scene->face_dtype == PF_F16_PACKED)
            hipLaunchKernelGGL((grp_grad_kernel<_Float16, PackedF16>), grid, dim3(256), 0, st, q);
        else if (scene->face_dtype == PF_F16)
            hipLaunchKernelGGL((grp_grad_kernel<_Float16, _Float16>), grid, dim3(256), 0, st, q);
        else hipLaunchKernelGGL((grp_grad_kernel<_Float16, float>), grid, dim3(256), 0, st, q);
    }
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}
