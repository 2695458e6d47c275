// pf_grouped.h — the cell-grouped walk backward's building blocks, shared by
// the walk backward (walk_grouped.hip) and the serving step (serve.hip).
//
// The recording forward stores, with each contributing segment, the
// transmittance before it and the ray's colour after it (the record's
// partials).  With those, a segment's gradients no longer depend on the
// other segments of its ray:
//   dL/d(delta_s) = g.(C - C_{<=s}) + g_logT - (g.c_s) T_s exp(delta_s)
// needs only the ray's incoming gradient g, g_logT and final colour C.  So
// the segments can be differentiated in any order, and they are sorted by
// bucket (a cell, or a (camera or request, cell) pair) with a counting sort
// on the device: count, scan, scatter.  The scatter writes each segment as a
// self-contained 64-byte entry (GroupEntry).  The gradient pass then takes a
// contiguous range of sorted entries per wave, 64 at a time with every lane
// busy, so runs of one bucket sit in neighbouring lanes: a segmented scan
// across the lanes (DPP row shifts and row broadcasts) sums each gradient
// term over its run, and the run's last lane adds the sum into the run's LDS
// row; a row reaches global memory once per run, as one 64-lane atomic add.
// A run that continues into the wave's next 64 entries keeps its row.  No
// LDS float atomic is issued per segment.
#pragma once
#include "pf_device.h"
#include "pf_walk.h"

namespace pf {

constexpr int kGroupRows = 16;              // LDS rows per wave (runs held at once)
constexpr int kGroupPitch = 8 + 7 * 8 + 1;  // K = 8: 64 terms, odd pitch
constexpr int kScanChunk = 4096;            // bucket offsets per scan block (256 x 16)

// One sorted segment, self-contained: its record entry, the two scalars of
// its ray that its gradients need, the ray's direction and group (camera or
// request), its ray and slot.  64 bytes, 64-byte aligned: the scatter writes
// whole 64-byte sectors (a partial sector costs HBM a read-modify-write).
struct alignas(64) GroupEntry {
    int4 seg;     // {cell, t_near bits, t_far bits, face ids}
    float4 coef;  // {T_s g, g.(C - C_<=s) + g_logT}
    float4 dirb;  // {ray direction, group (int bits)}
    int4 aux;     // {ray or pixel, slot, 0, 0}
};
static_assert(sizeof(GroupEntry) == 64, "sorted segment entry");

// The lanes holding bucket bk >= 0 agree on a leader with one ballot round
// per distinct bucket; rank is the lane's position among them, n their
// number.  Neighbouring rays record the same cells at the same step, so one
// atomic by the leader serves all of them.
struct BucketLead {
    int leader, rank, n;
};
__device__ __forceinline__ BucketLead bucket_lead(int bk) {
    const uint64_t lt = (1ull << __lane_id()) - 1;
    uint64_t pend = __ballot(bk >= 0);
    BucketLead r{0, 0, 0};
    while (pend) {
        const int lead = __ffsll((unsigned long long)pend) - 1;
        const int b = __builtin_amdgcn_readlane(bk, lead);
        const uint64_t mem = __ballot(bk == b);
        pend &= ~mem;
        if (bk == b) {
            r.leader = lead;
            r.rank = __builtin_popcountll(mem & lt);
            r.n = __builtin_popcountll(mem);
        }
    }
    return r;
}

// --------------------------------------------------------- bucket offsets --
// The bucket counts, turned in place into exclusive offsets (offsets[nb] is
// the total), and the scatter cursors.  nb is nb_fixed, or *nb_dev when the
// count is known only on the device (the serving step's table); the grids
// are then sized for the largest count and blocks past the chunks return.
struct BucketScan {
    int *offsets;  // nb + 1
    int *cursor;   // nb
    int *blk;      // chunk totals (chunks + 1)
    const int64_t *nb_dev;
    int64_t nb_fixed;
    __device__ __forceinline__ int64_t nb() const { return nb_dev ? *nb_dev : nb_fixed; }
    __device__ __forceinline__ int64_t chunks() const { return (nb() + 1 + kScanChunk - 1) / kScanChunk; }
};

// Zeroes the counts and the nc counters at c.  (The kernels here have
// internal linkage: two translation units include this header.)
static __global__ void bucket_zero_kernel(BucketScan b, int *c, int nc) {
    const int64_t n = b.nb() + 1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b.offsets[i] = 0;
    if (blockIdx.x == 0 && (int)threadIdx.x < nc) c[threadIdx.x] = 0;
}

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

// Exclusive scan in chunks of kScanChunk: chunk totals, a one-block scan of
// the totals, then each chunk from its prefix (also copied into the scatter
// cursors).
static __global__ __launch_bounds__(256) void bucket_scan_totals_kernel(BucketScan b) {
    __shared__ int sh[4];
    if ((int64_t)blockIdx.x >= b.chunks()) return;
    const int64_t n = b.nb() + 1;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    int s = 0;
    for (int k = threadIdx.x; k < kScanChunk; k += 256) {
        const int64_t i = base + k;
        s += i < n ? b.offsets[i] : 0;
    }
    s = block_incl_scan(s, sh);
    if (threadIdx.x == 255) b.blk[blockIdx.x] = s;
}

static __global__ __launch_bounds__(256) void bucket_scan_blocks_kernel(BucketScan b) {
    __shared__ int sh[4];
    __shared__ int carry;
    const int nblk = (int)b.chunks();
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nblk; base += 256) {
        const int i = base + (int)threadIdx.x;
        const int v = i < nblk ? b.blk[i] : 0;
        const int incl = block_incl_scan(v, sh);
        const int c = carry;
        if (i < nblk) b.blk[i] = c + incl - v;
        __syncthreads();
        if (threadIdx.x == 255) carry = c + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) b.blk[nblk] = carry;
}

static __global__ __launch_bounds__(256) void bucket_scan_chunks_kernel(BucketScan b) {
    __shared__ int sh[4];
    if ((int64_t)blockIdx.x >= b.chunks()) return;
    const int64_t nb = b.nb(), n = nb + 1;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    // 16 consecutive entries per thread
    int v[16];
    int s = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int64_t i = base + threadIdx.x * 16 + k;
        v[k] = i < n ? b.offsets[i] : 0;
        s += v[k];
    }
    const int incl = block_incl_scan(s, sh);
    int run = b.blk[blockIdx.x] + incl - s;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int64_t i = base + threadIdx.x * 16 + k;
        if (i < n) {
            b.offsets[i] = run;
            if (i < nb) b.cursor[i] = run;
        }
        run += v[k];
    }
}

// Launches the three scan kernels; chunks_max: grid of the chunk kernels.
static inline void bucket_scan(const BucketScan &b, unsigned chunks_max, hipStream_t st) {
    hipLaunchKernelGGL(bucket_scan_totals_kernel, dim3(chunks_max), dim3(256), 0, st, b);
    hipLaunchKernelGGL(bucket_scan_blocks_kernel, dim3(1), dim3(256), 0, st, b);
    hipLaunchKernelGGL(bucket_scan_chunks_kernel, dim3(chunks_max), dim3(256), 0, st, b);
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

// Adds the lanes' neighbour-face terms (one side) to the neighbours'
// sphere gradients: one round per distinct neighbour among the lanes, whose
// terms are summed across the wave and added by one lane.  Neighbouring
// rays through a cell leave it by few faces, so a wave has few distinct
// neighbours per side and the atomics no longer pile up on one address.
// Each round sums the four terms over the wave at once: halving by lane ^ 32
// and lane ^ 16 (v_permlane32_swap / v_permlane16_swap) leaves term k's
// partial sums in row k (16 lanes), which four DPP row shifts add up in the
// row's last lane, and lanes 15, 31, 47 and 63 make the atomics (14 VALU
// instructions a round, where four whole-wave sums took 28).
__device__ __forceinline__ void add_face_terms(int nb, f3 dc, float dr, float *spheres) {
    uint64_t pend = __ballot(nb >= 0);
    const int lane = __lane_id();
    while (pend) {
        const int lead = __ffsll((unsigned long long)pend) - 1;
        const int tgt = __builtin_amdgcn_readlane(nb, lead);
        const bool m = nb == tgt;
        pend &= ~__ballot(m);
        const float x0 = m ? dc.x : 0.0f, x1 = m ? dc.y : 0.0f, x2 = m ? dc.z : 0.0f, x3 = m ? dr : 0.0f;
        // lanes < 32 keep terms 0 and 1, lanes >= 32 terms 2 and 3
        const auto p0 = __builtin_amdgcn_permlane32_swap(__float_as_uint(x0), __float_as_uint(x2), false, false);
        const auto p1 = __builtin_amdgcn_permlane32_swap(__float_as_uint(x1), __float_as_uint(x3), false, false);
        const float a0 = __uint_as_float(p0[0]) + __uint_as_float(p0[1]);
        const float a1 = __uint_as_float(p1[0]) + __uint_as_float(p1[1]);
        // rows 0 and 2 keep a0 (terms 0, 2), rows 1 and 3 a1 (terms 1, 3):
        // row k now holds term k
        const auto q = __builtin_amdgcn_permlane16_swap(__float_as_uint(a0), __float_as_uint(a1), false, false);
        float t = __uint_as_float(q[0]) + __uint_as_float(q[1]);
        t += dpp_f<0x111, 0xF>(t);
        t += dpp_f<0x112, 0xF>(t);
        t += dpp_f<0x114, 0xF>(t);
        t += dpp_f<0x118, 0xF>(t);
        if ((lane & 15) == 15 && t != 0.0f) atomic_add(spheres + 4 * (int64_t)tgt + (lane >> 4), t);
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

// Adds a run's row to its cell's gradients G (the texel colours of group b
// at b * rgb_stride) and clears it: lane k moves term k (sphere 0-3,
// [n, sigma] 4-7, sites 8-31, colours 32-55, heights 56-63).
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

// A contiguous range [lo, hi) of the sorted entries, 64 at a time.
//
// Ctx supplies what an entry's group (dirb.w) selects:
//   Ctx::kOneGroup  true: a chunk holds the entries of one group only (its
//                   scene then stays in scalar registers; the serving step,
//                   whose groups are requests on different scenes), false:
//                   the lanes of a chunk may hold different groups of one
//                   scene (the walk's cameras);
//   view(b0)        the chunk's view: s (SceneDev<A>), eye(b), colors(b),
//                   grads(b) and key(cell, b), the entry's bucket;
//   flush(row, cell, b)  adds a run's row to group b's gradients of cell;
//   exact           the exact-texel backward flag;
//   count_bytes(aux, bytes)  adds an entry's algorithmic bytes.
template <typename A, typename F, typename Ctx>
__device__ __forceinline__ void grouped_range(const Ctx &ctx, const GroupEntry *ent, int64_t lo, int64_t hi,
                                              float *rows) {
    const int lane = __lane_id();
    for (int k = lane; k < kGroupRows * kGroupPitch; k += 64) as_lds(rows)[k] = 0.0f;
    __builtin_amdgcn_wave_barrier();
    int carry = -1;  // bucket of the run held in row 0
    int carry_cell = 0, carry_b = 0;
    for (int64_t i0 = lo; i0 < hi;) {
        const int64_t i = i0 + lane;
        const bool inb = i < hi;
        const GroupEntry *en = ent + (inb ? i : i0);
        const int4 rq = inb ? en->seg : make_int4(0, 0, 0, 0);
        const float4 cf = inb ? en->coef : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 db = inb ? en->dirb : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
        const int ray_b = __float_as_int(db.w);
        int b0 = 0;
        bool act = inb;
        if constexpr (Ctx::kOneGroup) {
            // lane 0 holds an entry (i0 < hi); the chunk takes its group's
            // entries, a prefix of the lanes since groups own disjoint,
            // increasing bucket ranges
            b0 = __builtin_amdgcn_readfirstlane(ray_b);
            act = inb && ray_b == b0;
        }
        const int adv = Ctx::kOneGroup ? __builtin_popcountll(__ballot(act)) : 64;
        const auto v = ctx.view(b0);
        const SceneDev<A> &s = v.s;
        const int cell = act ? rq.x : 0;
        const int key = act ? v.key(cell, ray_b) : -2;
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
            ctx.flush(rows, carry_cell, carry_b);
            carry = -1;
        }
        const RunScan scan{-(int)(lane - 1 >= rs), -(int)(lane - 2 >= rs), -(int)(lane - 4 >= rs),
                           -(int)(lane - 8 >= rs), -(int)(lane >= 16 && (lane & ~15) - 1 >= rs),
                           -(int)(lane >= 32 && 31 >= rs)};

        // ---- the segment: recorded t values, partials, texel pass
        const f3 d = mk(db.x, db.y, db.z);
        const f3 o = v.eye(ray_b);
        const MemRGB<A> colors = v.colors(ray_b);
        const GradDev G = v.grads(ray_b);
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
                             G.rgb + 3 * tbase, G.height + tbase};
        FaceTerms ft;
        segment_grads_to<A, F, 8>(s, colors, o, d,
                                  Seg{cell, c, nrm, r, sigma, tn0, dt, e_dlt, alpha, 0.0f, mk(0, 0, 0), tn_id,
                                      tf_id, tn0_id, tf0_id, e0, n_adj, tbase},
                                  tx, mk(cf.x, cf.y, cf.z), cf.w, G, ctx.exact, sink, &ft);
        add_face_terms(act ? ft.nb[0] : -1, ft.dc[0], ft.dr[0], G.spheres);
        add_face_terms(act ? ft.nb[1] : -1, ft.dc[1], ft.dr[1], G.spheres);
        if (act) ctx.count_bytes(en, tn0_id, tf0_id, n_adj);
        __builtin_amdgcn_wave_barrier();

        // ---- complete runs reach global memory; the last one is carried
        const uint64_t actm = __ballot(act);
        uint64_t hm = heads & actm;
        const int nr = __builtin_popcountll(hm) + (cont ? 1 : 0);
        int r_idx = 0;
        if (cont) {
            if (nr > 1) {
                ctx.flush(rows, carry_cell, carry_b);
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
                if (r_idx < kGroupRows) ctx.flush(rows + r_idx * kGroupPitch, rc, rb);
            } else if (r_idx < kGroupRows) {
                // the chunk's last run: keep its row for the next chunk
                if (r_idx != 0) {
                    lds_float *src = as_lds(rows) + r_idx * kGroupPitch + lane;
                    as_lds(rows)[lane] = *src;
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
        i0 += adv;
    }
    if (carry >= 0) ctx.flush(rows, carry_cell, carry_b);
}

}  // namespace pf
