// pf_walk_deferred.h — the forward walk of a wave with deferred, dense
// texel shading (K = 8).
//
// Same results as walk_ray_fwd (pf_walk.h) bit for bit: visits, sequences,
// hashes, colour, log T, depth, status, the record and its partials.  What
// changes is when the texel passes run.  In walk_ray_fwd a wave runs the
// 8-texel height and colour passes at nearly every step, because some of
// its 64 lanes have a segment to shade, while only ~1 lane in 5 does.  Here
// a lane whose cell contributes queues a shading job in the wave's LDS and
// walks on; when a step's new jobs would take the queue past 64 (or a lane
// holds kJobsPerLane of them), the wave shades the queue with one job per
// lane and then applies each lane's jobs in walk order, with the
// arithmetic of walk_ray_fwd.
//
// Walking ahead is speculation: a queued segment may end the ray (T falls
// below the threshold), and the cells walked after it were never visited.
// Each job keeps the visit count, the sequence hash and the bytes up to its
// cell, so a ray that ends at a job takes those back and clears the
// sequence entries it wrote past its end.  The loop guard (Brent's method on
// (cell, log T), pf_walk.h CycleGuard) needs log T at every visit; a lane
// whose saved guard state is still unresolved, or that has queued jobs, and
// that meets the saved cell again waits for the next flush.  Termination is
// tested after each applied job, where the walk's loop test would see it.
#pragma once
#include "pf_walk.h"

namespace pf {

constexpr int kJobs = 64;        // queued shading jobs per wave: one round
constexpr int kJobsPerLane = 8;  // pending jobs per lane at most

// One wave's LDS (6.75 KB).
struct WaveShade {
    int4 seg[kJobs];    // the record entry: {cell, face-bounded t_near bits, t_far bits, face ids}
    uint4 snap[kJobs];  // {visits after the cell's visit | kFarBit, hash lo, hash hi, bytes since the lane's last job}
    float4 io[kJobs];    // in: {owner lane, -, -, -}; out: {t_surf, dp, sigma, rgb.x}
    float2 rgbyz[kJobs];  // out: rgb.y, rgb.z
    unsigned char list[64 * kJobsPerLane];  // each lane's pending jobs, in walk order
    float dir[64 * 3];  // each lane's ray direction
    // each lane's walk state while the wave shades ({cell, visits, hash lo,
    // hash hi}, {bytes since its last job, guard cell, guard power, guard
    // count}): parked here, the registers go to the texel passes, and the
    // walk keeps five waves per SIMD
    int4 park[64 * 2];
};
constexpr uint32_t kFarBit = 0x80000000u;  // the cell's sphere bound lies past 1e10 (REC: replay the ray)

template <typename A, typename F, typename RGB, bool REC>
__device__ __forceinline__ RayFwd walk_wave_fwd(const SceneDev<A> &s, const RGB &colors, f3 o, f3 d, int start,
                                                float thr, float depth_q, int max_steps, int *seq, int seq_cap,
                                                SegRecord rec, WaveShade &ws) {
    constexpr int K = 8;
    constexpr uint64_t kVisitBytes = WalkBytes<A, F>::visit, kFaceBytes = WalkBytes<A, F>::face,
                       kTexelBytes = WalkBytes<A, F>::texel;
    const int lane = __lane_id();
    const uint64_t lt_mask = (1ull << lane) - 1;
    // one camera per wave: the origin is the same in every lane
    o = mk(__int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(o.x))),
           __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(o.y))),
           __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(o.z))));
    const RowReader<F> rows(s.faces, s.E);
    ws.dir[3 * lane] = d.x;
    ws.dir[3 * lane + 1] = d.y;
    ws.dir[3 * lane + 2] = d.z;

    // committed state: every job so far applied
    RayFwd out{mk(0, 0, 0), 0.0f, 0.0f, 0, 0, kFnvOffset, 0, 0};
    bool depth_found = false;
    float trans_now = 1.0f;  // pf_expf(out.log_t), kept as the log T changes
    // speculative walk state
    int prim = start, visits = 0, spec_status = 0;
    uint64_t hash = kFnvOffset;
    uint32_t bytes_since = 0;  // between two jobs: well under 2^32
    // the loop guard; its saved log T is unresolved while jobs queued before
    // the save are pending: it is the log T after the first snap_jobs of them
    int g_cell = kNoCell, snap_jobs = 0;
    float g_logt = 0.0f;
    bool g_unres = false;
    uint32_t g_power = 1, g_lam = 1;
    int npend = 0;
    bool term = false;  // ended by a job: its visits and hash are the ray's
    // the walk loop's first test (T = 1 before any segment)
    bool active = !(pf_expf(0.0f) < thr);
    int count = 0;  // queued jobs (the same in every lane)

    while (true) {
        bool stalled = false, push = false;
        int4 jseg = make_int4(0, 0, 0, 0);
        uint32_t jfar = 0;
        if (active) {
            if (prim == kNoCell) {
                active = false;
            } else if ((unsigned)prim >= (unsigned)s.N || visits >= max_steps) {
                spec_status |= PF_STATUS_STEP_CAP;
                active = false;
            } else if (prim == g_cell && (npend > 0 || g_unres)) {
                stalled = true;  // log T at this visit is not known yet
            } else if (prim == g_cell && out.log_t == g_logt) {
                spec_status |= PF_STATUS_CYCLE;
                active = false;
            } else {
                if (g_lam == g_power) {
                    g_cell = prim;
                    g_logt = out.log_t;
                    g_unres = npend > 0;
                    snap_jobs = npend;
                    g_power <<= 1;
                    g_lam = 0;
                }
                ++g_lam;
                if (seq && visits < seq_cap) seq[visits] = prim;
                ++visits;
                hash = (hash ^ (uint64_t)(uint32_t)prim) * kFnvPrime;

                const float4 sp = s.spheres[prim];
                const float sigma = (float)s.nsig[4 * (int64_t)prim + 3];
                const f3 c = mk(sp.x, sp.y, sp.z);
                const float r = sp.w;
                SphereHit sh = ray_sphere(o, d, c, r);
                if (length(c - o) < 4.0f * r) sh.hit = false;
                float tn = sh.tn, tf = sh.tf;
                int tn_id = 0, tf_id = 0;
                if constexpr (REC) {
                    tn_id = tn < 1e-8f ? -2 : -1;
                    tf_id = tf < 1e-8f ? -2 : -1;
                }
                const int e0 = s.off[prim];
                const int e1 = s.off[prim + 1];
                bytes_since += (uint32_t)(kVisitBytes + (uint64_t)(e1 - e0) * kFaceBytes);
                FaceFold ff;
                rows.scan(e0, e1, [&](int e, f3 fn, float fw, bool valid) { ff.add(e, o, d, fn, fw, valid); });
                const int exit_e = ff.exit_face();
                const float pt_far = ff.exit_t();
                if constexpr (REC) tn_id = ff.near_id(tn, e0, tn_id);
                tf = ff.t_far(tf);
                tn = ff.t_near(tn);
                const int next = exit_e >= 0 ? face_neighbour<F>(s.faces, s.adj, exit_e) : kNoCell;
                if (sh.hit && !(tn > tf) && !(sigma < 1e-3f)) {
                    if constexpr (REC) {
                        // the far face is the exit face when the exit comes
                        // first (walk_ray_fwd)
                        if (exit_e >= 0 && pt_far < sh.tf) tf_id = exit_e - e0;
                    }
                    bytes_since += (uint32_t)(K * kTexelBytes);
                    jseg = pack_segment(prim, tn, tf, tn_id, tf_id);
                    jfar = !(sh.tf <= 1e10f) ? kFarBit : 0u;
                    push = true;
                }
                prim = next;
            }
        }
        // shade the queue (at most one job per lane of the wave) before this
        // step's jobs if they would overfill it or a lane's list, if a lane
        // waits for log T, or if no lane walks on; a lane that the queue ends
        // drops its new job with its walk
        const uint64_t pm = __ballot(push);
        if (count + __builtin_popcountll(pm) > kJobs || __ballot(push && npend == kJobsPerLane) != 0 ||
            __ballot(stalled) != 0 || (__ballot(active) == 0 && count > 0)) {
            ws.park[2 * lane] = make_int4(prim, visits, (int)(uint32_t)hash, (int)(uint32_t)(hash >> 32));
            ws.park[2 * lane + 1] = make_int4((int)bytes_since, g_cell, (int)g_power, (int)g_lam);
            __asm__ volatile("" ::: "memory");  // reloaded below, not kept in registers
            __builtin_amdgcn_wave_barrier();
            // shade the queue, one job per lane of the wave
            const uint64_t ex = __ballot(true);
            const int nlanes = __builtin_popcountll(ex);
            const int rank = __builtin_popcountll(ex & lt_mask);
            for (int j0 = 0; j0 < count; j0 += nlanes) {
                const int j = j0 + rank;
                if (j < count) {
                    const int cell = ws.seg[j].x;
                    const float tn_face = __int_as_float(ws.seg[j].y);
                    const int owner = __float_as_int(ws.io[j].x);
                    const f3 dj = mk(ws.dir[3 * owner], ws.dir[3 * owner + 1], ws.dir[3 * owner + 2]);
                    const float4 sp = s.spheres[cell];
                    f3 nrm;
                    float sg;
                    ld4(s.nsig + 4 * (int64_t)cell, nrm, sg);
                    const int64_t tbase = (int64_t)cell * K;
                    const TexFwdT<K> tx = texel_fwd<A, K>(o, dj, tn_face, mk(sp.x, sp.y, sp.z), nrm, sp.w,
                                                          s.sites + 3 * tbase, colors.cell(cell, K),
                                                          s.height + tbase, K);
                    ws.io[j] = make_float4(tx.t_surf, tx.dp, sg, tx.rgb.x);
                    ws.rgbyz[j] = make_float2(tx.rgb.y, tx.rgb.z);
                }
            }
            __builtin_amdgcn_wave_barrier();
            // apply each lane's jobs in walk order, with walk_ray_fwd's
            // arithmetic
            for (int k = 0; __ballot(k < npend) != 0; ++k) {
                if (k < npend && !term) {
                    const int j = ws.list[lane * kJobsPerLane + k];
                    const int4 q = ws.seg[j];
                    const uint4 sn = ws.snap[j];
                    const float4 io = ws.io[j];
                    const float2 yz = ws.rgbyz[j];
                    const f3 rgb = mk(io.w, yz.x, yz.y);
                    const int cell = q.x;
                    const float tn_face = __int_as_float(q.y), tf_face = __int_as_float(q.z);
                    const float trans = trans_now;
                    const float tf = io.y >= 0.0f ? fmin_(io.x, tf_face) : tf_face;
                    const float tn = io.y < 0.0f ? fmax_(io.x, tn_face) : tn_face;
                    out.bytes += sn.w;
                    const float dt = tf - tn;
                    if (dt > 0.0f) {
                        int64_t rec_at = -1;
                        if constexpr (REC) {
                            if (sn.x & kFarBit) out.segs = out.segs > rec.cap ? out.segs : rec.cap;
                            if (out.segs < rec.cap) {
                                rec_at = rec_index(out.segs, rec.slot, rec.S);
                                rec.rec[rec_at] = q;
                                out.bytes += kSegRecordBytes;
                            }
                            ++out.segs;
                        }
                        const float sigma = io.z;
                        const float dlt = (-sigma) * dt;
                        const float alpha = 1.0f - pf_expf(dlt);
                        out.rgb.x = __builtin_fmaf(rgb.x * alpha, trans, out.rgb.x);
                        out.rgb.y = __builtin_fmaf(rgb.y * alpha, trans, out.rgb.y);
                        out.rgb.z = __builtin_fmaf(rgb.z * alpha, trans, out.rgb.z);
                        out.log_t += dlt;
                        if constexpr (REC) {
                            if (rec.part && rec_at >= 0) {
                                rec.part[rec_at] = make_float4(trans, out.rgb.x, out.rgb.y, out.rgb.z);
                                out.bytes += kSegPartBytes;
                            }
                        }
                        const float next_trans = pf_expf(out.log_t);
                        trans_now = next_trans;
                        if (!depth_found && next_trans < depth_q) {
                            out.depth = tn + pf_logf(trans / depth_q) / sigma;
                            depth_found = true;
                        }
                        if (next_trans < thr) {  // the walk ends after this cell
                            term = true;
                            active = false;
                            // the cells walked past it were never visited
                            const int v_end = (int)(sn.x & ~kFarBit);
                            if (seq)
                                for (int v = v_end; v < ws.park[2 * lane].y && v < seq_cap; ++v) seq[v] = -1;
                            ws.park[2 * lane].y = v_end;
                            ws.park[2 * lane].z = (int)sn.y;
                            ws.park[2 * lane].w = (int)sn.z;
                        }
                    }
                    if (g_unres && snap_jobs == k + 1) {
                        g_logt = out.log_t;
                        g_unres = false;
                    }
                }
            }
            count = 0;
            npend = 0;
            __builtin_amdgcn_wave_barrier();
            __asm__ volatile("" ::: "memory");
            const int4 pk0 = ws.park[2 * lane], pk1 = ws.park[2 * lane + 1];
            prim = pk0.x;
            visits = pk0.y;
            hash = (uint64_t)(uint32_t)pk0.z | ((uint64_t)(uint32_t)pk0.w << 32);
            bytes_since = (uint32_t)pk1.x;
            g_cell = pk1.y;
            g_power = (uint32_t)pk1.z;
            g_lam = (uint32_t)pk1.w;
            d = mk(ws.dir[3 * lane], ws.dir[3 * lane + 1], ws.dir[3 * lane + 2]);
        }
        const uint64_t qm = __ballot(push && !term);
        if (push && !term) {
            const int j = count + __builtin_popcountll(qm & lt_mask);
            ws.seg[j] = jseg;
            ws.snap[j] = make_uint4((uint32_t)visits | jfar, (uint32_t)hash, (uint32_t)(hash >> 32),
                                    (uint32_t)bytes_since);
            ws.io[j] = make_float4(__int_as_float(lane), 0.0f, 0.0f, 0.0f);
            ws.list[lane * kJobsPerLane + npend] = (unsigned char)j;
            ++npend;
            bytes_since = 0;
        }
        count += __builtin_popcountll(qm);
        if (__ballot(active) == 0 && count == 0) break;
    }
    if (term) {  // visits and hash were set back to the ending job's
        out.visits = visits;
        out.hash = hash;
    } else {
        out.visits = visits;
        out.hash = hash;
        out.status = spec_status;
        out.bytes += bytes_since;
    }
    return out;
}

}  // namespace pf
