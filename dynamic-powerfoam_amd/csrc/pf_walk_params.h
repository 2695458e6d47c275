// pf_walk_params.h — launch parameters of the walk kernels and the map
// from a wave's tile to its lanes' pixels, shared by walk.hip and
// walk_grouped.hip.
#pragma once
#include "pf_device.h"
#include "pf_hip.h"
#include "pf_walk.h"

namespace pf {

struct WalkParams {
    pf_scene s;
    const float *cams;  // (B, 12) pf_camera
    int B, W, H, tiles_h, tiles_per_cam;
    // a launch covers tiles [tile0, tile0 + total_tiles) of the cameras'
    // tile order; record slots count from the launch's first tile
    int64_t total_tiles, tile0;
    float thr, depth_q;
    int max_steps;
    pf_walk_out out;
    // backward only
    const float *ck_color, *ck_log_t, *g_color, *g_log_t;
    pf_grads grads;
    int *status;
    bool exact_texel;
    uint64_t *alg_bytes;  // backward per-ray byte counter (may be null)
    // record of contributing segments, indexed by lane slot (tile * 64 +
    // lane): the forward writes it, the recorded backward reads it
    int4 *rec_segs;
    int *rec_counts;
    int rec_cap;
    float4 *rec_part;  // optional partials (T before, colour after)
    const int *rec_header;  // optional stamp; [7] holds the backward's verdict
};

// The recording forward over a launch's tiles (walk.hip), for callers in
// other files that replay a walk band by band.
int walk_record_launch(const WalkParams &p, hipStream_t st);

struct TileLane {
    int b, pix_i, pix_j;
    bool valid;
    int64_t ray, slot;
};

// block: the workgroup's index in the XCD-aware numbering of its tiles.
__device__ __forceinline__ TileLane tile_lane(const WalkParams &p, int waves_per_block, int block) {
    TileLane tl;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t t = (int64_t)block * waves_per_block + wave;
    tl.valid = t < p.total_tiles;
    const int64_t tt = tl.valid ? t : 0;
    const int64_t tg = p.tile0 + tt;
    tl.b = (int)(tg / p.tiles_per_cam);
    const int tile = (int)(tg % p.tiles_per_cam);
    const int tile_i = tile % p.tiles_h;
    const int tile_j = tile / p.tiles_h;
    tl.pix_i = tile_i * 8 + (lane & 7);
    tl.pix_j = tile_j * 8 + (lane >> 3);
    tl.valid = tl.valid && tl.pix_i < p.H && tl.pix_j < p.W;
    tl.ray = ((int64_t)tl.b * p.H + tl.pix_i) * p.W + tl.pix_j;
    tl.slot = tt * 64 + lane;
    return tl;
}

}  // namespace pf
