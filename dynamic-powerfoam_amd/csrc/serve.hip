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
// also replay the walk for the gradient of sum_pixels |C - target|^2
// (checkpoint: the colour just computed, kept in registers).
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

template <int DEG>
__global__ __launch_bounds__(256) void serve_kernel(const pf_serve_scene *scenes, const pf_serve_request *reqs,
                                                    const int64_t *desc, float thr, int max_steps, int *status) {
    const int nreq = (int)desc[0];
    const int64_t step_tiles = desc[1];
    const int64_t g = (int64_t)xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (g >= step_tiles || nreq <= 0) return;
    const int lane = threadIdx.x & 63;
    const int ri = __builtin_amdgcn_readfirstlane(find_request(reqs, nreq, g));
    const pf_serve_request &rq = reqs[ri];
    const int W = rq.cam.width, H = rq.cam.height;
    const int tile = rq.tile_begin + (int)(g - rq.step_tile0);
    const int tiles_h = (H + 7) / 8;
    const int pi = (tile % tiles_h) * 8 + (lane & 7);
    const int pj = (tile / tiles_h) * 8 + (lane >> 3);
    if (pi >= H || pj >= W) return;
    const int64_t pix = (int64_t)pi * W + pj;
    const pf_serve_scene &sc = scenes[rq.scene];
    Cam cam = load_cam(reinterpret_cast<const float *>(&rq.cam));
    if (rq.start_cell) cam.start = __builtin_amdgcn_readfirstlane(*rq.start_cell);
    const f3 d = ray_dir(cam, pi, pj);
    const SceneDev<float> s = scene_dev<float>(sc.scene);
    int st = 0;
    if (sc.sh) {
        const SHRGB<DEG> colors{sc.sh, cam.eye};
        const RayFwd r = walk_ray_fwd<float, _Float16, 8>(s, colors, cam.eye, d, cam.start, thr, 0.5f, max_steps,
                                                           nullptr, 0);
        rq.color[3 * pix] = r.rgb.x;
        rq.color[3 * pix + 1] = r.rgb.y;
        rq.color[3 * pix + 2] = r.rgb.z;
        if (rq.log_t) rq.log_t[pix] = r.log_t;
        st |= r.status;
        if (rq.mode == 1) {
            const f3 tg = mk(rq.target[3 * pix], rq.target[3 * pix + 1], rq.target[3 * pix + 2]);
            const f3 gc = 2.0f * (r.rgb - tg);
            const GradDev G{rq.grads.spheres, rq.grads.nsigmas, rq.grads.texel_sites, rq.grads.texel_rgb,
                            rq.grads.texel_height};
            uint64_t bytes = 0;
            st |= walk_ray_bwd<float, _Float16, 8>(s, colors, cam.eye, d, cam.start, thr, max_steps, r.rgb, gc, 0.0f,
                                                   G, false, bytes);
        }
    } else {
        const MemRGB<float> colors{reinterpret_cast<const float *>(sc.scene.texel_rgb)};
        const RayFwd r = walk_ray_fwd<float, _Float16, 8>(s, colors, cam.eye, d, cam.start, thr, 0.5f, max_steps,
                                                           nullptr, 0);
        rq.color[3 * pix] = r.rgb.x;
        rq.color[3 * pix + 1] = r.rgb.y;
        rq.color[3 * pix + 2] = r.rgb.z;
        if (rq.log_t) rq.log_t[pix] = r.log_t;
        st |= r.status;
        if (rq.mode == 1) {
            const f3 tg = mk(rq.target[3 * pix], rq.target[3 * pix + 1], rq.target[3 * pix + 2]);
            const f3 gc = 2.0f * (r.rgb - tg);
            const GradDev G{rq.grads.spheres, rq.grads.nsigmas, rq.grads.texel_sites, rq.grads.texel_rgb,
                            rq.grads.texel_height};
            uint64_t bytes = 0;
            st |= walk_ray_bwd<float, _Float16, 8>(s, colors, cam.eye, d, cam.start, thr, max_steps, r.rgb, gc, 0.0f,
                                                   G, false, bytes);
        }
    }
    if (st && status) atomicOr(status, st);
}

}  // namespace pf

using namespace pf;

extern "C" int pf_serve_step(const pf_serve_scene *scenes, const pf_serve_request *requests, const int64_t *desc,
                             int64_t tile_capacity, float transmittance_threshold, int32_t max_steps,
                             int32_t *status, void *stream) {
    if (!scenes || !requests || !desc || tile_capacity <= 0 || max_steps <= 0) return PF_ERR_ARG;
    const unsigned blocks = (unsigned)((tile_capacity + 3) / 4);
    hipLaunchKernelGGL(serve_kernel<3>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, scenes, requests, desc,
                       transmittance_threshold, max_steps, status);
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}
