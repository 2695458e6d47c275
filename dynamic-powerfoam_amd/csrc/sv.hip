// sv.hip — spherical-Voronoi colour pre-pass (SphericalVoronoi,
// powerfoam/color_fn.py:54-184) for gfx950: one thread per texel site.
//
// Forward (color_fn.py:54-101): direction = normalize(site - eye); outside
// the field of view (direction . forward < cutoff) the value is 0.5 and the
// weight sum 0; otherwise value = sum_i w_i v_i / sum_i w_i + 0.5 clamped at
// 0, w_i = exp(-temp_i |direction - axis_i|).  Backward (:105-184): the clamp
// masks the incoming gradient, lobes with w_i / sum < 1e-3 are skipped, and
// gradients go to the lobe axes, values and temperatures (the site
// gradient is left out, as in the reference, :180-183).
// Lobe arrays are lobe-major: (sv_dof, N*K, 3) and (sv_dof, N*K).
#include "pf_device.h"
#include "pf_hip.h"

namespace pf {

// Direction of a site from the eye; a zero vector stays zero (wp.normalize).
__device__ __forceinline__ f3 unit_or_zero(f3 v) {
    const float l = length(v);
    return l > 0.0f ? divs(v, l) : mk(0.0f, 0.0f, 0.0f);
}

template <typename A>
__global__ __launch_bounds__(256) void sv_fwd_kernel(const float *sites, int64_t NK, float3 eye, float3 fwd,
                                                     float cutoff, const A *axes, const A *vals, const A *temps,
                                                     int dof, A *value_out, A *wsum_out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NK) return;
    const f3 e = mk(eye.x, eye.y, eye.z);
    const f3 dir = unit_or_zero(mk(sites[3 * t], sites[3 * t + 1], sites[3 * t + 2]) - e);
    if (dot(dir, mk(fwd.x, fwd.y, fwd.z)) < cutoff) {
        value_out[3 * t] = (A)0.5f;
        value_out[3 * t + 1] = (A)0.5f;
        value_out[3 * t + 2] = (A)0.5f;
        wsum_out[t] = (A)0.0f;
        return;
    }
    float ws = 0.0f;
    f3 vs = mk(0, 0, 0);
    for (int i = 0; i < dof; ++i) {
        const int64_t q = (int64_t)i * NK + t;
        const f3 ax = mk((float)axes[3 * q], (float)axes[3 * q + 1], (float)axes[3 * q + 2]);
        const f3 v = mk((float)vals[3 * q], (float)vals[3 * q + 1], (float)vals[3 * q + 2]);
        const float w = pf_expf(-(float)temps[q] * length(dir - ax));
        ws = ws + w;
        vs = macc(vs, w, v);
    }
    wsum_out[t] = (A)ws;
    const f3 val = divs(vs, ws) + mk(0.5f, 0.5f, 0.5f);
    value_out[3 * t] = (A)fmax_(val.x, 0.0f);
    value_out[3 * t + 1] = (A)fmax_(val.y, 0.0f);
    value_out[3 * t + 2] = (A)fmax_(val.z, 0.0f);
}

template <typename A>
__global__ __launch_bounds__(256) void sv_bwd_kernel(const float *sites, int64_t NK, float3 eye, float3 fwd,
                                                     float cutoff, const A *axes, const A *vals, const A *temps,
                                                     int dof, const A *wsum_in, const A *value_in, const A *grad_in,
                                                     A *g_axes, A *g_vals, A *g_temps) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NK) return;
    const f3 e = mk(eye.x, eye.y, eye.z);
    const f3 dir = unit_or_zero(mk(sites[3 * t], sites[3 * t + 1], sites[3 * t + 2]) - e);
    if (dot(dir, mk(fwd.x, fwd.y, fwd.z)) < cutoff) return;
    const float ws = (float)wsum_in[t];
    f3 value = mk((float)value_in[3 * t], (float)value_in[3 * t + 1], (float)value_in[3 * t + 2]);
    f3 g = mk((float)grad_in[3 * t], (float)grad_in[3 * t + 1], (float)grad_in[3 * t + 2]);
    if (value.x <= 0.0f) g.x = 0.0f;
    if (value.y <= 0.0f) g.y = 0.0f;
    if (value.z <= 0.0f) g.z = 0.0f;
    value = value - mk(0.5f, 0.5f, 0.5f);
    const float dp_out = dot(g, value);
    for (int i = 0; i < dof; ++i) {
        const int64_t q = (int64_t)i * NK + t;
        const f3 ax = mk((float)axes[3 * q], (float)axes[3 * q + 1], (float)axes[3 * q + 2]);
        const f3 v = mk((float)vals[3 * q], (float)vals[3 * q + 1], (float)vals[3 * q + 2]);
        const float temp = (float)temps[q];
        const f3 dv = dir - ax;
        const float dist = length(dv);
        const float w = pf_expf(-temp * dist);
        if (w / ws < 1e-3f) continue;
        const f3 gval = divs(w * g, ws);
        g_vals[3 * q] = (A)gval.x;
        g_vals[3 * q + 1] = (A)gval.y;
        g_vals[3 * q + 2] = (A)gval.z;
        const f3 dwds = divs((temp * w) * dv, dist);
        const float dwdt = -dist * w;
        const float dp = dot(g, v);
        const f3 gax = divs(dwds * (dp - dp_out), ws);
        g_axes[3 * q] = (A)gax.x;
        g_axes[3 * q + 1] = (A)gax.y;
        g_axes[3 * q + 2] = (A)gax.z;
        g_temps[q] = (A)((dwdt * (dp - dp_out)) / ws);
    }
}

}  // namespace pf

using namespace pf;

extern "C" int pf_sv_forward(const float *texel_sites, int64_t num_sites, const float *eye, const float *forward,
                             float fov_cos_cutoff, const void *axes, const void *values, const void *temps,
                             int32_t sv_dof, int32_t dtype, void *value_out, void *wsum_out, void *stream) {
    if (!texel_sites || num_sites <= 0 || !eye || !forward || !axes || !values || !temps || sv_dof <= 0 ||
        !value_out || !wsum_out)
        return PF_ERR_ARG;
    const float3 e = make_float3(eye[0], eye[1], eye[2]), f = make_float3(forward[0], forward[1], forward[2]);
    const dim3 grid((unsigned)((num_sites + 255) / 256));
    hipStream_t st = (hipStream_t)stream;
    if (dtype == PF_F32)
        hipLaunchKernelGGL(sv_fwd_kernel<float>, grid, dim3(256), 0, st, texel_sites, num_sites, e, f, fov_cos_cutoff,
                           (const float *)axes, (const float *)values, (const float *)temps, sv_dof,
                           (float *)value_out, (float *)wsum_out);
    else if (dtype == PF_F16)
        hipLaunchKernelGGL(sv_fwd_kernel<_Float16>, grid, dim3(256), 0, st, texel_sites, num_sites, e, f,
                           fov_cos_cutoff, (const _Float16 *)axes, (const _Float16 *)values, (const _Float16 *)temps,
                           sv_dof, (_Float16 *)value_out, (_Float16 *)wsum_out);
    else return PF_ERR_DTYPE;
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}

extern "C" int pf_sv_backward(const float *texel_sites, int64_t num_sites, const float *eye, const float *forward,
                              float fov_cos_cutoff, const void *axes, const void *values, const void *temps,
                              int32_t sv_dof, int32_t dtype, const void *wsum, const void *value, const void *grad_value,
                              void *grad_axes, void *grad_values, void *grad_temps, void *stream) {
    if (!texel_sites || num_sites <= 0 || !eye || !forward || !axes || !values || !temps || sv_dof <= 0 || !wsum ||
        !value || !grad_value || !grad_axes || !grad_values || !grad_temps)
        return PF_ERR_ARG;
    const float3 e = make_float3(eye[0], eye[1], eye[2]), f = make_float3(forward[0], forward[1], forward[2]);
    const dim3 grid((unsigned)((num_sites + 255) / 256));
    hipStream_t st = (hipStream_t)stream;
    if (dtype == PF_F32)
        hipLaunchKernelGGL(sv_bwd_kernel<float>, grid, dim3(256), 0, st, texel_sites, num_sites, e, f, fov_cos_cutoff,
                           (const float *)axes, (const float *)values, (const float *)temps, sv_dof,
                           (const float *)wsum, (const float *)value, (const float *)grad_value, (float *)grad_axes,
                           (float *)grad_values, (float *)grad_temps);
    else if (dtype == PF_F16)
        hipLaunchKernelGGL(sv_bwd_kernel<_Float16>, grid, dim3(256), 0, st, texel_sites, num_sites, e, f,
                           fov_cos_cutoff, (const _Float16 *)axes, (const _Float16 *)values, (const _Float16 *)temps,
                           sv_dof, (const _Float16 *)wsum, (const _Float16 *)value, (const _Float16 *)grad_value,
                           (_Float16 *)grad_axes, (_Float16 *)grad_values, (_Float16 *)grad_temps);
    else return PF_ERR_DTYPE;
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}
