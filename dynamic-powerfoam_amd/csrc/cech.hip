// cech.hip — Čech-complex adjacency (pairs of overlapping spheres) for
// gfx950, the graph the rasterizer clips with.
//
// Semantics follow AABBTree.build_cech_complex (powerfoam/bvh.py:7-123): a
// sphere is stored as its AABB [c - r, c + r] and recovered as
// c = 0.5 (lo + hi), r = 0.5 (hi.x - lo.x) (:17-18, :29-30); j != i is a
// neighbour of i when |c_i - c_j| < r_i + r_j (:31-33).  The reference finds
// candidates with a Warp BVH; here a uniform grid of cell size 2 r_max
// (every overlapping pair lies in adjacent grid cells) does it with
// counting-sort buckets, and each CSR row is sorted ascending, which fixes
// the row order the BVH traversal left unspecified.
#include "pf_device.h"
#include "pf_hip.h"

namespace pf {

struct Ball { f3 c; float r; };

// bvh.py:77-82 then :17-18
__device__ __forceinline__ Ball recover(const float *centers, const float *radii, int64_t i) {
    const float r = radii[i];
    const f3 c = mk(centers[3 * i], centers[3 * i + 1], centers[3 * i + 2]);
    const f3 lo = mk(c.x - r, c.y - r, c.z - r), hi = mk(c.x + r, c.y + r, c.z + r);
    return Ball{0.5f * (lo + hi), 0.5f * (hi.x - lo.x)};
}

__device__ __forceinline__ int ord_f(float x) {
    const int u = __float_as_int(x);
    return u >= 0 ? u : (u ^ 0x7fffffff);
}
__device__ __forceinline__ float unord_f(int u) { return __int_as_float(u >= 0 ? u : (u ^ 0x7fffffff)); }

__global__ void bounds_init_kernel(int *b) {
    const int t = threadIdx.x;
    if (t < 3) b[t] = 0x7fffffff;
    else if (t < 7) b[t] = (int)0x80000000;
}

__global__ __launch_bounds__(256) void bounds_kernel(const float *centers, const float *radii, int64_t N, int *b) {
    f3 lo = mk(__builtin_inff(), __builtin_inff(), __builtin_inff());
    f3 hi = mk(-__builtin_inff(), -__builtin_inff(), -__builtin_inff());
    float rmax = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const Ball s = recover(centers, radii, i);
        lo = mk(fmin_(lo.x, s.c.x), fmin_(lo.y, s.c.y), fmin_(lo.z, s.c.z));
        hi = mk(fmax_(hi.x, s.c.x), fmax_(hi.y, s.c.y), fmax_(hi.z, s.c.z));
        rmax = fmax_(rmax, s.r);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo.x = fmin_(lo.x, __shfl_xor(lo.x, off, 64));
        lo.y = fmin_(lo.y, __shfl_xor(lo.y, off, 64));
        lo.z = fmin_(lo.z, __shfl_xor(lo.z, off, 64));
        hi.x = fmax_(hi.x, __shfl_xor(hi.x, off, 64));
        hi.y = fmax_(hi.y, __shfl_xor(hi.y, off, 64));
        hi.z = fmax_(hi.z, __shfl_xor(hi.z, off, 64));
        rmax = fmax_(rmax, __shfl_xor(rmax, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(b + 0, ord_f(lo.x));
        atomicMin(b + 1, ord_f(lo.y));
        atomicMin(b + 2, ord_f(lo.z));
        atomicMax(b + 3, ord_f(hi.x));
        atomicMax(b + 4, ord_f(hi.y));
        atomicMax(b + 5, ord_f(hi.z));
        atomicMax(b + 6, ord_f(rmax));
    }
}

__global__ void bounds_finish_kernel(const int *b, float *out) {
    const int t = threadIdx.x;
    if (t < 7) out[t] = unord_f(b[t]);
}

struct Grid {
    float ox, oy, oz, cell;
    int nx, ny, nz;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ void cell_of(const Grid &g, f3 c, int &ix, int &iy, int &iz) {
    ix = clampi((int)__builtin_floorf((c.x - g.ox) / g.cell), 0, g.nx - 1);
    iy = clampi((int)__builtin_floorf((c.y - g.oy) / g.cell), 0, g.ny - 1);
    iz = clampi((int)__builtin_floorf((c.z - g.oz) / g.cell), 0, g.nz - 1);
}

__global__ __launch_bounds__(256) void grid_count_kernel(const float *centers, const float *radii, int64_t N, Grid g,
                                                         int *gcount, int *key) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const Ball s = recover(centers, radii, i);
    int ix, iy, iz;
    cell_of(g, s.c, ix, iy, iz);
    const int k = (iz * g.ny + iy) * g.nx + ix;
    key[i] = k;
    atomicAdd(gcount + k + 1, 1);
}

__global__ __launch_bounds__(256) void grid_scatter_kernel(int64_t N, const int *key, const int64_t *gstart,
                                                           int *gfill, int *items) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int k = key[i];
    items[gstart[k] + atomicAdd(gfill + k, 1)] = (int)i;
}

// Visit every j != i whose ball overlaps ball i (bvh.py:23-33 / :57-69).
template <typename Fn>
__device__ __forceinline__ void for_overlaps(const float *centers, const float *radii, const Grid &g,
                                             const int64_t *gstart, const int *items, int64_t i, Fn fn) {
    const Ball s = recover(centers, radii, i);
    int ix, iy, iz;
    cell_of(g, s.c, ix, iy, iz);
    for (int z = iz - 1; z <= iz + 1; ++z) {
        if (z < 0 || z >= g.nz) continue;
        for (int y = iy - 1; y <= iy + 1; ++y) {
            if (y < 0 || y >= g.ny) continue;
            for (int x = ix - 1; x <= ix + 1; ++x) {
                if (x < 0 || x >= g.nx) continue;
                const int k = (z * g.ny + y) * g.nx + x;
                for (int64_t q = gstart[k]; q < gstart[k + 1]; ++q) {
                    const int j = items[q];
                    if (j == i) continue;
                    const Ball t = recover(centers, radii, j);
                    if (length(s.c - t.c) < s.r + t.r) fn(j);
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) void nbr_count_kernel(const float *centers, const float *radii, int64_t N, Grid g,
                                                        const int64_t *gstart, const int *items, int *counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int n = 0;
    for_overlaps(centers, radii, g, gstart, items, i, [&](int) { ++n; });
    counts[i + 1] = n;
}

__global__ __launch_bounds__(256) void nbr_write_kernel(const float *centers, const float *radii, int64_t N, Grid g,
                                                        const int64_t *gstart, const int *items,
                                                        const int64_t *row_off, int *adj, int *off32) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > N) return;
    off32[i] = (int)row_off[i];
    if (i == N) return;
    const int64_t b = row_off[i], e = row_off[i + 1];
    int64_t w = b;
    for_overlaps(centers, radii, g, gstart, items, i, [&](int j) {
        if (w < e) adj[w++] = j;
    });
    for (int64_t a = b + 1; a < e; ++a) {  // rows are short: insertion sort
        const int v = adj[a];
        int64_t q = a - 1;
        while (q >= b && adj[q] > v) {
            adj[q + 1] = adj[q];
            --q;
        }
        adj[q + 1] = v;
    }
}

// Inclusive prefix sum of n ints into int64, in three launches over chunks
// of kScanChunk, with no workspace: each chunk's total is parked in its
// last output slot, one block scans the totals into each chunk's first
// slot, and each block then scans its chunk from that prefix.  Reads and
// writes are coalesced (a single block walking per-thread ranges, as this
// replaced, took 15.8 ms for 16M entries).
constexpr int kScanChunk = 4096;
constexpr int kScanThreads = 256;

__device__ __forceinline__ int64_t block_scan_incl(int64_t v, int64_t *sh) {
    // inclusive scan over the block's kScanThreads values
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int64_t u = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += u;
        __syncthreads();
    }
    const int64_t r = sh[t];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanThreads) void scan_totals_kernel(const int *in, int64_t n, int64_t *out) {
    __shared__ int64_t sh[kScanThreads];
    const int64_t b = (int64_t)blockIdx.x * kScanChunk;
    const int64_t e = b + kScanChunk < n ? b + kScanChunk : n;
    int64_t s = 0;
    for (int64_t i = b + threadIdx.x; i < e; i += kScanThreads) s += in[i];
    s = block_scan_incl(s, sh);
    if (threadIdx.x == kScanThreads - 1) out[e - 1] = s;
}

__device__ __forceinline__ int64_t chunk_last(int64_t c, int64_t n) {
    return (c + 1) * kScanChunk < n ? (c + 1) * kScanChunk - 1 : n - 1;
}

__global__ __launch_bounds__(1024) void scan_prefixes_kernel(int64_t n, int64_t *out) {
    __shared__ int64_t part[1024];
    const int64_t nch = (n + kScanChunk - 1) / kScanChunk;
    const int t = threadIdx.x;
    const int64_t per = (nch + 1023) / 1024;
    const int64_t c0 = t * per, c1 = c0 + per < nch ? c0 + per : nch;
    int64_t s = 0;
    for (int64_t c = c0; c < c1; ++c) s += out[chunk_last(c, n)];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t > 0 ? part[t - 1] : 0;
    for (int64_t c = c0; c < c1; ++c) {
        // a chunk's first and last slots coincide only for a one-entry
        // chunk, whose total this thread reads before writing the prefix
        const int64_t tot = out[chunk_last(c, n)];
        out[c * kScanChunk] = run;
        run += tot;
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_chunks64_kernel(const int *in, int64_t n, int64_t *out) {
    __shared__ int64_t sh[kScanThreads];
    __shared__ int64_t base;
    const int64_t b = (int64_t)blockIdx.x * kScanChunk;
    const int64_t e = b + kScanChunk < n ? b + kScanChunk : n;
    if (threadIdx.x == 0) base = out[b];
    __syncthreads();
    int64_t carry = base;
    for (int64_t r = b; r < e; r += kScanThreads) {
        const int64_t i = r + threadIdx.x;
        const int64_t incl = block_scan_incl(i < e ? in[i] : 0, sh);
        if (i < e) out[i] = carry + incl;
        if (threadIdx.x == kScanThreads - 1) sh[0] = incl;  // this round's total
        __syncthreads();
        carry += sh[0];
        __syncthreads();
    }
}

static void scan64(const int *in, int64_t n, int64_t *out, hipStream_t st) {
    const unsigned nch = (unsigned)((n + kScanChunk - 1) / kScanChunk);
    hipLaunchKernelGGL(scan_totals_kernel, dim3(nch), dim3(kScanThreads), 0, st, in, n, out);
    hipLaunchKernelGGL(scan_prefixes_kernel, dim3(1), dim3(1024), 0, st, n, out);
    hipLaunchKernelGGL(scan_chunks64_kernel, dim3(nch), dim3(kScanThreads), 0, st, in, n, out);
}

}  // namespace pf

using namespace pf;

static int ok_() { return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP; }

extern "C" int pf_cech_bounds(const float *centers, const float *radii, int64_t num_cells, int32_t *workspace,
                              float *bounds, void *stream) {
    if (!centers || !radii || num_cells <= 0 || !workspace || !bounds) return PF_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(bounds_init_kernel, dim3(1), dim3(64), 0, st, workspace);
    int blocks = (int)((num_cells + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(bounds_kernel, dim3(blocks), dim3(256), 0, st, centers, radii, num_cells, workspace);
    hipLaunchKernelGGL(bounds_finish_kernel, dim3(1), dim3(64), 0, st, workspace, bounds);
    return ok_();
}

static Grid make_grid(const pf_cech_grid *g) {
    return Grid{g->origin[0], g->origin[1], g->origin[2], g->cell, g->dims[0], g->dims[1], g->dims[2]};
}

extern "C" int pf_cech_count(const float *centers, const float *radii, int64_t num_cells, const pf_cech_grid *grid,
                             int32_t *grid_counts, int32_t *grid_fill, int32_t *cell_key, int64_t *grid_start,
                             int32_t *grid_items, int32_t *row_counts, int64_t *row_offsets, void *stream) {
    if (!centers || !radii || num_cells <= 0 || !grid || !grid_counts || !grid_fill || !cell_key || !grid_start ||
        !grid_items || !row_counts || !row_offsets)
        return PF_ERR_ARG;
    if (!(grid->cell > 0.0f) || grid->dims[0] <= 0 || grid->dims[1] <= 0 || grid->dims[2] <= 0) return PF_ERR_ARG;
    const Grid g = make_grid(grid);
    const int64_t G = (int64_t)g.nx * g.ny * g.nz;
    hipStream_t st = (hipStream_t)stream;
    const dim3 nb((unsigned)((num_cells + 255) / 256));
    hipLaunchKernelGGL(grid_count_kernel, nb, dim3(256), 0, st, centers, radii, num_cells, g, grid_counts, cell_key);
    scan64(grid_counts, G + 1, grid_start, st);
    hipLaunchKernelGGL(grid_scatter_kernel, nb, dim3(256), 0, st, num_cells, cell_key, grid_start, grid_fill,
                       grid_items);
    hipLaunchKernelGGL(nbr_count_kernel, nb, dim3(256), 0, st, centers, radii, num_cells, g, grid_start, grid_items,
                       row_counts);
    scan64(row_counts, num_cells + 1, row_offsets, st);
    return ok_();
}

extern "C" int pf_cech_write(const float *centers, const float *radii, int64_t num_cells, const pf_cech_grid *grid,
                             const int64_t *grid_start, const int32_t *grid_items, const int64_t *row_offsets,
                             int32_t *adjacency, int32_t *adjacency_offsets, void *stream) {
    if (!centers || !radii || num_cells <= 0 || !grid || !grid_start || !grid_items || !row_offsets ||
        !adjacency_offsets)
        return PF_ERR_ARG;
    const Grid g = make_grid(grid);
    hipLaunchKernelGGL(nbr_write_kernel, dim3((unsigned)((num_cells + 256) / 256)), dim3(256), 0,
                       (hipStream_t)stream, centers, radii, num_cells, g, grid_start, grid_items, row_offsets,
                       adjacency, adjacency_offsets);
    return ok_();
}
