// cech.hip — Čech-complex adjacency (pairs of overlapping spheres) for
// gfx950, the graph the rasterizer clips with.
//
// Semantics follow AABBTree.build_cech_complex (powerfoam/bvh.py:7-123): a
// sphere is stored as its AABB [c - r, c + r] and recovered as
// c = 0.5 (lo + hi), r = 0.5 (hi.x - lo.x) (:17-18, :29-30); j != i is a
// neighbour of i when |c_i - c_j| < r_i + r_j (:31-33).  The reference finds
// candidates with a Warp BVH; here a grid of levels does it with
// counting-sort buckets: level l has cells of size cell * 2^l, and a sphere
// is listed once, in the finest level whose cells are at least its
// diameter (the last level's are at least 2 r_max).  A query for sphere i
// scans, at every level, the cells within r_i + (that level's largest
// radius) of its centre.  With one level this is the uniform grid of cell
// size 2 r_max (adjacent cells); with a heavy-tailed radius distribution the
// small spheres keep small cells, where one grid sized by the largest
// radius would crowd them into few cells and make the build quadratic.
// Each CSR row is sorted ascending (rocPRIM's segmented radix sort over the
// rows, on the bits the cell count needs), which fixes the row order the BVH
// traversal left unspecified; a row of a large sphere can hold tens of
// thousands of neighbours, which a per-thread insertion sort took seconds
// over.
#include "pf_device.h"
#include "pf_hip.h"

#include <rocprim/device/device_segmented_radix_sort.hpp>

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

__global__ void bounds_init_kernel(int *b, float *out) {
    const int t = threadIdx.x;
    if (t < 3) b[t] = 0x7fffffff;
    else if (t < 7) b[t] = (int)0x80000000;
    if (t == 7) out[7] = 0.0f;  // radius sum
}

__global__ __launch_bounds__(256) void bounds_kernel(const float *centers, const float *radii, int64_t N, int *b,
                                                     float *out) {
    f3 lo = mk(__builtin_inff(), __builtin_inff(), __builtin_inff());
    f3 hi = mk(-__builtin_inff(), -__builtin_inff(), -__builtin_inff());
    float rmax = 0.0f, rsum = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const Ball s = recover(centers, radii, i);
        lo = mk(fmin_(lo.x, s.c.x), fmin_(lo.y, s.c.y), fmin_(lo.z, s.c.z));
        hi = mk(fmax_(hi.x, s.c.x), fmax_(hi.y, s.c.y), fmax_(hi.z, s.c.z));
        rmax = fmax_(rmax, s.r);
        rsum += s.r;
    }
    for (int off = 32; off > 0; off >>= 1) rsum += __shfl_xor(rsum, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out + 7, rsum);
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

__global__ void bounds_finish_kernel(const int *b, float *out, int64_t N) {
    const int t = threadIdx.x;
    if (t < 7) out[t] = unord_f(b[t]);
    if (t == 7) out[7] = out[7] / (float)N;  // mean radius
}

constexpr int kMaxLevels = 16;

// Level l: cells of size cell * 2^l from the common origin, dims halved
// from level 0 (rounding up), buckets base[l] .. base[l + 1].
struct Grid {
    float ox, oy, oz, cell;
    int levels;
    int nx[kMaxLevels], ny[kMaxLevels], nz[kMaxLevels];
    int64_t base[kMaxLevels + 1];
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ float level_cell(const Grid &g, int l) { return g.cell * (float)(1 << l); }

// The finest level whose cells are at least the sphere's diameter.
__device__ __forceinline__ int level_of(const Grid &g, float r) {
    int l = 0;
    while (l < g.levels - 1 && 2.0f * r > level_cell(g, l)) ++l;
    return l;
}

__device__ __forceinline__ int coord(float c, float o, float cell, int n) {
    return clampi((int)__builtin_floorf((c - o) / cell), 0, n - 1);
}

__global__ __launch_bounds__(256) void grid_count_kernel(const float *centers, const float *radii, int64_t N, Grid g,
                                                         int *gcount, int *key) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const Ball s = recover(centers, radii, i);
    const int l = level_of(g, s.r);
    const float cl = level_cell(g, l);
    const int ix = coord(s.c.x, g.ox, cl, g.nx[l]), iy = coord(s.c.y, g.oy, cl, g.ny[l]),
              iz = coord(s.c.z, g.oz, cl, g.nz[l]);
    const int64_t k = g.base[l] + ((int64_t)iz * g.ny[l] + iy) * g.nx[l] + ix;
    key[i] = (int)k;
    atomicAdd(gcount + k + 1, 1);
}

__global__ __launch_bounds__(256) void grid_scatter_kernel(int64_t N, const int *key, const int64_t *gstart,
                                                           int *gfill, int *items) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int k = key[i];
    items[gstart[k] + atomicAdd(gfill + k, 1)] = (int)i;
}

// Visit every j != i whose ball overlaps ball i (bvh.py:23-33 / :57-69):
// at level l the spheres listed have r <= half the level's cell, so a
// neighbour's centre lies within r_i + cell_l / 2 of c_i (plus a hundredth
// of a cell for the float bucketing).
template <typename Fn>
__device__ __forceinline__ void for_overlaps(const float *centers, const float *radii, const Grid &g,
                                             const int64_t *gstart, const int *items, int64_t i, Fn fn) {
    const Ball s = recover(centers, radii, i);
    for (int l = 0; l < g.levels; ++l) {
        const float cl = level_cell(g, l);
        const float R = s.r + 0.51f * cl;
        const int x0 = coord(s.c.x - R, g.ox, cl, g.nx[l]), x1 = coord(s.c.x + R, g.ox, cl, g.nx[l]);
        const int y0 = coord(s.c.y - R, g.oy, cl, g.ny[l]), y1 = coord(s.c.y + R, g.oy, cl, g.ny[l]);
        const int z0 = coord(s.c.z - R, g.oz, cl, g.nz[l]), z1 = coord(s.c.z + R, g.oz, cl, g.nz[l]);
        for (int z = z0; z <= z1; ++z)
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    const int64_t k = g.base[l] + ((int64_t)z * g.ny[l] + y) * g.nx[l] + x;
                    for (int64_t q = gstart[k]; q < gstart[k + 1]; ++q) {
                        const int j = items[q];
                        if (j == i) continue;
                        const Ball t = recover(centers, radii, j);
                        if (length(s.c - t.c) < s.r + t.r) fn(j);
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

// Each row's neighbours in the order the grid lists them (unsorted).
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
    hipLaunchKernelGGL(bounds_init_kernel, dim3(1), dim3(64), 0, st, workspace, bounds);
    int blocks = (int)((num_cells + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(bounds_kernel, dim3(blocks), dim3(256), 0, st, centers, radii, num_cells, workspace, bounds);
    hipLaunchKernelGGL(bounds_finish_kernel, dim3(1), dim3(64), 0, st, workspace, bounds, num_cells);
    return ok_();
}

// The levels of a pf_cech_grid (levels 0 means 1); false when out of range.
static bool make_grid(const pf_cech_grid *in, Grid &g) {
    if (!in || !(in->cell > 0.0f) || in->dims[0] <= 0 || in->dims[1] <= 0 || in->dims[2] <= 0) return false;
    const int levels = in->levels <= 0 ? 1 : in->levels;
    if (levels > kMaxLevels) return false;
    g.ox = in->origin[0];
    g.oy = in->origin[1];
    g.oz = in->origin[2];
    g.cell = in->cell;
    g.levels = levels;
    g.base[0] = 0;
    for (int l = 0; l < levels; ++l) {
        g.nx[l] = ((in->dims[0] - 1) >> l) + 1;
        g.ny[l] = ((in->dims[1] - 1) >> l) + 1;
        g.nz[l] = ((in->dims[2] - 1) >> l) + 1;
        g.base[l + 1] = g.base[l] + (int64_t)g.nx[l] * g.ny[l] * g.nz[l];
    }
    return g.base[levels] < 0x7fffffff;  // bucket keys are int32
}

extern "C" int64_t pf_cech_grid_size(const pf_cech_grid *grid) {
    Grid g;
    return make_grid(grid, g) ? g.base[g.levels] : -1;
}

extern "C" int pf_cech_count(const float *centers, const float *radii, int64_t num_cells, const pf_cech_grid *grid,
                             int32_t *grid_counts, int32_t *grid_fill, int32_t *cell_key, int64_t *grid_start,
                             int32_t *grid_items, int32_t *row_counts, int64_t *row_offsets, void *stream) {
    if (!centers || !radii || num_cells <= 0 || !grid || !grid_counts || !grid_fill || !cell_key || !grid_start ||
        !grid_items || !row_counts || !row_offsets)
        return PF_ERR_ARG;
    Grid g;
    if (!make_grid(grid, g)) return PF_ERR_ARG;
    const int64_t G = g.base[g.levels];
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

// Bits a sort key (a cell index below num_cells) needs.
static int key_bits(int64_t num_cells) {
    int bits = 1;
    while (bits < 31 && (int64_t(1) << bits) < num_cells) ++bits;
    return bits;
}

// Workspace of pf_cech_write: the unsorted rows (num_edges int32, 256-byte
// aligned) and rocPRIM's temporary storage for sorting them.
static int64_t sort_storage(int64_t num_edges, int64_t num_cells) {
    size_t bytes = 0;
    if (rocprim::segmented_radix_sort_keys(nullptr, bytes, (const unsigned *)nullptr, (unsigned *)nullptr,
                                           (unsigned)num_edges, (unsigned)num_cells, (const int *)nullptr,
                                           (const int *)nullptr, 0, (unsigned)key_bits(num_cells)) != hipSuccess)
        return -1;
    return (int64_t)bytes;
}

extern "C" int64_t pf_cech_write_workspace_bytes(int64_t num_edges, int64_t num_cells) {
    if (num_edges < 0 || num_edges >= 0x7fffffff || num_cells <= 0 || num_cells >= 0x7fffffff) return -1;
    const int64_t sort = sort_storage(num_edges, num_cells);
    if (sort < 0) return -1;
    return ((4 * num_edges + 255) & ~int64_t(255)) + sort + 256;
}

extern "C" int pf_cech_write(const float *centers, const float *radii, int64_t num_cells, const pf_cech_grid *grid,
                             const int64_t *grid_start, const int32_t *grid_items, const int64_t *row_offsets,
                             int64_t num_edges, int32_t *adjacency, int32_t *adjacency_offsets, void *workspace,
                             int64_t workspace_bytes, void *stream) {
    if (!centers || !radii || num_cells <= 0 || !grid || !grid_start || !grid_items || !row_offsets ||
        !adjacency_offsets || !workspace || num_edges < 0 || (num_edges > 0 && !adjacency))
        return PF_ERR_ARG;
    Grid g;
    if (!make_grid(grid, g)) return PF_ERR_ARG;
    const int64_t need = pf_cech_write_workspace_bytes(num_edges, num_cells);
    if (need < 0) return PF_ERR_ARG;
    if (workspace_bytes < need) return PF_ERR_OVERFLOW;
    hipStream_t st = (hipStream_t)stream;
    int *unsorted = reinterpret_cast<int *>(workspace);
    void *temp = reinterpret_cast<char *>(workspace) + ((4 * num_edges + 255) & ~int64_t(255));
    hipLaunchKernelGGL(nbr_write_kernel, dim3((unsigned)((num_cells + 256) / 256)), dim3(256), 0, st, centers, radii,
                       num_cells, g, grid_start, grid_items, row_offsets, unsorted, adjacency_offsets);
    if (num_edges > 0) {
        size_t bytes = (size_t)(workspace_bytes - ((4 * num_edges + 255) & ~int64_t(255)));
        if (rocprim::segmented_radix_sort_keys(temp, bytes, reinterpret_cast<const unsigned *>(unsorted),
                                               reinterpret_cast<unsigned *>(adjacency), (unsigned)num_edges,
                                               (unsigned)num_cells, adjacency_offsets, adjacency_offsets + 1, 0,
                                               (unsigned)key_bits(num_cells), st) != hipSuccess)
            return PF_ERR_HIP;
    }
    return ok_();
}
