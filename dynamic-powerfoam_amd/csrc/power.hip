// power.hip — power-diagram adjacency on the GPU (the graph that
// build_power_adjacency, benchmark.py:25-83, takes from a Qhull lift).
//
// One wavefront per cell i (power_cells_wave_kernel) clips a box by the
// power bisectors
//   y . d_j <= 0.5 (L_j - L_i) - p_i . d_j,   d_j = p_j - p_i,  L = |p|^2 - r^2,
// in FP64 coordinates y centred at p_i.  The cell is a convex polytope kept
// as its dual triangulation: each triangle is a vertex (three planes), with
// the vertex position cached.  A clip drops the triangles whose vertex lies
// beyond the new plane and fans new triangles from the new plane to the
// boundary edges of the dropped region.
//
// Order of the clips: first the kNear sites nearest p_i (found by grid
// rings), nearest first, so the cell shrinks to near its final size within
// a few clips.  Then each vertex v is certified: a site j cuts v only if
//   |v - d_j|^2 < |v|^2 + r_j^2 - r_i^2,
// i.e. only if p_j lies in the ball around v of radius
// rho^2 = |v|^2 + r_max^2 - r_i^2 (the empty-sphere test of the weighted
// Delaunay tetrahedron v is dual to).  If that ball lies inside the radius
// of the kNear nearest sites it holds no unclipped site; otherwise the sites
// in the ball are looked up in an occupancy pyramid over the buckets, and
// the first that cuts v is clipped.  The cell is final when every vertex is
// certified.  The lookups cost little for a cell in a cluster and stay
// local for a cell that reaches into empty space, whose vertex balls are
// large but empty.
//
// The box B is the sites' bounding box with a margin: the result is every
// face that meets B.  Cells whose final polytope touches B's boundary are
// flagged; the faces lying wholly outside B are faces of the power diagram of
// the flagged cells alone, which power_hull_wave_kernel completes.
#include "pf_device.h"
#include "pf_hip.h"

namespace pf {

constexpr int kMaxPlanes = 64;     // box + cutting bisectors alive at once
constexpr int kMaxTris = 160;      // 2 F - 4 for F <= 64, plus one clip's fan
constexpr int kBoxPlane = -1;
constexpr int kNear = 32;          // sites clipped nearest-first
constexpr int kMaxLevels = 12;     // occupancy pyramid levels above the buckets
constexpr int kMaxDim = 1 << 20;   // buckets per axis (packed traversal keys)
constexpr int kMaxRounds = 1 << 14;  // certification rounds per cell

struct d3 { double x, y, z; };
__device__ __forceinline__ d3 dd(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 dcross(d3 a, d3 b) {
    return dd(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

struct Plane { d3 n; double d; };

// Buckets (level 0) of size cell from origin, and an occupancy pyramid:
// level l >= 1 halves the dims of level l - 1 (rounding up) and holds one
// byte per bucket, non-zero when a site lies in it, at occ + base[l].
struct PowerGrid {
    float ox, oy, oz, cell;
    int levels;
    int nx[kMaxLevels + 1], ny[kMaxLevels + 1], nz[kMaxLevels + 1];
    int64_t base[kMaxLevels + 1];
};

__host__ __device__ inline int64_t grid_size(const PowerGrid &g, int l) {
    return (int64_t)g.nx[l] * g.ny[l] * g.nz[l];
}

// Fills the pyramid dims: halve until every axis has at most two buckets.
// Returns the occupancy bytes the levels above the buckets need.
__host__ inline int64_t pyramid(PowerGrid &g) {
    g.levels = 0;
    g.base[0] = 0;
    int64_t bytes = 0;
    while (g.levels < kMaxLevels && (g.nx[g.levels] > 2 || g.ny[g.levels] > 2 || g.nz[g.levels] > 2)) {
        const int l = ++g.levels;
        g.nx[l] = (g.nx[l - 1] + 1) / 2;
        g.ny[l] = (g.ny[l - 1] + 1) / 2;
        g.nz[l] = (g.nz[l - 1] + 1) / 2;
        g.base[l] = bytes;
        bytes += grid_size(g, l);
    }
    return bytes;
}

// True when n . y <= d holds on the whole ball |y|^2 <= r2, so that the
// plane cannot cut the cell.  The relative margin covers the rounding of the
// vertex dot products clip() compares.
__device__ __forceinline__ bool misses_ball(d3 n, double d, double r2) {
    return d > 0.0 && d * d > r2 * ddot(n, n) * (1.0 + 1e-9);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ int bucket_of(float p, float o, float cell, int n) {
    return clampi((int)floorf((p - o) / cell), 0, n - 1);
}

// Calls fn(j) for every site j in a bucket that meets the ball |x - c| <= rho
// (world coordinates), descending the occupancy pyramid from the level at
// which the ball spans a few buckets.  Returns false as soon as fn does.
template <class F>
__device__ bool ball_query(const PowerGrid &g, const unsigned char *occ, const int64_t *gstart, const int *items,
                           d3 c, double rho, F &&fn) {
    const double c0 = (double)g.cell;
    const double pad = 0.01 * c0;  // float bucketing of the sites
    const double rho2 = rho * rho;
    int top = 0;
    while (top < g.levels && 2.0 * rho > 4.0 * c0 * (double)(1 << top)) ++top;
    const double ct = c0 * (double)(1 << top);
    auto range = [&](double cc, float o, int n, int &a, int &b) {
        a = max((int)floor((cc - rho - pad - (double)o) / ct), 0);
        b = min((int)floor((cc + rho + pad - (double)o) / ct), n - 1);
    };
    int x0, x1, y0, y1, z0, z1;
    range(c.x, g.ox, g.nx[top], x0, x1);
    range(c.y, g.oy, g.ny[top], y0, y1);
    range(c.z, g.oz, g.nz[top], z0, z1);
    // depth-first over (level, x, y, z) packed as 4 + 3 x 20 bits
    unsigned long long stack[8 * kMaxLevels + 8];
    for (int tz = z0; tz <= z1; ++tz)
        for (int ty = y0; ty <= y1; ++ty)
            for (int tx = x0; tx <= x1; ++tx) {
                int sp = 0;
                stack[sp++] = ((unsigned long long)top << 60) | ((unsigned long long)tz << 40) |
                              ((unsigned long long)ty << 20) | (unsigned long long)tx;
                while (sp > 0) {
                    const unsigned long long e = stack[--sp];
                    const int l = (int)(e >> 60);
                    const int z = (int)((e >> 40) & 0xfffff), y = (int)((e >> 20) & 0xfffff), x = (int)(e & 0xfffff);
                    const double cs = c0 * (double)(1 << l);
                    auto gap = [&](double cc, float o, int k) {
                        const double lo = (double)o + k * cs - pad, hi = (double)o + (k + 1) * cs + pad;
                        return cc < lo ? lo - cc : (cc > hi ? cc - hi : 0.0);
                    };
                    const double gx = gap(c.x, g.ox, x), gy = gap(c.y, g.oy, y), gz = gap(c.z, g.oz, z);
                    if (gx * gx + gy * gy + gz * gz > rho2) continue;
                    if (l == 0) {
                        const int64_t b = ((int64_t)z * g.ny[0] + y) * g.nx[0] + x;
                        for (int64_t q = gstart[b]; q < gstart[b + 1]; ++q)
                            if (!fn(items[q])) return false;
                        continue;
                    }
                    if (!occ[g.base[l] + ((int64_t)z * g.ny[l] + y) * g.nx[l] + x]) continue;
                    for (int cz = 2 * z; cz <= min(2 * z + 1, g.nz[l - 1] - 1); ++cz)
                        for (int cy = 2 * y; cy <= min(2 * y + 1, g.ny[l - 1] - 1); ++cy)
                            for (int cx = 2 * x; cx <= min(2 * x + 1, g.nx[l - 1] - 1); ++cx)
                                stack[sp++] = ((unsigned long long)(l - 1) << 60) |
                                              ((unsigned long long)cz << 40) | ((unsigned long long)cy << 20) |
                                              (unsigned long long)cx;
                }
            }
    return true;
}

// ------------------------------------------------------ one wave per cell --
// power_cells_wave_kernel: one 64-lane wavefront per cell, the polytope in
// the wave's LDS (9 KB).  (One thread per cell kept a 7 KB polytope per
// thread in scratch memory: 300 GB of traffic and 355 ms at 1M cells.)
//   * A clip tests up to three triangles per lane, marks the dropped
//     triangles' directed edges in a 64 x 64 bit matrix (a boundary edge is
//     a dropped edge whose twin is not), places the new fan triangles by a
//     prefix sum over the lanes in the serial clip's order and compacts the
//     kept triangles in order: the polytope after each clip is the serial
//     clip's, triangle for triangle.
//   * The kNear nearest sites: each grid shell's sites, 64 at a time, sorted
//     across the lanes (bitonic, by (|d|^2, index)) and merged into the
//     sorted best list by rank.
//   * Certification: the cheap ball tests for every uncertified vertex at
//     once; the ball queries of the rest one vertex per lane, and the sites
//     they find clipped one after the other (a clip that no longer cuts is
//     skipped), until every vertex is certified.
struct WavePoly {
    double pl[kMaxPlanes][4];           // n.x n.y n.z d
    int nb[kMaxPlanes];
    unsigned int tr[kMaxTris];           // a | b << 8 | c << 16 | certified << 24
    double v[kMaxTris][3];
    unsigned long long edge[kMaxPlanes];  // dropped directed edges: bit b of row a
    double kd[kNear];                    // nearest sites, ascending (|d|^2, j)
    int kj[kNear];
    double cd[64];                       // a sorted chunk of candidates
    int cj[64];
    int cnt[256];                        // shell segment: sites per position (prefix)
};

struct WaveState {
    int np, nt;
    double r2;
    bool overflow;
};

__device__ __forceinline__ int tri_a(unsigned t) { return (int)(t & 0xff); }
__device__ __forceinline__ int tri_b(unsigned t) { return (int)((t >> 8) & 0xff); }
__device__ __forceinline__ int tri_c(unsigned t) { return (int)((t >> 16) & 0xff); }
__device__ __forceinline__ unsigned tri_pack(int a, int b, int c, int cert) {
    return (unsigned)a | ((unsigned)b << 8) | ((unsigned)c << 16) | ((unsigned)cert << 24);
}
__device__ __forceinline__ d3 wv(const WavePoly &P, int t) { return dd(P.v[t][0], P.v[t][1], P.v[t][2]); }

__device__ __forceinline__ double wave_max_d(double x) {
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    return x;
}
__device__ __forceinline__ int wave_excl_scan(int v, int &total) {
    const int lane = __lane_id();
    int incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    total = __shfl(incl, 63, 64);
    return incl - v;
}

__device__ __forceinline__ d3 wvertex_of(const WavePoly &P, int a, int b, int c) {
    const d3 An = dd(P.pl[a][0], P.pl[a][1], P.pl[a][2]), Bn = dd(P.pl[b][0], P.pl[b][1], P.pl[b][2]),
             Cn = dd(P.pl[c][0], P.pl[c][1], P.pl[c][2]);
    const double Ad = P.pl[a][3], Bd = P.pl[b][3], Cd = P.pl[c][3];
    const d3 bc = dcross(Bn, Cn), ca = dcross(Cn, An), ab = dcross(An, Bn);
    const double det = ddot(An, bc);
    const double s = 1.0 / det;
    return dd((Ad * bc.x + Bd * ca.x + Cd * ab.x) * s, (Ad * bc.y + Bd * ca.y + Cd * ab.y) * s,
              (Ad * bc.z + Bd * ca.z + Cd * ab.z) * s);
}

__device__ __forceinline__ void wradius(const WavePoly &P, WaveState &S) {
    double m = 0.0;
    for (int t = __lane_id(); t < S.nt; t += 64) m = fmax(m, ddot(wv(P, t), wv(P, t)));
    S.r2 = wave_max_d(m);
}

// init_box, by the lanes: the same planes and triangles in the same order.
__device__ __forceinline__ void winit_box(WavePoly &P, WaveState &S, d3 lo, d3 hi) {
    const int lane = __lane_id();
    if (lane < 6) {
        const int ax = lane >> 1;
        const bool neg = lane & 1;
        const double hv = ax == 0 ? hi.x : (ax == 1 ? hi.y : hi.z), lv = ax == 0 ? lo.x : (ax == 1 ? lo.y : lo.z);
        P.pl[lane][0] = ax == 0 ? (neg ? -1.0 : 1.0) : 0.0;
        P.pl[lane][1] = ax == 1 ? (neg ? -1.0 : 1.0) : 0.0;
        P.pl[lane][2] = ax == 2 ? (neg ? -1.0 : 1.0) : 0.0;
        P.pl[lane][3] = neg ? -lv : hv;
        P.nb[lane] = kBoxPlane;
    }
    if (lane < 8) {
        const int sx = lane >> 2, sy = (lane >> 1) & 1, sz = lane & 1;
        const int X = sx ? 1 : 0, Y = sy ? 3 : 2, Z = sz ? 5 : 4;
        const bool positive = ((sx + sy + sz) & 1) == 0;
        P.tr[lane] = tri_pack(positive ? X : Y, positive ? Y : X, Z, 0);
        P.v[lane][0] = sx ? lo.x : hi.x;
        P.v[lane][1] = sy ? lo.y : hi.y;
        P.v[lane][2] = sz ? lo.z : hi.z;
    }
    if (lane < kMaxPlanes) P.edge[lane] = 0ull;
    S.np = 6;
    S.nt = 8;
    S.overflow = false;
    __syncthreads();
    wradius(P, S);
}

// compact_planes: drops planes no triangle references.
__device__ void wcompact_planes(WavePoly &P, WaveState &S) {
    const int lane = __lane_id();
    unsigned long long used = 0;
    for (int t = lane; t < S.nt; t += 64) {
        const unsigned tr = P.tr[t];
        used |= (1ull << tri_a(tr)) | (1ull << tri_b(tr)) | (1ull << tri_c(tr));
    }
    for (int off = 32; off > 0; off >>= 1) used |= __shfl_xor(used, off, 64);
    double pl[4] = {0, 0, 0, 0};
    int nb = 0;
    const bool mine = lane < S.np && ((used >> lane) & 1);
    if (mine) {
        for (int q = 0; q < 4; ++q) pl[q] = P.pl[lane][q];
        nb = P.nb[lane];
    }
    __syncthreads();
    if (mine) {
        const int to = __builtin_popcountll(used & ((1ull << lane) - 1));
        for (int q = 0; q < 4; ++q) P.pl[to][q] = pl[q];
        P.nb[to] = nb;
    }
    for (int t = lane; t < S.nt; t += 64) {
        const unsigned tr = P.tr[t];
        auto re = [&](int p) { return __builtin_popcountll(used & ((1ull << p) - 1)); };
        P.tr[t] = tri_pack(re(tri_a(tr)), re(tri_b(tr)), re(tri_c(tr)), (int)(tr >> 24));
    }
    S.np = __builtin_popcountll(used);
    __syncthreads();
}

// clip(): n . y <= d, plane neighbour nbr.  All lanes call it.
__device__ bool wclip(WavePoly &P, WaveState &S, d3 n, double d, int nbr) {
    const int lane = __lane_id();
    const int nt = S.nt;
    unsigned long long out[3];
    int nout = 0;
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        const int t = 64 * w + lane;
        const bool o = t < nt && ddot(n, wv(P, t)) > d;
        out[w] = __ballot(o);
        nout += __builtin_popcountll(out[w]);
    }
    if (nout == 0) return false;
    if (nout == nt) {  // nothing of the cell survives: a hidden site
        S.nt = 0;
        return true;
    }
    if (S.np == kMaxPlanes) wcompact_planes(P, S);
    if (S.np == kMaxPlanes) {
        S.overflow = true;
        return false;
    }
    const int p = S.np;
    if (lane == 0) {
        P.pl[p][0] = n.x;
        P.pl[p][1] = n.y;
        P.pl[p][2] = n.z;
        P.pl[p][3] = d;
        P.nb[p] = nbr;
    }
    S.np = p + 1;
    // the dropped triangles' directed edges
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        if ((out[w] >> lane) & 1) {
            const unsigned tr = P.tr[64 * w + lane];
            const int a = tri_a(tr), b = tri_b(tr), c = tri_c(tr);
            atomicOr(&P.edge[a], 1ull << b);
            atomicOr(&P.edge[b], 1ull << c);
            atomicOr(&P.edge[c], 1ull << a);
        }
    }
    __syncthreads();
    // boundary edges (twin not dropped), in the serial order: triangle, then edge
    int base = nt;
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        const bool dropped = (out[w] >> lane) & 1;
        int a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, cnt = 0, bm = 0;
        if (dropped) {
            const unsigned tr = P.tr[64 * w + lane];
            const int x[3] = {tri_a(tr), tri_b(tr), tri_c(tr)};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                a[k] = x[k];
                b[k] = x[(k + 1) % 3];
                if (!((P.edge[b[k]] >> a[k]) & 1ull)) {
                    bm |= 1 << k;
                    ++cnt;
                }
            }
        }
        int total = 0;
        int at = base + wave_excl_scan(cnt, total);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if ((bm >> k) & 1) {
                if (at < kMaxTris) {
                    P.tr[at] = tri_pack(a[k], b[k], p, 0);
                    const d3 vv = wvertex_of(P, a[k], b[k], p);
                    P.v[at][0] = vv.x;
                    P.v[at][1] = vv.y;
                    P.v[at][2] = vv.z;
                }
                ++at;
            }
        }
        base += total;
    }
    __syncthreads();
    // clear the edge rows
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        if ((out[w] >> lane) & 1) {
            const unsigned tr = P.tr[64 * w + lane];
            P.edge[tri_a(tr)] = 0ull;
            P.edge[tri_b(tr)] = 0ull;
            P.edge[tri_c(tr)] = 0ull;
        }
    }
    if (base > kMaxTris) {
        __syncthreads();
        S.overflow = true;
        return false;
    }
    // drop the clipped triangles, keeping the order of the rest
    unsigned tr_keep[3];
    double vk[3][3];
    int to[3];
    bool keep[3];
    const int below0 = __builtin_popcountll(out[0]), below1 = below0 + __builtin_popcountll(out[1]);
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        const int t = 64 * w + lane;
        keep[w] = t < base && !((out[w] >> lane) & 1);
        to[w] = 0;
        if (keep[w]) {
            tr_keep[w] = P.tr[t];
            vk[w][0] = P.v[t][0];
            vk[w][1] = P.v[t][1];
            vk[w][2] = P.v[t][2];
            const int before = (w == 0 ? 0 : (w == 1 ? below0 : below1)) +
                               __builtin_popcountll(out[w] & ((1ull << lane) - 1));
            to[w] = t - before;
        }
    }
    __syncthreads();
    double m = 0.0;
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        if (keep[w]) {
            P.tr[to[w]] = tr_keep[w];
            P.v[to[w]][0] = vk[w][0];
            P.v[to[w]][1] = vk[w][1];
            P.v[to[w]][2] = vk[w][2];
            m = fmax(m, vk[w][0] * vk[w][0] + vk[w][1] * vk[w][1] + vk[w][2] * vk[w][2]);
        }
    }
    __syncthreads();
    S.nt = base - nout;
    S.r2 = wave_max_d(m);
    return true;
}

// (d2, j) strictly before (e2, k)
__device__ __forceinline__ bool key_less(double d2, int j, double e2, int k) {
    return d2 < e2 || (d2 == e2 && j < k);
}

// Sorts one key per lane ascending across the wave (bitonic).
__device__ __forceinline__ void wave_sort64(double &d2, int &j) {
    const int lane = __lane_id();
    for (int k = 2; k <= 64; k <<= 1)
        for (int s = k >> 1; s > 0; s >>= 1) {
            const double od = __shfl_xor(d2, s, 64);
            const int oj = __shfl_xor(j, s, 64);
            const bool up = (lane & k) == 0;
            const bool lower = (lane & s) == 0;
            const bool other_less = key_less(od, oj, d2, j);
            // ascending block: the lower lane keeps the smaller key
            const bool take = (up == lower) ? other_less : key_less(d2, j, od, oj);
            if (take) {
                d2 = od;
                j = oj;
            }
        }
}

// Merges a chunk of candidates (one per lane, invalid = +inf) into the
// sorted best list of kc <= kNear entries.
__device__ void wmerge_near(WavePoly &P, int &kc, double d2, int j) {
    const int lane = __lane_id();
    wave_sort64(d2, j);
    P.cd[lane] = d2;
    P.cj[lane] = j;
    double bd = 0.0;
    int bj = 0;
    const bool has = lane < kc;
    if (has) {
        bd = P.kd[lane];
        bj = P.kj[lane];
    }
    __syncthreads();
    // candidate lane: entries of the best list before it (binary search)
    int lo = 0, hi = kc;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (key_less(P.kd[mid], P.kj[mid], d2, j)) lo = mid + 1;
        else hi = mid;
    }
    const int cpos = lane + lo;
    // best entry: candidates before it
    int lo2 = 0, hi2 = 64;
    if (has) {
        while (lo2 < hi2) {
            const int mid = (lo2 + hi2) >> 1;
            if (key_less(P.cd[mid], P.cj[mid], bd, bj)) lo2 = mid + 1;
            else hi2 = mid;
        }
    }
    const int bpos = lane + lo2;
    const int nvalid = __builtin_popcountll(__ballot(d2 < INFINITY));
    __syncthreads();
    if (has && bpos < kNear) {
        P.kd[bpos] = bd;
        P.kj[bpos] = bj;
    }
    if (d2 < INFINITY && cpos < kNear) {
        P.kd[cpos] = d2;
        P.kj[cpos] = j;
    }
    kc = min(kc + nvalid, kNear);
    __syncthreads();
}

// Position q of the Chebyshev shell of radius k (k >= 1): the two z faces,
// then the rings of the layers between.
__device__ __forceinline__ void shell_pos(int k, int q, int &x, int &y, int &z) {
    const int s = 2 * k + 1;
    if (q < 2 * s * s) {
        const int r = q % (s * s);
        z = q < s * s ? -k : k;
        x = r % s - k;
        y = r / s - k;
        return;
    }
    q -= 2 * s * s;
    z = q / (8 * k) - k + 1;
    const int r = q % (8 * k);
    if (r < s) {
        y = -k;
        x = r - k;
    } else if (r < 2 * s) {
        y = k;
        x = r - s - k;
    } else {
        const int r2 = r - 2 * s;
        x = r2 < 2 * k - 1 ? -k : k;
        y = r2 % (2 * k - 1) - k + 1;
    }
}
__device__ __forceinline__ int shell_size(int k) { return k == 0 ? 1 : 2 * (2 * k + 1) * (2 * k + 1) + 8 * k * (2 * k - 1); }

// Writes the neighbours of the planes in mask (box planes left out) to row,
// ascending, at most max_nb of them, and their number to *count.  Returns
// flag bits 0 (a box plane is in mask) and 1 (more than max_nb).
__device__ int wwrite_row(const WavePoly &P, const WaveState &S, unsigned long long mask, int max_nb, int *row,
                          int *count) {
    const int lane = __lane_id();
    const bool mine = lane < S.np && ((mask >> lane) & 1);
    const int nbv = mine ? P.nb[lane] : kBoxPlane;
    int flag = __any(mine && nbv == kBoxPlane) ? 1 : 0;
    const bool val = mine && nbv != kBoxPlane;
    const unsigned long long valm = __ballot(val);
    int rank = 0;
    for (unsigned long long m = valm; m;) {
        const int src = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        rank += __shfl(nbv, src, 64) < nbv ? 1 : 0;
    }
    const int c = __builtin_popcountll(valm);
    if (val && rank < max_nb) row[rank] = nbv;
    if (lane == 0) *count = c < max_nb ? c : max_nb;
    return flag | (c > max_nb ? 2 : 0);
}

// One wavefront per cell.  lists (N, max_nb) int32 sorted ascending, counts
// (N), flags (N): bit 0 the cell touches the box boundary, bit 1 capacity
// overflow (the row is incomplete), bit 2 the cell is empty (hidden site).
// lifted (N): the lifted coordinate L = |p|^2 - r^2 as the caller computed
// it (benchmark.py:29-32 computes it in f32).  slack >= |L - (|p|^2 - r^2)|
// bounds that rounding.
__global__ __launch_bounds__(64) void power_cells_wave_kernel(const float *pts, const float *lifted, int64_t N,
                                                              PowerGrid g, const int64_t *gstart, const int *items,
                                                              const unsigned char *occ, float3 blo, float3 bhi,
                                                              float rmax, float slack, int max_nb, int *lists,
                                                              int *counts, int *flags) {
    __shared__ WavePoly P;
    WaveState S;
    const int64_t i = blockIdx.x;
    if (i >= N) return;
    const int lane = __lane_id();
    const d3 pi = dd(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    const double Li = (double)lifted[i];
    winit_box(P, S, dd(blo.x - pi.x, blo.y - pi.y, blo.z - pi.z), dd(bhi.x - pi.x, bhi.y - pi.y, bhi.z - pi.z));
    const int ix = bucket_of(pts[3 * i], g.ox, g.cell, g.nx[0]);
    const int iy = bucket_of(pts[3 * i + 1], g.oy, g.cell, g.ny[0]);
    const int iz = bucket_of(pts[3 * i + 2], g.oz, g.cell, g.nz[0]);
    const int kmax = max(g.nx[0], max(g.ny[0], g.nz[0]));
    auto offset = [&](int j) {
        return dd((double)pts[3 * j] - pi.x, (double)pts[3 * j + 1] - pi.y, (double)pts[3 * j + 2] - pi.z);
    };
    auto rhs_of = [&](int j, d3 dj) { return 0.5 * ((double)lifted[j] - Li) - ddot(pi, dj); };

    // 1. the kNear nearest sites in the order (|d_j|^2, j), ring by ring
    int kc = 0;
    for (int k = 0; k <= kmax; ++k) {
        const int ns = shell_size(k);
        for (int q0 = 0; q0 < ns; q0 += 256) {
            // sites per shell position of this segment, prefix in LDS
            int c4 = 0, cq[4];
            int64_t bq[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int q = q0 + 4 * lane + m;
                cq[m] = 0;
                bq[m] = 0;
                if (q < ns) {
                    int x = 0, y = 0, z = 0;
                    if (k > 0) shell_pos(k, q, x, y, z);
                    x += ix;
                    y += iy;
                    z += iz;
                    if (x >= 0 && x < g.nx[0] && y >= 0 && y < g.ny[0] && z >= 0 && z < g.nz[0]) {
                        bq[m] = ((int64_t)z * g.ny[0] + y) * g.nx[0] + x;
                        cq[m] = (int)(gstart[bq[m] + 1] - gstart[bq[m]]);
                    }
                }
                c4 += cq[m];
            }
            int tot = 0;
            int run = wave_excl_scan(c4, tot);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                P.cnt[4 * lane + m] = run;  // first site of position 4 lane + m
                run += cq[m];
            }
            __syncthreads();
            for (int s0 = 0; s0 < tot; s0 += 64) {
                const int s = s0 + lane;
                double d2 = INFINITY;
                int j = 0x7fffffff;
                if (s < tot) {
                    // the position holding site s: last prefix <= s
                    int lo = 0, hi = 255;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (P.cnt[mid] <= s) lo = mid;
                        else hi = mid - 1;
                    }
                    const int q = q0 + lo;
                    int x = 0, y = 0, z = 0;
                    if (k > 0) shell_pos(k, q, x, y, z);
                    const int64_t b = ((int64_t)(z + iz) * g.ny[0] + (y + iy)) * g.nx[0] + (x + ix);
                    const int jj = items[gstart[b] + (s - P.cnt[lo])];
                    if (jj != i) {
                        const d3 dj = offset(jj);
                        const double e2 = ddot(dj, dj);
                        if (e2 > 1e-24 && (kc < kNear || key_less(e2, jj, P.kd[kNear - 1], P.kj[kNear - 1]))) {
                            d2 = e2;
                            j = jj;
                        }
                    }
                }
                if (__any(d2 < INFINITY)) wmerge_near(P, kc, d2, j);
            }
            __syncthreads();
        }
        const double D = ((double)k - 0.01) * (double)g.cell;
        if (kc == kNear && k >= 1 && D * D > P.kd[kNear - 1]) break;
    }
    // 2. clip by them nearest first; later planes mostly miss the cell's ball
    for (int t = 0; t < kc && S.nt > 0 && !S.overflow; ++t) {
        const int j = P.kj[t];
        const d3 dj = offset(j);
        const double rhs = rhs_of(j, dj);
        if (!misses_ball(dj, rhs, S.r2)) wclip(P, S, dj, rhs, j);
    }
    // 3. certify the vertices.  With fewer than kNear other sites every site
    // has been clipped; otherwise the sites not clipped lie at least
    // sqrt(kd[kNear - 1]) from p_i, and a site that cuts vertex v lies within
    // rho of v, rho^2 = |v|^2 + r_max^2 - r_i^2 (+ the lift's rounding).
    const double rm2 = (double)rmax * (double)rmax + 2.0 * (double)slack;
    const double ri2 = ddot(pi, pi) - Li;
    const double dK = kc == kNear ? sqrt(P.kd[kNear - 1]) : 0.0;
    const double kd_last = kc == kNear ? P.kd[kNear - 1] : 0.0;
    const int kj_last = kc == kNear ? P.kj[kNear - 1] : 0;
    for (int round = 0; kc == kNear && S.nt > 0 && !S.overflow; ++round) {
        if (round == kMaxRounds) {
            S.overflow = true;
            break;
        }
        // every uncertified vertex: the cheap tests, else its ball's first cutting site
        int found[3] = {-1, -1, -1};
        bool any_open = false;
        for (int w = 0; w < 3; ++w) {
            const int t = 64 * w + lane;
            if (t >= S.nt) continue;
            const unsigned tr = P.tr[t];
            if (tr >> 24) continue;
            const d3 v = wv(P, t);
            const double v2 = ddot(v, v);
            const double rho2 = (v2 + rm2 - ri2) * (1.0 + 1e-9) + 1e-18;
            bool cert = rho2 <= 0.0;
            double rho = 0.0;
            if (!cert) {
                rho = sqrt(rho2);
                cert = (sqrt(v2) + rho) * (1.0 + 1e-9) < dK;
            }
            if (!cert) {
                int cut = -1;
                ball_query(g, occ, gstart, items, dd(pi.x + v.x, pi.y + v.y, pi.z + v.z), rho, [&](int j) {
                    if (j == i) return true;
                    const d3 dj = offset(j);
                    const double d2 = ddot(dj, dj);
                    if (d2 <= 1e-24) return true;
                    if (d2 < kd_last || (d2 == kd_last && j <= kj_last)) return true;  // clipped in 2.
                    const double rhs = rhs_of(j, dj);
                    if (!(ddot(dj, v) > rhs)) return true;
                    for (int p = 0; p < S.np; ++p)
                        if (P.nb[p] == j) return true;  // applied already: rounding only
                    cut = j;
                    return false;
                });
                cert = cut < 0;
                found[w] = cut;
            }
            if (cert) P.tr[t] = tr | (1u << 24);
            else any_open = true;
        }
        __syncthreads();
        if (!__any(any_open)) break;
        // clip by the sites found, lane by lane (the polytope changes between clips)
        for (int w = 0; w < 3; ++w) {
            unsigned long long pend = __ballot(found[w] >= 0);
            while (pend && S.nt > 0 && !S.overflow) {
                const int src = __ffsll((unsigned long long)pend) - 1;
                pend &= pend - 1;
                const int j = __shfl(found[w], src, 64);
                bool applied = false;
                for (int p = 0; p < S.np; ++p) applied |= P.nb[p] == j;
                if (applied) continue;
                const d3 dj = offset(j);
                wclip(P, S, dj, rhs_of(j, dj), j);
            }
        }
    }
    int flag = S.overflow ? 2 : 0;
    if (S.nt == 0) flag |= 4;
    // neighbours: planes referenced by surviving triangles, ascending
    unsigned long long used = 0;
    for (int t = lane; t < S.nt; t += 64) {
        const unsigned tr = P.tr[t];
        used |= (1ull << tri_a(tr)) | (1ull << tri_b(tr)) | (1ull << tri_c(tr));
    }
    for (int off = 32; off > 0; off >>= 1) used |= __shfl_xor(used, off, 64);
    flag |= wwrite_row(P, S, used, max_nb, lists + i * (int64_t)max_nb, counts + i);
    if (lane == 0) flags[i] = flag;
}

// Faces outside the box.  Every point outside box B is owned by a site whose
// cell meets B's boundary, so the diagram outside B is that of those sites
// (hull) alone.  One wavefront per hull site h clips a far box by the
// bisectors of the other hull sites and keeps the planes whose face has a
// vertex outside B (a convex face with every vertex inside B lies inside
// B).  The list neighbours come first, alternating sides: the hull list is
// in cell order, and with cells in Morton order neighbours in the list are
// near in space, so the cell shrinks early and later planes miss its ball.
// The lanes test 64 candidates at a time against the cell's ball; those
// that may cut are tested again with the current ball (which only shrinks:
// a clip's new vertices lie on the old edges) and clipped in the list
// order.  Output as power_cells_wave_kernel (lists (nh, max_nb) of global
// ids).
__global__ __launch_bounds__(64) void power_hull_wave_kernel(const float *pts, const float *lifted, const int *hull,
                                                             int nh, float3 blo, float3 bhi, float far, int max_nb,
                                                             int *lists, int *counts, int *flags) {
    __shared__ WavePoly P;
    WaveState S;
    const int h = blockIdx.x;
    if (h >= nh) return;
    const int lane = __lane_id();
    const int i = hull[h];
    const d3 pi = dd(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    const double Li = (double)lifted[i];
    winit_box(P, S, dd(-far, -far, -far), dd(far, far, far));
    for (int k0 = 1; k0 < nh && S.nt > 0 && !S.overflow; k0 += 64) {
        const int k = k0 + lane;
        d3 dj = dd(0, 0, 0);
        double rhs = 0.0;
        int j = -1;
        bool maybe = false;
        if (k < nh) {
            const int step = (k + 1) >> 1;
            int q = (k & 1) ? h + step : h - step;
            if (q >= nh) q -= nh;
            if (q < 0) q += nh;
            j = hull[q];
            dj = dd((double)pts[3 * j] - pi.x, (double)pts[3 * j + 1] - pi.y, (double)pts[3 * j + 2] - pi.z);
            rhs = 0.5 * ((double)lifted[j] - Li) - ddot(pi, dj);
            maybe = j != i && ddot(dj, dj) > 1e-24 && !misses_ball(dj, rhs, S.r2);
        }
        unsigned long long pend = __ballot(maybe);
        while (pend && S.nt > 0 && !S.overflow) {
            const int src = __ffsll((unsigned long long)pend) - 1;
            pend &= pend - 1;
            const d3 n = dd(__shfl(dj.x, src, 64), __shfl(dj.y, src, 64), __shfl(dj.z, src, 64));
            const double d = __shfl(rhs, src, 64);
            const int js = __shfl(j, src, 64);
            if (misses_ball(n, d, S.r2)) continue;
            wclip(P, S, n, d, js);
        }
    }
    const d3 lo = dd(blo.x - pi.x, blo.y - pi.y, blo.z - pi.z), hi = dd(bhi.x - pi.x, bhi.y - pi.y, bhi.z - pi.z);
    unsigned long long outside = 0;  // planes with a face vertex outside B
    for (int t = lane; t < S.nt; t += 64) {
        const d3 v = wv(P, t);
        if (v.x < lo.x || v.y < lo.y || v.z < lo.z || v.x > hi.x || v.y > hi.y || v.z > hi.z) {
            const unsigned tr = P.tr[t];
            outside |= (1ull << tri_a(tr)) | (1ull << tri_b(tr)) | (1ull << tri_c(tr));
        }
    }
    for (int off = 32; off > 0; off >>= 1) outside |= __shfl_xor(outside, off, 64);
    const int f = wwrite_row(P, S, outside, max_nb, lists + (int64_t)h * max_nb, counts + h);
    if (lane == 0) flags[h] = (S.overflow ? 2 : 0) | (f & 2);
}

__global__ __launch_bounds__(256) void pgrid_count_kernel(const float *pts, int64_t N, PowerGrid g, int *gcount,
                                                          int *key) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int ix = bucket_of(pts[3 * i], g.ox, g.cell, g.nx[0]);
    const int iy = bucket_of(pts[3 * i + 1], g.oy, g.cell, g.ny[0]);
    const int iz = bucket_of(pts[3 * i + 2], g.oz, g.cell, g.nz[0]);
    const int64_t k = ((int64_t)iz * g.ny[0] + iy) * g.nx[0] + ix;
    key[i] = (int)k;
    atomicAdd(gcount + k + 1, 1);
}

// Marks the pyramid levels above the buckets (occ zeroed before).
__global__ __launch_bounds__(256) void pgrid_mark_kernel(const float *pts, int64_t N, PowerGrid g,
                                                         unsigned char *occ) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int ix = bucket_of(pts[3 * i], g.ox, g.cell, g.nx[0]);
    const int iy = bucket_of(pts[3 * i + 1], g.oy, g.cell, g.ny[0]);
    const int iz = bucket_of(pts[3 * i + 2], g.oz, g.cell, g.nz[0]);
    for (int l = 1; l <= g.levels; ++l)
        occ[g.base[l] + ((int64_t)(iz >> l) * g.ny[l] + (iy >> l)) * g.nx[l] + (ix >> l)] = 1;
}

__global__ __launch_bounds__(256) void zero_bytes_kernel(unsigned char *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Inclusive scan of n int32 into int64 in three launches: each block sums
// a chunk of kScanChunk values, one block scans the chunk sums (exclusive,
// in place), and each block rescans its chunk from that offset.
constexpr int kScanChunk = 4096;  // values per block: 256 threads x 16

__host__ __device__ inline int64_t scan_chunks(int64_t G) { return (G + 1 + kScanChunk - 1) / kScanChunk; }

__device__ __forceinline__ int64_t block_sum_256(int64_t v, int64_t *part) {
    const int t = threadIdx.x;
    part[t] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) part[t] += part[t + off];
        __syncthreads();
    }
    const int64_t total = part[0];
    __syncthreads();
    return total;
}

__global__ __launch_bounds__(256) void scan_chunk_sums_kernel(const int *in, int64_t n, int64_t *sums) {
    __shared__ int64_t part[256];
    const int64_t b = (int64_t)blockIdx.x * kScanChunk;
    int64_t v = 0;
    for (int k = threadIdx.x; k < kScanChunk; k += 256)
        if (b + k < n) v += in[b + k];
    const int64_t total = block_sum_256(v, part);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of m chunk sums in place, by one block of 1024 threads.
__global__ __launch_bounds__(1024) void scan_sums_kernel(int64_t *sums, int64_t m) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (m + 1023) / 1024;
    const int64_t b = t * per, e = b + per < m ? b + per : m;
    int64_t s = 0;
    for (int64_t i = b; i < e; ++i) s += sums[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t > 0 ? part[t - 1] : 0;
    for (int64_t i = b; i < e; ++i) {
        const int64_t x = sums[i];
        sums[i] = run;
        run += x;
    }
}

// Each thread scans 16 consecutive values of the block's chunk; the block
// scans the 256 thread totals in LDS.
__global__ __launch_bounds__(256) void scan_chunks_kernel(const int *in, int64_t n, const int64_t *sums,
                                                          int64_t *out) {
    __shared__ int64_t part[256];
    const int t = threadIdx.x;
    const int64_t b = (int64_t)blockIdx.x * kScanChunk + (int64_t)t * 16;
    int64_t v = 0;
    for (int k = 0; k < 16; ++k)
        if (b + k < n) v += in[b + k];
    part[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int64_t x = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    int64_t run = sums[blockIdx.x] + (t > 0 ? part[t - 1] : 0);
    for (int k = 0; k < 16; ++k)
        if (b + k < n) {
            run += in[b + k];
            out[b + k] = run;
        }
}

__global__ __launch_bounds__(256) void pgrid_scatter_kernel(int64_t N, const int *key, const int64_t *gstart,
                                                            int *gfill, int *items) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int k = key[i];
    items[gstart[k] + atomicAdd(gfill + k, 1)] = (int)i;
}

// PowerGrid of a pf_cech_grid; false when its dims are out of range.
inline bool make_grid(const pf_cech_grid *in, PowerGrid &g, int64_t &occ_bytes) {
    if (!(in->cell > 0.0f)) return false;
    for (int k = 0; k < 3; ++k)
        if (in->dims[k] <= 0 || in->dims[k] > kMaxDim) return false;
    g.ox = in->origin[0];
    g.oy = in->origin[1];
    g.oz = in->origin[2];
    g.cell = in->cell;
    g.nx[0] = in->dims[0];
    g.ny[0] = in->dims[1];
    g.nz[0] = in->dims[2];
    if (grid_size(g, 0) >= 0x7fffffff) return false;
    occ_bytes = pyramid(g);
    return true;
}

}  // namespace pf

using namespace pf;

// Workspace layout: the occupancy pyramid, then (8-byte aligned) the bucket
// scan's chunk sums.
static int64_t workspace_layout(const PowerGrid &g, int64_t occ_bytes, int64_t &sums_offset) {
    sums_offset = (occ_bytes + 7) & ~(int64_t)7;
    return sums_offset + 8 * scan_chunks(grid_size(g, 0));
}

extern "C" int64_t pf_power_workspace_bytes(const pf_cech_grid *grid) {
    PowerGrid g;
    int64_t occ = 0, sums_offset = 0;
    if (!grid || !make_grid(grid, g, occ)) return -1;
    return workspace_layout(g, occ, sums_offset);
}

extern "C" int pf_power_cells(const float *points, const float *lifted, int64_t num_cells, const pf_cech_grid *grid,
                              int32_t *grid_counts, int32_t *grid_fill, int32_t *cell_key, int64_t *grid_start,
                              int32_t *grid_items, void *workspace, int64_t workspace_bytes, const float *box,
                              float r_max, float lift_slack, int32_t max_nb, int32_t *lists, int32_t *counts,
                              int32_t *flags, void *stream) {
    if (!points || !lifted || num_cells <= 0 || num_cells >= 0x7fffffff || !grid || !grid_counts || !grid_fill ||
        !cell_key || !grid_start || !grid_items || !workspace || !box || !lists || !counts || !flags)
        return PF_ERR_ARG;
    PowerGrid g;
    int64_t need = 0, sums_offset = 0;
    if (max_nb <= 0 || max_nb > kMaxPlanes || !make_grid(grid, g, need) ||
        workspace_bytes < workspace_layout(g, need, sums_offset))
        return PF_ERR_ARG;
    unsigned char *occupancy = static_cast<unsigned char *>(workspace);
    int64_t *scan_sums = reinterpret_cast<int64_t *>(occupancy + sums_offset);
    const int64_t G = grid_size(g, 0);
    hipStream_t st = (hipStream_t)stream;
    const unsigned nb = (unsigned)((num_cells + 255) / 256);
    if (need > 0) {
        const int64_t zblocks = (need + 255) / 256;
        const unsigned zb = (unsigned)(zblocks < 4096 ? zblocks : 4096);
        hipLaunchKernelGGL(zero_bytes_kernel, dim3(zb), dim3(256), 0, st, occupancy, need);
    }
    hipLaunchKernelGGL(pgrid_count_kernel, dim3(nb), dim3(256), 0, st, points, num_cells, g, grid_counts, cell_key);
    hipLaunchKernelGGL(pgrid_mark_kernel, dim3(nb), dim3(256), 0, st, points, num_cells, g, occupancy);
    const int64_t chunks = scan_chunks(G);
    hipLaunchKernelGGL(scan_chunk_sums_kernel, dim3((unsigned)chunks), dim3(256), 0, st, grid_counts, G + 1,
                       scan_sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, st, scan_sums, chunks);
    hipLaunchKernelGGL(scan_chunks_kernel, dim3((unsigned)chunks), dim3(256), 0, st, grid_counts, G + 1, scan_sums,
                       grid_start);
    hipLaunchKernelGGL(pgrid_scatter_kernel, dim3(nb), dim3(256), 0, st, num_cells, cell_key, grid_start, grid_fill,
                       grid_items);
    hipLaunchKernelGGL(power_cells_wave_kernel, dim3((unsigned)num_cells), dim3(64), 0, st, points, lifted,
                       num_cells, g, grid_start, grid_items, occupancy, make_float3(box[0], box[1], box[2]),
                       make_float3(box[3], box[4], box[5]), r_max, lift_slack, max_nb, lists, counts, flags);
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}

extern "C" int pf_power_hull(const float *points, const float *lifted, const int32_t *hull, int32_t num_hull,
                             const float *box, float far, int32_t max_nb, int32_t *lists, int32_t *counts,
                             int32_t *flags, void *stream) {
    if (!points || !lifted || (!hull && num_hull > 0) || num_hull < 0 || !box || !(far > 0.0f) || !lists ||
        !counts || !flags)
        return PF_ERR_ARG;
    if (max_nb <= 0 || max_nb > kMaxPlanes) return PF_ERR_ARG;
    if (num_hull == 0) return PF_OK;
    hipLaunchKernelGGL(power_hull_wave_kernel, dim3((unsigned)num_hull), dim3(64), 0, (hipStream_t)stream, points,
                       lifted, hull, num_hull, make_float3(box[0], box[1], box[2]),
                       make_float3(box[3], box[4], box[5]), far, max_nb, lists, counts, flags);
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}
