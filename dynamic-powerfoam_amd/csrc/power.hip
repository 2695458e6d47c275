// power.hip — power-diagram adjacency on the GPU (the graph that
// build_power_adjacency, benchmark.py:25-83, takes from a Qhull lift).
//
// One thread per cell i clips a box by the power bisectors
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
// the flagged cells alone, which power_hull_kernel completes.
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

struct Poly {
    Plane pl[kMaxPlanes];
    int nb[kMaxPlanes];
    unsigned char tr[kMaxTris][3];
    unsigned char cert[kMaxTris];  // vertex certified against every site
    d3 v[kMaxTris];
    double r2;  // max |v|^2 over the vertices: the cell lies in the ball of radius sqrt(r2)
    int np, nt;
    bool overflow;
};

__device__ __forceinline__ void update_radius(Poly &P) {
    double r2 = 0.0;
    for (int t = 0; t < P.nt; ++t) r2 = fmax(r2, ddot(P.v[t], P.v[t]));
    P.r2 = r2;
}

// True when n . y <= d holds on the whole ball |y|^2 <= r2, so that the
// plane cannot cut the cell.  The relative margin covers the rounding of the
// vertex dot products clip() compares.
__device__ __forceinline__ bool misses_ball(d3 n, double d, double r2) {
    return d > 0.0 && d * d > r2 * ddot(n, n) * (1.0 + 1e-9);
}

__device__ __forceinline__ d3 vertex_of(const Poly &P, int a, int b, int c) {
    const Plane &A = P.pl[a], &B = P.pl[b], &C = P.pl[c];
    const d3 bc = dcross(B.n, C.n), ca = dcross(C.n, A.n), ab = dcross(A.n, B.n);
    const double det = ddot(A.n, bc);
    const double s = 1.0 / det;
    return dd((A.d * bc.x + B.d * ca.x + C.d * ab.x) * s, (A.d * bc.y + B.d * ca.y + C.d * ab.y) * s,
              (A.d * bc.z + B.d * ca.z + C.d * ab.z) * s);
}

// The box [lo, hi] (local coordinates) as an octahedron of 8 oriented
// triangles: corner (sx, sy, sz) is planes (X, Y, Z), counter-clockwise
// seen from outside when sx sy sz > 0, swapped otherwise.
__device__ __forceinline__ void init_box(Poly &P, d3 lo, d3 hi) {
    P.pl[0] = Plane{dd(1, 0, 0), hi.x};
    P.pl[1] = Plane{dd(-1, 0, 0), -lo.x};
    P.pl[2] = Plane{dd(0, 1, 0), hi.y};
    P.pl[3] = Plane{dd(0, -1, 0), -lo.y};
    P.pl[4] = Plane{dd(0, 0, 1), hi.z};
    P.pl[5] = Plane{dd(0, 0, -1), -lo.z};
    for (int k = 0; k < 6; ++k) P.nb[k] = kBoxPlane;
    P.np = 6;
    P.nt = 0;
    for (int sx = 0; sx < 2; ++sx)
        for (int sy = 0; sy < 2; ++sy)
            for (int sz = 0; sz < 2; ++sz) {
                const int X = sx ? 1 : 0, Y = sy ? 3 : 2, Z = sz ? 5 : 4;
                const bool positive = ((sx + sy + sz) & 1) == 0;
                const int t = P.nt++;
                P.tr[t][0] = (unsigned char)(positive ? X : Y);
                P.tr[t][1] = (unsigned char)(positive ? Y : X);
                P.tr[t][2] = (unsigned char)Z;
                P.cert[t] = 0;
                P.v[t] = dd(sx ? lo.x : hi.x, sy ? lo.y : hi.y, sz ? lo.z : hi.z);
            }
    P.overflow = false;
    update_radius(P);
}

// Drops planes no triangle references (they were clipped away).
__device__ __forceinline__ void compact_planes(Poly &P) {
    unsigned long long used = 0;
    for (int t = 0; t < P.nt; ++t)
        for (int k = 0; k < 3; ++k) used |= 1ull << P.tr[t][k];
    unsigned char remap[kMaxPlanes];
    int n = 0;
    for (int p = 0; p < P.np; ++p) {
        if ((used >> p) & 1) {
            P.pl[n] = P.pl[p];
            P.nb[n] = P.nb[p];
            remap[p] = (unsigned char)n++;
        }
    }
    P.np = n;
    for (int t = 0; t < P.nt; ++t)
        for (int k = 0; k < 3; ++k) P.tr[t][k] = remap[P.tr[t][k]];
}

// Clips by n . y <= d.  Returns true when the plane cut the cell.  Kept
// triangles keep their order and certification; new ones are appended.
__device__ bool clip(Poly &P, d3 n, double d, int nbr) {
    unsigned long long out[(kMaxTris + 63) / 64] = {0, 0, 0};
    int nout = 0;
    for (int t = 0; t < P.nt; ++t)
        if (ddot(n, P.v[t]) > d) {
            out[t >> 6] |= 1ull << (t & 63);
            ++nout;
        }
    if (nout == 0) return false;
    if (nout == P.nt) {  // nothing of the cell survives: a hidden site
        P.nt = 0;
        return true;
    }
    if (P.np == kMaxPlanes) compact_planes(P);
    if (P.np == kMaxPlanes) {
        P.overflow = true;
        return false;
    }
    const int p = P.np++;
    P.pl[p] = Plane{n, d};
    P.nb[p] = nbr;
    // boundary edges: an edge a->b of a dropped triangle whose twin b->a
    // belongs to a kept triangle; each gives the new triangle (a, b, p)
    int nt = P.nt;
    for (int t = 0; t < P.nt; ++t) {
        if (!((out[t >> 6] >> (t & 63)) & 1)) continue;
        for (int k = 0; k < 3; ++k) {
            const int a = P.tr[t][k], b = P.tr[t][(k + 1) % 3];
            bool twin_dropped = false;
            for (int u = 0; u < P.nt && !twin_dropped; ++u) {
                if (u == t || !((out[u >> 6] >> (u & 63)) & 1)) continue;
                for (int m = 0; m < 3; ++m)
                    if (P.tr[u][m] == b && P.tr[u][(m + 1) % 3] == a) twin_dropped = true;
            }
            if (twin_dropped) continue;
            if (nt == kMaxTris) {
                P.overflow = true;
                return false;
            }
            P.tr[nt][0] = (unsigned char)a;
            P.tr[nt][1] = (unsigned char)b;
            P.tr[nt][2] = (unsigned char)p;
            P.cert[nt] = 0;
            P.v[nt] = vertex_of(P, a, b, p);
            ++nt;
        }
    }
    // drop the clipped triangles (keep order of the rest)
    int w = 0;
    for (int t = 0; t < nt; ++t) {
        const bool dropped = t < P.nt && ((out[t >> 6] >> (t & 63)) & 1);
        if (dropped) continue;
        if (w != t) {
            P.tr[w][0] = P.tr[t][0];
            P.tr[w][1] = P.tr[t][1];
            P.tr[w][2] = P.tr[t][2];
            P.cert[w] = P.cert[t];
            P.v[w] = P.v[t];
        }
        ++w;
    }
    P.nt = w;
    update_radius(P);
    return true;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ int bucket_of(float p, float o, float cell, int n) {
    return clampi((int)floorf((p - o) / cell), 0, n - 1);
}

// Calls fn(j) for every site j bucketed in the shell of Chebyshev radius k
// around bucket (ix, iy, iz).  Returns false as soon as fn does.
template <class F>
__device__ __forceinline__ bool for_shell(const PowerGrid &g, const int64_t *gstart, const int *items, int ix, int iy,
                                          int iz, int k, F &&fn) {
    for (int z = max(iz - k, 0); z <= min(iz + k, g.nz[0] - 1); ++z)
        for (int y = max(iy - k, 0); y <= min(iy + k, g.ny[0] - 1); ++y) {
            const bool yz_edge = abs(z - iz) == k || abs(y - iy) == k;
            // inside the shell only x = ix +- k belongs to it
            const int step = yz_edge ? 1 : 2 * k;
            for (int x = ix - k; x <= ix + k; x += step) {
                if (x < 0 || x >= g.nx[0]) continue;
                const int64_t b = ((int64_t)z * g.ny[0] + y) * g.nx[0] + x;
                for (int64_t q = gstart[b]; q < gstart[b + 1]; ++q)
                    if (!fn(items[q])) return false;
            }
        }
    return true;
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

// Writes the planes selected by keep into row (sorted ascending, box planes
// left out).  Returns the flag bits 0 (a box plane was selected) and 1 (row
// capacity).
__device__ int write_row(const Poly &P, unsigned long long keep, int max_nb, int *row, int &count) {
    int flag = 0, c = 0;
    for (int p = 0; p < P.np; ++p) {
        if (!((keep >> p) & 1)) continue;
        if (P.nb[p] == kBoxPlane) {
            flag |= 1;
            continue;
        }
        if (c == max_nb) {
            flag |= 2;
            break;
        }
        // insertion into the sorted row
        int v = P.nb[p], w = c++;
        while (w > 0 && row[w - 1] > v) {
            row[w] = row[w - 1];
            --w;
        }
        row[w] = v;
    }
    count = c;
    return flag;
}

// One cell per thread.  lists (N, max_nb) int32 sorted ascending, counts (N),
// flags (N): bit 0 the cell touches the box boundary, bit 1 capacity
// overflow (the row is incomplete), bit 2 the cell is empty (hidden site).
// lifted (N): the lifted coordinate L = |p|^2 - r^2 as the caller computed
// it (benchmark.py:29-32 computes it in f32).  slack >= |L - (|p|^2 - r^2)|
// bounds that rounding.
__global__ __launch_bounds__(64) void power_cells_kernel(const float *pts, const float *lifted, int64_t N,
                                                         PowerGrid g, const int64_t *gstart, const int *items,
                                                         const unsigned char *occ, float3 blo, float3 bhi,
                                                         float rmax, float slack, int max_nb, int *lists,
                                                         int *counts, int *flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    Poly P;
    const d3 pi = dd(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    const double Li = (double)lifted[i];
    init_box(P, dd(blo.x - pi.x, blo.y - pi.y, blo.z - pi.z), dd(bhi.x - pi.x, bhi.y - pi.y, bhi.z - pi.z));
    // the same bucketing as pgrid_count_kernel
    const int ix = bucket_of(pts[3 * i], g.ox, g.cell, g.nx[0]);
    const int iy = bucket_of(pts[3 * i + 1], g.oy, g.cell, g.ny[0]);
    const int iz = bucket_of(pts[3 * i + 2], g.oz, g.cell, g.nz[0]);
    const int kmax = max(g.nx[0], max(g.ny[0], g.nz[0]));
    auto offset = [&](int j) {
        return dd((double)pts[3 * j] - pi.x, (double)pts[3 * j + 1] - pi.y, (double)pts[3 * j + 2] - pi.z);
    };
    auto rhs_of = [&](int j, d3 dj) { return 0.5 * ((double)lifted[j] - Li) - ddot(pi, dj); };
    // 1. the kNear nearest sites in the order (|d_j|^2, j), by grid rings:
    // after ring k every site closer than (k - 0.01) buckets has been seen
    // (a hundredth of a bucket is left for the float bucketing)
    double kd[kNear];
    int kj[kNear];
    int kc = 0;
    for (int k = 0; k <= kmax; ++k) {
        for_shell(g, gstart, items, ix, iy, iz, k, [&](int j) {
            if (j == i) return true;
            const d3 dj = offset(j);
            const double d2 = ddot(dj, dj);
            if (d2 <= 1e-24) return true;  // coincident sites: no bisector
            if (kc == kNear && !(d2 < kd[kNear - 1] || (d2 == kd[kNear - 1] && j < kj[kNear - 1]))) return true;
            int w = kc < kNear ? kc++ : kNear - 1;
            for (; w > 0 && (kd[w - 1] > d2 || (kd[w - 1] == d2 && kj[w - 1] > j)); --w) {
                kd[w] = kd[w - 1];
                kj[w] = kj[w - 1];
            }
            kd[w] = d2;
            kj[w] = j;
            return true;
        });
        const double D = ((double)k - 0.01) * (double)g.cell;
        if (kc == kNear && k >= 1 && D * D > kd[kNear - 1]) break;
    }
    // 2. clip by them nearest first; later planes mostly miss the cell's ball
    for (int t = 0; t < kc && P.nt > 0 && !P.overflow; ++t) {
        const d3 dj = offset(kj[t]);
        const double rhs = rhs_of(kj[t], dj);
        if (!misses_ball(dj, rhs, P.r2)) clip(P, dj, rhs, kj[t]);
    }
    // 3. certify the vertices.  With fewer than kNear other sites every site
    // has been clipped; otherwise the sites not clipped lie at least
    // sqrt(kd[kNear - 1]) from p_i, and a site that cuts vertex v lies within
    // rho of v, rho^2 = |v|^2 + r_max^2 - r_i^2 (+ the lift's rounding).
    const double rm2 = (double)rmax * (double)rmax + 2.0 * (double)slack;
    const double ri2 = ddot(pi, pi) - Li;
    const double dK = kc == kNear ? sqrt(kd[kNear - 1]) : 0.0;
    for (int round = 0; kc == kNear && P.nt > 0 && !P.overflow; ++round) {
        int t = 0;
        while (t < P.nt && P.cert[t]) ++t;
        if (t == P.nt) break;
        if (round == kMaxRounds) {
            P.overflow = true;
            break;
        }
        const d3 v = P.v[t];
        const double v2 = ddot(v, v);
        const double rho2 = (v2 + rm2 - ri2) * (1.0 + 1e-9) + 1e-18;
        if (rho2 <= 0.0) {  // no site reaches v
            P.cert[t] = 1;
            continue;
        }
        const double rho = sqrt(rho2);
        if ((sqrt(v2) + rho) * (1.0 + 1e-9) < dK) {  // the ball holds clipped sites only
            P.cert[t] = 1;
            continue;
        }
        bool cut = false;
        ball_query(g, occ, gstart, items, dd(pi.x + v.x, pi.y + v.y, pi.z + v.z), rho, [&](int j) {
            if (j == i) return true;
            const d3 dj = offset(j);
            const double d2 = ddot(dj, dj);
            if (d2 <= 1e-24) return true;
            if (d2 < kd[kNear - 1] || (d2 == kd[kNear - 1] && j <= kj[kNear - 1])) return true;  // clipped in 2.
            const double rhs = rhs_of(j, dj);
            if (!(ddot(dj, v) > rhs)) return true;  // the same test as clip()'s
            for (int p = 0; p < P.np; ++p)
                if (P.nb[p] == j) return true;  // applied already: rounding only
            clip(P, dj, rhs, j);
            cut = true;
            return false;
        });
        if (!cut) P.cert[t] = 1;
    }
    int flag = P.overflow ? 2 : 0;
    if (P.nt == 0) flag |= 4;
    // neighbours: planes referenced by surviving triangles
    unsigned long long used = 0;
    for (int t = 0; t < P.nt; ++t)
        for (int m = 0; m < 3; ++m) used |= 1ull << P.tr[t][m];
    int c = 0;
    flag |= write_row(P, used, max_nb, lists + i * (int64_t)max_nb, c);
    counts[i] = c;
    flags[i] = flag;
}

// Faces outside the box.  Every point outside box B is owned by a site whose
// cell meets B's boundary, so the diagram outside B is that of those sites
// (hull) alone.  One thread per hull site h clips a far box by the
// bisectors of all other hull sites and keeps the planes whose face has a
// vertex outside B (a convex face with every vertex inside B lies inside
// B).  Output as power_cells_kernel (lists (nh, max_nb) of global ids).
__global__ __launch_bounds__(64) void power_hull_kernel(const float *pts, const float *lifted, const int *hull,
                                                        int nh, float3 blo, float3 bhi, float far, int max_nb,
                                                        int *lists, int *counts, int *flags) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const int i = hull[h];
    Poly P;
    const d3 pi = dd(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    const double Li = (double)lifted[i];
    init_box(P, dd(-far, -far, -far), dd(far, far, far));
    // list neighbours first, alternating sides: the hull list is in cell
    // order, and with cells in Morton order neighbours in the list are near
    // in space, so the cell shrinks early and later planes miss its ball
    for (int k = 1; k < nh && P.nt > 0 && !P.overflow; ++k) {
        const int step = (k + 1) >> 1;
        int q = (k & 1) ? h + step : h - step;
        if (q >= nh) q -= nh;
        if (q < 0) q += nh;
        const int j = hull[q];
        if (j == i) continue;
        const d3 dj = dd((double)pts[3 * j] - pi.x, (double)pts[3 * j + 1] - pi.y, (double)pts[3 * j + 2] - pi.z);
        if (ddot(dj, dj) <= 1e-24) continue;
        const double rhs = 0.5 * ((double)lifted[j] - Li) - ddot(pi, dj);
        if (misses_ball(dj, rhs, P.r2)) continue;
        clip(P, dj, rhs, j);
    }
    const d3 lo = dd(blo.x - pi.x, blo.y - pi.y, blo.z - pi.z), hi = dd(bhi.x - pi.x, bhi.y - pi.y, bhi.z - pi.z);
    unsigned long long outside = 0;  // planes with a face vertex outside B
    for (int t = 0; t < P.nt; ++t) {
        const d3 v = P.v[t];
        if (v.x < lo.x || v.y < lo.y || v.z < lo.z || v.x > hi.x || v.y > hi.y || v.z > hi.z)
            for (int m = 0; m < 3; ++m) outside |= 1ull << P.tr[t][m];
    }
    int c = 0;
    const int flag = (P.overflow ? 2 : 0) | (write_row(P, outside, max_nb, lists + (int64_t)h * max_nb, c) & 2);
    counts[h] = c;
    flags[h] = flag;
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
    hipLaunchKernelGGL(power_cells_kernel, dim3((unsigned)((num_cells + 63) / 64)), dim3(64), 0, st, points, lifted,
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
    hipLaunchKernelGGL(power_hull_kernel, dim3((unsigned)((num_hull + 63) / 64)), dim3(64), 0, (hipStream_t)stream,
                       points, lifted, hull, num_hull, make_float3(box[0], box[1], box[2]),
                       make_float3(box[3], box[4], box[5]), far, max_nb, lists, counts, flags);
    return hipGetLastError() == hipSuccess ? PF_OK : PF_ERR_HIP;
}
