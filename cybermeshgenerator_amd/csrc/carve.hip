// carve.hip — point-in-domain classification for the PLC carve: which tetrahedra of a
// Delaunay mesh (or which points) lie inside a closed boundary surface.
//
// Two classifiers, both bit-exact against the CPU restatements:
//   rays   (src/constrained/tetrahedralize_plc.cpp:39-168, CarveGrid): three fixed
//          generic ray directions; per direction a projection grid of r x r cells
//          (r = ceil(sqrt(#triangles)) capped at 256) whose cell list names the triangles
//          a query tests; Moller-Trumbore hits with the reference's epsilons; inside iff
//          at least two of the three crossing counts are odd.
//   parity (src/voxelize/voxelize.cpp:78-195 at an arbitrary point): the vertical line
//          through the query's (x, y) against the xy bucket of its cell in
//          make_grid(points, resolution, 1), exact crossing tests with the compat
//          orient3d, the kNudge retries, and the parity of the crossings strictly above
//          the query's z.
//
// Layout in HBM: the triangles' vertices packed 9 doubles per triangle (72 B, one
// contiguous read per test), the cell lists as one CSR (start: cells + 1 unsigned,
// items: triangle indices) over all grids. Set-up builds the lists on the device
// (per-triangle cell rectangles, counting, one scan, an atomic fill); the lists are
// unordered within a cell: both crossing counts are order-free, and the one
// order-dependent case of the parity rule (every nudge grazes: the reference keeps the
// crossings found before the first grazing triangle of its ascending bucket) is
// evaluated as "crossings of triangles with a smaller index than the smallest grazing
// one".
//
// The query kernels are FP64 VALU and gather-latency bound: one lane per query, every
// list entry one 72 B gather and ~50 FP64 operations (rays) or 2-5 naive FP64 orient3d
// determinants (parity).
//
// Compiled with -ffp-contract=off: every FP64 expression rounds as the reference's
// x86-64 build rounds it.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "carve_internal.hpp"
#include "cmg_hip_internal.hpp"
#include "device_common.hpp"
#include "occ_device.hpp"

namespace cmg::hip {
namespace {

// tetrahedralize_plc.cpp:63-65
constexpr double kRayDirs[3][3] = {{0.5773269, 0.5774341, 0.5771914},
                                   {0.3251097, -0.7211021, 0.6119274},
                                   {-0.8017412, 0.2673021, 0.5346203}};

struct RayDir {
    double d[3];     // the ray direction as the reference passes it (not normalised)
    double u[3], v[3];
    double min_u, min_v, inv_u, inv_v;
};

struct RayParams {
    RayDir k[3];
    int r;
};

// Cell rectangle of one triangle on one grid: a in [a0, a1], b in [b0, b1]; the cell id
// is base + a * sa + b * sb.
struct Rect {
    int a0, a1, b0, b1;
};

struct GridMap {
    int base, sa, sb;
};

// static_cast<int>(double) as the reference's x86-64 build executes it (cvttsd2si):
// truncation toward zero, INT_MIN for NaN and for values outside the int range.
__device__ __forceinline__ int x86_trunc(double v) {
    return (v > -2147483649.0 && v < 2147483648.0) ? (int)v : INT_MIN;
}

__device__ __forceinline__ int clampi(int i, int r) { return i < 0 ? 0 : (i >= r ? r - 1 : i); }

// tetrahedralize_plc.cpp:120-122 (a[0] * x + a[1] * y) + a[2] * z
__device__ __forceinline__ double proj(const double* a, double x, double y, double z) {
    return a[0] * x + a[1] * y + a[2] * z;
}

// tetrahedralize_plc.cpp:123-128
__device__ __forceinline__ int cell_u(const RayDir& g, int r, double x, double y, double z) {
    return clampi(x86_trunc((proj(g.u, x, y, z) - g.min_u) * g.inv_u), r);
}
__device__ __forceinline__ int cell_v(const RayDir& g, int r, double x, double y, double z) {
    return clampi(x86_trunc((proj(g.v, x, y, z) - g.min_v) * g.inv_v), r);
}

// tetrahedralize_plc.cpp:39-58, with the triangle's vertices at T[0..8]
__device__ __forceinline__ bool ray_hits(double ox, double oy, double oz, const double* d,
                                         const double* T) {
    const double e10 = T[3] - T[0], e11 = T[4] - T[1], e12 = T[5] - T[2];
    const double e20 = T[6] - T[0], e21 = T[7] - T[1], e22 = T[8] - T[2];
    const double p0 = d[1] * e22 - d[2] * e21;
    const double p1 = d[2] * e20 - d[0] * e22;
    const double p2 = d[0] * e21 - d[1] * e20;
    const double det = e10 * p0 + e11 * p1 + e12 * p2;
    if (fabs(det) < 1e-12) return false;
    const double inv = 1.0 / det;
    const double t0 = ox - T[0], t1 = oy - T[1], t2 = oz - T[2];
    const double u = (t0 * p0 + t1 * p1 + t2 * p2) * inv;
    if (u < 0 || u > 1) return false;
    const double q0 = t1 * e12 - t2 * e11;
    const double q1 = t2 * e10 - t0 * e12;
    const double q2 = t0 * e11 - t1 * e10;
    const double v = (d[0] * q0 + d[1] * q1 + d[2] * q2) * inv;
    if (v < 0 || u + v > 1) return false;
    const double t = (e20 * q0 + e21 * q1 + e22 * q2) * inv;
    return t > 1e-9;
}

// ---- set-up ---------------------------------------------------------------------------

// Per-block partial bounds: RAYS: (min u, max u, min v, max v) of the projections on each
// direction (12 values); else (min, max) of each coordinate (6 values). Min and max are
// exact, so the order of the reduction does not matter.
template <bool RAYS>
__global__ __launch_bounds__(256) void bounds_kernel(const double* xyz, long long npts,
                                                     RayParams P, double* part) {
    constexpr int NV = RAYS ? 12 : 6;
    double acc[NV];
    for (int j = 0; j < NV; j += 2) acc[j] = 1e300, acc[j + 1] = -1e300;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < npts;
         p += (long long)gridDim.x * blockDim.x) {
        const double x = xyz[3 * p], y = xyz[3 * p + 1], z = xyz[3 * p + 2];
        if constexpr (RAYS) {
            for (int k = 0; k < 3; ++k) {
                const double pu = proj(P.k[k].u, x, y, z), pv = proj(P.k[k].v, x, y, z);
                acc[4 * k] = fmin(acc[4 * k], pu);
                acc[4 * k + 1] = fmax(acc[4 * k + 1], pu);
                acc[4 * k + 2] = fmin(acc[4 * k + 2], pv);
                acc[4 * k + 3] = fmax(acc[4 * k + 3], pv);
            }
        } else {
            const double c[3] = {x, y, z};
            for (int a = 0; a < 3; ++a) {
                acc[2 * a] = fmin(acc[2 * a], c[a]);
                acc[2 * a + 1] = fmax(acc[2 * a + 1], c[a]);
            }
        }
    }
    __shared__ double s[NV][256];
    for (int j = 0; j < NV; ++j) s[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int j = 0; j < NV; j += 2) {
                s[j][threadIdx.x] = fmin(s[j][threadIdx.x], s[j][threadIdx.x + o]);
                s[j + 1][threadIdx.x] = fmax(s[j + 1][threadIdx.x], s[j + 1][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if ((int)threadIdx.x < NV) part[(size_t)blockIdx.x * NV + threadIdx.x] = s[threadIdx.x][0];
}

// Triangle vertex indices outside [0, npts), counted before any kernel gathers points
// through them (one atomic per wavefront that found some).
__global__ __launch_bounds__(256) void tri_check_kernel(const int* tris, long long ntris,
                                                        long long npts,
                                                        unsigned long long* bad) {
    unsigned long long c = 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < 3 * ntris;
         i += (long long)gridDim.x * blockDim.x) {
        const int v = tris[i];
        c += (v < 0 || v >= npts) ? 1ull : 0ull;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(bad, c);
}

// Packs the triangles' vertices (9 doubles each) and computes every triangle's cell
// rectangle on each grid, and sums the cells they cover (the list total) in 64 bits.
template <bool RAYS>
__global__ __launch_bounds__(256) void rect_kernel(const double* xyz, const int* tris,
                                                   long long ntris, RayParams P, GridSpec g,
                                                   double* tv, Rect* rect,
                                                   unsigned long long* entries) {
    constexpr int NG = RAYS ? 3 : 1;
    unsigned long long area = 0;  // list entries of this lane's rectangles, in 64 bits
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < ntris;
         t += (long long)gridDim.x * blockDim.x) {
        double V[3][3];
        for (int j = 0; j < 3; ++j) {
            const long long pi = tris[3 * t + j];
            for (int a = 0; a < 3; ++a) {
                V[j][a] = xyz[3 * pi + a];
                tv[9 * t + 3 * j + a] = V[j][a];
            }
        }
        for (int k = 0; k < NG; ++k) {
            Rect R;
            if constexpr (RAYS) {
                // tetrahedralize_plc.cpp:146-157
                const int r = P.r;
                R = Rect{r, 0, r, 0};
                for (int j = 0; j < 3; ++j) {
                    const int iu = cell_u(P.k[k], r, V[j][0], V[j][1], V[j][2]);
                    const int iv = cell_v(P.k[k], r, V[j][0], V[j][1], V[j][2]);
                    R.a0 = min(R.a0, iu), R.a1 = max(R.a1, iu);
                    R.b0 = min(R.b0, iv), R.b1 = max(R.b1, iv);
                }
            } else {
                // voxelize.cpp:90-97
                const int4 xr = occdev::xy_range(V, g);
                R = Rect{xr.x, xr.y, xr.z, xr.w};
            }
            rect[k * ntris + t] = R;
            if (R.a1 >= R.a0 && R.b1 >= R.b0)
                area += (unsigned long long)(R.a1 - R.a0 + 1) *
                        (unsigned long long)(R.b1 - R.b0 + 1);
        }
    }
    for (int o = 32; o > 0; o >>= 1) area += __shfl_xor(area, o);
    if (__lane_id() == 0 && area) atomicAdd(entries, area);
}

// Per-cell list lengths from the rectangles (run only once their total is known to fit).
template <bool RAYS>
__global__ __launch_bounds__(256) void count_kernel(long long ntris, RayParams P, GridSpec g,
                                                    const Rect* rect, unsigned* count) {
    constexpr int NG = RAYS ? 3 : 1;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < ntris;
         t += (long long)gridDim.x * blockDim.x) {
        for (int k = 0; k < NG; ++k) {
            const Rect R = rect[k * ntris + t];
            const GridMap M = RAYS ? GridMap{k * P.r * P.r, P.r, 1} : GridMap{0, 1, g.nx};
            for (int a = R.a0; a <= R.a1; ++a)
                for (int b = R.b0; b <= R.b1; ++b)
                    atomicAdd(&count[M.base + a * M.sa + b * M.sb], 1u);
        }
    }
}

template <bool RAYS>
__global__ __launch_bounds__(256) void fill_kernel(long long ntris, RayParams P, GridSpec g,
                                                   const Rect* rect, unsigned* cursor,
                                                   int* items) {
    constexpr int NG = RAYS ? 3 : 1;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < ntris;
         t += (long long)gridDim.x * blockDim.x) {
        for (int k = 0; k < NG; ++k) {
            const Rect R = rect[k * ntris + t];
            const GridMap M = RAYS ? GridMap{k * P.r * P.r, P.r, 1} : GridMap{0, 1, g.nx};
            for (int a = R.a0; a <= R.a1; ++a)
                for (int b = R.b0; b <= R.b1; ++b)
                    items[atomicAdd(&cursor[M.base + a * M.sa + b * M.sb], 1u)] = (int)t;
        }
    }
}

// ---- queries --------------------------------------------------------------------------

// The query point of item i: the point itself, or the centroid of tet i
// (tetrahedralize_plc.cpp:170-174: (a + b + c + d) / 4 per axis, summed left to right).
// False for a tet with an index outside [0, npts).
template <bool TETS>
__device__ __forceinline__ bool query_point(const double* q, const int* tets,
                                            const double* xyz, long long npts, long long i,
                                            double& x, double& y, double& z) {
    if constexpr (!TETS) {
        x = q[3 * i], y = q[3 * i + 1], z = q[3 * i + 2];
        return true;
    } else {
        const int4 t = *reinterpret_cast<const int4*>(tets + 4 * i);
        const int v[4] = {t.x, t.y, t.z, t.w};
        for (int j = 0; j < 4; ++j)
            if (v[j] < 0 || v[j] >= npts) return false;
        double c[3];
        for (int a = 0; a < 3; ++a)
            c[a] = (xyz[3 * (long long)v[0] + a] + xyz[3 * (long long)v[1] + a] +
                    xyz[3 * (long long)v[2] + a] + xyz[3 * (long long)v[3] + a]) / 4;
        x = c[0], y = c[1], z = c[2];
        return true;
    }
}

// Query set-up: the query point of item i (stored for tets), and one (cell, i) pair per
// grid: key = global cell id, ncell for a tet with an index out of range (answered 0,
// counted, sorted after every valid key). Sorting the pairs by key orders the queries by
// cell, grid by grid.
template <bool TETS, bool RAYS>
__global__ __launch_bounds__(256) void query_prep_kernel(const double* q, const int* tets,
                                                         const double* xyz, long long npts,
                                                         long long n, RayParams P, GridSpec g,
                                                         unsigned ncell, double* qbuf,
                                                         unsigned* key, unsigned* val,
                                                         uint8_t* out, unsigned* nbad,
                                                         unsigned long long* bad) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double x, y, z;
        const bool ok = query_point<TETS>(q, tets, xyz, npts, i, x, y, z);
        if (TETS) {
            qbuf[3 * i] = x, qbuf[3 * i + 1] = y, qbuf[3 * i + 2] = z;
        }
        if (!ok) {
            out[i] = 0;
            atomicAdd(bad, 1ull);
            atomicAdd(nbad, 1u);
        }
        if constexpr (RAYS) {
            const int r = P.r;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                key[k * n + i] = ok ? (unsigned)(k * r * r + cell_u(P.k[k], r, x, y, z) * r +
                                                 cell_v(P.k[k], r, x, y, z))
                                    : ncell;
                val[k * n + i] = (unsigned)i;
            }
        } else {
            unsigned c = ncell;
            if (ok) {
                const int ci = occdev::clamp_cell(x, g.corner[0], g.spacing, g.nx);
                const int cj = occdev::clamp_cell(y, g.corner[1], g.spacing, g.ny);
                c = (unsigned)((long long)cj * g.nx + ci);
            }
            key[i] = c;
            val[i] = (unsigned)i;
        }
    }
}

// One ray direction (tetrahedralize_plc.cpp:82-90) for the queries of grid K, taken in
// cell order so that the lanes of a wavefront mostly share one cell list (uniform trip
// counts, one triangle record per step for the whole wavefront). odd[K n + i] = parity
// of the crossings of query i.
template <int K>
__global__ __launch_bounds__(256) void rays_pass_kernel(const double* qpts, long long n,
                                                        RayParams P, const unsigned* nbad,
                                                        const unsigned* skey,
                                                        const unsigned* sval,
                                                        const unsigned* start,
                                                        const int* items, const double* tv,
                                                        uint8_t* odd) {
    const long long m = n - *nbad;  // valid queries: each has one pair per grid
    const long long base = K * m;
    const RayDir& G = P.k[K];
    for (long long j = xcd_block() * blockDim.x + threadIdx.x; j < m;
         j += (long long)gridDim.x * blockDim.x) {
        const unsigned i = sval[base + j];
        const unsigned c = skey[base + j];
        const double x = qpts[3 * (long long)i], y = qpts[3 * (long long)i + 1],
                     z = qpts[3 * (long long)i + 2];
        const unsigned e0 = start[c], e1 = start[c + 1];
        bool o = false;
        if (e0 < e1) {
            // software pipeline: the next record and the index after it are loaded while
            // the current test computes
            double cur[9], nxt[9];
            int ib = e0 + 1 < e1 ? items[e0 + 1] : 0;
            const double* T0 = tv + 9 * (long long)items[e0];
#pragma unroll
            for (int a = 0; a < 9; ++a) cur[a] = T0[a];
            for (unsigned e = e0; e < e1; ++e) {
                const bool more = e + 1 < e1;
                const int ic = e + 2 < e1 ? items[e + 2] : 0;
                if (more) {
                    const double* Tn = tv + 9 * (long long)ib;
#pragma unroll
                    for (int a = 0; a < 9; ++a) nxt[a] = Tn[a];
                }
                o ^= ray_hits(x, y, z, G.d, cur);
#pragma unroll
                for (int a = 0; a < 9; ++a) cur[a] = nxt[a];
                ib = ic;
            }
        }
        odd[K * n + i] = o;
    }
}

// tetrahedralize_plc.cpp:91-92: inside iff at least two of the three parities are odd
__global__ __launch_bounds__(256) void rays_vote_kernel(const unsigned* key, unsigned ncell,
                                                        long long n, const uint8_t* odd,
                                                        uint8_t* out) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        if (key[i] != ncell) out[i] = (uint8_t)(odd[i] + odd[n + i] + odd[2 * n + i] >= 2);
}

// One attempt of voxelize.cpp:121-143 for the line through (x, y): false if a sign is
// exactly zero (gmin = the smallest grazing triangle index), else above = the number of
// crossings strictly above z.
__device__ bool parity_attempt(double x, double y, double zmin, double zmax, double z,
                               const int* items, unsigned e0, unsigned e1, const double* tv,
                               int& above, int& gmin) {
    const double P[3] = {x, y, zmin};
    const double Q[3] = {x, y, zmax};
    above = 0;
    gmin = INT_MAX;
    for (unsigned e = e0; e < e1; ++e) {
        const int t = items[e];
        const double* T = tv + 9 * (long long)t;
        const int sa = occdev::o3_compat(P, T, T + 3, T + 6);
        const int sb = occdev::o3_compat(Q, T, T + 3, T + 6);
        if (sa == 0 || sb == 0) {
            gmin = min(gmin, t);
            continue;
        }
        if (sa == sb) continue;
        const int s1 = occdev::o3_compat(P, Q, T, T + 3);
        const int s2 = occdev::o3_compat(P, Q, T + 3, T + 6);
        const int s3 = occdev::o3_compat(P, Q, T + 6, T);
        if (s1 == 0 || s2 == 0 || s3 == 0) {
            gmin = min(gmin, t);
            continue;
        }
        if (s1 != s2 || s2 != s3) continue;
        above += occdev::plane_z(T, T + 3, T + 6, x, y) > z;
    }
    return gmin == INT_MAX;
}

// The crossings above z among triangles with index < gmin: what the reference's
// column_crossings has appended when it stops at its first grazing triangle.
__device__ int parity_partial(double x, double y, double zmin, double zmax, double z,
                              const int* items, unsigned e0, unsigned e1, const double* tv,
                              int gmin) {
    const double P[3] = {x, y, zmin};
    const double Q[3] = {x, y, zmax};
    int above = 0;
    for (unsigned e = e0; e < e1; ++e) {
        const int t = items[e];
        if (t >= gmin) continue;
        const double* T = tv + 9 * (long long)t;
        const int sa = occdev::o3_compat(P, T, T + 3, T + 6);
        const int sb = occdev::o3_compat(Q, T, T + 3, T + 6);
        if (sa == sb) continue;  // (no zero signs below gmin)
        const int s1 = occdev::o3_compat(P, Q, T, T + 3);
        const int s2 = occdev::o3_compat(P, Q, T + 3, T + 6);
        const int s3 = occdev::o3_compat(P, Q, T + 6, T);
        if (s1 != s2 || s2 != s3) continue;
        above += occdev::plane_z(T, T + 3, T + 6, x, y) > z;
    }
    return above;
}

// The column rule (voxelize.cpp:148-195) for the queries in column order: the lanes of a
// wavefront mostly share one bucket.
__global__ __launch_bounds__(256) void parity_pass_kernel(const double* qpts, long long n,
                                                          GridSpec g, const unsigned* nbad,
                                                          const unsigned* skey,
                                                          const unsigned* sval,
                                                          const unsigned* start,
                                                          const int* items, const double* tv,
                                                          uint8_t* out) {
    // voxelize.cpp:176-177
    const double zmin = g.corner[2] - g.spacing;
    const double zmax = g.corner[2] + (g.nz + 1) * g.spacing;
    const long long m = n - *nbad;
    for (long long j = xcd_block() * blockDim.x + threadIdx.x; j < m;
         j += (long long)gridDim.x * blockDim.x) {
        const unsigned i = sval[j];
        const double x = qpts[3 * (long long)i], y = qpts[3 * (long long)i + 1],
                     z = qpts[3 * (long long)i + 2];
        const unsigned col = skey[j];
        const unsigned e0 = start[col], e1 = start[col + 1];
        int above = 0, gmin = INT_MAX;
        // voxelize.cpp:154-165
        if (!parity_attempt(x, y, zmin, zmax, z, items, e0, e1, tv, above, gmin)) {
            bool ok = false;
            double nxp = x, nyp = y;
#pragma unroll 1
            for (int d = 0; d < 6 && !ok; ++d) {
                nxp = x + occdev::kNudge[d][0] * g.spacing;
                nyp = y + occdev::kNudge[d][1] * g.spacing;
                ok = parity_attempt(nxp, nyp, zmin, zmax, z, items, e0, e1, tv, above, gmin);
            }
            if (!ok) above = parity_partial(nxp, nyp, zmin, zmax, z, items, e0, e1, tv, gmin);
        }
        out[i] = (uint8_t)(above & 1);
    }
}

int carve_resolution(long long ntris) {
    // tetrahedralize_plc.cpp:73-75
    long long r = 1;
    while (r * r < ntris) ++r;
    return (int)(r < 256 ? r : 256);
}

// tetrahedralize_plc.cpp:100-119, host side (same FP64 operations, libm sqrt)
void carve_basis(const double* d, double* u, double* v) {
    double n[3] = {d[0], d[1], d[2]};
    double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int i = 0; i < 3; ++i) n[i] /= len;
    double h[3] = {1, 0, 0};
    if (std::fabs(n[0]) > 0.9) {
        h[0] = 0;
        h[1] = 1;
    }
    u[0] = n[1] * h[2] - n[2] * h[1];
    u[1] = n[2] * h[0] - n[0] * h[2];
    u[2] = n[0] * h[1] - n[1] * h[0];
    double ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int i = 0; i < 3; ++i) u[i] /= ul;
    v[0] = n[1] * u[2] - n[2] * u[1];
    v[1] = n[2] * u[0] - n[0] * u[2];
    v[2] = n[0] * u[1] - n[1] * u[0];
}

}  // namespace

struct CarveEngine::Impl {
    int mode = kCarveRays;
    const double* xyz = nullptr;
    long long npts = 0, ntris = 0;
    RayParams P{};
    GridSpec g{};
    long long ncell = 0;
    unsigned long long entries = 0;
    double* tv = nullptr;
    unsigned* start = nullptr;
    int* items = nullptr;
    unsigned long long* bad = nullptr;
    std::vector<void*> owned;

    template <class T>
    T* dalloc(size_t n) {
        void* p = nullptr;
        CMG_HIP(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
        owned.push_back(p);
        return static_cast<T*>(p);
    }
    ~Impl() {
        for (void* p : owned) (void)hipFree(p);
    }
};

CarveEngine::CarveEngine(const double* d_xyz, long long npts, const int* d_tris,
                         long long ntris, int mode, int resolution, hipStream_t st)
    : p_(std::make_unique<Impl>()) {
    Impl& I = *p_;
    if (mode != kCarveRays && mode != kCarveParity)
        throw InvalidInput("carve: mode must be CMG_CARVE_RAYS (0) or CMG_CARVE_PARITY (1)");
    if (npts < 1) throw InvalidInput("carve: no points");
    if (ntris < 1) throw InvalidInput("PLC has facets but no boundary triangles");
    if (ntris >= (1LL << 31) || npts >= (1LL << 31))
        throw InvalidInput("carve: more than 2^31 points or triangles");
    if (!d_xyz || !d_tris) throw InvalidInput("carve: null pointer");
    I.mode = mode;
    I.xyz = d_xyz;
    I.npts = npts;
    I.ntris = ntris;
    const bool rays = mode == kCarveRays;
    const int r = carve_resolution(ntris);
    I.P.r = r;
    for (int k = 0; k < 3; ++k) {
        for (int a = 0; a < 3; ++a) I.P.k[k].d[a] = kRayDirs[k][a];
        carve_basis(kRayDirs[k], I.P.k[k].u, I.P.k[k].v);
    }

    // bounds (projections or coordinates) from per-block partials, reduced on the host
    const unsigned nb = grid_for((size_t)npts, 256, 4);
    const int nv = rays ? 12 : 6;
    double* d_part = I.dalloc<double>((size_t)nb * nv);
    if (rays)
        bounds_kernel<true><<<nb, 256, 0, st>>>(d_xyz, npts, I.P, d_part);
    else
        bounds_kernel<false><<<nb, 256, 0, st>>>(d_xyz, npts, I.P, d_part);
    CMG_HIP_LAUNCH();
    unsigned long long* d_badtri = I.dalloc<unsigned long long>(1);
    CMG_HIP(hipMemsetAsync(d_badtri, 0, sizeof(unsigned long long), st));
    tri_check_kernel<<<grid_for((size_t)(3 * ntris), 256, 4), 256, 0, st>>>(d_tris, ntris,
                                                                           npts, d_badtri);
    CMG_HIP_LAUNCH();
    std::vector<double> part((size_t)nb * nv);
    CMG_HIP(hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(),
                           hipMemcpyDeviceToHost, st));
    unsigned long long badtri = 0;
    CMG_HIP(hipMemcpyAsync(&badtri, d_badtri, sizeof badtri, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    if (badtri)
        throw InvalidInput("carve: " + std::to_string(badtri) +
                           " triangle vertex indices outside the point array");
    std::vector<double> b(nv);
    for (int j = 0; j < nv; j += 2) b[j] = 1e300, b[j + 1] = -1e300;
    for (unsigned blk = 0; blk < nb; ++blk)
        for (int j = 0; j < nv; j += 2) {
            b[j] = std::fmin(b[j], part[(size_t)blk * nv + j]);
            b[j + 1] = std::fmax(b[j + 1], part[(size_t)blk * nv + j + 1]);
        }
    if (rays) {
        // tetrahedralize_plc.cpp:139-144
        for (int k = 0; k < 3; ++k) {
            const double lou = b[4 * k], hiu = b[4 * k + 1], lov = b[4 * k + 2],
                         hiv = b[4 * k + 3];
            const double eu = hiu - lou, ev = hiv - lov;
            I.P.k[k].min_u = lou;
            I.P.k[k].min_v = lov;
            I.P.k[k].inv_u = r / (eu > 0 ? eu * 1.0000001 : 1.0);
            I.P.k[k].inv_v = r / (ev > 0 ? ev * 1.0000001 : 1.0);
        }
        I.ncell = 3LL * r * r;
    } else {
        // make_grid over the two corners reproduces make_grid over all points
        const double two[6] = {b[0], b[2], b[4], b[1], b[3], b[5]};
        const int res = resolution > 0 ? resolution : r;
        if (!make_grid(two, 2, res, 1, I.g))
            throw InvalidInput("carve: degenerate (zero-extent) point set");
        if ((long long)I.g.nx * I.g.ny > (1LL << 24))
            throw InvalidInput("carve: parity grid has more than 2^24 columns");
        I.ncell = (long long)I.g.nx * I.g.ny;
    }

    // cell lists: rectangles and counts, scan, fill
    const int ng = rays ? 3 : 1;
    I.tv = I.dalloc<double>((size_t)ntris * 9);
    Rect* d_rect = I.dalloc<Rect>((size_t)ntris * ng);
    unsigned* d_count = I.dalloc<unsigned>((size_t)I.ncell);
    I.start = I.dalloc<unsigned>((size_t)I.ncell + 1);
    I.bad = I.dalloc<unsigned long long>(1);
    unsigned long long* d_entries = I.dalloc<unsigned long long>(1);
    CMG_HIP(hipMemsetAsync(d_count, 0, sizeof(unsigned) * (size_t)I.ncell, st));
    CMG_HIP(hipMemsetAsync(I.bad, 0, sizeof(unsigned long long), st));
    CMG_HIP(hipMemsetAsync(d_entries, 0, sizeof(unsigned long long), st));
    const unsigned tb = grid_for((size_t)ntris, 256, 8);
    if (rays)
        rect_kernel<true><<<tb, 256, 0, st>>>(d_xyz, d_tris, ntris, I.P, I.g, I.tv, d_rect,
                                              d_entries);
    else
        rect_kernel<false><<<tb, 256, 0, st>>>(d_xyz, d_tris, ntris, I.P, I.g, I.tv, d_rect,
                                               d_entries);
    CMG_HIP_LAUNCH();
    // The list total is summed over the rectangles in 64 bits and checked before any
    // per-cell count, scan or fill: the counts and list offsets are 32-bit, so more than
    // 2^31 entries in all is rejected.
    CMG_HIP(hipMemcpyAsync(&I.entries, d_entries, sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    if (I.entries >= (1ULL << 31))
        throw InvalidInput("carve: cell lists exceed 2^31 entries");
    if (rays)
        count_kernel<true><<<tb, 256, 0, st>>>(ntris, I.P, I.g, d_rect, d_count);
    else
        count_kernel<false><<<tb, 256, 0, st>>>(ntris, I.P, I.g, d_rect, d_count);
    CMG_HIP_LAUNCH();
    // start = exclusive prefix of the counts, start[ncell] = the total
    size_t scan_tmp = 0;
    CMG_HIP(rocprim::exclusive_scan(nullptr, scan_tmp, d_count, I.start, 0u, (size_t)I.ncell,
                                    rocprim::plus<unsigned>(), st));
    void* tmp = I.dalloc<char>(scan_tmp);
    CMG_HIP(rocprim::exclusive_scan(tmp, scan_tmp, d_count, I.start, 0u, (size_t)I.ncell,
                                    rocprim::plus<unsigned>(), st));
    {
        const unsigned tot = (unsigned)I.entries;
        CMG_HIP(hipMemcpyAsync(I.start + I.ncell, &tot, sizeof(unsigned),
                               hipMemcpyHostToDevice, st));
        CMG_HIP(hipStreamSynchronize(st));
    }
    I.items = I.dalloc<int>((size_t)I.entries);
    CMG_HIP(hipMemcpyAsync(d_count, I.start, sizeof(unsigned) * (size_t)I.ncell,
                           hipMemcpyDeviceToDevice, st));
    if (rays)
        fill_kernel<true><<<tb, 256, 0, st>>>(ntris, I.P, I.g, d_rect, d_count, I.items);
    else
        fill_kernel<false><<<tb, 256, 0, st>>>(ntris, I.P, I.g, d_rect, d_count, I.items);
    CMG_HIP_LAUNCH();
    CMG_HIP(hipStreamSynchronize(st));
}

CarveEngine::~CarveEngine() = default;

int CarveEngine::mode() const { return p_->mode; }
int CarveEngine::cells_per_axis() const {
    return p_->mode == kCarveRays ? p_->P.r : p_->g.nx;
}
long long CarveEngine::list_entries() const { return (long long)p_->entries; }

// Orders the queries by cell (a radix sort of (cell, query) pairs) and runs the pass
// kernels. Work buffers come from the stream-ordered allocator: calls stay
// asynchronous.
template <bool TETS>
void classify(const CarveEngine::Impl& I, const double* d_q, const int* d_tets, long long n,
              uint8_t* d_out, hipStream_t st);

void CarveEngine::points(const double* d_q, long long nq, uint8_t* d_inside,
                         hipStream_t st) {
    if (nq < 0) throw InvalidInput("carve: negative count");
    if (nq == 0) return;
    if (!d_q || !d_inside) throw InvalidInput("carve: null pointer");
    classify<false>(*p_, d_q, nullptr, nq, d_inside, st);
}

void CarveEngine::tets(const int* d_tets, long long nt, uint8_t* d_keep, hipStream_t st) {
    if (nt < 0) throw InvalidInput("carve: negative count");
    if (nt == 0) return;
    if (!d_tets || !d_keep) throw InvalidInput("carve: null pointer");
    if (reinterpret_cast<uintptr_t>(d_tets) % 16)
        throw InvalidInput("carve: tets must be 16-byte aligned");
    classify<true>(*p_, nullptr, d_tets, nt, d_keep, st);
}

template <bool TETS>
void classify(const CarveEngine::Impl& I, const double* d_q, const int* d_tets, long long n,
              uint8_t* d_out, hipStream_t st) {
    if (n >= (1LL << 30)) throw InvalidInput("carve: more than 2^30 queries per call");
    const bool rays = I.mode == kCarveRays;
    const int grids = rays ? 3 : 1;
    const unsigned ncell = (unsigned)I.ncell;
    const long long pairs = grids * n;
    int bits = 1;
    while ((1ull << bits) <= ncell) ++bits;  // keys 0..ncell
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t sort_tmp = 0;
    CMG_HIP(rocprim::radix_sort_pairs(nullptr, sort_tmp, (const unsigned*)nullptr,
                                      (unsigned*)nullptr, (const unsigned*)nullptr,
                                      (unsigned*)nullptr, (size_t)pairs, 0, bits, st));
    const size_t b_q = TETS ? al(sizeof(double) * 3 * n) : 0;
    const size_t b_pair = al(sizeof(unsigned) * pairs);
    const size_t b_odd = rays ? al((size_t)3 * n) : 0;
    const size_t total = b_q + 4 * b_pair + b_odd + al(8) + al(sort_tmp);
    char* base = nullptr;
    CMG_HIP(hipMallocAsync(reinterpret_cast<void**>(&base), total, st));
    char* cur = base;
    auto take = [&](size_t b) {
        char* p = cur;
        cur += b;
        return p;
    };
    double* qbuf = TETS ? reinterpret_cast<double*>(take(b_q)) : nullptr;
    unsigned* key = reinterpret_cast<unsigned*>(take(b_pair));
    unsigned* val = reinterpret_cast<unsigned*>(take(b_pair));
    unsigned* skey = reinterpret_cast<unsigned*>(take(b_pair));
    unsigned* sval = reinterpret_cast<unsigned*>(take(b_pair));
    uint8_t* odd = rays ? reinterpret_cast<uint8_t*>(take(b_odd)) : nullptr;
    unsigned* nbad = reinterpret_cast<unsigned*>(take(al(8)));
    void* tmp = take(al(sort_tmp));
    const double* qpts = TETS ? qbuf : d_q;
    // 20 waves per CU: the ray passes' registers allow 5 per SIMD to hide their dependent
    // loads; the parity pass (2 per SIMD) keeps a grid of 8 per CU, measured faster
    // (multiples of 8: the passes renumber their workgroups by XCD, xcd_block)
    const unsigned nb = (grid_for((size_t)n, 256, 20) + 7u) & ~7u;
    const unsigned nb_parity = (grid_for((size_t)n, 256, 8) + 7u) & ~7u;
    {
        TimedLaunch tl("carve_query", st);
        CMG_HIP(hipMemsetAsync(nbad, 0, sizeof(unsigned), st));
        if (rays)
            query_prep_kernel<TETS, true><<<nb, 256, 0, st>>>(d_q, d_tets, I.xyz, I.npts, n,
                                                              I.P, I.g, ncell, qbuf, key, val,
                                                              d_out, nbad, I.bad);
        else
            query_prep_kernel<TETS, false><<<nb, 256, 0, st>>>(d_q, d_tets, I.xyz, I.npts, n,
                                                               I.P, I.g, ncell, qbuf, key, val,
                                                               d_out, nbad, I.bad);
        CMG_HIP_LAUNCH();
        CMG_HIP(rocprim::radix_sort_pairs(tmp, sort_tmp, key, skey, val, sval, (size_t)pairs,
                                          0, bits, st));
        if (rays) {
            rays_pass_kernel<0><<<nb, 256, 0, st>>>(qpts, n, I.P, nbad, skey, sval, I.start,
                                                    I.items, I.tv, odd);
            rays_pass_kernel<1><<<nb, 256, 0, st>>>(qpts, n, I.P, nbad, skey, sval, I.start,
                                                    I.items, I.tv, odd);
            rays_pass_kernel<2><<<nb, 256, 0, st>>>(qpts, n, I.P, nbad, skey, sval, I.start,
                                                    I.items, I.tv, odd);
            CMG_HIP_LAUNCH();
            rays_vote_kernel<<<nb, 256, 0, st>>>(key, ncell, n, odd, d_out);
        } else {
            parity_pass_kernel<<<nb_parity, 256, 0, st>>>(qpts, n, I.g, nbad, skey, sval,
                                                          I.start, I.items, I.tv, d_out);
        }
        CMG_HIP_LAUNCH();
    }
    CMG_HIP(hipFreeAsync(base, st));
}

unsigned long long CarveEngine::bad_tets(hipStream_t st) {
    unsigned long long b = 0;
    CMG_HIP(hipMemcpyAsync(&b, p_->bad, sizeof b, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    if (b) CMG_HIP(hipMemsetAsync(p_->bad, 0, sizeof b, st));
    return b;
}

}  // namespace cmg::hip
