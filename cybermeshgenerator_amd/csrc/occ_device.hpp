// occ_device.hpp — device-side pieces of the occupancy path (batch.hip): bucket ranges,
// the compat orient3d, plane heights and the cell count below a crossing, each following
// src/voxelize/voxelize.cpp operation for operation. Every TU that includes it is
// compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

#include "cmg_hip_internal.hpp"
#include "expansion.hpp"

namespace cmg::hip {
namespace occdev {

// voxelize.cpp:148-151
static __constant__ double kNudge[6][2] = {
    {1.0e-6, 3.0e-7},   {-7.0e-7, 5.0e-7}, {4.0e-7, -9.0e-7},
    {-2.0e-6, -1.0e-6}, {8.0e-7, 6.0e-7},  {-5.0e-7, -4.0e-6},
};

constexpr int kTile = 8;    // 8x8 columns per workgroup
constexpr int kChunk = 64;  // triangles staged in LDS per step

__device__ __forceinline__ int clamp_cell(double v, double corner, double spacing, int n) {
    // voxelize.cpp:82-89 (floor, then clamp into the grid)
    double f = floor((v - corner) / spacing);
    if (!(f >= 0.0)) return 0;          // negative (or NaN)
    if (f >= (double)n) return n - 1;
    return (int)f;
}

// Exact sign for the rare lanes whose naive determinant is exactly 0; kept out of line
// so the expansion arrays do not weigh on the common path.
__device__ __noinline__ int o3_exact_slow(const double* pa, const double* pb,
                                          const double* pc, const double* pd) {
    int s = orient3d_exact_sign<40>(pa, pb, pc, pd);
    if (s == kOverflow) s = orient3d_exact_sign<320>(pa, pb, pc, pd);
    return s == kOverflow ? 0 : s;
}

// Sign of cmg::robust::orient3d(pa, pb, pc, pd) as shipped: the naive FP64 determinant
// (predicates.cpp:2201-2222) when nonzero, else the exact sign.
__device__ __forceinline__ int o3_compat(const double* pa, const double* pb, const double* pc,
                                         const double* pd) {
    double adx = pa[0] - pd[0], ady = pa[1] - pd[1], adz = pa[2] - pd[2];
    double bdx = pb[0] - pd[0], bdy = pb[1] - pd[1], bdz = pb[2] - pd[2];
    double cdx = pc[0] - pd[0], cdy = pc[1] - pd[1], cdz = pc[2] - pd[2];
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    double cdxady = cdx * ady, adxcdy = adx * cdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    if (det > 0.0) return 1;
    if (det < 0.0) return -1;
    return o3_exact_slow(pa, pb, pc, pd);
}

// voxelize.cpp:107-114
__device__ __forceinline__ double plane_z(const double* a, const double* b, const double* c,
                                          double x, double y) {
    double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    double vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    double nx = uy * vz - uz * vy;
    double ny = uz * vx - ux * vz;
    double nz = ux * vy - uy * vx;
    return a[2] - (nx * (x - a[0]) + ny * (y - a[1])) / nz;
}

__device__ __forceinline__ double zc_of(const GridSpec& g, int k) {
    return g.corner[2] + (k + 0.5) * g.spacing;  // voxelize.cpp:187
}

// Number of cells k in [0, nz) with zc_k < h (the cells a crossing at h lies above).
__device__ int cells_below(const GridSpec& g, double h) {
    const int nz = g.nz;
    if (!(h > zc_of(g, 0))) return 0;  // also NaN
    if (h > zc_of(g, nz - 1)) return nz;
    double f = (h - g.corner[2]) / g.spacing - 0.5;
    int k = f < 0.0 ? 0 : (f > (double)(nz - 1) ? nz - 1 : (int)f);
    while (k + 1 < nz && zc_of(g, k + 1) < h) ++k;
    while (k > 0 && !(zc_of(g, k) < h)) --k;
    return k + 1;
}


// xy bucket range of one triangle (voxelize.cpp:90-97): i0, i1, j0, j1.
__device__ inline int4 xy_range(const double (&P)[3][3], const GridSpec& g) {
    int i0 = g.nx, i1 = 0, j0 = g.ny, j1 = 0;
    for (int v = 0; v < 3; ++v) {
        int i = clamp_cell(P[v][0], g.corner[0], g.spacing, g.nx);
        int j = clamp_cell(P[v][1], g.corner[1], g.spacing, g.ny);
        i0 = min(i0, i);
        i1 = max(i1, i);
        j0 = min(j0, j);
        j1 = max(j1, j);
    }
    return make_int4(i0, i1, j0, j1);
}

}  // namespace occdev
}  // namespace cmg::hip
