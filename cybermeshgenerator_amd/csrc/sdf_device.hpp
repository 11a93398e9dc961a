// sdf_device.hpp — device-side pieces of the signed-distance path (batch.hip): the FP32
// triangle record with its bounding disk, voxel-leaf geometry, the bounding-cylinder
// lower bound, the FP32 point-triangle distance and closest-point barycentrics, and the
// reference's FP64 distance (src/voxelize/voxelize.cpp:197-247).
// Every TU that includes it is compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "cmg_hip_internal.hpp"

namespace cmg::hip {
namespace sdfdev {

constexpr int kLeaf = 4;          // 4^3 voxels per wavefront
constexpr float kU = 0x1p-24f;    // FP32 unit roundoff

// FP32 triangle record, coordinates relative to the grid corner (128 bytes): the vertex
// a and edge vectors for the distance, the dual basis and unit normal for its interior
// test, and an oriented bounding box (axes n, u along the longest edge, w = n x u;
// centre c; half-extents en, eu, ew) for the lower bound.
struct __align__(16) TriRec {
    float dn, du, dw, en;      // box centre along n, u, w; half-extent along n
    float ax, ay, az, ub;      // vertex a; ub = 1 (or -1: no interior, degenerate)
    float abx, aby, abz, eu;   // b - a; half-extent along u
    float acx, acy, acz, ew;   // c - a; half-extent along w
    float ux, uy, uz, nx;      // box axis u; unit normal x
    float wx, wy, wz, ny;      // box axis w; unit normal y
    float g1x, g1y, g1z, nz;   // dual basis: v = (p - a).g1, w = (p - a).g2; normal z
    float g2x, g2y, g2z, iab;  // (degenerate triangle: n = (0,0,1), u = x, w = y)
    float iac, ibc, pad0, pad1; // iab, iac, ibc: 1 / squared edge length (0 for no length)
};
static_assert(sizeof(TriRec) == 144, "record size");
constexpr int kTriRecF4 = (int)(sizeof(TriRec) / 16);  // float4s per record

struct __align__(8) TriRec64 {  // FP64 vertices for the near-surface re-evaluation
    double a[3], b[3], c[3];
};

struct Level {
    int s;           // brick edge in voxels along x and y
    int sz;          // brick edge along z
    int bx, by, bz;  // bricks per axis
};

__host__ __device__ inline Level make_level(const GridSpec& g, int s, int sz = 0) {
    if (sz == 0) sz = s;
    return Level{s, sz, (g.nx + s - 1) / s, (g.ny + s - 1) / s, (g.nz + sz - 1) / sz};
}

// Centre (relative to the corner) and half-diagonal of brick b's voxel centres.
__device__ inline void brick_geom(const GridSpec& g, const Level& L, int b, float& cx,
                                  float& cy, float& cz, float& hd) {
    const int ix = b % L.bx, iy = (b / L.bx) % L.by, iz = b / (L.bx * L.by);
    const int lx = ix * L.s, ly = iy * L.s, lz = iz * L.sz;
    const int hx = min(lx + L.s, g.nx) - 1, hy = min(ly + L.s, g.ny) - 1,
              hz = min(lz + L.sz, g.nz) - 1;
    const float h = (float)g.spacing;
    cx = (0.5f * (lx + hx) + 0.5f) * h;
    cy = (0.5f * (ly + hy) + 0.5f) * h;
    cz = (0.5f * (lz + hz) + 0.5f) * h;
    const float ex = 0.5f * (hx - lx) * h, ey = 0.5f * (hy - ly) * h, ez = 0.5f * (hz - lz) * h;
    hd = sqrtf(ex * ex + ey * ey + ez * ez) * (1.0f + 4.0f * kU);
}

__device__ __forceinline__ float fma3(float a, float b, float c, float d, float e, float f) {
    return __builtin_fmaf(a, b, __builtin_fmaf(c, d, e * f));
}

// a . p - d with three roundings: the signed distance of p from the plane a . x = d
__device__ __forceinline__ float proj(float ax, float ay, float az, float d, float px, float py,
                                      float pz) {
    return __builtin_fmaf(ax, px, __builtin_fmaf(ay, py, __builtin_fmaf(az, pz, -d)));
}

// Largest grid dimension times the spacing: every cell centre p (relative to the grid
// corner) has 0 <= p_i <= grid_extent.
__host__ __device__ inline double grid_extent(const GridSpec& g) {
    const int m = g.nx > g.ny ? (g.nx > g.nz ? g.nx : g.nz) : (g.ny > g.nz ? g.ny : g.nz);
    return m * g.spacing;
}

__device__ __forceinline__ float clamp01(float x) {
    return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f);
}

// At least sqrt(x) (1 + 4u), from the hardware square root (within 1 ulp) and one product:
// the scale of the FP32 search bounds.
__device__ __forceinline__ float bound_root(float x) {
    return __builtin_amdgcn_sqrtf(x) * (1.0f + 8.0f * kU);
}

// Squared FP32 distance from p (relative to the corner) to the triangle.
__device__ __forceinline__ float tri_d2(const TriRec& T, float px, float py, float pz) {
    const float apx = px - T.ax, apy = py - T.ay, apz = pz - T.az;
    const float bcx = T.acx - T.abx, bcy = T.acy - T.aby, bcz = T.acz - T.abz;
    // edge ab
    float t = clamp01(fma3(apx, T.abx, apy, T.aby, apz, T.abz) * T.iab);
    float ex = __builtin_fmaf(-t, T.abx, apx), ey = __builtin_fmaf(-t, T.aby, apy),
          ez = __builtin_fmaf(-t, T.abz, apz);
    float d = fma3(ex, ex, ey, ey, ez, ez);
    // edge ac
    t = clamp01(fma3(apx, T.acx, apy, T.acy, apz, T.acz) * T.iac);
    ex = __builtin_fmaf(-t, T.acx, apx);
    ey = __builtin_fmaf(-t, T.acy, apy);
    ez = __builtin_fmaf(-t, T.acz, apz);
    d = fminf(d, fma3(ex, ex, ey, ey, ez, ez));
    // edge bc, from b
    const float bpx = apx - T.abx, bpy = apy - T.aby, bpz = apz - T.abz;
    t = clamp01(fma3(bpx, bcx, bpy, bcy, bpz, bcz) * T.ibc);
    ex = __builtin_fmaf(-t, bcx, bpx);
    ey = __builtin_fmaf(-t, bcy, bpy);
    ez = __builtin_fmaf(-t, bcz, bpz);
    d = fminf(d, fma3(ex, ex, ey, ey, ez, ez));
    // interior: plane distance when the projection falls inside
    const float v = fma3(apx, T.g1x, apy, T.g1y, apz, T.g1z);
    const float w = fma3(apx, T.g2x, apy, T.g2y, apz, T.g2z);
    const float hn = fma3(apx, T.nx, apy, T.ny, apz, T.nz);
    const bool inside = v >= 0.0f && w >= 0.0f && v + w <= T.ub;
    return inside ? fminf(d, hn * hn) : d;
}

// Ericson closest-point barycentrics (voxelize.cpp:202-247 regions) in FP32, for the
// adjoint. Returns (v, w); the weight of a is 1 - v - w.
__device__ inline void tri_bary(const TriRec& T, float px, float py, float pz, float& v,
                                float& w) {
    const float apx = px - T.ax, apy = py - T.ay, apz = pz - T.az;
    const float d1 = fma3(T.abx, apx, T.aby, apy, T.abz, apz);
    const float d2 = fma3(T.acx, apx, T.acy, apy, T.acz, apz);
    v = 0.0f;
    w = 0.0f;
    if (d1 <= 0.0f && d2 <= 0.0f) return;  // vertex a
    const float bpx = apx - T.abx, bpy = apy - T.aby, bpz = apz - T.abz;
    const float d3 = fma3(T.abx, bpx, T.aby, bpy, T.abz, bpz);
    const float d4 = fma3(T.acx, bpx, T.acy, bpy, T.acz, bpz);
    if (d3 >= 0.0f && d4 <= d3) {  // vertex b
        v = 1.0f;
        return;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {  // edge ab
        v = d1 / (d1 - d3);
        return;
    }
    const float cpx = apx - T.acx, cpy = apy - T.acy, cpz = apz - T.acz;
    const float d5 = fma3(T.abx, cpx, T.aby, cpy, T.abz, cpz);
    const float d6 = fma3(T.acx, cpx, T.acy, cpy, T.acz, cpz);
    if (d6 >= 0.0f && d5 <= d6) {  // vertex c
        w = 1.0f;
        return;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {  // edge ac
        w = d2 / (d2 - d6);
        return;
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {  // edge bc
        w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        v = 1.0f - w;
        return;
    }
    const float denom = 1.0f / (va + vb + vc);  // interior
    v = vb * denom;
    w = vc * denom;
}

// The reference's FP64 squared distance, same operation order (voxelize.cpp:197-247).
__device__ inline double dot3d(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ inline double dist2_to(const double* p, const double* q) {
    double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
    return dx * dx + dy * dy + dz * dz;
}

// With bary given, also the barycentric weights (of a, b, c) of the closest point.
__device__ double point_tri_dist2_ref(const double* p, const double* a, const double* b,
                                      const double* c, double* bary = nullptr) {
    double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double d1 = dot3d(ab, ap), d2 = dot3d(ac, ap);
    auto set = [&](double u, double v, double w) {
        if (bary) {
            bary[0] = u;
            bary[1] = v;
            bary[2] = w;
        }
    };
    if (d1 <= 0 && d2 <= 0) {
        set(1.0, 0.0, 0.0);
        return dist2_to(p, a);
    }
    double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    double d3 = dot3d(ab, bp), d4 = dot3d(ac, bp);
    if (d3 >= 0 && d4 <= d3) {
        set(0.0, 1.0, 0.0);
        return dist2_to(p, b);
    }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        double v = d1 / (d1 - d3);
        double q[3] = {a[0] + v * ab[0], a[1] + v * ab[1], a[2] + v * ab[2]};
        set(1.0 - v, v, 0.0);
        return dist2_to(p, q);
    }
    double cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    double d5 = dot3d(ab, cp), d6 = dot3d(ac, cp);
    if (d6 >= 0 && d5 <= d6) {
        set(0.0, 0.0, 1.0);
        return dist2_to(p, c);
    }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        double w = d2 / (d2 - d6);
        double q[3] = {a[0] + w * ac[0], a[1] + w * ac[1], a[2] + w * ac[2]};
        set(1.0 - w, 0.0, w);
        return dist2_to(p, q);
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        double q[3] = {b[0] + w * (c[0] - b[0]), b[1] + w * (c[1] - b[1]),
                       b[2] + w * (c[2] - b[2])};
        set(0.0, 1.0 - w, w);
        return dist2_to(p, q);
    }
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom, w = vc * denom;
    double q[3] = {a[0] + ab[0] * v + ac[0] * w, a[1] + ab[1] * v + ac[1] * w,
                   a[2] + ab[2] * v + ac[2] * w};
    set(1.0 - v - w, v, w);
    return dist2_to(p, q);
}

// Loads of records every lane needs (uniform index t) through the constant address
// space, so that they become scalar loads into SGPRs (broadcast operands of the VALU)
// instead of 64 identical vector loads. The records are read-only during every kernel.
typedef const __attribute__((address_space(4))) float* ConstF32Ptr;
typedef const __attribute__((address_space(4))) double* ConstF64Ptr;

__device__ __forceinline__ TriRec load_uniform(const TriRec* base, int t) {
    ConstF32Ptr p = (ConstF32Ptr)(base + t);
    TriRec r;
    float* d = reinterpret_cast<float*>(&r);
#pragma unroll
    for (int i = 0; i < 36; ++i) d[i] = p[i];
    return r;
}

__device__ __forceinline__ TriRec64 load_uniform(const TriRec64* base, int t) {
    ConstF64Ptr p = (ConstF64Ptr)(base + t);
    TriRec64 r;
    double* d = reinterpret_cast<double*>(&r);
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] = p[i];
    return r;
}


// Fills the FP32 and FP64 records of one triangle (vertices P[3][3], world coordinates).
// The FP32 record is in the mesh's search units: coordinates relative to the grid corner
// times sc, a power of two (exact).
__device__ inline void prep_triangle(const double (&P)[3][3], const GridSpec& g, double sc,
                                     TriRec& r, TriRec64& r64) {
    for (int a = 0; a < 3; ++a) {
        r64.a[a] = P[0][a];
        r64.b[a] = P[1][a];
        r64.c[a] = P[2][a];
    }
    double A[3], AB[3], AC[3], BC[3];
    for (int a = 0; a < 3; ++a) {
        A[a] = (P[0][a] - g.corner[a]) * sc;
        AB[a] = (P[1][a] - P[0][a]) * sc;
        AC[a] = (P[2][a] - P[0][a]) * sc;
        BC[a] = (P[2][a] - P[1][a]) * sc;
    }
    const double a11 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    const double a22 = AC[0] * AC[0] + AC[1] * AC[1] + AC[2] * AC[2];
    const double a33 = BC[0] * BC[0] + BC[1] * BC[1] + BC[2] * BC[2];
    const double a12 = AB[0] * AC[0] + AB[1] * AC[1] + AB[2] * AC[2];
    const double det = a11 * a22 - a12 * a12;
    double N[3] = {AB[1] * AC[2] - AB[2] * AC[1], AB[2] * AC[0] - AB[0] * AC[2],
                   AB[0] * AC[1] - AB[1] * AC[0]};
    const double nn = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    // degenerate (no interior) when the area is negligible against the edge lengths
    const bool degen = !(det > 1e-24 * a11 * a22) || !(nn > 0.0);
    r.ax = (float)A[0];
    r.ay = (float)A[1];
    r.az = (float)A[2];
    r.ub = degen ? -1.0f : 1.0f;
    r.abx = (float)AB[0];
    r.aby = (float)AB[1];
    r.abz = (float)AB[2];
    r.acx = (float)AC[0];
    r.acy = (float)AC[1];
    r.acz = (float)AC[2];
    // 1 / |e|^2 of the float edges (e = ab, ac and bc = ac - ab as the distance forms
    // them), rounded once; 0 when the edge has no length
    {
        const double e[3][3] = {{r.abx, r.aby, r.abz},
                                {r.acx, r.acy, r.acz},
                                {(double)(r.acx - r.abx), (double)(r.acy - r.aby),
                                 (double)(r.acz - r.abz)}};
        float* inv[3] = {&r.iab, &r.iac, &r.ibc};
        for (int q = 0; q < 3; ++q) {
            const double l = e[q][0] * e[q][0] + e[q][1] * e[q][1] + e[q][2] * e[q][2];
            *inv[q] = l > 0.0 ? (float)fmin(1.0 / l, (double)FLT_MAX) : 0.0f;
        }
    }
    r.pad0 = r.pad1 = 0.0f;
    // box axes: the unit normal, the longest edge's direction and their cross product;
    // the coordinate axes for a degenerate triangle
    double n[3] = {0.0, 0.0, 1.0}, u[3] = {1.0, 0.0, 0.0}, w[3] = {0.0, 1.0, 0.0};
    if (degen) {
        r.g1x = r.g1y = r.g1z = r.g2x = r.g2y = r.g2z = 0.0f;
    } else {
        for (int a = 0; a < 3; ++a) {
            (&r.g1x)[a] = (float)((a22 * AB[a] - a12 * AC[a]) / det);
            (&r.g2x)[a] = (float)((a11 * AC[a] - a12 * AB[a]) / det);
            n[a] = N[a] / nn;
        }
        const double* E = a11 >= a22 ? (a11 >= a33 ? AB : BC) : (a22 >= a33 ? AC : BC);
        const double el = sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
        for (int a = 0; a < 3; ++a) u[a] = E[a] / el;
        w[0] = n[1] * u[2] - n[2] * u[1];
        w[1] = n[2] * u[0] - n[0] * u[2];
        w[2] = n[0] * u[1] - n[1] * u[0];
    }
    // extents measured along the float axes (so the float box holds the vertices), grown
    // by the rounding of the vertices and of the centre and by the rounding error of the
    // test's projection fma(n, p, -dn) for any point p of the grid (|p|_1 <= pmax)
    float nf[3], uf[3], wf[3];
    for (int a = 0; a < 3; ++a) {
        nf[a] = (float)n[a];
        uf[a] = (float)u[a];
        wf[a] = (float)w[a];
    }
    const float* axf[3] = {nf, uf, wf};
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double ext = 0.0;
    for (int v = 0; v < 3; ++v) {
        double q[3];
        for (int a = 0; a < 3; ++a) {
            q[a] = (P[v][a] - g.corner[a]) * sc;
            ext = fmax(ext, fabs(q[a]));
        }
        for (int k = 0; k < 3; ++k) {
            const double d = q[0] * axf[k][0] + q[1] * axf[k][1] + q[2] * axf[k][2];
            lo[k] = fmin(lo[k], d);
            hi[k] = fmax(hi[k], d);
        }
    }
    const double pmax = 3.0 * grid_extent(g) * sc;
    double d[3], e[3];
    for (int k = 0; k < 3; ++k) {
        d[k] = 0.5 * (lo[k] + hi[k]);
        e[k] = 0.5 * (hi[k] - lo[k]) * (1.0 + 1e-6) +
               8.0 * 0x1p-24 * (ext + pmax + fabs(d[k]));
    }
    r.dn = (float)d[0];
    r.du = (float)d[1];
    r.dw = (float)d[2];
    r.en = (float)e[0];
    r.eu = (float)e[1];
    r.ew = (float)e[2];
    r.ux = uf[0];
    r.uy = uf[1];
    r.uz = uf[2];
    r.wx = wf[0];
    r.wy = wf[1];
    r.wz = wf[2];
    r.nx = nf[0];
    r.ny = nf[1];
    r.nz = nf[2];
}

// Squared distance from p to the triangle's oriented box: a lower bound on its distance.
// In search units every |projection| is below 1, so clamping to [0, 1] is max(., 0) and
// folds into the subtraction (one instruction per axis).
__device__ __forceinline__ float tri_lb2(const TriRec& T, float px, float py, float pz) {
    const float tn = clamp01(fabsf(proj(T.nx, T.ny, T.nz, T.dn, px, py, pz)) - T.en);
    const float tu = clamp01(fabsf(proj(T.ux, T.uy, T.uz, T.du, px, py, pz)) - T.eu);
    const float tw = clamp01(fabsf(proj(T.wx, T.wy, T.wz, T.dw, px, py, pz)) - T.ew);
    return fma3(tn, tn, tu, tu, tw, tw);
}

// Search units of a grid: 2^-e with 2^(e-2) >= grid extent, so every grid point, box
// centre and projection has magnitude below 1/2 * sqrt(3) * 2 < 1.
inline double search_scale(const GridSpec& g) {
    int e = 0;
    std::frexp(grid_extent(g), &e);  // grid extent < 2^e
    return std::ldexp(1.0, -(e + 2));
}
// Distances at or below this are zero up to the FP64 rounding of the coordinates: such
// a cell lies on the surface, where the unsigned distance has a kink, and carries no
// gradient (2^-40 of the grid extent).
inline double zero_distance(const GridSpec& g) {
    const int m = g.nx > g.ny ? (g.nx > g.nz ? g.nx : g.nz) : (g.ny > g.nz ? g.ny : g.nz);
    return std::ldexp(m * g.spacing, -40);
}

// FP32 error margin for distances measured from the grid corner: a few ulps of the
// largest coordinate involved (the grid extent), covering the rounding of the inputs to
// float and of every operation in tri_d2.
inline float distance_eps(const GridSpec& g) {
    const int m = g.nx > g.ny ? (g.nx > g.nz ? g.nx : g.nz) : (g.ny > g.nz ? g.ny : g.nz);
    const double ext = m * g.spacing * 1.7320508075688772;
    return (float)(64.0 * 0x1p-24 * ext + 1e-6 * g.spacing);
}

}  // namespace sdfdev
}  // namespace cmg::hip
