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

// FP32 triangle record, coordinates relative to the grid corner (128 bytes).
struct __align__(16) TriRec {
    float mx, my, mz, r;       // bounding sphere (centroid, radius rounded up)
    float ax, ay, az, ub;      // vertex a; ub = 1 (or -1: no interior, degenerate)
    float abx, aby, abz, iab;  // b - a and 1/|b - a|^2 (0 for a zero-length edge)
    float acx, acy, acz, iac;  // c - a and 1/|c - a|^2
    float bcx, bcy, bcz, ibc;  // c - b and 1/|c - b|^2
    float g1x, g1y, g1z, pad0; // dual basis: v = (p - a).g1, w = (p - a).g2
    float g2x, g2y, g2z, pad1;
    float nx, ny, nz, t;       // unit normal and half-thickness of the bounding disk
                               // (degenerate triangle: normal (0,0,1), t = r)
};
static_assert(sizeof(TriRec) == 128, "record size");

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

__device__ __forceinline__ float clamp01(float x) {
    return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f);
}

// Squared FP32 distance from p (relative to the corner) to the triangle.
__device__ __forceinline__ float tri_d2(const TriRec& T, float px, float py, float pz) {
    const float apx = px - T.ax, apy = py - T.ay, apz = pz - T.az;
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
    t = clamp01(fma3(bpx, T.bcx, bpy, T.bcy, bpz, T.bcz) * T.ibc);
    ex = __builtin_fmaf(-t, T.bcx, bpx);
    ey = __builtin_fmaf(-t, T.bcy, bpy);
    ez = __builtin_fmaf(-t, T.bcz, bpz);
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

// ---- wave helpers ------------------------------------------------------------------

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

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
    for (int i = 0; i < 32; ++i) d[i] = p[i];
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
__device__ inline void prep_triangle(const double (&P)[3][3], const GridSpec& g, TriRec& r,
                                     TriRec64& r64) {
    for (int a = 0; a < 3; ++a) {
        r64.a[a] = P[0][a];
        r64.b[a] = P[1][a];
        r64.c[a] = P[2][a];
    }
    double A[3], AB[3], AC[3], BC[3];
    for (int a = 0; a < 3; ++a) {
        A[a] = P[0][a] - g.corner[a];
        AB[a] = P[1][a] - P[0][a];
        AC[a] = P[2][a] - P[0][a];
        BC[a] = P[2][a] - P[1][a];
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
    double M[3];
    double rad = 0.0;
    for (int a = 0; a < 3; ++a) M[a] = (P[0][a] + P[1][a] + P[2][a]) / 3.0;
    for (int v = 0; v < 3; ++v) {
        double dx = P[v][0] - M[0], dy = P[v][1] - M[1], dz = P[v][2] - M[2];
        rad = fmax(rad, sqrt(dx * dx + dy * dy + dz * dz));
    }
    r.mx = (float)(M[0] - g.corner[0]);
    r.my = (float)(M[1] - g.corner[1]);
    r.mz = (float)(M[2] - g.corner[2]);
    // rounded up: covers the float rounding of the centre and of the radius itself
    double ext = fmax(fmax(fabs(M[0] - g.corner[0]), fabs(M[1] - g.corner[1])),
                      fabs(M[2] - g.corner[2]));
    r.r = (float)(rad * (1.0 + 1e-6) + 4.0 * 0x1p-24 * ext);
    r.ax = (float)A[0];
    r.ay = (float)A[1];
    r.az = (float)A[2];
    r.ub = degen ? -1.0f : 1.0f;
    r.abx = (float)AB[0];
    r.aby = (float)AB[1];
    r.abz = (float)AB[2];
    r.iab = a11 > 0.0 ? (float)(1.0 / a11) : 0.0f;
    r.acx = (float)AC[0];
    r.acy = (float)AC[1];
    r.acz = (float)AC[2];
    r.iac = a22 > 0.0 ? (float)(1.0 / a22) : 0.0f;
    r.bcx = (float)BC[0];
    r.bcy = (float)BC[1];
    r.bcz = (float)BC[2];
    r.ibc = a33 > 0.0 ? (float)(1.0 / a33) : 0.0f;
    // bounding disk: centre m, radius r, unit normal; its half-thickness covers the float
    // rounding of the vertices and of the normal (a degenerate triangle gets a cylinder
    // around its bounding sphere instead)
    if (degen) {
        r.g1x = r.g1y = r.g1z = r.g2x = r.g2y = r.g2z = 0.0f;
        r.nx = r.ny = 0.0f;
        r.nz = 1.0f;
        r.t = r.r;
    } else {
        for (int a = 0; a < 3; ++a) {
            (&r.g1x)[a] = (float)((a22 * AB[a] - a12 * AC[a]) / det);
            (&r.g2x)[a] = (float)((a11 * AC[a] - a12 * AB[a]) / det);
            (&r.nx)[a] = (float)(N[a] / nn);
        }
        r.t = (float)(8.0 * 0x1p-24 * (ext + rad));
    }
    r.pad0 = r.pad1 = 0.0f;
}

// Squared distance from p to the solid cylinder with centre c, unit axis n, radius r and
// half-height t (a disk when t is 0): a lower bound on the distance to anything inside.
// The lateral distance comes from |n x v| (no cancellation).
__device__ __forceinline__ float cyl_lb2(float cx, float cy, float cz, float nx, float ny,
                                         float nz, float r, float t, float px, float py,
                                         float pz) {
    const float vx = px - cx, vy = py - cy, vz = pz - cz;
    const float h = fma3(vx, nx, vy, ny, vz, nz);
    const float wx = __builtin_fmaf(vy, nz, -vz * ny);
    const float wy = __builtin_fmaf(vz, nx, -vx * nz);
    const float wz = __builtin_fmaf(vx, ny, -vy * nx);
    const float rho = sqrtf(fma3(wx, wx, wy, wy, wz, wz));
    const float dh = fmaxf(fabsf(h) - t, 0.0f), dr = fmaxf(rho - r, 0.0f);
    return __builtin_fmaf(dh, dh, dr * dr);
}

__device__ __forceinline__ float tri_lb2(const TriRec& T, float px, float py, float pz) {
    return cyl_lb2(T.mx, T.my, T.mz, T.nx, T.ny, T.nz, T.r, T.t, px, py, pz);
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
