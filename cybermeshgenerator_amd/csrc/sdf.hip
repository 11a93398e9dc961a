// sdf.hip — signed-distance field and its adjoint on gfx950.
//
// The magnitude the reference stores (src/voxelize/voxelize.cpp:331-358) is the FP64
// minimum over all triangles of the Ericson point-triangle distance (:197-247), rounded
// once to float, signed by occupancy. Here:
//
// Cull (built on the GPU, no host search). Cells are grouped in cubic bricks of 16^3,
// 8^3 and 4^3 voxels. For a brick with centre c and half-diagonal hd (over its voxel
// centres), any voxel p in it satisfies |dist(p,T) - dist(c,T)| <= hd, so a triangle can
// be nearest to some voxel of the brick only if dist(c,T) <= min_T' dist(c,T') + 2 hd.
// The 16^3 lists are filtered from all triangles, the 8^3 lists from their parent's,
// and each 4^3 list is filtered inside the voxel kernel into LDS. Distances in the cull
// are FP32 with an explicit error margin, so no candidate that could be nearest is lost.
//
// Voxel pass. One wavefront per 4^3 brick, one lane per voxel; every lane visits the
// same triangle at the same time (broadcast scalar loads of a 128-byte FP32 record).
// A bounding-sphere test (wave-uniform ballot) skips triangles that cannot beat any
// lane's current bound; the rest get the full FP32 distance (three clamped edge
// projections and the plane distance inside the triangle). Voxels whose FP32 distance
// is within one cell of the surface are re-evaluated in FP64 with the reference's own
// operation order over the same candidate list, which makes them bit-identical to the
// reference; the rest keep the FP32 value.
//
// Adjoint. The backward pass recomputes each voxel's closest triangle (same cull and
// prune), takes the closest point's barycentric weights (u, v, w), and scatters
//   dL/da += -g s u (p - q)/d,   likewise for b (v) and c (w)
// (d = |p - q|, s = -1 inside, +1 outside, held constant; d = 0 carries nothing) with
// wavefront-aggregated float atomics: lanes sharing a triangle are summed in registers
// and one lane adds the nine components.
//
// This TU is compiled with -ffp-contract=off; FP32 FMAs are written explicitly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "sdf_device.hpp"
#include "voxelize_internal.hpp"

namespace cmg::hip {

namespace {

using namespace sdfdev;

// ---- triangle preparation ---------------------------------------------------------

__global__ __launch_bounds__(256) void tri_prep_kernel(const double* __restrict__ xyz,
                                                       const int* __restrict__ tris, int ntris,
                                                       GridSpec g, TriRec* rec,
                                                       TriRec64* rec64) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntris) return;
    double P[3][3];
    for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a) P[v][a] = xyz[3 * (size_t)tris[3 * t + v] + a];
    TriRec r;
    TriRec64 r64;
    prep_triangle(P, g, r, r64);
    rec[t] = r;
    rec64[t] = r64;
}

// ---- cull: brick lists ---------------------------------------------------------------
//
// A level's lists are built in three passes over (brick, candidate) pairs, one
// workgroup per brick: the minimum centre distance, the count of survivors, and (after a
// scan gives each brick its offset) the survivors themselves. Candidates are all
// triangles for the first level and the parent brick's list below it.

struct CullArgs {
    const TriRec* rec;
    int ntris;
    GridSpec g;
    Level L;                 // this level
    Level P;                 // parent level (when parent lists are used)
    const int* par_off;      // parent CSR (nullptr: every triangle is a candidate)
    const int* par_items;
    float eps;               // FP32 distance error margin
    float* dmin;             // per brick
    int* count;              // per brick
    const int* off;          // per brick (exclusive scan of count)
    int* items;
};

__device__ inline void cand_range(const CullArgs& A, int b, int& beg, int& end) {
    if (!A.par_off) {
        beg = 0;
        end = A.ntris;
        return;
    }
    const int ix = b % A.L.bx, iy = (b / A.L.bx) % A.L.by, iz = b / (A.L.bx * A.L.by);
    const int r = A.L.s == 0 ? 1 : A.P.s / A.L.s;
    const int pb = ((iz / r) * A.P.by + (iy / r)) * A.P.bx + (ix / r);
    beg = A.par_off[pb];
    end = A.par_off[pb + 1];
}

__device__ __forceinline__ int cand_at(const CullArgs& A, int k) {
    return A.par_off ? A.par_items[k] : k;
}

template <int PASS>
__global__ __launch_bounds__(256) void cull_kernel(CullArgs A) {
    const int b = blockIdx.x;
    float cx, cy, cz, hd;
    brick_geom(A.g, A.L, b, cx, cy, cz, hd);
    int beg, end;
    cand_range(A, b, beg, end);
    __shared__ float sred[4];
    __shared__ int scount[4];
    __shared__ int sbase;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (PASS == 0) {
        float m = FLT_MAX;
        for (int k = beg + (int)threadIdx.x; k < end; k += blockDim.x)
            m = fminf(m, tri_d2(A.rec[cand_at(A, k)], cx, cy, cz));
        m = wave_min(m);
        if (lane == 0) sred[wave] = m;
        __syncthreads();
        if (threadIdx.x == 0)
            A.dmin[b] = sqrtf(fminf(fminf(sred[0], sred[1]), fminf(sred[2], sred[3])));
        return;
    }
    const float thr = A.dmin[b] + 2.0f * hd + A.eps;
    const float thr2 = thr * thr;
    if (PASS == 1) {
        int c = 0;
        for (int k = beg + (int)threadIdx.x; k < end; k += blockDim.x)
            c += tri_d2(A.rec[cand_at(A, k)], cx, cy, cz) <= thr2;
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
        if (lane == 0) scount[wave] = c;
        __syncthreads();
        if (threadIdx.x == 0) A.count[b] = scount[0] + scount[1] + scount[2] + scount[3];
        return;
    }
    // PASS 2: write survivors at off[b] (block-level compaction per iteration)
    if (threadIdx.x == 0) sbase = A.off[b];
    __syncthreads();
    for (int k0 = beg; k0 < end; k0 += blockDim.x) {
        const int k = k0 + threadIdx.x;
        int t = -1;
        bool keep = false;
        if (k < end) {
            t = cand_at(A, k);
            keep = tri_d2(A.rec[t], cx, cy, cz) <= thr2;
        }
        unsigned long long mask = __ballot(keep);
        if (lane == 0) scount[wave] = __popcll(mask);
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; ++w) before += scount[w];
        const int base = sbase;
        if (keep) {
            unsigned long long below = mask & ((1ull << lane) - 1ull);
            A.items[base + before + __popcll(below)] = t;
        }
        __syncthreads();
        if (threadIdx.x == 0) sbase = base + scount[0] + scount[1] + scount[2] + scount[3];
        __syncthreads();
    }
}

// ---- voxel pass ----------------------------------------------------------------------

struct VoxArgs {
    const TriRec* rec;
    const TriRec64* rec64;
    const int* tris;
    GridSpec g;
    Level L0, L1;
    const int* off1;     // 8^3 lists
    const int* items1;
    float eps;
    const uint8_t* occ;
    float* dist;         // forward output
    const float* gsdf;   // backward input
    float* grad;         // backward output (3 per point); null: no scatter
    int* nearest;        // optional: the triangle the adjoint uses, per cell
    unsigned long long* pairs;  // full FP32 evaluations (statistics)
};

template <bool BWD>
__global__ __launch_bounds__(64) void sdf_voxel_kernel(VoxArgs A) {
    __shared__ int slist[kLdsList];
    const int b = blockIdx.x;
    const int lane = threadIdx.x;
    const GridSpec& g = A.g;
    const int ix = b % A.L0.bx, iy = (b / A.L0.bx) % A.L0.by, iz = b / (A.L0.bx * A.L0.by);
    const int i = ix * kLeaf + (lane & 3), j = iy * kLeaf + ((lane >> 2) & 3),
              k = iz * kLeaf + (lane >> 4);
    const bool valid = i < g.nx && j < g.ny && k < g.nz;
    const float h = (float)g.spacing;
    const float px = (i + 0.5f) * h, py = (j + 0.5f) * h, pz = (k + 0.5f) * h;
    float cx, cy, cz, hd;
    brick_geom(g, A.L0, b, cx, cy, cz, hd);
    const int pb = ((iz >> 1) * A.L1.by + (iy >> 1)) * A.L1.bx + (ix >> 1);
    const int beg = A.off1[pb], end = A.off1[pb + 1];
    // 4^3 list: minimum centre distance, then the survivors into LDS
    float m = FLT_MAX;
    for (int q = beg + lane; q < end; q += 64) m = fminf(m, tri_d2(A.rec[A.items1[q]], cx, cy, cz));
    const float dmin = sqrtf(wave_min(m));
    const float thr = dmin + 2.0f * hd + A.eps;
    const float thr2 = thr * thr;
    int n = 0;
    for (int q0 = beg; q0 < end; q0 += 64) {
        const int q = q0 + lane;
        int t = -1;
        bool keep = false;
        if (q < end) {
            t = A.items1[q];
            keep = tri_d2(A.rec[t], cx, cy, cz) <= thr2;
        }
        unsigned long long mask = __ballot(keep);
        if (keep) {
            const int pos = n + __popcll(mask & ((1ull << lane) - 1ull));
            if (pos < kLdsList) slist[pos] = t;
        }
        n += __popcll(mask);
    }
    __syncthreads();
    const bool spill = n > kLdsList;  // too many for LDS: walk the 8^3 list instead
    const int count = spill ? end - beg : n;
    // upper bound on every voxel's nearest distance in this brick
    const float cap = dmin + hd + A.eps;
    float best = FLT_MAX;  // squared FP32 distance
    int arg = -1;
    float bound = cap;
    unsigned long long evals = 0;
    for (int q = 0; q < count; ++q) {
        const int t = uniform_int(spill ? A.items1[beg + q] : slist[q]);
        const TriRec T = load_uniform(A.rec, t);
        const float dx = px - T.mx, dy = py - T.my, dz = pz - T.mz;
        const float rr = T.r + bound;
        const bool need = valid && fma3(dx, dx, dy, dy, dz, dz) <= rr * rr;
        if (__ballot(need) == 0) continue;
        ++evals;
        const float d2 = tri_d2(T, px, py, pz);
        if (d2 < best) {
            best = d2;
            arg = t;
        }
        bound = fminf(cap, sqrtf(best) * (1.0f + 4.0f * kU) + A.eps);
    }
    if (lane == 0 && A.pairs) atomicAdd(A.pairs, evals);
    const size_t idx = valid ? ((size_t)k * g.ny + j) * g.nx + i : 0;
    if (!BWD) {
        // within one cell of the surface: FP64 over the same candidates, reference order
        const float d32 = sqrtf(best);
        const bool near = valid && d32 <= h * (1.0f + 1e-3f) + 2.0f * A.eps;
        double best64 = DBL_MAX;
        if (__ballot(near) != 0) {
            const double p64[3] = {g.corner[0] + (i + 0.5) * g.spacing,
                                   g.corner[1] + (j + 0.5) * g.spacing,
                                   g.corner[2] + (k + 0.5) * g.spacing};
            for (int q = 0; q < count; ++q) {
                const int t = uniform_int(spill ? A.items1[beg + q] : slist[q]);
                const TriRec64 T = load_uniform(A.rec64, t);
                if (near) {
                    const double d = point_tri_dist2_ref(p64, T.a, T.b, T.c);
                    if (d < best64) best64 = d;  // std::min(best, d) (voxelize.cpp:317)
                }
            }
        }
        if (valid) {
            if (near) {
                const double d = sqrt(best64);
                A.dist[idx] = (float)(A.occ[idx] ? -d : d);  // voxelize.cpp:354-355
            } else {
                A.dist[idx] = A.occ[idx] ? -d32 : d32;
            }
        }
        return;
    }
    // ---- adjoint ----
    if (A.nearest && valid) A.nearest[idx] = arg;
    float c[9];
    bool act = false;
    if (valid && arg >= 0 && A.grad) {
        const TriRec T = A.rec[arg];
        float v, w;
        tri_bary(T, px, py, pz, v, w);
        const float u = 1.0f - v - w;
        const float qx = T.ax + v * T.abx + w * T.acx;
        const float qy = T.ay + v * T.aby + w * T.acy;
        const float qz = T.az + v * T.abz + w * T.acz;
        const float ex = px - qx, ey = py - qy, ez = pz - qz;
        const float d = sqrtf(ex * ex + ey * ey + ez * ez);
        const float gv = A.gsdf[idx];
        if (d > 0.0f && gv != 0.0f) {
            const float s = A.occ[idx] ? -1.0f : 1.0f;
            const float coef = -gv * s / d;
            c[0] = coef * u * ex; c[1] = coef * u * ey; c[2] = coef * u * ez;
            c[3] = coef * v * ex; c[4] = coef * v * ey; c[5] = coef * v * ez;
            c[6] = coef * w * ex; c[7] = coef * w * ey; c[8] = coef * w * ez;
            act = true;
        }
    }
    // wavefront-aggregated scatter: one group of lanes per distinct triangle
    while (true) {
        const unsigned long long mask = __ballot(act);
        if (mask == 0) break;
        const int leader = __ffsll((long long)mask) - 1;
        const int lt = __shfl(arg, leader);
        const bool match = act && arg == lt;
        float sum[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) sum[e] = wave_sum(match ? c[e] : 0.0f);
        if (lane == leader) {
            const int* tv = A.tris + 3 * (size_t)lt;
#pragma unroll
            for (int vtx = 0; vtx < 3; ++vtx) {
                float* gp = A.grad + 3 * (size_t)tv[vtx];
                atomicAdd(gp + 0, sum[3 * vtx + 0]);
                atomicAdd(gp + 1, sum[3 * vtx + 1]);
                atomicAdd(gp + 2, sum[3 * vtx + 2]);
            }
        }
        act = act && !match;
    }
}

__global__ __launch_bounds__(1024) void scan_kernel(const int* __restrict__ in,
                                                    int* __restrict__ out, int n) {
    __shared__ int s[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        int i = base + threadIdx.x;
        int v = i < n ? in[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int add = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < n) out[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = carry;
}

struct SdfState {
    TriRec* rec;
    TriRec64* rec64;
    Level L0, L1, L2;
    int* off1;
    int* items1;
    float eps;
};

// Builds the triangle records and the 16^3 and 8^3 brick lists.
SdfState build_cull(const double* d_xyz, const int* d_tris, int ntris, const GridSpec& g,
                    DeviceArena& arena, hipStream_t st) {
    SdfState S;
    S.rec = arena.alloc<TriRec>(ntris);
    S.rec64 = arena.alloc<TriRec64>(ntris);
    tri_prep_kernel<<<(ntris + 255) / 256, 256, 0, st>>>(d_xyz, d_tris, ntris, g, S.rec, S.rec64);
    CMG_HIP_LAUNCH();
    S.L0 = make_level(g, kLeaf);
    S.L1 = make_level(g, 2 * kLeaf);
    S.L2 = make_level(g, 4 * kLeaf);
    S.eps = distance_eps(g);
    int* off2 = nullptr;
    int* items2 = nullptr;
    // level 2 from all triangles, then level 1 from level 2
    for (int lvl = 2; lvl >= 1; --lvl) {
        const Level& L = lvl == 2 ? S.L2 : S.L1;
        const int nb = L.bx * L.by * L.bz;
        float* dmin = arena.alloc<float>(nb);
        int* count = arena.alloc<int>(nb);
        int* off = arena.alloc<int>(nb + 1);
        CullArgs A{S.rec, ntris, g, L, S.L2, lvl == 2 ? nullptr : off2,
                   lvl == 2 ? nullptr : items2, S.eps, dmin, count, off, nullptr};
        cull_kernel<0><<<nb, 256, 0, st>>>(A);
        CMG_HIP_LAUNCH();
        cull_kernel<1><<<nb, 256, 0, st>>>(A);
        CMG_HIP_LAUNCH();
        scan_kernel<<<1, 1024, 0, st>>>(count, off, nb);
        CMG_HIP_LAUNCH();
        int total = 0;
        CMG_HIP(hipMemcpyAsync(&total, off + nb, sizeof(int), hipMemcpyDeviceToHost, st));
        CMG_HIP(hipStreamSynchronize(st));
        A.items = arena.alloc<int>(total > 0 ? total : 1);
        cull_kernel<2><<<nb, 256, 0, st>>>(A);
        CMG_HIP_LAUNCH();
        if (lvl == 2) {
            off2 = off;
            items2 = A.items;
        } else {
            S.off1 = off;
            S.items1 = A.items;
        }
    }
    return S;
}

}  // namespace

void launch_sdf_forward(const double* d_xyz, int npts, const int* d_tris, int ntris,
                        const GridSpec& g, const uint8_t* d_occ, float* d_dist,
                        DeviceArena& arena, hipStream_t st) {
    (void)npts;
    SdfState S = build_cull(d_xyz, d_tris, ntris, g, arena, st);
    VoxArgs A{S.rec, S.rec64, d_tris, g, S.L0, S.L1, S.off1, S.items1, S.eps, d_occ, d_dist,
              nullptr, nullptr, nullptr};
    const int nb = S.L0.bx * S.L0.by * S.L0.bz;
    TimedLaunch timed("sdf_voxels", st);
    sdf_voxel_kernel<false><<<nb, 64, 0, st>>>(A);
    CMG_HIP_LAUNCH();
}

void launch_sdf_backward(const double* d_xyz, int npts, const int* d_tris, int ntris,
                         const GridSpec& g, const uint8_t* d_occ, const float* d_gsdf,
                         float* d_grad, int* d_nearest, DeviceArena& arena, hipStream_t st) {
    (void)npts;
    SdfState S = build_cull(d_xyz, d_tris, ntris, g, arena, st);
    VoxArgs A{S.rec, S.rec64, d_tris, g, S.L0, S.L1, S.off1, S.items1, S.eps, d_occ, nullptr,
              d_gsdf, d_grad, d_nearest, nullptr};
    const int nb = S.L0.bx * S.L0.by * S.L0.bz;
    TimedLaunch timed("sdf_backward", st);
    sdf_voxel_kernel<true><<<nb, 64, 0, st>>>(A);
    CMG_HIP_LAUNCH();
}

}  // namespace cmg::hip
