// voxelize.hip — occupancy voxelization on gfx950, bit-exact with
// src/voxelize/voxelize.cpp.
//
// Occupancy (voxelize.cpp:78-195). The xy bucket of a column is reproduced exactly:
// triangle t belongs to column (i, j) iff i0 <= i <= i1 and j0 <= j <= j1, with the
// reference's clamp(floor((x - corner) / spacing)) ranges over the three vertices.
// Columns are processed in 8x8 tiles, one 64-lane workgroup per tile, one lane per
// column. Each tile's triangle list (every triangle whose xy range touches the tile) is
// streamed through LDS in chunks of 64 triangles, stored as structure-of-arrays FP64.
// A lane runs the reference's crossing test (compat orient3d, FP64 in the reference's
// operation order) on the triangles of its own bucket.
//
// Heights are not stored or sorted: the reference counts, per cell, the crossing heights
// strictly above the cell centre zc_k and keeps the parity (voxelize.cpp:185-191). A
// crossing at height h flips exactly the cells with zc_k < h, i.e. cells 0 .. K(h)-1, so
// each lane keeps one bit per cell boundary ("flip below K") in a per-column bitmask and
// the final parity of cell k is the XOR of the bits above k.
//
// Grazing (an orient3d exactly 0) restarts the column with the next kNudge offset
// (voxelize.cpp:148-165). If the last nudge grazes too, the reference keeps the
// crossings found before the first grazing triangle of the bucket, which lists triangles
// in ascending index order; pass 6 records the smallest grazing index and a final pass
// replays that nudge counting only triangles below it.
//
// Compiled with -ffp-contract=off: every FP64 expression rounds as in the reference.
#include <hip/hip_runtime.h>

#include <memory>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "cmg_hip_internal.hpp"
#include "device_common.hpp"
#include "expansion.hpp"
#include "occ_device.hpp"
#include "batch_internal.hpp"
#include "voxelize_internal.hpp"

namespace cmg::hip {

bool make_grid(const double* xyz, size_t npts, int resolution, int pad, GridSpec& g) {
    // voxelize.cpp:52-74
    double lo[3] = {xyz[0], xyz[1], xyz[2]}, hi[3] = {xyz[0], xyz[1], xyz[2]};
    for (size_t i = 0; i < npts; ++i)
        for (int a = 0; a < 3; ++a) {
            double v = xyz[3 * i + a];
            if (v < lo[a]) lo[a] = v;
            if (v > hi[a]) hi[a] = v;
        }
    double ext = 0;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, hi[a] - lo[a]);
    g = GridSpec{};
    g.resolution = resolution;
    g.pad = pad;
    g.spacing = ext / resolution;
    if (!(g.spacing > 0)) return false;
    int* dims[3] = {&g.nx, &g.ny, &g.nz};
    for (int a = 0; a < 3; ++a) {
        double e = hi[a] - lo[a];
        int core = std::max(1, static_cast<int>(std::ceil(e / g.spacing)));
        double covered = core * g.spacing;
        g.corner[a] = lo[a] - (covered - e) / 2.0 - pad * g.spacing;
        *dims[a] = core + 2 * pad;
    }
    return true;
}

namespace {

using namespace occdev;

struct OccArgs {
    const double* xyz;
    const int* tris;
    const int4* xyr;        // per triangle: i0, i1, j0, j1
    const int* tile_off;    // CSR over tiles
    const int* tile_items;
    int ntx, nty;
    GridSpec g;
    unsigned long long* masks;  // per column: words x 64 bits of "flip below K"
    int words;
    uint8_t* occ;
    unsigned* stats;        // [0] columns nudged, [1] columns where every nudge grazed
};

__global__ __launch_bounds__(256) void tri_xy_range_kernel(const double* __restrict__ xyz,
                                                           const int* __restrict__ tris,
                                                           int ntris, GridSpec g, int4* xyr,
                                                           int* tile_count, int ntx) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntris) return;
    double P[3][3];
    for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a) P[v][a] = xyz[3 * (size_t)tris[3 * t + v] + a];
    const int4 r = xy_range(P, g);
    const int i0 = r.x, i1 = r.y, j0 = r.z, j1 = r.w;
    xyr[t] = r;
    for (int tj = j0 / kTile; tj <= j1 / kTile; ++tj)
        for (int ti = i0 / kTile; ti <= i1 / kTile; ++ti) atomicAdd(&tile_count[tj * ntx + ti], 1);
}

__global__ __launch_bounds__(256) void tile_fill_kernel(const int4* __restrict__ xyr, int ntris,
                                                        const int* __restrict__ tile_off,
                                                        int* cursor, int* items, int ntx) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntris) return;
    int4 r = xyr[t];
    for (int tj = r.z / kTile; tj <= r.w / kTile; ++tj)
        for (int ti = r.x / kTile; ti <= r.y / kTile; ++ti) {
            int tile = tj * ntx + ti;
            int slot = atomicAdd(&cursor[tile], 1);
            items[tile_off[tile] + slot] = t;
        }
}

// Single-workgroup exclusive scan of n ints (n + 1 outputs; out[n] = total).
__global__ __launch_bounds__(1024) void exclusive_scan_kernel(const int* __restrict__ in,
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

__global__ __launch_bounds__(64) void occupancy_kernel(OccArgs a) {
    const GridSpec& g = a.g;
    const int tile = blockIdx.x;
    const int tx = tile % a.ntx, ty = tile / a.ntx;
    const int lane = threadIdx.x;
    const int ci = tx * kTile + (lane % kTile), cj = ty * kTile + (lane / kTile);
    const bool valid = ci < g.nx && cj < g.ny;
    __shared__ double sv[9][kChunk];
    __shared__ int4 sr[kChunk];
    __shared__ int st[kChunk];
    const int beg = a.tile_off[tile], end = a.tile_off[tile + 1];
    unsigned long long* mask =
        a.masks + ((size_t)(valid ? cj : 0) * g.nx + (valid ? ci : 0)) * a.words;
    // voxelize.cpp:175-181
    const double zmin = g.corner[2] - g.spacing;
    const double zmax = g.corner[2] + (g.nz + 1) * g.spacing;
    const double cx = g.corner[0] + (ci + 0.5) * g.spacing;
    const double cy = g.corner[1] + (cj + 0.5) * g.spacing;

    bool active = valid;
    int cutoff = INT_MAX;
    for (int pass = 0; pass < 8; ++pass) {
        if (!__syncthreads_or(active)) break;
        double x = cx, y = cy;
        if (pass > 0) {
            const int d = (pass < 7 ? pass : 6) - 1;
            x = cx + kNudge[d][0] * g.spacing;
            y = cy + kNudge[d][1] * g.spacing;
        }
        if (active)
            for (int w = 0; w < a.words; ++w) mask[w] = 0ull;
        bool grazed = false;
        int first_graze = INT_MAX;
        const double P[3] = {x, y, zmin};
        const double Q[3] = {x, y, zmax};
        for (int c0 = beg; c0 < end; c0 += kChunk) {
            __syncthreads();
            const int k = c0 + lane;
            if (k < end) {
                const int t = a.tile_items[k];
                st[lane] = t;
                sr[lane] = a.xyr[t];
#pragma unroll
                for (int v = 0; v < 3; ++v) {
                    const double* p = a.xyz + 3 * (size_t)a.tris[3 * t + v];
                    sv[3 * v][lane] = p[0];
                    sv[3 * v + 1][lane] = p[1];
                    sv[3 * v + 2][lane] = p[2];
                }
            }
            __syncthreads();
            if (!active) continue;
            const int cnt = min(kChunk, end - c0);
            for (int q = 0; q < cnt; ++q) {
                const int4 r = sr[q];
                if (ci < r.x || ci > r.y || cj < r.z || cj > r.w) continue;  // not in bucket
                const int t = st[q];
                if (pass == 7 && t >= cutoff) continue;
                if (grazed && pass < 6) break;  // this attempt is already lost
                const double A[3] = {sv[0][q], sv[1][q], sv[2][q]};
                const double B[3] = {sv[3][q], sv[4][q], sv[5][q]};
                const double C[3] = {sv[6][q], sv[7][q], sv[8][q]};
                // voxelize.cpp:131-140
                const int sa = o3_compat(P, A, B, C);
                const int sb = o3_compat(Q, A, B, C);
                if (sa == 0 || sb == 0) {
                    grazed = true;
                    first_graze = min(first_graze, t);
                    continue;
                }
                if (sa == sb) continue;
                const int s1 = o3_compat(P, Q, A, B);
                const int s2 = o3_compat(P, Q, B, C);
                const int s3 = o3_compat(P, Q, C, A);
                if (s1 == 0 || s2 == 0 || s3 == 0) {
                    grazed = true;
                    first_graze = min(first_graze, t);
                    continue;
                }
                if (s1 != s2 || s2 != s3) continue;
                const int K = cells_below(g, plane_z(A, B, C, x, y));
                mask[K >> 6] ^= 1ull << (K & 63);
            }
        }
        if (active) {
            if (pass == 0 && grazed) atomicAdd(&a.stats[0], 1u);
            if (pass < 6) {
                active = grazed;
            } else if (pass == 6) {
                if (grazed) {
                    atomicAdd(&a.stats[1], 1u);
                    cutoff = first_graze;  // replay counting only triangles below it
                }
                active = grazed;
            } else {
                active = false;
            }
        }
    }
    if (!valid) return;
    write_column(mask, a.words, g, ci, cj, a.occ);
}

}  // namespace

// ---- host orchestration ---------------------------------------------------------------

OccPlan occupancy_plan(const GridSpec& g) {
    OccPlan p;
    p.ntx = (g.nx + kTile - 1) / kTile;
    p.nty = (g.ny + kTile - 1) / kTile;
    p.words = (g.nz + 1 + 63) / 64;
    return p;
}

void launch_occupancy(const double* d_xyz, const int* d_tris, int ntris, const GridSpec& g,
                      uint8_t* d_occ, DeviceArena& arena, hipStream_t st, unsigned* h_stats) {
    OccPlan p = occupancy_plan(g);
    const int ntiles = p.ntx * p.nty;
    int4* xyr = arena.alloc<int4>(ntris);
    int* count = arena.alloc<int>(ntiles + 1);
    int* off = arena.alloc<int>(ntiles + 1);
    int* cursor = arena.alloc<int>(ntiles + 1);
    unsigned* stats = arena.alloc<unsigned>(2);
    unsigned long long* masks = arena.alloc<unsigned long long>((size_t)g.nx * g.ny * p.words);
    CMG_HIP(hipMemsetAsync(count, 0, sizeof(int) * (ntiles + 1), st));
    CMG_HIP(hipMemsetAsync(cursor, 0, sizeof(int) * (ntiles + 1), st));
    CMG_HIP(hipMemsetAsync(stats, 0, sizeof(unsigned) * 2, st));
    const unsigned tb = 256, tg = (unsigned)((ntris + tb - 1) / tb);
    tri_xy_range_kernel<<<tg, tb, 0, st>>>(d_xyz, d_tris, ntris, g, xyr, count, p.ntx);
    CMG_HIP_LAUNCH();
    exclusive_scan_kernel<<<1, 1024, 0, st>>>(count, off, ntiles);
    CMG_HIP_LAUNCH();
    int total = 0;
    CMG_HIP(hipMemcpyAsync(&total, off + ntiles, sizeof(int), hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    int* items = arena.alloc<int>(total > 0 ? total : 1);
    tile_fill_kernel<<<tg, tb, 0, st>>>(xyr, ntris, off, cursor, items, p.ntx);
    CMG_HIP_LAUNCH();
    OccArgs a{d_xyz, d_tris, xyr, off, items, p.ntx, p.nty, g, masks, p.words, d_occ, stats};
    {
        TimedLaunch timed("occupancy_columns", st);
        occupancy_kernel<<<ntiles, 64, 0, st>>>(a);
        CMG_HIP_LAUNCH();
    }
    if (h_stats) {
        CMG_HIP(hipMemcpyAsync(h_stats, stats, sizeof(unsigned) * 2, hipMemcpyDeviceToHost, st));
        CMG_HIP(hipStreamSynchronize(st));
    }
}

namespace {

// One mesh through the voxelization engine (device copies of the points and triangles
// live in the arena; the engine holds them until it is destroyed, before the arena).
struct SingleMesh {
    DeviceArena arena;
    double* d_xyz;
    int* d_tris;
    std::unique_ptr<BatchEngine> engine;

    SingleMesh(const double* xyz, size_t npts, const int* tris, size_t ntris,
               const GridSpec& g, hipStream_t st)
        : arena(st) {
        d_xyz = arena.alloc<double>(3 * npts);
        d_tris = arena.alloc<int>(3 * ntris);
        CMG_HIP(hipMemcpyAsync(d_xyz, xyz, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, st));
        CMG_HIP(hipMemcpyAsync(d_tris, tris, sizeof(int) * 3 * ntris, hipMemcpyHostToDevice, st));
        const long long ps[2] = {0, (long long)npts}, ts[2] = {0, (long long)ntris};
        const int res = g.resolution, pad = g.pad;
        engine = std::make_unique<BatchEngine>(1, ps, ts, &res, &pad, d_xyz, d_tris, st);
        const GridSpec& e = engine->mesh(0).g;
        if (e.nx != g.nx || e.ny != g.ny || e.nz != g.nz || e.spacing != g.spacing ||
            e.corner[0] != g.corner[0] || e.corner[1] != g.corner[1] ||
            e.corner[2] != g.corner[2])
            throw HipError("voxelize: device grid differs from the host grid");
    }
    ~SingleMesh() {
        (void)hipStreamSynchronize(arena.stream());
        engine.reset();
    }
};

}  // namespace

void voxelize_host(const double* xyz, size_t npts, const int* tris, size_t ntris,
                   const GridSpec& g, bool sdf, uint8_t* occ, float* dist) {
    const size_t ncell = (size_t)g.nx * g.ny * g.nz;
    hipStream_t st = nullptr;
    if (!sdf) {
        DeviceArena arena(st);
        double* d_xyz = arena.alloc<double>(3 * npts);
        int* d_tris = arena.alloc<int>(3 * ntris);
        uint8_t* d_occ = arena.alloc<uint8_t>(ncell);
        CMG_HIP(hipMemcpyAsync(d_xyz, xyz, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, st));
        CMG_HIP(hipMemcpyAsync(d_tris, tris, sizeof(int) * 3 * ntris, hipMemcpyHostToDevice, st));
        launch_occupancy(d_xyz, d_tris, (int)ntris, g, d_occ, arena, st, nullptr);
        CMG_HIP(hipMemcpyAsync(occ, d_occ, ncell, hipMemcpyDeviceToHost, st));
        CMG_HIP(hipStreamSynchronize(st));
        return;
    }
    SingleMesh one(xyz, npts, tris, ntris, g, st);
    float* d_dist = one.arena.alloc<float>(ncell);
    one.engine->run(kBatchForward, one.d_xyz, d_dist, nullptr, nullptr, st);
    CMG_HIP(hipMemcpyAsync(occ, one.engine->occupancy(), ncell, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipMemcpyAsync(dist, d_dist, sizeof(float) * ncell, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
}

void sdf_backward_host(const double* xyz, size_t npts, const int* tris, size_t ntris,
                       const GridSpec& g, const uint8_t* occ, const float* grad_sdf,
                       float* grad_xyz, int* nearest) {
    // the occupancy of the grid is recomputed by the forward (it is a function of the
    // same points and triangles)
    (void)occ;
    const size_t ncell = (size_t)g.nx * g.ny * g.nz;
    const bool adj = grad_sdf && grad_xyz;
    hipStream_t st = nullptr;
    SingleMesh one(xyz, npts, tris, ntris, g, st);
    float* d_dist = one.arena.alloc<float>(ncell);
    float* d_g = adj ? one.arena.alloc<float>(ncell) : nullptr;
    float* d_grad = adj ? one.arena.alloc<float>(3 * npts) : nullptr;
    if (adj)
        CMG_HIP(hipMemcpyAsync(d_g, grad_sdf, sizeof(float) * ncell, hipMemcpyHostToDevice, st));
    one.engine->run(kBatchForward | (adj ? kBatchBackward : 0), one.d_xyz, d_dist, d_g, d_grad,
                    st);
    if (adj)
        CMG_HIP(hipMemcpyAsync(grad_xyz, d_grad, sizeof(float) * 3 * npts,
                               hipMemcpyDeviceToHost, st));
    if (nearest) {
        int* d_near = one.arena.alloc<int>(ncell);
        one.engine->nearest(d_near, st);
        CMG_HIP(hipMemcpyAsync(nearest, d_near, sizeof(int) * ncell, hipMemcpyDeviceToHost, st));
    }
    CMG_HIP(hipStreamSynchronize(st));
}

}  // namespace cmg::hip
