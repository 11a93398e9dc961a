// voxelize.hip — host side of single-mesh voxelization: make_grid exactly as
// src/voxelize/voxelize.cpp:52-74, the column-tile plan, and the host-pointer entry
// points, which run one mesh through the voxelization engine (batch.hip).
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

OccPlan occupancy_plan(const GridSpec& g) {
    OccPlan p;
    p.ntx = (g.nx + occdev::kTile - 1) / occdev::kTile;
    p.nty = (g.ny + occdev::kTile - 1) / occdev::kTile;
    p.words = (g.nz + 1 + 63) / 64;
    return p;
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
    SingleMesh one(xyz, npts, tris, ntris, g, st);
    if (!sdf) {
        one.engine->run(kBatchOccupancy, one.d_xyz, nullptr, nullptr, nullptr, st);
        CMG_HIP(hipMemcpyAsync(occ, one.engine->occupancy(), ncell, hipMemcpyDeviceToHost, st));
        CMG_HIP(hipStreamSynchronize(st));
        return;
    }
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
