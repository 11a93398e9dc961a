// cmg_hip_internal.hpp — declarations shared between the kernel TUs and the C ABI layer
// of libcmg_hip (not installed; the public surface is include/cmg_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace cmg::hip {

constexpr int kModeExact = 0;   // exact sign (Shewchuk orient3dexact / insphereexact)
constexpr int kModeCompat = 1;  // what cmg::robust::orient3d / insphere return as shipped

// ---- batched predicates ---------------------------------------------------------------

struct PredCounters {
    unsigned long long fp32;    // lanes decided by the FP32 filter
    unsigned long long fp64;    // lanes decided by the FP64 stage
    unsigned long long exact;   // lanes sent to expansion arithmetic
    unsigned int wide_n;        // lanes re-run with wide expansion capacity
    unsigned int failed;        // lanes whose expansions outgrew even the wide capacity
    unsigned long long inline_; // lanes decided exactly inside the streaming kernel
    unsigned long long fixed;   // queued lanes decided by the fixed-point exact stage
    unsigned int rest_n;        // queued lanes the register-only stages left (rest list)
    unsigned int pad;
};
static_assert(sizeof(PredCounters) <= 256, "counter block");

// Launch shape of a filter pass: `iters` grid-stride iterations per thread, one queue
// segment of `seg` entries per wave.
struct PredPlan {
    unsigned grid = 1, block = 256, iters = 0, nwaves = 0, seg = 0;
};

struct PredWorkspace {
    PredCounters* ctr = nullptr;
    uint32_t* wave_q = nullptr;   // per-wave queued-lane count
    uint32_t* wave_32 = nullptr;  // per-wave FP32-decided count
    uint32_t* wave_64 = nullptr;  // per-wave FP64-decided count
    uint32_t* wave_x = nullptr;   // per-wave count decided by the inline exact stage
    uint32_t* queue = nullptr;    // nwaves segments of seg entries
    uint32_t* qpref = nullptr;    // exclusive prefix of wave_q (nwaves + 1 entries)
    uint32_t* wide = nullptr;     // wide-capacity queue (n entries)
    uint32_t* rest = nullptr;     // lanes left after the register-only stages (n entries)
    unsigned seg = 0, nwaves = 0;
};

PredPlan pred_plan(size_t items, unsigned per_thread);
size_t pred_workspace_bytes(const PredPlan& p, size_t n);
PredWorkspace pred_workspace(void* base, const PredPlan& p, size_t n);

// npts = 4 (orient3d, 12 planes) or 5 (insphere, 15 planes); plane q at soa + q*ld.
// Plan with pred_plan(n, 2).
void launch_signs_soa(int npts, int mode, const double* soa, size_t n, size_t ld,
                      int8_t* out, const PredPlan& p, const PredWorkspace& ws,
                      hipStream_t stream);
// Index records of npts ints per tuple into AoS points. Plan with pred_plan(n, 1).
void launch_signs_indexed(int npts, int mode, const double* xyz, const int* idx, size_t n,
                          int8_t* out, const PredPlan& p, const PredWorkspace& ws,
                          hipStream_t stream);
// Values (npts 4 = orient3d, 5 = insphere): the naive FP64 determinant when nonzero,
// else the exact determinant's leading component. Plan with pred_plan(n, 1).
void launch_values_indexed(int npts, const double* xyz, const int* idx, size_t n,
                           double* out, const PredPlan& p, const PredWorkspace& ws,
                           hipStream_t stream);
// Fills ctr->fp32/fp64/exact from the per-wave counts.
void launch_pred_stats(const PredWorkspace& ws, hipStream_t stream);

// ---- voxelization ------------------------------------------------------------------

// Grid geometry (src/voxelize/voxelize.cpp:44-48): corner = world coordinate of the MIN
// corner of cell (0,0,0); cell (i,j,k) has centre corner + (i+0.5)*spacing per axis.
struct GridSpec {
    double corner[3] = {0, 0, 0};
    double spacing = 0;
    int nx = 0, ny = 0, nz = 0;
    int resolution = 0, pad = 0;  // the make_grid arguments that produced it
};

// make_grid (voxelize.cpp:52-74), host side, same FP64 operations. False for a
// zero-extent point set.
bool make_grid(const double* xyz, size_t npts, int resolution, int pad, GridSpec& g);

// Whole voxelize on the GPU from host arrays: occupancy bytes into occ (always), and in
// SDF mode the signed distance into dist. Synchronous.
void voxelize_host(const double* xyz, size_t npts, const int* tris, size_t ntris,
                   const GridSpec& g, bool sdf, uint8_t* occ, float* dist);

// SDF adjoint from host arrays (occ = the grid's inside bits). Synchronous.
// grad_sdf/grad_xyz (both or neither) give the adjoint; nearest (optional) receives the
// triangle the adjoint uses for every cell (-1 where there is none).
void sdf_backward_host(const double* xyz, size_t npts, const int* tris, size_t ntris,
                       const GridSpec& g, const uint8_t* occ, const float* grad_sdf,
                       float* grad_xyz, int* nearest);

}  // namespace cmg::hip
