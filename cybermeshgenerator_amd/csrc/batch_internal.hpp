// batch_internal.hpp — the many-mesh voxelization engine (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>

#include "cmg_hip_internal.hpp"
#include "sdf_device.hpp"
#include "voxelize_internal.hpp"

namespace cmg::hip {

// Per-mesh description, uploaded once at creation.
struct MeshDesc {
    GridSpec g;
    long long vbase;      // first point of the mesh in the concatenated point array
    long long tbase;      // first triangle
    int ntris;
    long long cell_base;  // first cell in the concatenated outputs
    long long mask_base;  // first occupancy bit-mask word
    int words, ntx, nty;  // occupancy: mask words per column, column tiles per axis
    int tile_base;        // first column tile in the global tile arrays
    sdfdev::Level L0, L0b, L1, L2;  // L0: 4x4x4 leaves, L0b: 4x4x8 leaves
    int l2_base, l1_base; // first 16^3 / 8^3 brick in the global brick arrays
    float eps;            // FP32 distance error margin for this grid
};

enum BatchPhase { kPhTile = 0, kPhL2 = 1, kPhL1 = 2, kPhL0 = 3, kPhL0b = 4, kNumPhase = 5 };

struct BatchCounters {
    unsigned long long queue[10];  // work queues of the persistent kernels
    unsigned long long tile_used, l2_used, l1_used;  // pool cursors
    unsigned long long pairs;      // full FP32 distance evaluations (statistics)
    unsigned overflow;             // bit 0 tiles, bit 1 16^3 lists, bit 2 8^3 lists
    unsigned pad;
};

enum : int { kBatchForward = 1, kBatchBackward = 2, kBatchGraph = 4 };

// A voxel-pass variant: voxels per lane (1: 4x4x4 leaves, 2: 4x4x8) and the LDS tile
// depth (0: records read with scalar loads, 32/64: staged through LDS in tiles).
struct BatchVariant {
    int vpl, depth;
    float ms;  // forward + backward voxel passes, when measured by autotune
};

class BatchEngine {
  public:
    // pt_start/tri_start: host prefix arrays (nm + 1 entries) into the concatenated
    // device arrays d_xyz (3 doubles per point) and d_tris (3 local point indices per
    // triangle). Plans every grid and sizes the list pools (synchronises `st`).
    BatchEngine(int nm, const long long* pt_start, const long long* tri_start,
                const int* resolution, const int* pad, const double* d_xyz,
                const int* d_tris, hipStream_t st);
    ~BatchEngine();
    BatchEngine(const BatchEngine&) = delete;
    BatchEngine& operator=(const BatchEngine&) = delete;

    long long total_cells() const;
    long long total_points() const;
    int meshes() const;
    const MeshDesc& mesh(int m) const;
    const uint8_t* occupancy() const;  // device, total_cells bytes (after a forward)

    // One step on stream st: kBatchForward writes the SDF of every mesh to d_sdf
    // (total_cells floats); kBatchBackward (after a forward) writes dL/dxyz to d_grad
    // (3 floats per point) for the upstream d_gsdf. kBatchGraph captures the sequence
    // into a hipGraph on first use (per buffer set) and replays it.
    void run(int flags, const double* d_xyz, float* d_sdf, const float* d_gsdf, float* d_grad,
             hipStream_t st);

    // Times every voxel-pass variant on this batch (forward + backward voxel passes, HIP
    // events on st, after one forward builds the lists) and keeps the fastest.
    std::vector<BatchVariant> autotune(const float* d_gsdf, float* d_sdf, float* d_grad,
                                       hipStream_t st);
    void set_variant(int vpl, int depth);
    BatchVariant variant() const;

    // Synchronises st and reads the overflow flags and the evaluation count.
    void status(hipStream_t st, unsigned* overflow, unsigned long long* pairs);

  private:
    struct Impl;
    std::unique_ptr<Impl> p_;
};

}  // namespace cmg::hip
