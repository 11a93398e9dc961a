// batch_internal.hpp — the voxelization engine (batch.hip): one or many meshes per step.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "cmg_hip_internal.hpp"
#include "sdf_device.hpp"
#include "voxelize_internal.hpp"

namespace cmg::hip {

// Levels of the per-mesh triangle tree: level 0 are the triangles (in a median-split order
// of their centroids), a level-l node covers 8^l consecutive triangles; a mesh has at most
// 2^30 = 8^10 triangles, so at most 10 levels above the triangles.
constexpr int kMaxLevels = 10;

// Oriented bounding box of a tree node (coordinates relative to the grid corner): the
// first axis is the patch's aligned mean normal, the other two its principal in-plane
// directions; dn/du/dw are the box centre's coordinates along n/u/w and en/eu/ew the
// half-extents. link = (first << 4) | count of the children: on level 1 (bit 31 set) the
// node's tree leaf index in its mesh and its triangle count, above it the global index of
// the first child node. The box centre itself is kept apart (Ctx::nctr).
struct __align__(16) NodeRec {
    float nx, ny, nz, dn;
    float ux, uy, uz, du;
    float wx, wy, wz, dw;
    float en, eu, ew;
    unsigned link;
};
constexpr unsigned kLinkLeaf = 1u << 31;
constexpr long long kMaxNodes = 1LL << 27;  // global node indices and links hold 27 bits
static_assert(sizeof(NodeRec) == 64, "node size");

// Per-mesh description, uploaded once at creation.
struct MeshDesc {
    GridSpec g;
    long long vbase;      // first point of the mesh in the concatenated point array
    long long tbase;      // first triangle (and first tree slot)
    int ntris;
    int nlev;             // root level (>= 1)
    long long cell_base;  // first cell in the concatenated outputs
    long long blk_base;   // first entry of the mesh's blocked per-cell arrays (64 per
                          // 4x4x4 block of L0)
    long long mask_base;  // first occupancy bit-mask word
    int words, ntx, nty;  // occupancy: mask words per column, column tiles per axis
    int tile_base;        // first column tile in the global tile arrays
    sdfdev::Level L0, L0b, L0c;    // voxel leaves: 4x4x4, 4x4x8 and 4x4x16
    int node_base[kMaxLevels + 1]; // first node of each level (1..nlev) in the node array
    int nnodes[kMaxLevels + 1];    // nodes per level (level 0: triangles)
    float eps;            // FP32 distance error margin for this grid, in search units
    double dzero;         // distances at or below this carry no gradient (world units)
    double sc;            // search units: (x - corner) * sc, a power of two
    float scf, iscf;      // sc and 1 / sc as floats (exact)
};

// Work items per mesh: column tiles, voxel leaves (4x4x4, 4x4x8, 4x4x16), runs of up to
// 64 4x4x4 leaves along x (the SDF pass's rows), and grid rows (the occupancy bytes).
enum BatchPhase {
    kPhTile = 0, kPhL0 = 1, kPhL0b = 2, kPhL0c = 3, kPhRow = 4, kPhOccRow = 5,
    kPhV1 = 6, kPhV2 = 7, kPhV4 = 8,  // voxel-pass items (1, 2, 4 voxels per lane), by super-tile
    kNumPhase = 9
};
constexpr int kRowRun = 64;  // 4x4x4 leaves per SDF-pass item
// Voxel-pass super-tiles: kSuper x kSuper leaves in x and y, 16 cells in z (kSuperZ / vpl
// leaves of 4 vpl cells); their items cover the grid's edge too and are skipped there.
constexpr int kSuper = 4, kSuperZ = 4;

struct BatchCounters {
    unsigned long long queue[4];   // work queues: 0 tile-list segments, 3 tiles redone
                                   // with exact predicates (1, 2 unused)
    unsigned long long tile_used;  // tile-list pool cursor
    unsigned long long seg_used;   // occupancy work items (tile list segments)
    unsigned long long redo_n;     // tiles queued for the exact occupancy pass
    unsigned long long xq[8];      // voxel pass: items taken from each XCD's range
    unsigned long long aq[8];      // adjoint: chunks taken from each XCD's range
    // work of the voxel passes, counted per wavefront: box tests of tree nodes, box
    // tests of triangles, FP32 point-triangle distances
    unsigned long long node_tests, tri_tests, tri_evals;
    unsigned overflow;             // bit 0: tile lists
    unsigned pad;
};

enum : int { kBatchForward = 1, kBatchBackward = 2, kBatchGraph = 4, kBatchOccupancy = 8 };

// A voxel-pass variant: voxels per lane (1: 4x4x4 leaves, 2: 4x4x8, 4: 4x4x16) and the
// LDS tile depth of a tree leaf's triangle records (0: read with scalar loads, 8: staged
// through LDS as one 1 KiB tile).
struct BatchVariant {
    int vpl, depth;
    float ms;  // forward voxel pass + adjoint pass, when measured by autotune
};

class BatchEngine {
  public:
    // pt_start/tri_start: host prefix arrays (nm + 1 entries) into the concatenated
    // device arrays d_xyz (3 doubles per point) and d_tris (3 local point indices per
    // triangle). Plans every grid, orders every mesh's triangles by median splits of
    // their centroids and sizes the tile pool (synchronises `st`).
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
    // (3 floats per point) for the upstream d_gsdf; kBatchOccupancy alone computes only
    // the occupancy (occupancy()). kBatchGraph captures the sequence into a hipGraph on
    // first use (per buffer set) and replays it.
    void run(int flags, const double* d_xyz, float* d_sdf, const float* d_gsdf, float* d_grad,
             hipStream_t st);

    // After a forward: for every cell, the mesh-local index of the triangle whose closest
    // point the adjoint uses (-1 where there is none), into d_tri (total_cells ints).
    void nearest(int* d_tri, hipStream_t st);

    // Times every voxel-pass variant on this batch (forward + backward voxel passes, HIP
    // events on st, after one forward) and keeps the fastest.
    std::vector<BatchVariant> autotune(const float* d_gsdf, float* d_sdf, float* d_grad,
                                       hipStream_t st);
    void set_variant(int vpl, int depth);
    BatchVariant variant() const;

    // Synchronises st and reads the counters of the last step.
    void status(hipStream_t st, unsigned* overflow, unsigned long long* evals);
    BatchCounters counters(hipStream_t st);

  private:
    struct Impl;
    std::unique_ptr<Impl> p_;
};

}  // namespace cmg::hip
