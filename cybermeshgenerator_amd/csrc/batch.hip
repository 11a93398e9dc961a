// batch.hip — voxelization engine for one or many meshes of uneven size in one pass.
//
// All meshes live in one set of device arrays (points, triangles, cells concatenated).
// A step is a fixed sequence of kernels; the persistent ones pull work items — (mesh,
// column tile) or (mesh, voxel leaf) — from a device work queue (one atomic counter per
// kernel), most expensive meshes first, so small meshes fill the gaps left by large ones
// instead of each occupying the GPU alone. The sequence is captured once into a hipGraph
// and replayed; a step needs no host round trip.
//
// Nearest-triangle search. At creation every mesh's triangles are put in a search order of
// their centroids built by median splits (every run of 8^l slots spatially compact). A
// level-l tree node covers 8^l consecutive triangles (an implicit
// 8-ary tree, no pointers) and carries an oriented bounding box whose first axis is the
// patch's aligned area-weighted normal and whose other two are its principal in-plane
// directions. For a surface patch seen from afar the box is a thin plate, and the
// distance to it bounds the distance to the patch far more tightly than a sphere or an
// axis-aligned box would; every triangle has its own oriented box. Each step rebuilds the
// boxes bottom-up from the current vertices. A wavefront owns 64 or more neighbouring
// voxels and walks the tree once for all of them: a node is opened when any voxel's lower
// bound beats that voxel's best distance so far, the child nearest the voxel block's
// centre first (a wave-uniform stack in one register). The minimum of the FP32 distance
// over the triangles that survive is the minimum over all triangles (every rejection is a
// rigorous lower bound), so results do not depend on the tree. Voxels within one cell of
// the surface walk again with the bound fixed at the FP32 minimum plus twice the FP32
// error margin and take the reference's FP64 distance over what survives.
//
// Phases of a forward step (all meshes at once):
//   reset        zero the queues, pool cursor, tile counters and occupancy masks (a fill
//                kernel: no memset nodes in the graph)
//   prep         per tree slot: FP32/FP64 triangle records; per triangle: xy bucket
//                range and tile counts
//   tile alloc   per tile: list offset from the tile pool, occupancy work items
//   tile fill    per triangle: append to the lists of the tiles its range touches
//   occupancy    per segment of an 8x8 column tile's list: exact column parity (below);
//                tiles where a naive determinant is exactly 0 are redone with exact
//                predicates
//   tree         level 1 from the triangles, then each level from the one below
//   voxels       per (mesh, voxel leaf): FP32 nearest, FP64 near the surface, signed;
//                records per cell (4 bytes) the triangle the distance comes from
// A backward step reuses the forward's occupancy and per-cell triangles: after zeroing
// the gradient, the adjoint pass recomputes each cell's closest point on its triangle
// (the closest feature and barycentric weights are not stored) and scatters.
//
// Compiled with -ffp-contract=off like the single-mesh kernels it shares code with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "batch_internal.hpp"
#include "device_common.hpp"
#include "occ_device.hpp"
#include "sdf_device.hpp"

namespace cmg::hip {

using namespace sdfdev;
using namespace occdev;

namespace {

constexpr int kSegTris = 16 * kChunk;  // tile-list entries per occupancy work item
constexpr int kXcds = 8;  // XCDs of an MI355X, each with its own L2 (power of two)

// The XCD this wave runs on: bits 3:0 of hardware register XCC_ID (register 20;
// s_getreg operand (size - 1) << 11 | offset << 6 | id).
__device__ __forceinline__ int xcd_id() {
    return (int)(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (kXcds - 1));
}

// ---- work queue -----------------------------------------------------------------------

// Global item -> position in processing order (binary search over the phase prefix).
__device__ __forceinline__ int find_pos(const long long* off, int nm, long long item) {
    int lo = 0, hi = nm;  // off[lo] <= item < off[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int find_int(const int* off, int nm, int item) {
    int lo = 0, hi = nm;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Work split by XCD (each XCD has its own L2), for the adjoint: chunk c of [0, nchunks)
// belongs to XCD c / per (per = ceil(nchunks / kXcds)), so the waves sharing an L2 walk
// one contiguous range of chunks; once its range is taken, a workgroup takes chunks from
// the next XCD's. ctrs[y] counts the chunks taken from XCD y's range; *yoff (starting at
// 0) is the range the caller takes from, relative to its own XCD x0. Returns the chunk, or
// -1 when every range is taken. Called by one lane (the value is not broadcast).
__device__ __forceinline__ long long xcd_take(unsigned long long* ctrs, int x0, int* yoff,
                                              long long nchunks) {
    const long long per = (nchunks + kXcds - 1) / kXcds;
    while (*yoff < kXcds) {
        const int y = (x0 + *yoff) & (kXcds - 1);
        const long long k = (long long)atomicAdd(&ctrs[y], 1ull);
        const long long c = (long long)y * per + k;
        if (k < per && c < nchunks) return c;
        ++*yoff;
    }
    return -1;
}

// Next item of queue q for this workgroup (thread 0 fetches, all threads read).
__device__ __forceinline__ long long next_item(unsigned long long* q) {
    __shared__ long long s_item;
    __syncthreads();
    if (threadIdx.x == 0) s_item = (long long)atomicAdd(q, 1ull);
    __syncthreads();
    return s_item;
}

// Orders LDS traffic between lanes of one wavefront (no workgroup barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct Ctx {  // kernel arguments shared by every phase
    const MeshDesc* mesh;
    int nm;
    const int* order;           // processing order -> mesh id
    const long long* item_off;  // [kNumPhase][nm + 1] prefix of items in processing order
    const long long* tri_start; // [nm + 1] first global triangle per mesh id
    const int* lvl_off;         // [kMaxLevels + 1][nm + 1] prefix of nodes per level
    long long total_tris;
    const double* xyz;
    const int* tris;            // local point indices of each mesh
    const int* perm;            // tree slot -> global triangle
    int4* svtx;                 // tree slot -> global point indices of its vertices (w: 0)
    BatchCounters* ctr;
    TriRec* rec;                // per tree slot
    TriRec64* rec64;
    NodeRec* node;
    float4* nctr;               // per node: box centre (x, y, z) and, as bits, its link
    int4* xyr;                  // per triangle
    int* tile_count;
    int* tile_off;
    int* tile_cursor;
    int* tile_pool;
    long long tile_cap;
    int2* seg;                  // occupancy work items: (tile item, segment of its list)
    long long seg_cap;
    int* tile_left;             // per tile: segments not yet finished
    int* tile_flag;             // per tile: 1 if a segment met an exactly-0 determinant
    unsigned long long* masks;
    long long total_words;      // mask words of all meshes
    int* redo;                  // tiles (global items) for the exact occupancy pass
    uint8_t* occ;
    int* cell_tri;              // per cell, blocked (64 cells of a 4x4x4 block in a row):
                                // the tree slot of the triangle the cell's distance comes
                                // from: s >= 0 (FP32 nearest), -2 - s (FP64 near the
                                // surface), -1 (none)
    float* blk_dist;            // per cell, blocked like cell_tri: its unsigned distance
    const long long* cell_start;  // [nm + 1] first cell of each mesh id (concatenated)
    float* sdf;
    const float* gsdf;
    float* grad;
    float* xgrad;               // backward: per-XCD partial gradients (kXcds arrays of
                                // 3 floats per point, xstride floats apart)
    long long xstride;
    int* nearest;
    int total_tiles;
};

__device__ __forceinline__ const long long* phase_off(const Ctx& C, int ph) {
    return C.item_off + (size_t)ph * (C.nm + 1);
}

__device__ __forceinline__ int mesh_of_tri(const Ctx& C, long long t) {
    int lo = 0, hi = C.nm;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (C.tri_start[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

typedef const __attribute__((address_space(4))) float* ConstNodePtr;

__device__ __forceinline__ NodeRec load_node(const NodeRec* base, int i) {
    ConstNodePtr p = (ConstNodePtr)(base + i);
    NodeRec r;
    float* d = reinterpret_cast<float*>(&r);
#pragma unroll
    for (int k = 0; k < 16; ++k) d[k] = p[k];
    return r;
}

// Squared distance from p to the oriented box: a lower bound on the distance to anything
// inside (the float axes are orthonormal to a few ulps; the distance error margin covers
// the difference).
__device__ __forceinline__ float node_lb2(const NodeRec& N, float px, float py, float pz) {
    // search units: |projection| < 1, so clamping to [0, 1] is max(., 0)
    const float tn = clamp01(fabsf(proj(N.nx, N.ny, N.nz, N.dn, px, py, pz)) - N.en);
    const float tu = clamp01(fabsf(proj(N.ux, N.uy, N.uz, N.du, px, py, pz)) - N.eu);
    const float tw = clamp01(fabsf(proj(N.wx, N.wy, N.wz, N.dw, px, py, pz)) - N.ew);
    return fma3(tn, tn, tu, tu, tw, tw);
}

// true when the predicate holds on any lane of the wavefront
__device__ __forceinline__ bool wave_any(bool b) { return __builtin_amdgcn_ballot_w64(b) != 0; }

// One step of a DPP max reduction: x = max(x, x of the lane `ctrl` selects), on the rows
// row_mask enables (the other lanes keep x).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_max(float x) {
    const int xi = __builtin_bit_cast(int, x);
    const int y = __builtin_amdgcn_update_dpp(xi, xi, CTRL, ROW_MASK, 0xF, false);
    return fmaxf(x, __builtin_bit_cast(float, y));
}

// Largest x over the wavefront (every lane active), as a wave-uniform value: a max within
// quads (quad_perm [1,0,3,2], [2,3,0,1]), within rows of 16 (row_half_mirror,
// row_mirror), then across rows (row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2
// and 3), so lane 63 holds the maximum. Uniform by construction (readlane), so a branch
// on it keeps the code after it scalar.
__device__ __forceinline__ float wave_max(float x) {
    x = dpp_max<0xB1, 0xF>(x);
    x = dpp_max<0x4E, 0xF>(x);
    x = dpp_max<0x141, 0xF>(x);
    x = dpp_max<0x140, 0xF>(x);
    x = dpp_max<0x142, 0xA>(x);
    x = dpp_max<0x143, 0xC>(x);
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), 63));
}

// One step of a DPP sum: x plus the x of the lane `ctrl` selects, on the rows row_mask
// enables (the other rows add 0).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_add(float x) {
    const int y = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, ROW_MASK,
                                              0xF, false);
    return x + __builtin_bit_cast(float, y);
}

// Sum of x over the wavefront (every lane active), as a wave-uniform value, in VALU
// cross-lane steps (no LDS): wave_max's steps with additions, each adding disjoint partial
// sums — pairs, quads, halves of rows, rows; then row 0 into row 1 and row 2 into row 3,
// then rows 0-1 into row 3 — so lane 63 ends with the total.
__device__ __forceinline__ float wave_sum_dpp(float x) {
    x = dpp_add<0xB1, 0xF>(x);
    x = dpp_add<0x4E, 0xF>(x);
    x = dpp_add<0x141, 0xF>(x);
    x = dpp_add<0x140, 0xF>(x);
    x = dpp_add<0x142, 0xA>(x);
    x = dpp_add<0x143, 0xC>(x);
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), 63));
}

// Squared lower bound from the oriented box N to every point of the axis-aligned block
// centred at (cx, cy, cz) with half-sizes (hx, hy, hz): along each box axis the block
// projects to an interval of half-width hx|a_x| + hy|a_y| + hz|a_z|, and the squared
// separations along the three orthonormal axes add up to at most the squared distance.
// kBlockSlack (search units) covers the FP32 rounding of this bound and of node_lb2's, so
// the result never exceeds node_lb2 at a point of the block.
constexpr float kBlockSlack = 1e-5f;
__device__ __forceinline__ float block_lb2(const NodeRec& N, float cx, float cy, float cz,
                                           float hx, float hy, float hz) {
    const float rn = fma3(hx, fabsf(N.nx), hy, fabsf(N.ny), hz, fabsf(N.nz));
    const float ru = fma3(hx, fabsf(N.ux), hy, fabsf(N.uy), hz, fabsf(N.uz));
    const float rw = fma3(hx, fabsf(N.wx), hy, fabsf(N.wy), hz, fabsf(N.wz));
    const float tn = clamp01(fabsf(proj(N.nx, N.ny, N.nz, N.dn, cx, cy, cz)) - N.en - rn -
                             kBlockSlack);
    const float tu = clamp01(fabsf(proj(N.ux, N.uy, N.uz, N.du, cx, cy, cz)) - N.eu - ru -
                             kBlockSlack);
    const float tw = clamp01(fabsf(proj(N.wx, N.wy, N.wz, N.dw, cx, cy, cz)) - N.ew - rw -
                             kBlockSlack);
    return fma3(tn, tn, tu, tu, tw, tw);
}

// ---- phases ----------------------------------------------------------------------------

__global__ __launch_bounds__(256) void reset_kernel(Ctx C) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    if (tid == 0) *C.ctr = BatchCounters{};
    for (long long i = tid; i < C.total_tiles; i += stride) {
        C.tile_count[i] = 0;
        C.tile_cursor[i] = 0;
    }
    // the occupancy segments of a tile XOR their crossings into cleared columns
    for (long long i = tid; i < C.total_words; i += stride) C.masks[i] = 0ull;
}

__global__ __launch_bounds__(256) void zero_grad_kernel(float* g, long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        g[i] = 0.0f;
}

// Resets the work counters of one pass: q = 1 the voxel pass (its per-XCD counters),
// q = 2 the adjoint.
__global__ void reset_queue_kernel(BatchCounters* ctr, int q) {
    if (threadIdx.x == 0) ctr->queue[q] = 0;
    if (q == 1 && threadIdx.x < kXcds) ctr->xq[threadIdx.x] = 0;
    if (q == 2 && threadIdx.x < kXcds) ctr->aq[threadIdx.x] = 0;
}

// Centroid of each triangle relative to its mesh's grid corner (for the search order).
__global__ __launch_bounds__(256) void centroid_kernel(Ctx C, float4* cen) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < C.total_tris;
         t += stride) {
        const MeshDesc& M = C.mesh[mesh_of_tri(C, t)];
        const double* xyz = C.xyz + 3 * M.vbase;
        float q[3];
        for (int a = 0; a < 3; ++a) {
            double c = 0.0;
            for (int v = 0; v < 3; ++v) c += xyz[3 * (size_t)C.tris[3 * t + v] + a];
            q[a] = (float)(c / 3.0 - M.g.corner[a]);
        }
        cen[t] = make_float4(q[0], q[1], q[2], 0.0f);
    }
}

// Per tree slot: the triangle records. Per triangle: the xy bucket range and tile counts.
// ctr[key] += 1 for every active lane that has a key, with one atomic per distinct key in
// the wavefront; returns the lane's slot (the counter's old value plus its rank among the
// lanes with that key). Triangles in slot order are spatially close, so the lanes of a
// wavefront mostly share a column tile.
__device__ __forceinline__ int wave_increment(int* ctr, int key, bool has) {
    const int lane = (int)__lane_id();
    int slot = 0;
    while (true) {
        const unsigned long long m = __builtin_amdgcn_ballot_w64(has);
        if (m == 0) break;
        const int leader = __ffsll((long long)m) - 1;
        const int lk = __shfl(key, leader);
        const bool match = has && key == lk;
        const unsigned long long mm = __builtin_amdgcn_ballot_w64(match);
        int base = 0;
        if (lane == leader) base = atomicAdd(&ctr[lk], __popcll(mm));
        base = __shfl(base, leader);
        if (match) {
            slot = base + __popcll(mm & ((1ull << lane) - 1ull));
            has = false;
        }
    }
    return slot;
}

__global__ __launch_bounds__(256) void prep_kernel(Ctx C) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < C.total_tris;
         s += stride) {
        const MeshDesc& M = C.mesh[mesh_of_tri(C, s)];
        const double* xyz = C.xyz + 3 * M.vbase;
        const int t = C.perm[s];
        double P[3][3];
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) P[v][a] = xyz[3 * (size_t)C.tris[3 * (size_t)t + v] + a];
        TriRec r;
        TriRec64 r64;
        prep_triangle(P, M.g, M.sc, r, r64);
        C.rec[s] = r;
        C.rec64[s] = r64;
        const int4 x = xy_range(P, M.g);
        C.xyr[t] = x;
        // the r-th column tile of this triangle, r = 0, 1, ... (all lanes together)
        const int tx0 = x.x / kTile, nx = x.y / kTile - tx0 + 1;
        const int ty0 = x.z / kTile, ntiles = nx * (x.w / kTile - ty0 + 1);
        for (int r = 0; __builtin_amdgcn_ballot_w64(r < ntiles) != 0; ++r) {
            const bool has = r < ntiles;
            const int gt = has ? M.tile_base + (ty0 + r / nx) * M.ntx + tx0 + r % nx : 0;
            wave_increment(C.tile_count, gt, has);
        }
    }
}

// Per tree slot: the global point indices of its triangle's vertices (once at creation:
// the adjoint's scatter then takes one load instead of the slot's triangle and its
// indices).
__global__ __launch_bounds__(256) void slot_vtx_kernel(Ctx C) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < C.total_tris;
         s += stride) {
        const MeshDesc& M = C.mesh[mesh_of_tri(C, s)];
        const int* tv = C.tris + 3 * (size_t)C.perm[s];
        C.svtx[s] = make_int4((int)(M.vbase + tv[0]), (int)(M.vbase + tv[1]),
                              (int)(M.vbase + tv[2]), 0);
    }
}

// Per tile (in processing order): its list's offset in the pool and its occupancy work
// items, one per kSegTris list entries (none for an empty tile: its columns stay clear).
__global__ __launch_bounds__(256) void tile_alloc_kernel(Ctx C) {
    const long long* off = phase_off(C, kPhTile);
    const long long total = off[C.nm];
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long item = (long long)blockIdx.x * blockDim.x + threadIdx.x; item < total;
         item += stride) {
        const int pos = find_pos(off, C.nm, item);
        const int gt = C.mesh[C.order[pos]].tile_base + (int)(item - off[pos]);
        const int c = C.tile_count[gt];
        const long long at = (long long)atomicAdd(&C.ctr->tile_used, (unsigned long long)c);
        C.tile_off[gt] = (int)at;
        if (at + c > C.tile_cap) {
            atomicOr(&C.ctr->overflow, 1u);  // reported; the list is not read
            continue;
        }
        if (c == 0 || !C.seg) continue;
        const int nseg = (c + kSegTris - 1) / kSegTris;
        C.tile_left[gt] = nseg;
        C.tile_flag[gt] = 0;
        const long long s0 = (long long)atomicAdd(&C.ctr->seg_used, (unsigned long long)nseg);
        for (int k = 0; k < nseg; ++k)
            if (s0 + k < C.seg_cap) C.seg[s0 + k] = make_int2((int)item, k);
    }
}

__global__ __launch_bounds__(256) void tile_fill_kernel(Ctx C) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < C.total_tris;
         s += stride) {
        const MeshDesc& M = C.mesh[mesh_of_tri(C, s)];
        const int t = C.perm[s];  // slot order: neighbouring lanes share tiles
        const int4 x = C.xyr[t];
        const int tx0 = x.x / kTile, nx = x.y / kTile - tx0 + 1;
        const int ty0 = x.z / kTile, ntiles = nx * (x.w / kTile - ty0 + 1);
        for (int r = 0; __builtin_amdgcn_ballot_w64(r < ntiles) != 0; ++r) {
            const bool has = r < ntiles;
            const int gt = has ? M.tile_base + (ty0 + r / nx) * M.ntx + tx0 + r % nx : 0;
            const int slot = wave_increment(C.tile_cursor, gt, has);
            const long long at = has ? (long long)C.tile_off[gt] + slot : 0;
            if (has && at < C.tile_cap) C.tile_pool[at] = t;
        }
    }
}

// Occupancy (voxelize.cpp:78-195). The xy bucket of a column is reproduced exactly:
// triangle t belongs to column (i, j) iff i0 <= i <= i1 and j0 <= j <= j1, with the
// reference's clamp(floor((x - corner) / spacing)) ranges over the three vertices. One
// 64-lane workgroup per 8x8 column tile, one lane per column; the tile's triangle list
// (every triangle whose xy range touches the tile) is streamed through LDS in chunks of
// 64 triangles as structure-of-arrays FP64, and a lane runs the reference's crossing
// test (compat orient3d in the reference's operation order) on its own bucket.
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

// Turns a column's crossing bits (a crossing at height h sets bit K & 63 of word K >> 6,
// K = cells_below(h)) into its occupancy bits in place: cell k is inside iff the number of
// crossings above its centre is odd (voxelize.cpp:186-190), the XOR of the crossing bits
// above k (a suffix XOR from the top word). COHERENT reads the words with atomic loads
// (the crossings were added by other lanes' atomics), which bypass the non-coherent L1.
template <bool COHERENT>
__device__ inline void finish_column(unsigned long long* mask, int words) {
    unsigned long long carry = 0;
    for (int w = words - 1; w >= 0; --w) {
        const unsigned long long x =
            COHERENT ? __hip_atomic_load(mask + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                     : mask[w];
        unsigned long long s = x;
        s ^= s >> 1;
        s ^= s >> 2;
        s ^= s >> 4;
        s ^= s >> 8;
        s ^= s >> 16;
        s ^= s >> 32;  // bit b: XOR of bits b..63
        mask[w] = (s >> 1) ^ (carry ? ~0ull : 0ull);  // bit b: cell 64 w + b
        carry ^= s & 1ull;
    }
}

// The occupancy bytes in flat order (idx = (k ny + j) nx + i per mesh) from the columns'
// occupancy bits. One item per (mesh, grid row j, mask word w, run of 64 columns): lane l
// loads word w of column (64 r + l, j) and writes its up to 64 cells k = 64 w + b, one
// store instruction of 64 consecutive bytes per cell plane. (A tile pass writing its 8x8
// columns cell by cell stored 8-byte row pieces, 5x the bytes in partial lines.)
__global__ __launch_bounds__(256) void occ_rows_kernel(Ctx C) {
    const long long* off = phase_off(C, kPhOccRow);
    const long long total = off[C.nm];
    const int lane = threadIdx.x & 63;
    const long long nwave = (long long)gridDim.x * (blockDim.x / 64);
    int pos = 0;
    long long p_lo = 0, p_hi = 0;  // items of mesh position pos
    for (long long item = (long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
         item < total; item += nwave) {
        if (item < p_lo || item >= p_hi) {
            pos = find_pos(off, C.nm, item);
            p_lo = off[pos];
            p_hi = off[pos + 1];
        }
        const MeshDesc& M = C.mesh[C.order[pos]];
        const GridSpec& g = M.g;
        const int nrun = (g.nx + 63) / 64;
        long long r = item - p_lo;  // (j ny-major, w, run)
        const int run = (int)(r % nrun);
        r /= nrun;
        const int w = (int)(r % M.words), j = (int)(r / M.words);
        const int i = 64 * run + lane;
        const unsigned long long bits =
            i < g.nx ? C.masks[M.mask_base + ((size_t)j * g.nx + i) * M.words + w] : 0ull;
        const int k0 = 64 * w, k1 = min(g.nz, k0 + 64);
        uint8_t* col = C.occ + M.cell_base + (size_t)j * g.nx + i;
        const size_t plane = (size_t)g.nx * g.ny;
        if (i < g.nx)
            for (int k = k0; k < k1; ++k) col[(size_t)k * plane] = (uint8_t)((bits >> (k - k0)) & 1ull);
    }
}

// Sign of the naive FP64 orient3d determinant (predicates.cpp:2201-2222), 0 when it is
// exactly 0 — then only the exact sign can tell, and the tile is redone by the exact pass.
__device__ __forceinline__ int o3_naive_sign(const double* pa, const double* pb,
                                             const double* pc, const double* pd) {
    double adx = pa[0] - pd[0], ady = pa[1] - pd[1], adz = pa[2] - pd[2];
    double bdx = pb[0] - pd[0], bdy = pb[1] - pd[1], bdz = pb[2] - pd[2];
    double cdx = pc[0] - pd[0], cdy = pc[1] - pd[1], cdz = pc[2] - pd[2];
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    double cdxady = cdx * ady, adxcdy = adx * cdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    return det > 0.0 ? 1 : (det < 0.0 ? -1 : 0);
}

// Fast occupancy pass. A work item is a segment of up to kSegTris entries of one 8x8
// column tile's triangle list, so that a tile under a dense patch (tens of thousands of
// list entries) is shared by many workgroups. The lanes work on (column, triangle) pairs
// instead of columns, so that the FP64 predicates run on full wavefronts (a triangle
// usually covers only a few of a tile's 64 columns): each LDS chunk of 64 triangles is
// expanded into the pairs of its triangles' column ranges clipped to the tile (a wave
// prefix sum gives every pair a lane), and a crossing flips its column's "below K" bit
// with an atomic XOR into the cleared masks. Every predicate is the naive FP64 sign. The
// last segment of a tile to finish (a per-tile countdown) turns the crossing bits into
// occupancy bits, or, if any segment met a determinant that is exactly 0, queues the tile
// for the exact pass (occ_kernel), which settles zeros with the exact sign, nudges and
// replays as the reference does.
__global__ __launch_bounds__(64) void occ_pairs_kernel(Ctx C) {
    __shared__ double sv[9][kChunk];
    __shared__ int s_i0[kChunk], s_j0[kChunk], s_ni[kChunk];
    __shared__ int s_off[kChunk + 1];
    const long long* off = phase_off(C, kPhTile);
    const long long total = (long long)min(C.ctr->seg_used, (unsigned long long)C.seg_cap);
    const int lane = threadIdx.x;
    while (true) {
        const long long sitem = next_item(&C.ctr->queue[0]);
        if (sitem >= total) break;
        const int2 sg = C.seg[sitem];
        const long long item = sg.x;
        const int pos = find_pos(off, C.nm, item);
        const MeshDesc& M = C.mesh[C.order[pos]];
        const GridSpec g = M.g;
        const int tile = (int)(item - off[pos]);
        const int tx0 = (tile % M.ntx) * kTile, ty0 = (tile / M.ntx) * kTile;
        const int gt = M.tile_base + tile;
        const long long tbeg = C.tile_off[gt];
        const long long beg = tbeg + (long long)sg.y * kSegTris;
        const long long end = min(tbeg + C.tile_count[gt], beg + kSegTris);
        const double* xyz = C.xyz + 3 * M.vbase;
        unsigned long long* const masks = C.masks + M.mask_base;
        const double zmin = g.corner[2] - g.spacing;
        const double zmax = g.corner[2] + (g.nz + 1) * g.spacing;
        bool unsure = false;
        for (long long c0 = beg; c0 < end; c0 += kChunk) {
            const int cnt = (int)min((long long)kChunk, end - c0);
            int npair = 0;
            __syncthreads();
            if (lane < cnt) {
                const int t = C.tile_pool[c0 + lane];
                const int4 r = C.xyr[t];  // a list triangle's range meets the tile
                const int i0 = max(r.x, tx0), i1 = min(r.y, tx0 + kTile - 1);
                const int j0 = max(r.z, ty0), j1 = min(r.w, ty0 + kTile - 1);
                s_i0[lane] = i0;
                s_j0[lane] = j0;
                s_ni[lane] = i1 - i0 + 1;
                npair = (i1 >= i0 && j1 >= j0) ? (i1 - i0 + 1) * (j1 - j0 + 1) : 0;
#pragma unroll
                for (int v = 0; v < 3; ++v) {
                    const double* p = xyz + 3 * (size_t)C.tris[3 * (size_t)t + v];
                    sv[3 * v][lane] = p[0];
                    sv[3 * v + 1][lane] = p[1];
                    sv[3 * v + 2][lane] = p[2];
                }
            }
            int incl = npair;  // inclusive prefix over the lanes
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            s_off[lane + 1] = incl;
            if (lane == 0) s_off[0] = 0;
            __syncthreads();
            const int npairs = s_off[cnt];
            for (int p = lane; p < npairs; p += 64) {
                int lo = 0, hi = cnt;  // s_off[lo] <= p < s_off[hi]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s_off[mid] <= p) lo = mid;
                    else hi = mid;
                }
                const int q = lo, r = p - s_off[q];
                const int i = s_i0[q] + r % s_ni[q], j = s_j0[q] + r / s_ni[q];
                const double x = g.corner[0] + (i + 0.5) * g.spacing;  // voxelize.cpp:180
                const double y = g.corner[1] + (j + 0.5) * g.spacing;  // voxelize.cpp:181
                const double A[3] = {sv[0][q], sv[1][q], sv[2][q]};
                const double B[3] = {sv[3][q], sv[4][q], sv[5][q]};
                const double Cc[3] = {sv[6][q], sv[7][q], sv[8][q]};
                const double P[3] = {x, y, zmin};
                const double Q[3] = {x, y, zmax};
                const int sa = o3_naive_sign(P, A, B, Cc);
                const int sb = o3_naive_sign(Q, A, B, Cc);
                if (sa == 0 || sb == 0) {
                    unsure = true;
                    continue;
                }
                if (sa == sb) continue;
                const int s1 = o3_naive_sign(P, Q, A, B);
                const int s2 = o3_naive_sign(P, Q, B, Cc);
                const int s3 = o3_naive_sign(P, Q, Cc, A);
                if (s1 == 0 || s2 == 0 || s3 == 0) {
                    unsure = true;
                    continue;
                }
                if (s1 != s2 || s2 != s3) continue;
                const int K = cells_below(g, plane_z(A, B, Cc, x, y));
                atomicXor(masks + ((size_t)j * g.nx + i) * M.words + (K >> 6), 1ull << (K & 63));
            }
        }
        unsure = __syncthreads_or(unsure);
        if (unsure && lane == 0) atomicOr(&C.tile_flag[gt], 1);
        __threadfence();  // this segment's crossings and flag land before its countdown
        int last = 0;
        if (lane == 0) last = atomicSub(&C.tile_left[gt], 1) == 1;
        if (!__shfl(last, 0)) continue;
        // every segment's crossings are visible: they are device-scope atomics (segments
        // of one tile may run on different XCDs), read back below with device-scope loads
        __threadfence();
        if (__hip_atomic_load(&C.tile_flag[gt], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            if (lane == 0) C.redo[atomicAdd(&C.ctr->redo_n, 1ull)] = (int)item;
            continue;
        }
        const int ci = tx0 + (lane % kTile), cj = ty0 + (lane / kTile);
        if (ci < g.nx && cj < g.ny)
            finish_column<true>(masks + ((size_t)cj * g.nx + ci) * M.words, M.words);
    }
}

// One 8x8 column tile per workgroup item. EXACT = false: every predicate is the naive
// FP64 sign; a tile in which any of them is exactly 0 (a possible graze, which only the
// exact sign can settle) writes nothing and is queued for the EXACT pass, which runs the
// compat predicate (naive sign, exact sign on 0) with the nudge and replay passes. Kept
// as two kernels so that the expansion arithmetic's registers do not limit the
// occupancy of the common one.
template <bool EXACT>
__global__ __launch_bounds__(64) void occ_kernel(Ctx C) {
    __shared__ double sv[9][kChunk];
    __shared__ int4 sr[kChunk];
    __shared__ int st[kChunk];
    const long long* off = phase_off(C, kPhTile);
    const long long total = EXACT ? 0 : off[C.nm];
    const int lane = threadIdx.x;
    while (true) {
        long long item = next_item(&C.ctr->queue[EXACT ? 3 : 0]);
        if (EXACT) {
            if (item >= (long long)C.ctr->redo_n) break;
            item = C.redo[item];
        } else if (item >= total) {
            break;
        }
        const int pos = find_pos(off, C.nm, item);
        const MeshDesc& M = C.mesh[C.order[pos]];
        const GridSpec g = M.g;
        const int tile = (int)(item - off[pos]);
        const int tx = tile % M.ntx, ty = tile / M.ntx;
        const int ci = tx * kTile + (lane % kTile), cj = ty * kTile + (lane / kTile);
        const bool valid = ci < g.nx && cj < g.ny;
        const int gt = M.tile_base + tile;
        const long long beg = C.tile_off[gt];
        const long long end = beg + C.tile_count[gt];
        const bool ok = end <= C.tile_cap;  // overflowed lists are reported, not read
        const double* xyz = C.xyz + 3 * M.vbase;
        unsigned long long* mask =
            C.masks + M.mask_base + ((size_t)(valid ? cj : 0) * g.nx + (valid ? ci : 0)) * M.words;
        const double zmin = g.corner[2] - g.spacing;
        const double zmax = g.corner[2] + (g.nz + 1) * g.spacing;
        const double cx = g.corner[0] + (ci + 0.5) * g.spacing;
        const double cy = g.corner[1] + (cj + 0.5) * g.spacing;
        bool active = valid && ok;
        bool unsure = false;  // EXACT = false: a naive determinant was exactly 0
        int cutoff = INT_MAX;
        auto o3 = [&](const double* a, const double* b, const double* c, const double* d) {
            if (EXACT) return o3_compat(a, b, c, d);
            const int sgn = o3_naive_sign(a, b, c, d);
            unsure |= sgn == 0;
            return sgn;
        };
        for (int pass = 0; pass < (EXACT ? 8 : 1); ++pass) {
            if (!__syncthreads_or(active)) break;
            double x = cx, y = cy;
            if (pass > 0) {
                const int d = (pass < 7 ? pass : 6) - 1;
                x = cx + kNudge[d][0] * g.spacing;
                y = cy + kNudge[d][1] * g.spacing;
            }
            if (active)
                for (int w = 0; w < M.words; ++w) mask[w] = 0ull;
            bool grazed = false;
            int first_graze = INT_MAX;
            const double P[3] = {x, y, zmin};
            const double Q[3] = {x, y, zmax};
            for (long long c0 = beg; c0 < end; c0 += kChunk) {
                __syncthreads();
                const long long k = c0 + lane;
                if (k < end) {
                    const int t = C.tile_pool[k];
                    st[lane] = t;
                    sr[lane] = C.xyr[t];
#pragma unroll
                    for (int v = 0; v < 3; ++v) {
                        const double* p = xyz + 3 * (size_t)C.tris[3 * (size_t)t + v];
                        sv[3 * v][lane] = p[0];
                        sv[3 * v + 1][lane] = p[1];
                        sv[3 * v + 2][lane] = p[2];
                    }
                }
                __syncthreads();
                if (!active) continue;
                const int cnt = (int)min((long long)kChunk, end - c0);
                for (int q = 0; q < cnt; ++q) {
                    const int4 r = sr[q];
                    if (ci < r.x || ci > r.y || cj < r.z || cj > r.w) continue;
                    const int t = st[q];  // ascending within a mesh iff ascending globally
                    if (pass == 7 && t >= cutoff) continue;
                    if (grazed && pass < 6) break;
                    const double A[3] = {sv[0][q], sv[1][q], sv[2][q]};
                    const double B[3] = {sv[3][q], sv[4][q], sv[5][q]};
                    const double Cc[3] = {sv[6][q], sv[7][q], sv[8][q]};
                    const int sa = o3(P, A, B, Cc);
                    const int sb = o3(Q, A, B, Cc);
                    if (sa == 0 || sb == 0) {
                        grazed = true;
                        first_graze = min(first_graze, t);
                        continue;
                    }
                    if (sa == sb) continue;
                    const int s1 = o3(P, Q, A, B);
                    const int s2 = o3(P, Q, B, Cc);
                    const int s3 = o3(P, Q, Cc, A);
                    if (s1 == 0 || s2 == 0 || s3 == 0) {
                        grazed = true;
                        first_graze = min(first_graze, t);
                        continue;
                    }
                    if (s1 != s2 || s2 != s3) continue;
                    const int K = cells_below(g, plane_z(A, B, Cc, x, y));
                    mask[K >> 6] ^= 1ull << (K & 63);
                }
            }
            if (active) {
                if (pass < 6) {
                    active = grazed;
                } else if (pass == 6) {
                    if (grazed) cutoff = first_graze;
                    active = grazed;
                } else {
                    active = false;
                }
            }
        }
        if (!EXACT && __syncthreads_or(unsure)) {
            if (lane == 0) C.redo[atomicAdd(&C.ctr->redo_n, 1ull)] = (int)item;
            continue;
        }
        if (valid && ok) finish_column<false>(mask, M.words);
    }
}

// ---- tree of oriented bounding boxes ---------------------------------------------------

__device__ __forceinline__ void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ double dot3(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ void normalize3(double* a) {
    const double l = sqrt(dot3(a, a));
    for (int k = 0; k < 3; ++k) a[k] /= l;
}

// The oriented box of a point set given by points(f) (f is called with every point, twice)
// and the aligned normal sum N: axis n = N / |N| (or z), u the principal direction of the
// points projected on the plane of n (closed-form 2x2 eigenvector), w = n x u. The axes
// are rounded to float first and the extents measured along the float axes, so the float
// box contains every point; the extents then grow by the rounding of the centre's
// coordinates and by the rounding error of the test's projection fma(n, p, -dn) for any
// point p with |p|_1 <= pmax. The centre goes to ctr.
template <class Points>
__device__ inline NodeRec make_box(const double* N, double pmax, float4& ctr, Points&& points) {
    double n[3] = {N[0], N[1], N[2]};
    if (dot3(n, n) > 0.0) {
        normalize3(n);
    } else {
        n[0] = n[1] = 0.0;
        n[2] = 1.0;
    }
    double a[3], b[3];
    const double tmp[3] = {fabs(n[0]) < 0.9 ? 1.0 : 0.0, fabs(n[0]) < 0.9 ? 0.0 : 1.0, 0.0};
    cross3(n, tmp, a);
    normalize3(a);
    cross3(n, a, b);
    // second moments of the points in the (a, b) plane
    double cnt = 0.0, mx = 0.0, my = 0.0, sxx = 0.0, syy = 0.0, sxy = 0.0;
    points([&](const double* P) {
        const double x = dot3(P, a), y = dot3(P, b);
        cnt += 1.0;
        mx += x;
        my += y;
        sxx += x * x;
        syy += y * y;
        sxy += x * y;
    });
    mx /= cnt;
    my /= cnt;
    sxx = sxx / cnt - mx * mx;
    syy = syy / cnt - my * my;
    sxy = sxy / cnt - mx * my;
    const double th = 0.5 * atan2(2.0 * sxy, sxx - syy);
    double u[3], w[3];
    for (int k = 0; k < 3; ++k) u[k] = cos(th) * a[k] + sin(th) * b[k];
    normalize3(u);
    cross3(n, u, w);
    normalize3(w);
    float nf[3], uf[3], wf[3];
    double ax[3][3];
    for (int k = 0; k < 3; ++k) {
        nf[k] = (float)n[k];
        uf[k] = (float)u[k];
        wf[k] = (float)w[k];
        ax[0][k] = nf[k];
        ax[1][k] = uf[k];
        ax[2][k] = wf[k];
    }
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    points([&](const double* P) {
        for (int k = 0; k < 3; ++k) {
            const double q = dot3(P, ax[k]);
            lo[k] = fmin(lo[k], q);
            hi[k] = fmax(hi[k], q);
        }
    });
    double c[3] = {0.0, 0.0, 0.0}, d[3], e[3];
    for (int k = 0; k < 3; ++k) {
        d[k] = 0.5 * (lo[k] + hi[k]);
        e[k] = 0.5 * (hi[k] - lo[k]);
        for (int j = 0; j < 3; ++j) c[j] += d[k] * ax[k][j];
    }
    ctr = make_float4((float)c[0], (float)c[1], (float)c[2], 0.0f);
    NodeRec B;
    B.dn = (float)d[0];
    B.du = (float)d[1];
    B.dw = (float)d[2];
    B.nx = nf[0];
    B.ny = nf[1];
    B.nz = nf[2];
    B.ux = uf[0];
    B.uy = uf[1];
    B.uz = uf[2];
    B.wx = wf[0];
    B.wy = wf[1];
    B.wz = wf[2];
    B.en = (float)(e[0] * (1.0 + 1e-6) + 8.0 * 0x1p-24 * (pmax + fabs(d[0])));
    B.eu = (float)(e[1] * (1.0 + 1e-6) + 8.0 * 0x1p-24 * (pmax + fabs(d[1])));
    B.ew = (float)(e[2] * (1.0 + 1e-6) + 8.0 * 0x1p-24 * (pmax + fabs(d[2])));
    B.link = 0u;
    return B;
}

// Level 1: the box of up to 8 consecutive triangles, from their FP64 vertices (relative
// to the grid corner); the normal is the sum of the triangles' area vectors, each turned
// to agree with the first.
__global__ __launch_bounds__(256) void node_leaf_kernel(Ctx C, int total) {
    const int* lo = C.lvl_off + (C.nm + 1);
    const int stride = gridDim.x * blockDim.x;
    for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi < total; gi += stride) {
        const int m = find_int(lo, C.nm, gi);
        const MeshDesc& M = C.mesh[m];
        const int loc = gi - lo[m];
        const long long s0 = M.tbase + 8LL * loc;
        const int cnt = min(8, M.ntris - 8 * loc);
        double N[3] = {0.0, 0.0, 0.0}, ref[3] = {0.0, 0.0, 0.0};
        bool have_ref = false;
        for (int k = 0; k < cnt; ++k) {
            const TriRec64& T = C.rec64[s0 + k];
            double e1[3], e2[3], nt[3];
            for (int a = 0; a < 3; ++a) {
                e1[a] = T.b[a] - T.a[a];
                e2[a] = T.c[a] - T.a[a];
            }
            cross3(e1, e2, nt);
            if (!have_ref && dot3(nt, nt) > 0.0) {
                for (int a = 0; a < 3; ++a) ref[a] = nt[a];
                have_ref = true;
            }
            const double sg = dot3(nt, ref) < 0.0 ? -1.0 : 1.0;
            for (int a = 0; a < 3; ++a) N[a] += sg * nt[a];
        }
        float4 ctr;
        NodeRec B = make_box(N, 3.0 * grid_extent(M.g) * M.sc, ctr, [&](auto&& f) {
            for (int k = 0; k < cnt; ++k) {
                const TriRec64& T = C.rec64[s0 + k];
                for (int v = 0; v < 3; ++v) {
                    const double* P = v == 0 ? T.a : (v == 1 ? T.b : T.c);
                    const double q[3] = {(P[0] - M.g.corner[0]) * M.sc,
                                         (P[1] - M.g.corner[1]) * M.sc,
                                         (P[2] - M.g.corner[2]) * M.sc};
                    f(q);
                }
            }
        });
        B.link = kLinkLeaf | ((unsigned)loc << 4) | (unsigned)cnt;
        ctr.w = __uint_as_float(B.link);
        C.node[M.node_base[1] + loc] = B;
        C.nctr[M.node_base[1] + loc] = ctr;
    }
}

// Level l >= 2: the box around the corners of up to 8 child boxes; the normal is the sum
// of the child normals weighted by their in-plane areas, each turned to agree with the
// first.
__global__ __launch_bounds__(256) void node_merge_kernel(Ctx C, int level, int total) {
    const int* lo = C.lvl_off + (size_t)level * (C.nm + 1);
    const int stride = gridDim.x * blockDim.x;
    for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi < total; gi += stride) {
        const int m = find_int(lo, C.nm, gi);
        const MeshDesc& M = C.mesh[m];
        const int loc = gi - lo[m];
        const int first = 8 * loc;
        const int cnt = min(8, M.nnodes[level - 1] - first);
        const int cfirst = M.node_base[level - 1] + first;
        const NodeRec* ch = C.node + cfirst;
        const double ref[3] = {ch[0].nx, ch[0].ny, ch[0].nz};
        double N[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < cnt; ++k) {
            const double nk[3] = {ch[k].nx, ch[k].ny, ch[k].nz};
            const double wgt = (double)ch[k].eu * ch[k].ew + 1e-30;
            const double sg = dot3(nk, ref) < 0.0 ? -1.0 : 1.0;
            for (int a = 0; a < 3; ++a) N[a] += sg * wgt * nk[a];
        }
        float4 ctr;
        NodeRec B = make_box(N, 3.0 * grid_extent(M.g) * M.sc, ctr, [&](auto&& f) {
            for (int k = 0; k < cnt; ++k) {
                const NodeRec K = ch[k];
                const float4 KC = C.nctr[cfirst + k];
                for (int corner = 0; corner < 8; ++corner) {
                    const double sn = (corner & 1) ? K.en : -(double)K.en;
                    const double su = (corner & 2) ? K.eu : -(double)K.eu;
                    const double sw = (corner & 4) ? K.ew : -(double)K.ew;
                    const double q[3] = {KC.x + sn * K.nx + su * K.ux + sw * K.wx,
                                         KC.y + sn * K.ny + su * K.uy + sw * K.wy,
                                         KC.z + sn * K.nz + su * K.uz + sw * K.wz};
                    f(q);
                }
            }
        });
        B.link = ((unsigned)cfirst << 4) | (unsigned)cnt;
        ctr.w = __uint_as_float(B.link);
        C.node[M.node_base[level] + loc] = B;
        C.nctr[M.node_base[level] + loc] = ctr;
    }
}

// ---- voxel pass ---------------------------------------------------------------------------

// Blocked index of cell (i, j, k) in a mesh's cell_tri array: the 64 cells of each
// 4x4x4 block are contiguous (one wavefront's cells in the 4x4x4 variant).
__device__ __forceinline__ size_t blocked_cell(const MeshDesc& M, int i, int j, int k) {
    const size_t blk = (size_t)(i >> 2) +
                       (size_t)M.L0.bx * ((size_t)(j >> 2) + (size_t)M.L0.by * (size_t)(k >> 2));
    return blk * 64 + (size_t)((i & 3) + 4 * (j & 3) + 16 * (k & 3));
}

// One voxel leaf per item: 4x4x(4 VPL) voxels, VPL per lane, 4 cells apart along z. D = 0
// reads a tree leaf's triangle records with scalar loads (SGPR operands); D = 8 stages the
// leaf's 8 records through LDS as one 1 KiB tile. Writes the signed distance and, for the
// backward pass, the triangle each cell's distance comes from.
// The register budget keeps at least 5 waves per SIMD for the 4x4x4 variant (at most 96
// VGPRs), 4 and 3 for the taller ones. (A 6-wave budget spills and measured slower.)
template <int VPL, int D>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(VPL == 1 ? 5 : (VPL == 2 ? 4 : 3)))) void vox_kernel(Ctx C) {
    __shared__ TriRec s_rec[D > 0 ? D : 1];
    const long long* off = phase_off(C, VPL == 1 ? kPhV1 : (VPL == 2 ? kPhV2 : kPhV4));
    const long long total = off[C.nm];
    const int lane = threadIdx.x;
    unsigned long long nvis = 0, ntest = 0, neval = 0;
    // Items run over super-tiles of 4 x 4 x SZ leaves (16^3 cells; super-tiles in x-fastest
    // order per mesh, meshes in processing order). XCD x takes super-tiles x, x + 8,
    // x + 16, ... from its own counter, so that the waves sharing one L2 walk a few
    // compact regions at a time; a wave whose XCD's super-tiles are all taken helps the
    // next XCD. (One contiguous range of super-tiles per XCD was 20 % slower: the ranges
    // differ in cost, and the XCDs that finish early all help the same neighbour.)
    constexpr int SZ = kSuperZ / VPL;
    constexpr int kPer = kSuper * kSuper * SZ;  // items per super-tile
    const long long nst = (total + kPer - 1) / kPer;
    int yoff = 0;  // the XCD counter this wave takes from: xcd_id() + yoff
    const int x0 = xcd_id();
    while (true) {
        long long item = -1;
        while (yoff < kXcds) {
            const int y = (x0 + yoff) & (kXcds - 1);
            const long long mine = (nst - y + kXcds - 1) / kXcds;  // super-tiles of XCD y
            unsigned long long k = 0;
            if (lane == 0) k = atomicAdd(&C.ctr->xq[y], 1ull);
            // wave-uniform (readfirstlane keeps it in scalar registers)
            k = (unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)k) |
                ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(k >> 32)) << 32);
            if ((long long)(k / kPer) < mine) {
                item = ((long long)(k / kPer) * kXcds + y) * kPer + (long long)(k % kPer);
                break;
            }
            ++yoff;
        }
        if (item < 0) break;
        if (item >= total) continue;  // the last super-tile is only partly in range
        const int pos = find_pos(off, C.nm, item);
        const MeshDesc& M = C.mesh[C.order[pos]];
        const GridSpec& g = M.g;
        const Level L0 = VPL == 1 ? M.L0 : (VPL == 2 ? M.L0b : M.L0c);
        const long long rel = item - off[pos];
        const int sbx = (L0.bx + kSuper - 1) / kSuper, sby = (L0.by + kSuper - 1) / kSuper;
        const long long st = rel / kPer;
        const int w = (int)(rel % kPer);
        const int ix = kSuper * (int)(st % sbx) + w % kSuper;
        const int iy = kSuper * (int)((st / sbx) % sby) + (w / kSuper) % kSuper;
        const int iz = SZ * (int)(st / ((long long)sbx * sby)) + w / (kSuper * kSuper);
        if (ix >= L0.bx || iy >= L0.by || iz >= L0.bz) continue;  // past the grid's edge
        const int i = ix * kLeaf + (lane & 3), j = iy * kLeaf + ((lane >> 2) & 3);
        int k[VPL];
        bool valid[VPL];
        float pz[VPL];
        const float h = (float)g.spacing * M.scf;  // the cell size in search units (exact)
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            k[v] = iz * L0.sz + (lane >> 4) + 4 * v;
            valid[v] = i < g.nx && j < g.ny && k[v] < g.nz;
            pz[v] = (k[v] + 0.5f) * h;
        }
        const float px = (i + 0.5f) * h, py = (j + 0.5f) * h;
        const float eps = M.eps;
        // centre of the wavefront's voxel block (relative to the grid corner)
        const float wcx = (ix * kLeaf + 2) * h, wcy = (iy * kLeaf + 2) * h,
                    wcz = (iz * L0.sz + 0.5f * L0.sz) * h;
        // half-sizes of the block of voxel centres (cells 0..3 around the centre 2, and
        // 0..sz-1 around sz/2)
        const float whx = 1.5f * h, why = 1.5f * h, whz = (0.5f * L0.sz - 0.5f) * h;

        // Walks the tree with the squared bounds b2[] (which on_leaf may tighten; lanes
        // that take no part hold a negative bound); on_leaf(loc, cnt) visits the cnt
        // triangles of tree leaf loc. The stack is wave-uniform and lives in three VGPRs,
        // entry e in lane e (at most 1 + 7 (nlev - 1) <= 64 entries, nlev <= 10): the
        // node's global index, its child link and the version it was pushed at.
        // `ver` counts the leaf visits that tightened some lane's bound; a node pushed at
        // the current version passed its test with the current bounds and is not re-tested.
        int ver = 0;
        const int root = M.node_base[M.nlev];
        const long long tbase = M.tbase;
        // the stack entry below a leaf being visited (its link), and the records of that
        // entry's leaf loaded ahead into registers (slot pf_s0, -1: none)
        unsigned next_link = 0u;
        long long pf_s0 = -1;
        float4 pf0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), pf1 = pf0;
        auto walk = [&](const float* b2, auto&& on_leaf) {
            float mb = 0.0f;  // largest bound of the wavefront, valid at version mb_ver
            int mb_ver = -1;
            int stk = root;  // entry 0: the root, never tested before
            int stl = 0;
            int stv = -1;
            int sp = 1;
            while (sp > 0) {
                --sp;
                const int node = __builtin_amdgcn_readlane(stk, sp);
                unsigned link;
                if (__builtin_amdgcn_readlane(stv, sp) != ver) {
                    // the bounds have tightened since this node was pushed
                    ++nvis;
                    const NodeRec N = load_node(C.node, node);
                    bool need = false;
#pragma unroll
                    for (int v = 0; v < VPL; ++v) need |= node_lb2(N, px, py, pz[v]) <= b2[v];
                    if (!wave_any(need)) continue;
                    link = N.link;
                } else {
                    link = (unsigned)__builtin_amdgcn_readlane(stl, sp);
                }
                const int cnt = (int)(link & 15u), first = (int)((link >> 4) & 0x7ffffffu);
                if (link & kLinkLeaf) {
                    next_link = sp > 0 ? (unsigned)__builtin_amdgcn_readlane(stl, sp - 1) : 0u;
                    on_leaf(first, cnt);
                    continue;
                }
                // order key of child `lane`: squared distance from the voxel block's centre
                // to the child's box centre (discriminates better than the lower bounds,
                // which are 0 for every box that overlaps the block); non-negative, so its
                // bits order as unsigned integers
                float ckey = 0.0f;
                int clink = 0;  // link of child `lane`
                // block-level test of child `lane`: a child whose box is farther from the
                // whole voxel block than the largest bound of the wavefront fails every
                // lane's test below, which is then skipped
                if (mb_ver != ver) {  // the bounds only change when `ver` does
                    mb = b2[0];
#pragma unroll
                    for (int v = 1; v < VPL; ++v) mb = fmaxf(mb, b2[v]);
                    mb = wave_max(mb);
                    mb_ver = ver;
                }
                bool bkeep = false;
                if (lane < cnt) {
                    const float4 cc = C.nctr[first + lane];
                    const float kx = cc.x - wcx, ky = cc.y - wcy, kz = cc.z - wcz;
                    ckey = fma3(kx, kx, ky, ky, kz, kz);
                    clink = __float_as_int(cc.w);
                    const float4* np = reinterpret_cast<const float4*>(C.node + first + lane);
                    NodeRec K;
                    float4* kd = reinterpret_cast<float4*>(&K);
#pragma unroll
                    for (int q4 = 0; q4 < 4; ++q4) kd[q4] = np[q4];
                    bkeep = block_lb2(K, wcx, wcy, wcz, whx, why, whz) <= mb;
                }
                const unsigned bmask = (unsigned)__builtin_amdgcn_ballot_w64(bkeep);
                unsigned mask = 0, near_k = ~0u;
                int near_c = -1;
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    if (c < cnt && ((bmask >> c) & 1u)) {
                        ++nvis;
                        const NodeRec K = load_node(C.node, first + c);
                        bool need = false;
#pragma unroll
                        for (int v = 0; v < VPL; ++v)
                            need |= node_lb2(K, px, py, pz[v]) <= b2[v];
                        if (wave_any(need)) {
                            mask |= 1u << c;
                            const unsigned key = (unsigned)__builtin_amdgcn_readlane(
                                __builtin_bit_cast(int, ckey), c);
                            if (key < near_k) {
                                near_k = key;
                                near_c = c;
                            }
                        }
                    }
                }
                // the nearest child last, so it is visited first
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (((mask >> c) & 1u) && c != near_c) {
                        const bool mine = lane == sp;
                        stk = mine ? first + c : stk;
                        stl = mine ? __builtin_amdgcn_readlane(clink, c) : stl;
                        stv = mine ? ver : stv;
                        ++sp;
                    }
                if (near_c >= 0) {
                    const bool mine = lane == sp;
                    stk = mine ? first + near_c : stk;
                    stl = mine ? __builtin_amdgcn_readlane(clink, near_c) : stl;
                    stv = mine ? ver : stv;
                    ++sp;
                }
            }
        };

        // triangles of one tree leaf, through LDS (D > 0) or scalar loads. With LDS, the
        // records of the leaf stacked below this one (usually a sibling, visited next) are
        // loaded into registers while this leaf's triangles are tested; a leaf that is
        // then visited takes them from there (records do not change during the pass).
        auto leaf_rec = [&](int loc, int cnt) -> long long {
            const long long s0 = tbase + 8LL * loc;
            if (D > 0) {
                static_assert(8 * kTriRecF4 <= 128, "a leaf's records fit two loads per lane");
                const int nq = cnt * kTriRecF4;
                wave_sync();
                float4* dst = reinterpret_cast<float4*>(s_rec);
                if (s0 == pf_s0) {
                    if (lane < nq) dst[lane] = pf0;
                    if (lane + 64 < nq) dst[lane + 64] = pf1;
                } else {
                    const float4* src = reinterpret_cast<const float4*>(C.rec + s0);
                    if (lane < nq) dst[lane] = src[lane];
                    if (lane + 64 < nq) dst[lane + 64] = src[lane + 64];
                }
                wave_sync();
                pf_s0 = -1;
                if (next_link & kLinkLeaf) {
                    pf_s0 = tbase + 8LL * (long long)((next_link >> 4) & 0x7ffffffu);
                    const int nn = (int)(next_link & 15u) * kTriRecF4;
                    const float4* src = reinterpret_cast<const float4*>(C.rec + pf_s0);
                    if (lane < nn) pf0 = src[lane];
                    if (lane + 64 < nn) pf1 = src[lane + 64];
                }
            }
            return s0;
        };

        // ---- FP32 nearest triangle
        float best[VPL], b2[VPL];
        int arg[VPL];
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            best[v] = FLT_MAX;
            b2[v] = valid[v] ? FLT_MAX : -1.0f;
            arg[v] = -1;
        }
        walk(b2, [&](int loc, int cnt) {
            const long long s0 = leaf_rec(loc, cnt);
            for (int c = 0; c < cnt; ++c) {
                const TriRec T = D > 0 ? s_rec[c] : load_uniform(C.rec, (int)(s0 + c));
                ++ntest;
                bool need = false;
#pragma unroll
                for (int v = 0; v < VPL; ++v)
                    need |= tri_lb2(T, px, py, pz[v]) <= b2[v];
                if (!wave_any(need)) continue;
                ++neval;
                bool tighter = false;
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    // equal distances go to the lower slot, so the choice does not
                    // depend on the order the tree is walked in
                    const float d2 = tri_d2(T, px, py, pz[v]);
                    const int s = (int)(s0 + c);
                    if (d2 < best[v] || (d2 == best[v] && s < arg[v])) {
                        best[v] = d2;
                        arg[v] = s;
                    }
                    const float bb = bound_root(best[v]) + eps;
                    const bool lower = bb * bb < b2[v];
                    tighter |= lower;
                    b2[v] = lower ? bb * bb : b2[v];
                }
                if (wave_any(tighter)) ++ver;
            }
        });

        // ---- within one cell of the surface: the reference's FP64 distance over every
        // triangle within the FP32 minimum plus twice the error margin
        bool near[VPL];
        bool any_near = false;
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            near[v] = valid[v] && sqrtf(best[v]) <= h * (1.0f + 1e-3f) + 2.0f * eps;
            any_near |= near[v];
        }
        double best64[VPL];
        int arg64[VPL];
        double p64[VPL][3];
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            best64[v] = DBL_MAX;
            arg64[v] = -1;
            p64[v][0] = g.corner[0] + (i + 0.5) * g.spacing;
            p64[v][1] = g.corner[1] + (j + 0.5) * g.spacing;
            p64[v][2] = g.corner[2] + (k[v] + 0.5) * g.spacing;
        }
        if (wave_any(any_near)) {
            ++ver;  // new bounds: nothing on the stack was tested against them
            float b2n[VPL], thr[VPL];
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                const float bb = bound_root(best[v]) + 2.0f * eps;
                b2n[v] = near[v] ? bb * bb : -1.0f;
                const float tb = bound_root(best[v]) + 3.0f * eps;
                thr[v] = tb * tb;
            }
            walk(b2n, [&](int loc, int cnt) {
                const long long s0 = leaf_rec(loc, cnt);
                for (int c = 0; c < cnt; ++c) {
                    const TriRec T = D > 0 ? s_rec[c] : load_uniform(C.rec, (int)(s0 + c));
                    bool need[VPL];
                    bool anyn = false;
#pragma unroll
                    for (int v = 0; v < VPL; ++v) {
                        need[v] = tri_lb2(T, px, py, pz[v]) <= b2n[v];
                        anyn |= need[v];
                    }
                    if (!wave_any(anyn)) continue;
                    // only a triangle within twice the FP32 margin of the FP32 minimum can
                    // be the FP64 minimum (or tie with it): its FP32 distance decides
                    // whether the FP64 one is needed
                    anyn = false;
#pragma unroll
                    for (int v = 0; v < VPL; ++v) {
                        need[v] = need[v] && tri_d2(T, px, py, pz[v]) <= thr[v];
                        anyn |= need[v];
                    }
                    if (!wave_any(anyn)) continue;
                    const int s = (int)(s0 + c);
                    const TriRec64 T64 = load_uniform(C.rec64, s);
#pragma unroll
                    for (int v = 0; v < VPL; ++v)
                        if (need[v]) {
                            const double d = point_tri_dist2_ref(p64[v], T64.a, T64.b, T64.c);
                            // std::min(best, d) (voxelize.cpp:317); ties between distinct
                            // triangles go to the lower input index
                            if (d < best64[v] ||
                                (d == best64[v] && C.perm[s] < C.perm[arg64[v]])) {
                                best64[v] = d;
                                arg64[v] = s;
                            }
                        }
                }
            });
        }
        // the cell's source triangle and its unsigned distance: the FP32 minimum back in
        // world units (exact scaling), or the reference's FP64 distance rounded once
        // (voxelize.cpp:354-355); sdf_kernel signs them in flat order
        int* const tri_out = C.cell_tri + M.blk_base;
        float* const dist_out = C.blk_dist + M.blk_base;
#pragma unroll
        for (int v = 0; v < VPL; ++v)
            if (valid[v]) {
                const size_t bc = blocked_cell(M, i, j, k[v]);
                tri_out[bc] = near[v] ? -2 - arg64[v] : arg[v];
                dist_out[bc] = near[v] ? (float)sqrt(best64[v]) : sqrtf(best[v]) * M.iscf;
            }
    }
    if (lane == 0) {
        if (nvis) atomicAdd(&C.ctr->node_tests, nvis);
        if (ntest) atomicAdd(&C.ctr->tri_tests, ntest);
        if (neval) atomicAdd(&C.ctr->tri_evals, neval);
    }
}

// ---- signed distance in flat order ---------------------------------------------------------
//
// The voxel pass leaves each cell's unsigned distance in 4x4x4 blocks (blocked like
// cell_tri); the SDF is the flat array idx = (k ny + j) nx + i per mesh, meshes back to
// back, signed by the occupancy (negation is exact, so this equals signing before the
// rounding). One item is a run of up to kRowRun blocks along x at one (block row, block
// slice): kRowRun x 64 consecutive floats of the blocked array, staged in LDS (row
// (j & 3) + 4 (k & 3), column x; one float of padding per row against bank conflicts)
// and written as 16 grid rows of up to 4 kRowRun consecutive floats. Reads and writes
// are whole lines.
__global__ __launch_bounds__(256) void sdf_kernel(Ctx C) {
    constexpr int kW = 4 * kRowRun;  // cells per tile row
    __shared__ __align__(16) float tile[16][kW + 4];
    const long long* off = phase_off(C, kPhRow);
    const long long total = off[C.nm];
    // a contiguous range of items per workgroup: neighbouring rows of one mesh, whose
    // position in processing order is looked up only when the range enters a new mesh
    const long long per = (total + gridDim.x - 1) / gridDim.x;
    const long long first = (long long)blockIdx.x * per;
    const long long last = min(total, first + per);
    int pos = 0;
    long long p_lo = 0, p_hi = 0;
    for (long long item = first; item < last; ++item) {
        if (item >= p_hi) {
            pos = find_pos(off, C.nm, item);
            p_lo = off[pos];
            p_hi = off[pos + 1];
        }
        const MeshDesc& M = C.mesh[C.order[pos]];
        const GridSpec& g = M.g;
        const Level& L = M.L0;
        const int nrun = (L.bx + kRowRun - 1) / kRowRun;
        long long r = item - p_lo;
        const int xr = (int)(r % nrun);
        r /= nrun;
        const int iy = (int)(r % L.by), iz = (int)(r / L.by);
        const int ix0 = xr * kRowRun;
        const int nb = min(kRowRun, L.bx - ix0);  // blocks in this run
        const float4* src = reinterpret_cast<const float4*>(
            C.blk_dist + M.blk_base +
            64 * ((size_t)ix0 + (size_t)L.bx * ((size_t)iy + (size_t)L.by * iz)));
        __syncthreads();  // the previous item's reads of the tile are done
        for (int e = threadIdx.x; e < nb * 16; e += blockDim.x) {
            // cells 4q..4q+3 of block e / 16, q = e % 16: x = 4 (e / 16) + 0..3 of tile row q
            const int q = e & 15;
            *reinterpret_cast<float4*>(&tile[q][4 * (e >> 4)]) = src[e];
        }
        __syncthreads();
        const int i0 = 4 * ix0;
        const int x = threadIdx.x;  // this thread's column of the tile (blockDim.x == kW)
        if (x < min(4 * nb, g.nx - i0)) {
            // the 16 occupancy bytes are loaded before any store: a byte load issued after
            // a store waits for it (one memory counter), and uint8_t may alias the SDF
            size_t idx[16];
            bool in[16];
            uint8_t inside[16];
#pragma unroll
            for (int row = 0; row < 16; ++row) {
                const int j = 4 * iy + (row & 3), k = 4 * iz + (row >> 2);
                in[row] = j < g.ny && k < g.nz;
                idx[row] = (size_t)M.cell_base + ((size_t)k * g.ny + j) * g.nx + i0 + x;
                inside[row] = in[row] ? C.occ[idx[row]] : 0;
            }
#pragma unroll
            for (int row = 0; row < 16; ++row)
                if (in[row]) {
                    const float u = tile[row][x];
                    C.sdf[idx[row]] = inside[row] ? -u : u;
                }
        }
    }
}

// ---- adjoint ------------------------------------------------------------------------------
//
// One 4x4x4 block per wavefront and item. The forward recorded, per cell, the triangle its
// distance comes from; the closest point on it (barycentric weights: Ericson's regions in
// FP32, or the reference's FP64 routine within one cell of the surface) is recomputed here
// and the cell's contribution d(dist)/d(vertex) = -w_m (p - q) / d, times the upstream
// gradient and the sign, is summed over the wavefront per triangle (one group of lanes per
// distinct triangle), then over the workgroup per vertex in an LDS table keyed by point
// (a vertex collects the sums of all the triangles around it: three atomics per point
// where a per-triangle table flushed nine per triangle), and only then added to the
// gradient with global atomics, into the copy of the gradient of
// the XCD the workgroup runs on (xgrad_sum_kernel adds the copies). A workgroup takes
// chunks of neighbouring blocks (neighbouring blocks share their nearest triangles: far
// from a coarse mesh whole regions share one), so the same vertices are not hammered by
// every wavefront. When C.nearest is set, the mesh-local triangle of every cell is written too.
constexpr int kAdjSlots = 1024;   // LDS table entries (points) per workgroup
constexpr int kAdjShift = 22;     // hash bits: the top log2(kAdjSlots)
static_assert((1 << (32 - kAdjShift)) == kAdjSlots, "table size is a power of two");
constexpr int kAdjProbe = 16;     // probes before a contribution goes straight to HBM
constexpr int kAdjGroup = 8;      // lanes sharing a triangle worth a wavefront sum


// Adds the 3 partial derivatives v3 of point `vtx` (global index) to gx, the copy of the
// gradient that belongs to the XCD the wave runs on. Each XCD adds into its own copy, so
// atomics from different XCDs never meet on one address; the atomics are device-scope, so
// the sum is right whichever copy a wave picks.
__device__ __forceinline__ void add_vtx_grad(float* gx, int vtx, const float* v3) {
    float* gp = gx + 3 * (size_t)vtx;
    atomicAdd(gp + 0, v3[0]);
    atomicAdd(gp + 1, v3[1]);
    atomicAdd(gp + 2, v3[2]);
}

// grad = the sum of the per-XCD copies of the gradient, in XCD order (one pass after
// the adjoint).
__global__ __launch_bounds__(256) void xgrad_sum_kernel(const float* xg, long long stride,
                                                       float* g) {
    const long long step = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < stride;
         i += step) {
        float s = 0.0f;
#pragma unroll
        for (int x = 0; x < kXcds; ++x) s += xg[x * stride + i];
        g[i] = s;
    }
}

// Adds v3 into the workgroup table entry of point `vtx` (or, when the table is crowded,
// to the gradient directly); a newly taken entry is listed in used[].
__device__ __forceinline__ void table_add(int vtx, const float* v3, int* key, float* val,
                                          int* used, int* nused, float* gx) {
    const unsigned hsh = (unsigned)vtx * 0x9E3779B1u;
    for (int pr = 0; pr < kAdjProbe; ++pr) {
        const int e = (int)((hsh >> kAdjShift) + (unsigned)pr) & (kAdjSlots - 1);
        const int cur = atomicCAS(&key[e], -1, vtx);
        if (cur == -1) used[atomicAdd(nused, 1)] = e;
        if (cur == -1 || cur == vtx) {
#pragma unroll
            for (int q = 0; q < 3; ++q) atomicAdd(&val[e * 3 + q], v3[q]);
            return;
        }
    }
    add_vtx_grad(gx, vtx, v3);
}

// Workgroups (4 wavefronts) take chunks of `chunk` neighbouring blocks from a work queue
// and walk each chunk four blocks at a time (one per wavefront); the table is added to
// the gradient and cleared whenever it is a quarter full and at the end (an iteration
// adds at most 3 x 256 entries: a table that fills up sends the rest straight to the
// gradient, so the sum is right in any case). Longer chunks keep more of a workgroup's
// cells in one region (fewer flushes of the same points); the host sizes them so that
// every workgroup still gets several.
constexpr int kAdjChunkMin = 16, kAdjChunkMax = 64;
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
void adjoint_kernel(Ctx C, int chunk) {  // register budget: 5 waves per SIMD (96 VGPRs)
    __shared__ int s_key[kAdjSlots];
    __shared__ float s_val[kAdjSlots * 3];
    __shared__ int s_used[kAdjSlots];  // entries taken since the last flush
    __shared__ int s_nused;
    const long long* off = phase_off(C, kPhL0);
    const long long total = off[C.nm];
    const int lane = threadIdx.x & 63;
    const int nwave = blockDim.x / 64;
    const bool scatter = C.grad != nullptr;
    const int x0 = xcd_id();
    float* const gx = scatter ? C.xgrad + (size_t)x0 * C.xstride : nullptr;
    __shared__ long long s_base;
    int yoff = 0;  // (thread 0's) XCD range to take chunks from, relative to x0
    const long long nchunk = (total + chunk - 1) / chunk;
    for (int e = threadIdx.x; e < kAdjSlots; e += blockDim.x) s_key[e] = -1;
    for (int e = threadIdx.x; e < kAdjSlots * 3; e += blockDim.x) s_val[e] = 0.0f;
    if (threadIdx.x == 0) s_nused = 0;
    __syncthreads();
    long long base = 0, last = 0;
    // this wavefront's current mesh: position pos in processing order, items [p_lo, p_hi)
    // (consecutive items mostly stay in one mesh: the search runs when they leave it)
    int pos = 0;
    long long p_lo = 0, p_hi = 0;
    long long pf_item = -1;  // the block whose cell triangles pf_code holds
    int pf_code = 0;
    while (true) {
        if (base >= last) {  // next chunk (workgroup-uniform), from this XCD's range
            __syncthreads();
            if (threadIdx.x == 0) {
                const long long c = xcd_take(C.ctr->aq, x0, &yoff, nchunk);
                s_base = c < 0 ? total : c * chunk;
            }
            __syncthreads();
            base = s_base;
            last = base + chunk < total ? base + chunk : total;
        }
        const bool done = base >= total;
        const long long item = base + (threadIdx.x >> 6);
        base += nwave;
        if (!done && item < last) {
            if (item < p_lo || item >= p_hi) {
                pos = find_pos(off, C.nm, item);
                p_lo = off[pos];
                p_hi = off[pos + 1];
            }
            const MeshDesc& M = C.mesh[C.order[pos]];
            const GridSpec& g = M.g;
            const int b = (int)(item - p_lo);
            const int ix = b % M.L0.bx, iy = (b / M.L0.bx) % M.L0.by,
                      iz = b / (M.L0.bx * M.L0.by);
            const int i = ix * kLeaf + (lane & 3), j = iy * kLeaf + ((lane >> 2) & 3),
                      k = iz * kLeaf + (lane >> 4);
            const bool valid = i < g.nx && j < g.ny && k < g.nz;
            const int code_b = item == pf_item ? pf_code
                                               : C.cell_tri[M.blk_base + (size_t)b * 64 + lane];
            const int code = valid ? code_b : -1;
            // the cell triangles of this wavefront's next block of the chunk (same mesh)
            // load while this block is processed
            pf_item = -1;
            if (item + nwave < last && item + nwave < p_hi) {
                pf_item = item + nwave;
                pf_code = C.cell_tri[M.blk_base + (size_t)(b + nwave) * 64 + lane];
            }
            const bool near = code <= -2;
            const int tri = near ? -2 - code : code;
            const size_t idx = valid ? ((size_t)k * g.ny + j) * g.nx + i : 0;
            if (valid && C.nearest)
                C.nearest[M.cell_base + idx] = tri >= 0 ? (int)(C.perm[tri] - M.tbase) : -1;
            float c[9];
            bool act = false;
            if (valid && tri >= 0 && scatter) {
                const float gv = C.gsdf[M.cell_base + idx];
                const float sg = C.occ[M.cell_base + idx] ? -1.0f : 1.0f;
                if (near) {  // FP64 closest point near the surface
                    const double p64[3] = {g.corner[0] + (i + 0.5) * g.spacing,
                                           g.corner[1] + (j + 0.5) * g.spacing,
                                           g.corner[2] + (k + 0.5) * g.spacing};
                    const TriRec64& T64 = C.rec64[tri];
                    double w[3];
                    const double d2 = point_tri_dist2_ref(p64, T64.a, T64.b, T64.c, w);
                    double e[3];
                    for (int a = 0; a < 3; ++a)
                        e[a] = p64[a] - (w[0] * T64.a[a] + w[1] * T64.b[a] + w[2] * T64.c[a]);
                    const double d = sqrt(d2);
                    if (d > M.dzero && gv != 0.0f) {
                        const double coef = -(double)gv * sg / d;
#pragma unroll
                        for (int m = 0; m < 3; ++m)
#pragma unroll
                            for (int a = 0; a < 3; ++a)
                                c[3 * m + a] = (float)(coef * w[m] * e[a]);
                        act = true;
                    }
                } else {
                    const float h = (float)g.spacing * M.scf;  // cell size, search units
                    const float px = (i + 0.5f) * h, py = (j + 0.5f) * h, pz = (k + 0.5f) * h;
                    const TriRec T = C.rec[tri];
                    float vv, ww;
                    tri_bary(T, px, py, pz, vv, ww);
                    const float u = 1.0f - vv - ww;
                    const float qx = T.ax + vv * T.abx + ww * T.acx;
                    const float qy = T.ay + vv * T.aby + ww * T.acy;
                    const float qz = T.az + vv * T.abz + ww * T.acz;
                    const float ex = px - qx, ey = py - qy, ez = pz - qz;
                    const float d = sqrtf(ex * ex + ey * ey + ez * ez);
                    if (d > 0.0f && gv != 0.0f) {
                        const float coef = -gv * sg / d;
                        c[0] = coef * u * ex; c[1] = coef * u * ey; c[2] = coef * u * ez;
                        c[3] = coef * vv * ex; c[4] = coef * vv * ey; c[5] = coef * vv * ez;
                        c[6] = coef * ww * ex; c[7] = coef * ww * ey; c[8] = coef * ww * ez;
                        act = true;
                    }
                }
            }
            // large groups of lanes on one triangle (far from the surface) are summed
            // across the wavefront first; the rest go into the table lane by lane (every
            // triangle of the wavefront is looked at once: a small group does not hide
            // a large one behind it)
            unsigned long long pending = __ballot(act);
            while (pending != 0ull) {
                const int leader = __ffsll((long long)pending) - 1;
                const int lt = __builtin_amdgcn_readlane(tri, leader);
                const bool match = act && tri == lt;
                const unsigned long long mm = __ballot(match);
                pending &= ~mm;
                if (__popcll(mm) < kAdjGroup) continue;
                float sum[9];
#pragma unroll
                for (int e = 0; e < 9; ++e) sum[e] = wave_sum_dpp(match ? c[e] : 0.0f);
                // lanes 0, 1, 2 add the sums of the triangle's three vertices
                const int4 tv = C.svtx[lt];
                if (lane < 3) {
                    const bool l0 = lane == 0, l1 = lane == 1;
                    const float v3[3] = {l0 ? sum[0] : (l1 ? sum[3] : sum[6]),
                                         l0 ? sum[1] : (l1 ? sum[4] : sum[7]),
                                         l0 ? sum[2] : (l1 ? sum[5] : sum[8])};
                    table_add(l0 ? tv.x : (l1 ? tv.y : tv.z), v3, s_key, s_val, s_used,
                              &s_nused, gx);
                }
                act = act && !match;
            }
            if (act) {
                const int4 tv = C.svtx[tri];
                table_add(tv.x, c, s_key, s_val, s_used, &s_nused, gx);
                table_add(tv.y, c + 3, s_key, s_val, s_used, &s_nused, gx);
                table_add(tv.z, c + 6, s_key, s_val, s_used, &s_nused, gx);
            }
        }
        if (!scatter) {
            if (done) break;
            continue;
        }
        __syncthreads();
        const int nu = s_nused;
        // every thread has read the count before any wavefront can add to it again: the
        // flush decision (and the barriers inside it) stays uniform
        __syncthreads();
        if (nu <= kAdjSlots / 4 && !done) continue;
        for (int u = threadIdx.x; u < nu; u += blockDim.x) {
            const int e = s_used[u];
            const int t = s_key[e];
            add_vtx_grad(gx, t, &s_val[e * 3]);
            s_key[e] = -1;
#pragma unroll
            for (int q = 0; q < 3; ++q) s_val[e * 3 + q] = 0.0f;
        }
        __syncthreads();
        if (threadIdx.x == 0) s_nused = 0;
        __syncthreads();
        if (done) break;
    }
}

__global__ __launch_bounds__(256) void bbox_kernel(const double* xyz, const long long* vstart,
                                                   int nm, double* lohi) {
    // one workgroup per mesh: min and max of every coordinate (exact, order-free)
    const int m = blockIdx.x;
    if (m >= nm) return;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (long long p = vstart[m] + threadIdx.x; p < vstart[m + 1]; p += blockDim.x)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin(lo[a], xyz[3 * p + a]);
            hi[a] = fmax(hi[a], xyz[3 * p + a]);
        }
    __shared__ double s[6][256];
    for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = lo[a];
        s[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fmin(s[a][threadIdx.x], s[a][threadIdx.x + o]);
                s[3 + a][threadIdx.x] = fmax(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) lohi[6 * m + threadIdx.x] = s[threadIdx.x][0];
}

}  // namespace

// ---- host side ---------------------------------------------------------------------------

namespace {

// Stable LSD radix sort of (code, index) pairs by the 30-bit code, 3 passes of 10 bits.
// Search order of one mesh's triangles idx[0, n) (global triangle ids, centroids cen):
// the implicit tree makes every run of 8^l consecutive slots a node, so the runs are made
// spatially compact: `parts` groups of `child` slots each, formed by binary splits at the
// group boundary nearest the middle, each along the longest extent of the range's
// centroids (a median split by count, ties by id); each group is then split the same way
// into 8 groups of child / 8 slots, down to single triangles.
static void split_order(int* idx, size_t n, size_t child, int parts, const float4* cen) {
    if (n <= 1) return;
    if (parts == 1) {
        if (child > 1) split_order(idx, n, child / 8, 8, cen);
        return;
    }
    const int lp = parts / 2;
    const size_t ln = std::min(n, (size_t)lp * child);
    if (ln < n) {
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (size_t k = 0; k < n; ++k) {
            const float4 c = cen[idx[k]];
            const float v[3] = {c.x, c.y, c.z};
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], v[a]);
                hi[a] = std::max(hi[a], v[a]);
            }
        }
        int ax = 0;
        for (int a = 1; a < 3; ++a)
            if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
        auto key = [&](int t) {
            const float4 c = cen[t];
            return ax == 0 ? c.x : (ax == 1 ? c.y : c.z);
        };
        std::nth_element(idx, idx + ln, idx + n, [&](int x, int y) {
            const float kx = key(x), ky = key(y);
            return kx != ky ? kx < ky : x < y;
        });
    }
    split_order(idx, ln, child, lp, cen);
    split_order(idx + ln, n - ln, child, parts - lp, cen);
}

}  // namespace

struct BatchEngine::Impl {
    hipStream_t stream = nullptr;
    int nm = 0;
    std::vector<MeshDesc> h_mesh;
    std::vector<int> h_order;
    std::vector<long long> h_item_off;
    std::vector<int> h_lvl_off;  // [kMaxLevels + 1][nm + 1]
    std::vector<long long> h_tri_start, h_pt_start, h_cell_start;
    long long total_tris = 0, total_pts = 0, total_cells = 0, total_tiles = 0;
    long long total_cols_words = 0;
    int max_lev = 1;
    // device buffers (owned)
    std::vector<void*> owned;
    Ctx ctx{};
    const double* xyz = nullptr;
    const int* tris = nullptr;
    // graph state
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture = nullptr;  // engine-owned stream the graph is recorded on
    int graph_flags = -1;
    const void* graph_key[4] = {nullptr, nullptr, nullptr, nullptr};
    unsigned grid64 = 0, grid256 = 0;

    template <class T>
    T* dalloc(size_t n) {
        void* p = nullptr;
        CMG_HIP(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 256)));
        owned.push_back(p);
        return static_cast<T*>(p);
    }

    // Frees a dalloc buffer before the engine ends (the caller has synchronised its uses).
    void dfree(void* p) {
        auto it = std::find(owned.begin(), owned.end(), p);
        if (it == owned.end()) return;
        owned.erase(it);
        CMG_HIP(hipFree(p));
    }

    ~Impl() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamSynchronize(stream);
        if (capture) (void)hipStreamDestroy(capture);
        for (void* p : owned) (void)hipFree(p);
    }

    int vpl = 1, depth = 0;  // voxel-pass variant (voxels per lane, LDS tile depth)
    bool have_forward = false;  // a forward was issued: occupancy and cell triangles exist

    void launch_forward(hipStream_t st);
    void launch_backward(hipStream_t st);
    void launch_tree(hipStream_t st);
    void launch_occupancy(hipStream_t st);
    void launch_occupancy_phases(hipStream_t st);

    // Persistent voxel passes get as many one-wave workgroups as fit on the device at
    // once (the tree walk is latency bound: every resident wave hides some of it).
    template <class K>
    unsigned resident_grid(K kernel) {
        int per_cu = 0;
        CMG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0));
        return (unsigned)cu_count() * (unsigned)std::max(per_cu, 1);
    }

    void launch_vox(hipStream_t st) {
#define CMG_VOX(V, DD)                                                 \
    if (vpl == V && depth == DD) {                                    \
        static unsigned grid = 0;                                      \
        if (!grid) grid = resident_grid(vox_kernel<V, DD>);            \
        vox_kernel<V, DD><<<grid, 64, 0, st>>>(ctx);                   \
        CMG_HIP_LAUNCH();                                              \
        return;                                                        \
    }
        CMG_VOX(1, 0) CMG_VOX(1, 8) CMG_VOX(2, 0) CMG_VOX(2, 8) CMG_VOX(4, 0) CMG_VOX(4, 8)
#undef CMG_VOX
        throw HipError("batch: unknown voxel-pass variant");
    }

    // Persistent workgroups, as many as fit on the device at once.
    void launch_adjoint(hipStream_t st) {
        static unsigned grid = 0;
        if (!grid) {
            int per_cu = 0;
            CMG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adjoint_kernel, 256, 0));
            grid = (unsigned)cu_count() * (unsigned)std::max(per_cu, 1);
        }
        reset_queue_kernel<<<1, 64, 0, st>>>(ctx.ctr, 2);
        CMG_HIP_LAUNCH();
        if (ctx.grad) {
            zero_grad_kernel<<<grid256, 256, 0, st>>>(ctx.xgrad, kXcds * ctx.xstride);
            CMG_HIP_LAUNCH();
        }
        // about four chunks per workgroup, within [kAdjChunkMin, kAdjChunkMax] blocks
        const long long items = h_item_off[(size_t)kPhL0 * (nm + 1) + nm];
        const int chunk = (int)std::min<long long>(
            kAdjChunkMax, std::max<long long>(kAdjChunkMin, items / (4LL * grid)));
        adjoint_kernel<<<grid, 256, 0, st>>>(ctx, chunk);
        CMG_HIP_LAUNCH();
        if (ctx.grad) {
            xgrad_sum_kernel<<<grid256, 256, 0, st>>>(ctx.xgrad, ctx.xstride, ctx.grad);
            CMG_HIP_LAUNCH();
        }
    }
};

void BatchEngine::Impl::launch_tree(hipStream_t st) {
    const int n1 = h_lvl_off[(size_t)1 * (nm + 1) + nm];
    node_leaf_kernel<<<grid256, 256, 0, st>>>(ctx, n1);
    CMG_HIP_LAUNCH();
    for (int l = 2; l <= max_lev; ++l) {
        const int nl = h_lvl_off[(size_t)l * (nm + 1) + nm];
        const unsigned g = (unsigned)std::min<long long>(grid256, (nl + 255) / 256);
        node_merge_kernel<<<g ? g : 1, 256, 0, st>>>(ctx, l, nl);
        CMG_HIP_LAUNCH();
    }
}

void BatchEngine::Impl::launch_occupancy(hipStream_t st) {
    static unsigned g_fast = 0, g_exact = 0;
    if (!g_fast) g_fast = resident_grid(occ_pairs_kernel);
    if (!g_exact) g_exact = resident_grid(occ_kernel<true>);
    TimedLaunch timed("batch_occupancy", st);
    occ_pairs_kernel<<<g_fast, 64, 0, st>>>(ctx);
    CMG_HIP_LAUNCH();
    occ_kernel<true><<<g_exact, 64, 0, st>>>(ctx);
    CMG_HIP_LAUNCH();
    occ_rows_kernel<<<grid256 * 4, 256, 0, st>>>(ctx);
    CMG_HIP_LAUNCH();
}

void BatchEngine::Impl::launch_occupancy_phases(hipStream_t st) {
    Ctx& C = ctx;
    reset_kernel<<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    prep_kernel<<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    tile_alloc_kernel<<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    tile_fill_kernel<<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    launch_occupancy(st);
}

void BatchEngine::Impl::launch_forward(hipStream_t st) {
    have_forward = true;
    launch_occupancy_phases(st);
    launch_tree(st);
    {
        TimedLaunch timed("batch_sdf", st);
        launch_vox(st);
        sdf_kernel<<<grid256 * 4, 256, 0, st>>>(ctx);
        CMG_HIP_LAUNCH();
    }
}

void BatchEngine::Impl::launch_backward(hipStream_t st) {
    {
        TimedLaunch timed("batch_sdf_backward", st);
        launch_adjoint(st);
    }
}

BatchEngine::BatchEngine(int nm, const long long* pt_start, const long long* tri_start,
                         const int* resolution, const int* pad, const double* d_xyz,
                         const int* d_tris, hipStream_t st)
    : p_(new Impl) {
    Impl& I = *p_;
    if (nm < 1) throw InvalidInput("batch: no meshes");
    I.stream = st;
    I.nm = nm;
    I.xyz = d_xyz;
    I.tris = d_tris;
    I.h_pt_start.assign(pt_start, pt_start + nm + 1);
    I.h_tri_start.assign(tri_start, tri_start + nm + 1);
    I.total_pts = pt_start[nm];
    I.total_tris = tri_start[nm];
    if (I.total_tris >= (1LL << 31)) throw InvalidInput("batch: more than 2^31 triangles");
    if (I.total_pts >= (1LL << 31)) throw InvalidInput("batch: more than 2^31 points");
    for (int m = 0; m < nm; ++m) {
        if (pt_start[m + 1] <= pt_start[m]) throw InvalidInput("batch: a mesh has no points");
        if (tri_start[m + 1] <= tri_start[m])
            throw InvalidInput("batch: a mesh has no triangles");
        // (at most 10 tree levels: the traversal stack fits one register)
        if (tri_start[m + 1] - tri_start[m] > (1LL << 30))
            throw InvalidInput("batch: a mesh has more than 2^30 triangles");
        if (resolution[m] < 1) throw InvalidInput("batch: resolution must be >= 1");
    }
    // grids from device bounding boxes (min/max are exact, so make_grid matches the
    // reference's host computation)
    long long* d_vs = I.dalloc<long long>(nm + 1);
    double* d_lohi = I.dalloc<double>(6 * nm);
    CMG_HIP(hipMemcpyAsync(d_vs, pt_start, sizeof(long long) * (nm + 1), hipMemcpyHostToDevice, st));
    bbox_kernel<<<nm, 256, 0, st>>>(d_xyz, d_vs, nm, d_lohi);
    CMG_HIP_LAUNCH();
    std::vector<double> lohi(6 * nm);
    CMG_HIP(hipMemcpyAsync(lohi.data(), d_lohi, sizeof(double) * 6 * nm, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    I.h_mesh.resize(nm);
    long long cells = 0, tiles = 0, words = 0, blocked = 0;
    std::vector<double> cost(nm);
    std::vector<long long> nodes_per_level(kMaxLevels + 1, 0);
    I.h_lvl_off.assign((size_t)(kMaxLevels + 1) * (nm + 1), 0);
    for (int m = 0; m < nm; ++m) {
        double two[6] = {lohi[6 * m], lohi[6 * m + 1], lohi[6 * m + 2],
                         lohi[6 * m + 3], lohi[6 * m + 4], lohi[6 * m + 5]};
        // make_grid over the two corners reproduces make_grid over all points
        GridSpec g;
        if (!make_grid(two, 2, resolution[m], pad[m], g))
            throw InvalidInput("batch: degenerate (zero-extent) mesh");
        if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0)
            throw InvalidInput("batch: grid has no cells (negative pad)");
        MeshDesc& M = I.h_mesh[m];
        M = MeshDesc{};
        M.g = g;
        M.vbase = pt_start[m];
        M.tbase = tri_start[m];
        M.ntris = (int)(tri_start[m + 1] - tri_start[m]);
        M.cell_base = cells;
        M.blk_base = blocked;
        OccPlan op = occupancy_plan(g);
        M.words = op.words;
        M.ntx = op.ntx;
        M.nty = op.nty;
        M.tile_base = (int)tiles;
        M.mask_base = words;
        M.L0 = make_level(g, kLeaf);
        M.L0b = make_level(g, kLeaf, 2 * kLeaf);
        M.L0c = make_level(g, kLeaf, 4 * kLeaf);
        blocked += 64LL * M.L0.bx * M.L0.by * M.L0.bz;
        M.sc = search_scale(g);
        M.scf = (float)M.sc;
        M.iscf = (float)(1.0 / M.sc);
        M.eps = distance_eps(g) * M.scf;
        M.dzero = zero_distance(g);
        // tree levels: level l has ceil(ntris / 8^l) nodes, up to the single root
        M.nnodes[0] = M.ntris;
        int l = 0;
        long long n = M.ntris;
        do {
            ++l;
            n = (n + 7) / 8;
            M.nnodes[l] = (int)n;
        } while (n > 1);
        M.nlev = l;
        I.max_lev = std::max(I.max_lev, l);
        const long long nc = (long long)g.nx * g.ny * g.nz;
        cells += nc;
        tiles += (long long)op.ntx * op.nty;
        words += (long long)g.nx * g.ny * op.words;
        cost[m] = (double)nc * std::log2(2.0 + M.ntris);
    }
    // node numbering: level by level, meshes in id order; lo[] holds each level's
    // per-mesh prefix counted from that level's first node
    long long node_total = 0;
    for (int l = 1; l <= kMaxLevels; ++l) {
        int* lo = &I.h_lvl_off[(size_t)l * (nm + 1)];
        long long rel = 0;
        for (int m = 0; m < nm; ++m) {
            MeshDesc& M = I.h_mesh[m];
            lo[m] = (int)rel;
            M.node_base[l] = (int)(node_total + rel);
            rel += l <= M.nlev ? M.nnodes[l] : 0;
        }
        lo[nm] = (int)rel;
        node_total += rel;
    }
    if (node_total >= kMaxNodes)
        throw InvalidInput("batch: more than 2^27 tree nodes (about 10^9 triangles) in one batch");
    I.total_cells = cells;
    I.total_tiles = tiles;
    I.total_cols_words = words;
    // processing order: most expensive meshes first
    I.h_order.resize(nm);
    std::iota(I.h_order.begin(), I.h_order.end(), 0);
    std::stable_sort(I.h_order.begin(), I.h_order.end(),
                     [&](int a, int b) { return cost[a] > cost[b]; });
    I.h_item_off.assign((size_t)kNumPhase * (nm + 1), 0);
    for (int ph = 0; ph < kNumPhase; ++ph) {
        long long* o = &I.h_item_off[(size_t)ph * (nm + 1)];
        for (int p = 0; p < nm; ++p) {
            const MeshDesc& M = I.h_mesh[I.h_order[p]];
            long long n = 0;
            if (ph == kPhTile) n = (long long)M.ntx * M.nty;
            if (ph == kPhL0) n = (long long)M.L0.bx * M.L0.by * M.L0.bz;
            if (ph == kPhL0b) n = (long long)M.L0b.bx * M.L0b.by * M.L0b.bz;
            if (ph == kPhL0c) n = (long long)M.L0c.bx * M.L0c.by * M.L0c.bz;
            if (ph == kPhV1 || ph == kPhV2 || ph == kPhV4) {
                const int vpl = ph == kPhV1 ? 1 : (ph == kPhV2 ? 2 : 4);
                const Level& L = vpl == 1 ? M.L0 : (vpl == 2 ? M.L0b : M.L0c);
                const int sz = kSuperZ / vpl;
                n = (long long)((L.bx + kSuper - 1) / kSuper) * ((L.by + kSuper - 1) / kSuper) *
                    ((L.bz + sz - 1) / sz) * (kSuper * kSuper * sz);
            }
            if (ph == kPhRow)
                n = (long long)((M.L0.bx + kRowRun - 1) / kRowRun) * M.L0.by * M.L0.bz;
            if (ph == kPhOccRow) n = (long long)M.g.ny * M.words * ((M.g.nx + 63) / 64);
            o[p + 1] = o[p] + n;
        }
    }
    // static device buffers
    Ctx& C = I.ctx;
    C.nm = nm;
    MeshDesc* d_mesh = I.dalloc<MeshDesc>(nm);
    int* d_order = I.dalloc<int>(nm);
    long long* d_item = I.dalloc<long long>(I.h_item_off.size());
    long long* d_ts = I.dalloc<long long>(nm + 1);
    int* d_lvl = I.dalloc<int>(I.h_lvl_off.size());
    CMG_HIP(hipMemcpyAsync(d_mesh, I.h_mesh.data(), sizeof(MeshDesc) * nm, hipMemcpyHostToDevice, st));
    CMG_HIP(hipMemcpyAsync(d_order, I.h_order.data(), sizeof(int) * nm, hipMemcpyHostToDevice, st));
    CMG_HIP(hipMemcpyAsync(d_item, I.h_item_off.data(), sizeof(long long) * I.h_item_off.size(),
                           hipMemcpyHostToDevice, st));
    CMG_HIP(hipMemcpyAsync(d_ts, tri_start, sizeof(long long) * (nm + 1), hipMemcpyHostToDevice, st));
    CMG_HIP(hipMemcpyAsync(d_lvl, I.h_lvl_off.data(), sizeof(int) * I.h_lvl_off.size(),
                           hipMemcpyHostToDevice, st));
    C.mesh = d_mesh;
    C.order = d_order;
    C.item_off = d_item;
    C.tri_start = d_ts;
    C.lvl_off = d_lvl;
    C.total_tris = I.total_tris;
    C.xyz = d_xyz;
    C.tris = d_tris;
    C.ctr = I.dalloc<BatchCounters>(1);
    C.rec = I.dalloc<TriRec>(I.total_tris);
    C.rec64 = I.dalloc<TriRec64>(I.total_tris);
    C.node = I.dalloc<NodeRec>(node_total);
    C.nctr = I.dalloc<float4>(node_total);
    C.xyr = I.dalloc<int4>(I.total_tris);
    C.tile_count = I.dalloc<int>(tiles);
    C.tile_off = I.dalloc<int>(tiles);
    C.tile_cursor = I.dalloc<int>(tiles);
    C.tile_left = I.dalloc<int>(tiles);
    C.tile_flag = I.dalloc<int>(tiles);
    C.masks = I.dalloc<unsigned long long>(words);
    C.total_words = words;
    C.redo = I.dalloc<int>(tiles);
    C.occ = I.dalloc<uint8_t>(cells);
    C.cell_tri = I.dalloc<int>(blocked);
    C.blk_dist = I.dalloc<float>(blocked);
    C.xstride = 3 * I.total_pts;
    C.xgrad = I.dalloc<float>((size_t)kXcds * C.xstride);
    I.h_cell_start.resize(nm + 1);
    for (int m = 0; m < nm; ++m) I.h_cell_start[m] = I.h_mesh[m].cell_base;
    I.h_cell_start[nm] = cells;
    {
        long long* d_cs = I.dalloc<long long>(nm + 1);
        CMG_HIP(hipMemcpyAsync(d_cs, I.h_cell_start.data(), sizeof(long long) * (nm + 1),
                               hipMemcpyHostToDevice, st));
        C.cell_start = d_cs;
    }
    C.total_tiles = (int)tiles;
    I.grid64 = (unsigned)cu_count() * 16;
    I.grid256 = (unsigned)cu_count() * 4;
    // search order of every mesh's triangles (once, from the creation-time positions),
    // meshes on host threads
    {
        float4* d_cen = I.dalloc<float4>(I.total_tris);
        centroid_kernel<<<I.grid256, 256, 0, st>>>(C, d_cen);
        CMG_HIP_LAUNCH();
        std::vector<float4> cen(I.total_tris);
        CMG_HIP(hipMemcpyAsync(cen.data(), d_cen, sizeof(float4) * I.total_tris,
                               hipMemcpyDeviceToHost, st));
        CMG_HIP(hipStreamSynchronize(st));
        std::vector<int> perm(I.total_tris);
        std::iota(perm.begin(), perm.end(), 0);
        std::atomic<int> next{0};
        auto work = [&] {
            for (int m; (m = next.fetch_add(1)) < nm;) {
                const long long t0 = tri_start[m], t1 = tri_start[m + 1];
                size_t cap = 1;
                while (cap < (size_t)(t1 - t0)) cap *= 8;
                split_order(perm.data() + t0, (size_t)(t1 - t0), cap / 8, 8, cen.data());
            }
        };
        const int nthr = (int)std::min<unsigned>(
            (unsigned)nm, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        std::vector<std::thread> pool;
        for (int k = 1; k < nthr; ++k) pool.emplace_back(work);
        work();
        for (auto& th : pool) th.join();
        I.dfree(d_cen);
        int* d_perm = I.dalloc<int>(I.total_tris);
        CMG_HIP(hipMemcpyAsync(d_perm, perm.data(), sizeof(int) * I.total_tris,
                               hipMemcpyHostToDevice, st));
        C.perm = d_perm;
        C.svtx = I.dalloc<int4>(I.total_tris);
        slot_vtx_kernel<<<I.grid256, 256, 0, st>>>(C);
        CMG_HIP_LAUNCH();
        CMG_HIP(hipStreamSynchronize(st));
    }
    // planning pass: the tile-list sizes
    C.tile_cap = 0;
    C.tile_pool = nullptr;
    C.seg = nullptr;
    C.seg_cap = 0;
    C.sdf = nullptr;
    reset_kernel<<<I.grid256, 256, 0, st>>>(C);
    prep_kernel<<<I.grid256, 256, 0, st>>>(C);
    tile_alloc_kernel<<<I.grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    BatchCounters h{};
    CMG_HIP(hipMemcpyAsync(&h, C.ctr, sizeof h, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    C.tile_cap = (long long)(h.tile_used + h.tile_used / 4 + 4096);
    C.tile_pool = I.dalloc<int>(C.tile_cap);
    // a tile of c <= pool entries has ceil(c / kSegTris) segments
    C.seg_cap = (long long)tiles + C.tile_cap / kSegTris + 1;
    C.seg = I.dalloc<int2>(C.seg_cap);
}

BatchEngine::~BatchEngine() = default;

long long BatchEngine::total_cells() const { return p_->total_cells; }
long long BatchEngine::total_points() const { return p_->total_pts; }
int BatchEngine::meshes() const { return p_->nm; }
const MeshDesc& BatchEngine::mesh(int m) const { return p_->h_mesh.at(m); }

void BatchEngine::run(int flags, const double* d_xyz, float* d_sdf, const float* d_gsdf,
                      float* d_grad, hipStream_t st) {
    Impl& I = *p_;
    if (d_xyz != I.xyz) throw InvalidInput("batch: the points must stay at the address used at creation");
    const bool fwd = flags & kBatchForward, bwd = flags & kBatchBackward;
    const bool occ_only = (flags & kBatchOccupancy) && !fwd;
    if (bwd && !fwd && (occ_only || !I.have_forward))
        throw InvalidInput("batch: backward needs a forward");
    if (fwd && !d_sdf) throw InvalidInput("batch: forward needs an sdf buffer");
    if (bwd && (!d_gsdf || !d_grad)) throw InvalidInput("batch: backward needs grad buffers");
    // an occupancy-only pass rewrites the occupancy from the current points but leaves
    // the last forward's per-cell triangles as they were: a backward needs a new forward
    // after it
    if (fwd) I.have_forward = true;
    else if (occ_only) I.have_forward = false;
    I.ctx.sdf = d_sdf;
    I.ctx.gsdf = d_gsdf;
    I.ctx.grad = d_grad;
    I.ctx.nearest = nullptr;
    if (!(flags & kBatchGraph)) {
        if (occ_only) I.launch_occupancy_phases(st);
        if (fwd) I.launch_forward(st);
        if (bwd) I.launch_backward(st);
        return;
    }
    // The graph is recorded on an engine-owned non-blocking stream (the caller's stream
    // may be the legacy default stream, which cannot be captured) and launched on st;
    // it depends only on the buffer addresses and the pass selection.
    const void* key[4] = {d_xyz, d_sdf, d_gsdf, d_grad};
    const int f = flags & (kBatchForward | kBatchBackward | kBatchOccupancy);
    if (!I.exec || I.graph_flags != f || std::memcmp(key, I.graph_key, sizeof key) != 0) {
        if (I.exec) {
            CMG_HIP(hipGraphExecDestroy(I.exec));
            I.exec = nullptr;
        }
        if (I.graph) {
            CMG_HIP(hipGraphDestroy(I.graph));
            I.graph = nullptr;
        }
        if (!I.capture) CMG_HIP(hipStreamCreateWithFlags(&I.capture, hipStreamNonBlocking));
        hipStream_t cs = I.capture;
        CMG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
        try {
            if (occ_only) I.launch_occupancy_phases(cs);
            if (fwd) I.launch_forward(cs);
            if (bwd) I.launch_backward(cs);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(cs, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        CMG_HIP(hipStreamEndCapture(cs, &I.graph));
        CMG_HIP(hipGraphInstantiate(&I.exec, I.graph, nullptr, nullptr, 0));
        I.graph_flags = f;
        std::memcpy(I.graph_key, key, sizeof key);
    }
    CMG_HIP(hipGraphLaunch(I.exec, st));
}

void BatchEngine::nearest(int* d_tri, hipStream_t st) {
    Impl& I = *p_;
    if (!I.have_forward) throw InvalidInput("batch: nearest needs a forward");
    Ctx saved = I.ctx;
    I.ctx.gsdf = nullptr;
    I.ctx.grad = nullptr;
    I.ctx.nearest = d_tri;
    I.launch_adjoint(st);
    I.ctx = saved;
}

std::vector<BatchVariant> BatchEngine::autotune(const float* d_gsdf, float* d_sdf,
                                                float* d_grad, hipStream_t st) {
    Impl& I = *p_;
    I.ctx.sdf = d_sdf;
    I.ctx.gsdf = d_gsdf;
    I.ctx.grad = d_grad;
    I.ctx.nearest = nullptr;
    I.launch_forward(st);  // occupancy and tree for this geometry
    std::vector<BatchVariant> out;
    hipEvent_t e0, e1;
    CMG_HIP(hipEventCreate(&e0));
    CMG_HIP(hipEventCreate(&e1));
    float best_ms = 1e30f;
    int best_v = I.vpl, best_d = I.depth;
    for (int v : {1, 2, 4})
        for (int d : {0, 8}) {
            I.vpl = v;
            I.depth = d;
            float total = 0;
            for (int rep = 0; rep < 3; ++rep) {
                reset_queue_kernel<<<1, 64, 0, st>>>(I.ctx.ctr, 1);
                zero_grad_kernel<<<I.grid256, 256, 0, st>>>(I.ctx.grad, 3 * I.total_pts);
                CMG_HIP(hipEventRecord(e0, st));
                I.launch_vox(st);
                I.launch_adjoint(st);
                CMG_HIP(hipEventRecord(e1, st));
                CMG_HIP(hipEventSynchronize(e1));
                float ms = 0;
                CMG_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) total += ms;  // first run warms caches
            }
            const float ms = total / 2;
            out.push_back(BatchVariant{v, d, ms});
            if (ms < best_ms) {
                best_ms = ms;
                best_v = v;
                best_d = d;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_variant(best_v, best_d);
    return out;
}

void BatchEngine::set_variant(int vpl, int depth) {
    if (!((vpl == 1 || vpl == 2 || vpl == 4) && (depth == 0 || depth == 8)))
        throw InvalidInput("batch: variant must have vpl in {1,2,4} and depth in {0,8}");
    Impl& I = *p_;
    if (I.vpl != vpl || I.depth != depth) {
        I.vpl = vpl;
        I.depth = depth;
        I.graph_flags = -1;  // recapture with the new kernels
    }
}

BatchVariant BatchEngine::variant() const { return BatchVariant{p_->vpl, p_->depth, 0.0f}; }

BatchCounters BatchEngine::counters(hipStream_t st) {
    BatchCounters h{};
    CMG_HIP(hipMemcpyAsync(&h, p_->ctx.ctr, sizeof h, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    return h;
}

void BatchEngine::status(hipStream_t st, unsigned* overflow, unsigned long long* evals) {
    BatchCounters h = counters(st);
    if (overflow) *overflow = h.overflow;
    if (evals) *evals = h.tri_evals;
}

const uint8_t* BatchEngine::occupancy() const { return p_->ctx.occ; }

}  // namespace cmg::hip
