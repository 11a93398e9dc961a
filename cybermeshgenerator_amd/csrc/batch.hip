// batch.hip — voxelization engine for many meshes of uneven size in one pass.
//
// All meshes live in one set of device arrays (points, triangles, cells concatenated).
// A step is a fixed sequence of persistent kernels; each kernel's workgroups pull work
// items — (mesh, tile) or (mesh, brick) — from a device work queue (one atomic counter
// per kernel), with the items of the most expensive meshes first, so that small meshes
// fill the gaps left by large ones instead of each mesh occupying the GPU alone.
// Brick and tile lists are allocated from device pools by atomic bump allocation, so a
// step needs no host round trip; the sequence is captured once into a hipGraph and
// replayed. Pool capacities are sized by a planning pass at creation; a step that would
// exceed them sets an overflow flag instead of writing out of bounds.
//
// Phases of a forward step (all meshes at once):
//   reset        zero the queues, pool cursors and tile counters (a fill kernel: no
//                memset nodes in the graph)
//   prep         per triangle: FP32/FP64 records, xy bucket range, tile counts
//   tile alloc   per tile: list offset from the tile pool
//   tile fill    per triangle: append to the lists of the tiles its range touches
//   occupancy    per (mesh, 8x8 column tile): exact column parity, as voxelize.hip
//   cull         per (mesh, 16^3 brick) from all the mesh's triangles, then per
//                (mesh, 8^3 brick) from the parent list: min, count+alloc, write
//   voxels       per (mesh, 4^3 brick): FP32 distance, FP64 near the surface, signed
// A backward step reuses the forward's lists and occupancy and runs the adjoint voxel
// pass (closest features recomputed) after zeroing the gradient.
//
// Compiled with -ffp-contract=off like the single-mesh kernels it shares code with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "batch_internal.hpp"
#include "device_common.hpp"
#include "occ_device.hpp"
#include "sdf_device.hpp"

namespace cmg::hip {

using namespace sdfdev;
using namespace occdev;

namespace {

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

// Next item of queue q for this workgroup (thread 0 fetches, all threads read).
__device__ __forceinline__ long long next_item(unsigned long long* q) {
    __shared__ long long s_item;
    __syncthreads();
    if (threadIdx.x == 0) s_item = (long long)atomicAdd(q, 1ull);
    __syncthreads();
    return s_item;
}

struct Ctx {  // kernel arguments shared by every phase
    const MeshDesc* mesh;
    int nm;
    const int* order;           // processing order -> mesh id
    const long long* item_off;  // [kNumPhase][nm + 1] prefix of items in processing order
    const long long* tri_start; // [nm + 1] first global triangle per mesh id
    long long total_tris;
    const double* xyz;
    const int* tris;            // local point indices of each mesh
    BatchCounters* ctr;
    TriRec* rec;
    TriRec64* rec64;
    int4* xyr;
    int* tile_count;
    int* tile_off;
    int* tile_cursor;
    int* tile_pool;
    long long tile_cap;
    unsigned long long* masks;
    uint8_t* occ;
    float* dmin2;
    int* cnt2;
    long long* off2;
    int* pool2;
    long long cap2;
    float* dmin1;
    int* cnt1;
    long long* off1;
    int* pool1;
    long long cap1;
    float* sdf;
    const float* gsdf;
    float* grad;
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

// ---- phases ----------------------------------------------------------------------------

__global__ __launch_bounds__(256) void reset_kernel(Ctx C, int zero_grad, long long npts3) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    if (tid == 0) *C.ctr = BatchCounters{};
    for (long long i = tid; i < C.total_tiles; i += stride) {
        C.tile_count[i] = 0;
        C.tile_cursor[i] = 0;
    }
    if (zero_grad)
        for (long long i = tid; i < npts3; i += stride) C.grad[i] = 0.0f;
}

__global__ __launch_bounds__(256) void zero_grad_kernel(float* g, long long n,
                                                        BatchCounters* ctr) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->queue[8] = 0;  // backward queue
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        g[i] = 0.0f;
}

__global__ void reset_queue_kernel(BatchCounters* ctr, int a, int b) {
    if (threadIdx.x == 0) {
        ctr->queue[a] = 0;
        ctr->queue[b] = 0;
    }
}

__global__ __launch_bounds__(256) void prep_kernel(Ctx C) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < C.total_tris;
         t += stride) {
        const MeshDesc& M = C.mesh[mesh_of_tri(C, t)];
        const double* xyz = C.xyz + 3 * M.vbase;
        double P[3][3];
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) P[v][a] = xyz[3 * (size_t)C.tris[3 * t + v] + a];
        TriRec r;
        TriRec64 r64;
        prep_triangle(P, M.g, r, r64);
        C.rec[t] = r;
        C.rec64[t] = r64;
        const int4 x = xy_range(P, M.g);
        C.xyr[t] = x;
        for (int tj = x.z / kTile; tj <= x.w / kTile; ++tj)
            for (int ti = x.x / kTile; ti <= x.y / kTile; ++ti)
                atomicAdd(&C.tile_count[M.tile_base + tj * M.ntx + ti], 1);
    }
}

__global__ __launch_bounds__(256) void tile_alloc_kernel(Ctx C) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < C.total_tiles;
         i += stride) {
        const int c = C.tile_count[i];
        const long long off = (long long)atomicAdd(&C.ctr->tile_used, (unsigned long long)c);
        C.tile_off[i] = (int)off;
        if (off + c > C.tile_cap) atomicOr(&C.ctr->overflow, 1u);
    }
}

__global__ __launch_bounds__(256) void tile_fill_kernel(Ctx C) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < C.total_tris;
         t += stride) {
        const MeshDesc& M = C.mesh[mesh_of_tri(C, t)];
        const int4 x = C.xyr[t];
        for (int tj = x.z / kTile; tj <= x.w / kTile; ++tj)
            for (int ti = x.x / kTile; ti <= x.y / kTile; ++ti) {
                const int gt = M.tile_base + tj * M.ntx + ti;
                const int slot = atomicAdd(&C.tile_cursor[gt], 1);
                const long long at = (long long)C.tile_off[gt] + slot;
                if (at < C.tile_cap) C.tile_pool[at] = (int)t;
            }
    }
}

// One 8x8 column tile per workgroup item; same algorithm as voxelize.hip's kernel.
__global__ __launch_bounds__(64) void occ_kernel(Ctx C) {
    __shared__ double sv[9][kChunk];
    __shared__ int4 sr[kChunk];
    __shared__ int st[kChunk];
    const long long* off = phase_off(C, kPhTile);
    const long long total = off[C.nm];
    const int lane = threadIdx.x;
    while (true) {
        const long long item = next_item(&C.ctr->queue[0]);
        if (item >= total) break;
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
                    const int sa = o3_compat(P, A, B, Cc);
                    const int sb = o3_compat(Q, A, B, Cc);
                    if (sa == 0 || sb == 0) {
                        grazed = true;
                        first_graze = min(first_graze, t);
                        continue;
                    }
                    if (sa == sb) continue;
                    const int s1 = o3_compat(P, Q, A, B);
                    const int s2 = o3_compat(P, Q, B, Cc);
                    const int s3 = o3_compat(P, Q, Cc, A);
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
        if (valid && ok) write_column(mask, M.words, g, ci, cj, C.occ + M.cell_base);
    }
}

// Brick lists. LEVEL 2: 16^3 bricks, candidates = the mesh's triangles; LEVEL 1: 8^3
// bricks, candidates = the parent 16^3 brick's list. PASS 0: min centre distance;
// PASS 1: survivors counted and their list allocated from the pool; PASS 2: written.
template <int LEVEL, int PASS>
__global__ __launch_bounds__(256) void cull_kernel(Ctx C) {
    __shared__ float sred[4];
    __shared__ int scount[4];
    __shared__ long long sbase;
    const int ph = LEVEL == 2 ? kPhL2 : kPhL1;
    const long long* off = phase_off(C, ph);
    const long long total = off[C.nm];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long* q = &C.ctr->queue[1 + (LEVEL == 2 ? 0 : 3) + PASS];
    float* dmin = LEVEL == 2 ? C.dmin2 : C.dmin1;
    int* cnt = LEVEL == 2 ? C.cnt2 : C.cnt1;
    long long* loff = LEVEL == 2 ? C.off2 : C.off1;
    int* pool = LEVEL == 2 ? C.pool2 : C.pool1;
    const long long cap = LEVEL == 2 ? C.cap2 : C.cap1;
    while (true) {
        const long long item = next_item(q);
        if (item >= total) break;
        const int pos = find_pos(off, C.nm, item);
        const MeshDesc& M = C.mesh[C.order[pos]];
        const int b = (int)(item - off[pos]);
        const Level& L = LEVEL == 2 ? M.L2 : M.L1;
        const int gb = (LEVEL == 2 ? M.l2_base : M.l1_base) + b;
        float cx, cy, cz, hd;
        brick_geom(M.g, L, b, cx, cy, cz, hd);
        long long beg, end;
        const int* cand = nullptr;  // nullptr: candidates are triangles tbase + k
        if (LEVEL == 2) {
            beg = M.tbase;
            end = M.tbase + M.ntris;
        } else {
            const int ix = b % L.bx, iy = (b / L.bx) % L.by, iz = b / (L.bx * L.by);
            const int pb = M.l2_base + ((iz >> 1) * M.L2.by + (iy >> 1)) * M.L2.bx + (ix >> 1);
            beg = C.off2[pb];
            end = beg + C.cnt2[pb];
            if (end > C.cap2) end = beg;  // parent list overflowed (flag already set)
            cand = C.pool2;
        }
        if (PASS == 0) {
            float m = FLT_MAX;
            for (long long k = beg + threadIdx.x; k < end; k += blockDim.x)
                m = fminf(m, tri_d2(C.rec[cand ? cand[k] : k], cx, cy, cz));
            m = wave_min(m);
            if (lane == 0) sred[wave] = m;
            __syncthreads();
            if (threadIdx.x == 0)
                dmin[gb] = sqrtf(fminf(fminf(sred[0], sred[1]), fminf(sred[2], sred[3])));
            continue;
        }
        const float thr = dmin[gb] + 2.0f * hd + M.eps;
        const float thr2 = thr * thr;
        if (PASS == 1) {
            int c = 0;
            for (long long k = beg + threadIdx.x; k < end; k += blockDim.x)
                c += tri_d2(C.rec[cand ? cand[k] : k], cx, cy, cz) <= thr2;
            for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
            if (lane == 0) scount[wave] = c;
            __syncthreads();
            if (threadIdx.x == 0) {
                const int total_c = scount[0] + scount[1] + scount[2] + scount[3];
                const long long at =
                    (long long)atomicAdd(LEVEL == 2 ? &C.ctr->l2_used : &C.ctr->l1_used,
                                         (unsigned long long)total_c);
                cnt[gb] = total_c;
                loff[gb] = at;
                if (at + total_c > cap) atomicOr(&C.ctr->overflow, LEVEL == 2 ? 2u : 4u);
            }
            continue;
        }
        // PASS 2
        if (threadIdx.x == 0) sbase = loff[gb];
        __syncthreads();
        const bool fits = loff[gb] + cnt[gb] <= cap;
        for (long long k0 = beg; k0 < end; k0 += blockDim.x) {
            const long long k = k0 + threadIdx.x;
            int t = -1;
            bool keep = false;
            if (k < end) {
                t = cand ? cand[k] : (int)k;
                keep = tri_d2(C.rec[t], cx, cy, cz) <= thr2;
            }
            const unsigned long long mask = __ballot(keep);
            if (lane == 0) scount[wave] = __popcll(mask);
            __syncthreads();
            int before = 0;
            for (int w = 0; w < wave; ++w) before += scount[w];
            const long long base = sbase;
            if (keep && fits) pool[base + before + __popcll(mask & ((1ull << lane) - 1ull))] = t;
            __syncthreads();
            if (threadIdx.x == 0) sbase = base + scount[0] + scount[1] + scount[2] + scount[3];
            __syncthreads();
        }
    }
}

// One brick per item: 4x4x4 voxels (VPL = 1 voxel per lane) or 4x4x8 (VPL = 2, the
// second voxel of a lane 4 cells above the first). D = 0 reads each candidate's record
// with scalar loads (SGPR operands); D > 0 stages D records at a time through LDS.
// Same algorithm as sdf.hip's voxel kernel otherwise.
template <bool BWD, int VPL, int D>
__global__ __launch_bounds__(64) void vox_kernel(Ctx C) {
    __shared__ int slist[kLdsList];
    __shared__ TriRec srec[D > 0 ? D : 1];
    const long long* off = phase_off(C, VPL == 1 ? kPhL0 : kPhL0b);
    const long long total = off[C.nm];
    const int lane = threadIdx.x;
    unsigned long long* q = &C.ctr->queue[BWD ? 8 : 7];
    unsigned long long evals = 0;
    while (true) {
        const long long item = next_item(q);
        if (item >= total) break;
        const int pos = find_pos(off, C.nm, item);
        const MeshDesc& M = C.mesh[C.order[pos]];
        const GridSpec& g = M.g;
        const Level L0 = VPL == 1 ? M.L0 : M.L0b;
        const int b = (int)(item - off[pos]);
        const int ix = b % L0.bx, iy = (b / L0.bx) % L0.by, iz = b / (L0.bx * L0.by);
        const int i = ix * kLeaf + (lane & 3), j = iy * kLeaf + ((lane >> 2) & 3);
        int k[VPL];
        bool valid[VPL];
        float pz[VPL];
        const float h = (float)g.spacing;
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            k[v] = iz * L0.sz + (lane >> 4) + 4 * v;
            valid[v] = i < g.nx && j < g.ny && k[v] < g.nz;
            pz[v] = (k[v] + 0.5f) * h;
        }
        const float px = (i + 0.5f) * h, py = (j + 0.5f) * h;
        float cx, cy, cz, hd;
        brick_geom(g, L0, b, cx, cy, cz, hd);
        const int pb = M.l1_base +
                       ((VPL == 1 ? iz >> 1 : iz) * M.L1.by + (iy >> 1)) * M.L1.bx + (ix >> 1);
        const long long beg = C.off1[pb];
        long long end = beg + C.cnt1[pb];
        if (end > C.cap1) end = beg;  // overflowed (flag already set)
        float m = FLT_MAX;
        for (long long qq = beg + lane; qq < end; qq += 64)
            m = fminf(m, tri_d2(C.rec[C.pool1[qq]], cx, cy, cz));
        const float dmin = sqrtf(wave_min(m));
        const float thr = dmin + 2.0f * hd + M.eps;
        const float thr2 = thr * thr;
        int n = 0;
        for (long long q0 = beg; q0 < end; q0 += 64) {
            const long long qq = q0 + lane;
            int t = -1;
            bool keep = false;
            if (qq < end) {
                t = C.pool1[qq];
                keep = tri_d2(C.rec[t], cx, cy, cz) <= thr2;
            }
            const unsigned long long mask = __ballot(keep);
            if (keep) {
                const int p = n + __popcll(mask & ((1ull << lane) - 1ull));
                if (p < kLdsList) slist[p] = t;
            }
            n += __popcll(mask);
        }
        __syncthreads();
        const bool spill = n > kLdsList;
        const int count = spill ? (int)(end - beg) : n;
        auto list_at = [&](int qq) { return spill ? C.pool1[beg + qq] : slist[qq]; };
        const float cap = dmin + hd + M.eps;
        float best[VPL], bound[VPL];
        int arg[VPL];
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            best[v] = FLT_MAX;
            bound[v] = cap;
            arg[v] = -1;
        }
        auto visit = [&](const TriRec& T, int t) {
            bool need = false;
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                const float dx = px - T.mx, dy = py - T.my, dz = pz[v] - T.mz;
                const float rr = T.r + bound[v];
                need |= valid[v] && fma3(dx, dx, dy, dy, dz, dz) <= rr * rr;
            }
            if (__ballot(need) == 0) return;
            ++evals;
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                const float d2 = tri_d2(T, px, py, pz[v]);
                if (d2 < best[v]) {
                    best[v] = d2;
                    arg[v] = t;
                }
                bound[v] = fminf(cap, sqrtf(best[v]) * (1.0f + 4.0f * kU) + M.eps);
            }
        };
        if (D == 0) {
            for (int qq = 0; qq < count; ++qq) {
                const int t = uniform_int(list_at(qq));
                visit(load_uniform(C.rec, t), t);
            }
        } else {
            for (int c0 = 0; c0 < count; c0 += D) {
                const int nq = min(D, count - c0);
                __syncthreads();
                for (int e = lane; e < nq * 8; e += 64) {
                    const int t = list_at(c0 + (e >> 3));
                    reinterpret_cast<float4*>(&srec[e >> 3])[e & 7] =
                        reinterpret_cast<const float4*>(C.rec + t)[e & 7];
                }
                __syncthreads();
                for (int qq = 0; qq < nq; ++qq) visit(srec[qq], uniform_int(list_at(c0 + qq)));
            }
        }
        const uint8_t* occ = C.occ + M.cell_base;
        if (!BWD) {
            bool near[VPL];
            bool any_near = false;
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                near[v] = valid[v] && sqrtf(best[v]) <= h * (1.0f + 1e-3f) + 2.0f * M.eps;
                any_near |= near[v];
            }
            double best64[VPL];
#pragma unroll
            for (int v = 0; v < VPL; ++v) best64[v] = DBL_MAX;
            if (__ballot(any_near) != 0) {
                double p64[VPL][3];
#pragma unroll
                for (int v = 0; v < VPL; ++v) {
                    p64[v][0] = g.corner[0] + (i + 0.5) * g.spacing;
                    p64[v][1] = g.corner[1] + (j + 0.5) * g.spacing;
                    p64[v][2] = g.corner[2] + (k[v] + 0.5) * g.spacing;
                }
                for (int qq = 0; qq < count; ++qq) {
                    const int t = uniform_int(list_at(qq));
                    const TriRec64 T = load_uniform(C.rec64, t);
#pragma unroll
                    for (int v = 0; v < VPL; ++v)
                        if (near[v]) {
                            const double d = point_tri_dist2_ref(p64[v], T.a, T.b, T.c);
                            if (d < best64[v]) best64[v] = d;
                        }
                }
            }
            float* out = C.sdf + M.cell_base;
#pragma unroll
            for (int v = 0; v < VPL; ++v) {
                if (!valid[v]) continue;
                const size_t idx = ((size_t)k[v] * g.ny + j) * g.nx + i;
                if (near[v]) {
                    const double d = sqrt(best64[v]);
                    out[idx] = (float)(occ[idx] ? -d : d);  // voxelize.cpp:354-355
                } else {
                    const float d32 = sqrtf(best[v]);
                    out[idx] = occ[idx] ? -d32 : d32;
                }
            }
            __syncthreads();
            continue;
        }
#pragma unroll
        for (int v = 0; v < VPL; ++v) {
            float c[9];
            bool act = false;
            const size_t idx = valid[v] ? ((size_t)k[v] * g.ny + j) * g.nx + i : 0;
            if (valid[v] && arg[v] >= 0) {
                const TriRec T = C.rec[arg[v]];
                float vv, ww;
                tri_bary(T, px, py, pz[v], vv, ww);
                const float u = 1.0f - vv - ww;
                const float qx = T.ax + vv * T.abx + ww * T.acx;
                const float qy = T.ay + vv * T.aby + ww * T.acy;
                const float qz = T.az + vv * T.abz + ww * T.acz;
                const float ex = px - qx, ey = py - qy, ez = pz[v] - qz;
                const float d = sqrtf(ex * ex + ey * ey + ez * ez);
                const float gv = C.gsdf[M.cell_base + idx];
                if (d > 0.0f && gv != 0.0f) {
                    const float s = occ[idx] ? -1.0f : 1.0f;
                    const float coef = -gv * s / d;
                    c[0] = coef * u * ex; c[1] = coef * u * ey; c[2] = coef * u * ez;
                    c[3] = coef * vv * ex; c[4] = coef * vv * ey; c[5] = coef * vv * ez;
                    c[6] = coef * ww * ex; c[7] = coef * ww * ey; c[8] = coef * ww * ez;
                    act = true;
                }
            }
            // wavefront-aggregated scatter: one group of lanes per distinct triangle
            while (true) {
                const unsigned long long mask = __ballot(act);
                if (mask == 0) break;
                const int leader = __ffsll((long long)mask) - 1;
                const int lt = __shfl(arg[v], leader);
                const bool match = act && arg[v] == lt;
                float sum[9];
#pragma unroll
                for (int e = 0; e < 9; ++e) sum[e] = wave_sum(match ? c[e] : 0.0f);
                if (lane == leader) {
                    const int* tv = C.tris + 3 * (size_t)lt;
                    float* gbase = C.grad + 3 * M.vbase;
#pragma unroll
                    for (int vtx = 0; vtx < 3; ++vtx) {
                        float* gp = gbase + 3 * (size_t)tv[vtx];
                        atomicAdd(gp + 0, sum[3 * vtx + 0]);
                        atomicAdd(gp + 1, sum[3 * vtx + 1]);
                        atomicAdd(gp + 2, sum[3 * vtx + 2]);
                    }
                }
                act = act && !match;
            }
        }
        __syncthreads();
    }
    if (lane == 0 && evals) atomicAdd(&C.ctr->pairs, evals);
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

struct BatchEngine::Impl {
    hipStream_t stream = nullptr;
    int nm = 0;
    std::vector<MeshDesc> h_mesh;
    std::vector<int> h_order;
    std::vector<long long> h_item_off;
    std::vector<long long> h_tri_start, h_pt_start, h_cell_start;
    long long total_tris = 0, total_pts = 0, total_cells = 0, total_tiles = 0;
    long long total_cols_words = 0, n_l2 = 0, n_l1 = 0;
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

    ~Impl() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamSynchronize(stream);
        if (capture) (void)hipStreamDestroy(capture);
        for (void* p : owned) (void)hipFree(p);
    }

    int vpl = 1, depth = 0;  // voxel-pass variant (voxels per lane, LDS tile depth)

    void launch_forward(hipStream_t st);
    void launch_backward(hipStream_t st);

    template <bool BWD>
    void launch_vox(hipStream_t st) {
#define CMG_VOX(V, DD)                                                 \
    if (vpl == V && depth == DD) {                                    \
        vox_kernel<BWD, V, DD><<<grid64, 64, 0, st>>>(ctx);            \
        CMG_HIP_LAUNCH();                                              \
        return;                                                        \
    }
        CMG_VOX(1, 0) CMG_VOX(1, 32) CMG_VOX(1, 64) CMG_VOX(2, 0) CMG_VOX(2, 32) CMG_VOX(2, 64)
#undef CMG_VOX
        throw HipError("batch: unknown voxel-pass variant");
    }
};

void BatchEngine::Impl::launch_forward(hipStream_t st) {
    Ctx& C = ctx;
    reset_kernel<<<grid256, 256, 0, st>>>(C, 0, 0);
    CMG_HIP_LAUNCH();
    prep_kernel<<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    tile_alloc_kernel<<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    tile_fill_kernel<<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    occ_kernel<<<grid64, 64, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    cull_kernel<2, 0><<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    cull_kernel<2, 1><<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    cull_kernel<2, 2><<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    cull_kernel<1, 0><<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    cull_kernel<1, 1><<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    cull_kernel<1, 2><<<grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    {
        TimedLaunch timed("batch_sdf", st);
        launch_vox<false>(st);
    }
}

void BatchEngine::Impl::launch_backward(hipStream_t st) {
    Ctx& C = ctx;
    zero_grad_kernel<<<grid256, 256, 0, st>>>(C.grad, 3 * total_pts, C.ctr);
    CMG_HIP_LAUNCH();
    {
        TimedLaunch timed("batch_sdf_backward", st);
        launch_vox<true>(st);
    }
}

BatchEngine::BatchEngine(int nm, const long long* pt_start, const long long* tri_start,
                         const int* resolution, const int* pad, const double* d_xyz,
                         const int* d_tris, hipStream_t st)
    : p_(new Impl) {
    Impl& I = *p_;
    I.stream = st;
    I.nm = nm;
    I.xyz = d_xyz;
    I.tris = d_tris;
    I.h_pt_start.assign(pt_start, pt_start + nm + 1);
    I.h_tri_start.assign(tri_start, tri_start + nm + 1);
    I.total_pts = pt_start[nm];
    I.total_tris = tri_start[nm];
    for (int m = 0; m < nm; ++m) {
        if (pt_start[m + 1] <= pt_start[m]) throw InvalidInput("batch: a mesh has no points");
        if (tri_start[m + 1] <= tri_start[m])
            throw InvalidInput("batch: a mesh has no triangles");
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
    long long cells = 0, tiles = 0, words = 0, l2 = 0, l1 = 0;
    std::vector<double> cost(nm);
    for (int m = 0; m < nm; ++m) {
        double box[6] = {lohi[6 * m], lohi[6 * m + 1], lohi[6 * m + 2],
                         lohi[6 * m + 3], lohi[6 * m + 4], lohi[6 * m + 5]};
        // make_grid over the two corners reproduces make_grid over all points
        double two[6] = {box[0], box[1], box[2], box[3], box[4], box[5]};
        GridSpec g;
        if (!make_grid(two, 2, resolution[m], pad[m], g))
            throw InvalidInput("batch: degenerate (zero-extent) mesh");
        if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0)
            throw InvalidInput("batch: grid has no cells (negative pad)");
        MeshDesc& M = I.h_mesh[m];
        M.g = g;
        M.vbase = pt_start[m];
        M.tbase = tri_start[m];
        M.ntris = (int)(tri_start[m + 1] - tri_start[m]);
        M.cell_base = cells;
        OccPlan op = occupancy_plan(g);
        M.words = op.words;
        M.ntx = op.ntx;
        M.nty = op.nty;
        M.tile_base = (int)tiles;
        M.mask_base = words;
        M.L0 = make_level(g, kLeaf);
        M.L0b = make_level(g, kLeaf, 2 * kLeaf);
        M.L1 = make_level(g, 2 * kLeaf);
        M.L2 = make_level(g, 4 * kLeaf);
        M.l2_base = (int)l2;
        M.l1_base = (int)l1;
        M.eps = distance_eps(g);
        const long long nc = (long long)g.nx * g.ny * g.nz;
        cells += nc;
        tiles += (long long)op.ntx * op.nty;
        words += (long long)g.nx * g.ny * op.words;
        l2 += (long long)M.L2.bx * M.L2.by * M.L2.bz;
        l1 += (long long)M.L1.bx * M.L1.by * M.L1.bz;
        cost[m] = (double)nc * std::log2(2.0 + M.ntris);
    }
    I.total_cells = cells;
    I.total_tiles = tiles;
    I.total_cols_words = words;
    I.n_l2 = l2;
    I.n_l1 = l1;
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
            if (ph == kPhL2) n = (long long)M.L2.bx * M.L2.by * M.L2.bz;
            if (ph == kPhL1) n = (long long)M.L1.bx * M.L1.by * M.L1.bz;
            if (ph == kPhL0) n = (long long)M.L0.bx * M.L0.by * M.L0.bz;
            if (ph == kPhL0b) n = (long long)M.L0b.bx * M.L0b.by * M.L0b.bz;
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
    CMG_HIP(hipMemcpyAsync(d_mesh, I.h_mesh.data(), sizeof(MeshDesc) * nm, hipMemcpyHostToDevice, st));
    CMG_HIP(hipMemcpyAsync(d_order, I.h_order.data(), sizeof(int) * nm, hipMemcpyHostToDevice, st));
    CMG_HIP(hipMemcpyAsync(d_item, I.h_item_off.data(), sizeof(long long) * I.h_item_off.size(),
                           hipMemcpyHostToDevice, st));
    CMG_HIP(hipMemcpyAsync(d_ts, tri_start, sizeof(long long) * (nm + 1), hipMemcpyHostToDevice, st));
    C.mesh = d_mesh;
    C.order = d_order;
    C.item_off = d_item;
    C.tri_start = d_ts;
    C.total_tris = I.total_tris;
    C.xyz = d_xyz;
    C.tris = d_tris;
    C.ctr = I.dalloc<BatchCounters>(1);
    C.rec = I.dalloc<TriRec>(I.total_tris);
    C.rec64 = I.dalloc<TriRec64>(I.total_tris);
    C.xyr = I.dalloc<int4>(I.total_tris);
    C.tile_count = I.dalloc<int>(tiles);
    C.tile_off = I.dalloc<int>(tiles);
    C.tile_cursor = I.dalloc<int>(tiles);
    C.masks = I.dalloc<unsigned long long>(words);
    C.occ = I.dalloc<uint8_t>(cells);
    C.dmin2 = I.dalloc<float>(l2);
    C.cnt2 = I.dalloc<int>(l2);
    C.off2 = I.dalloc<long long>(l2);
    C.dmin1 = I.dalloc<float>(l1);
    C.cnt1 = I.dalloc<int>(l1);
    C.off1 = I.dalloc<long long>(l1);
    C.total_tiles = (int)tiles;
    I.grid64 = (unsigned)cu_count() * 16;
    I.grid256 = (unsigned)cu_count() * 4;
    // planning pass: run the forward once with empty pools to learn the list sizes
    C.tile_cap = C.cap2 = C.cap1 = 0;
    C.tile_pool = C.pool2 = C.pool1 = nullptr;
    C.sdf = nullptr;
    // sizes come from the count passes; writes are skipped while the caps are 0
    reset_kernel<<<I.grid256, 256, 0, st>>>(C, 0, 0);
    prep_kernel<<<I.grid256, 256, 0, st>>>(C);
    tile_alloc_kernel<<<I.grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    BatchCounters h{};
    CMG_HIP(hipMemcpyAsync(&h, C.ctr, sizeof h, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    auto margin = [](unsigned long long n) { return (long long)(n + n / 4 + 4096); };
    C.tile_cap = margin(h.tile_used);
    C.tile_pool = I.dalloc<int>(C.tile_cap);
    // second planning step: full occupancy and level-2 counts
    reset_kernel<<<I.grid256, 256, 0, st>>>(C, 0, 0);
    prep_kernel<<<I.grid256, 256, 0, st>>>(C);
    tile_alloc_kernel<<<I.grid256, 256, 0, st>>>(C);
    tile_fill_kernel<<<I.grid256, 256, 0, st>>>(C);
    occ_kernel<<<I.grid64, 64, 0, st>>>(C);
    cull_kernel<2, 0><<<I.grid256, 256, 0, st>>>(C);
    cull_kernel<2, 1><<<I.grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    CMG_HIP(hipMemcpyAsync(&h, C.ctr, sizeof h, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    C.cap2 = margin(h.l2_used);
    C.pool2 = I.dalloc<int>(C.cap2);
    cull_kernel<2, 2><<<I.grid256, 256, 0, st>>>(C);
    cull_kernel<1, 0><<<I.grid256, 256, 0, st>>>(C);
    cull_kernel<1, 1><<<I.grid256, 256, 0, st>>>(C);
    CMG_HIP_LAUNCH();
    CMG_HIP(hipMemcpyAsync(&h, C.ctr, sizeof h, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    C.cap1 = margin(h.l1_used);
    C.pool1 = I.dalloc<int>(C.cap1);
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
    if (fwd && !d_sdf) throw InvalidInput("batch: forward needs an sdf buffer");
    if (bwd && (!d_gsdf || !d_grad)) throw InvalidInput("batch: backward needs grad buffers");
    I.ctx.sdf = d_sdf;
    I.ctx.gsdf = d_gsdf;
    I.ctx.grad = d_grad;
    if (!(flags & kBatchGraph)) {
        if (fwd) I.launch_forward(st);
        if (bwd) I.launch_backward(st);
        return;
    }
    // The graph is recorded on an engine-owned non-blocking stream (the caller's stream
    // may be the legacy default stream, which cannot be captured) and launched on st;
    // it depends only on the buffer addresses and the pass selection.
    const void* key[4] = {d_xyz, d_sdf, d_gsdf, d_grad};
    const int f = flags & (kBatchForward | kBatchBackward);
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

std::vector<BatchVariant> BatchEngine::autotune(const float* d_gsdf, float* d_sdf,
                                                float* d_grad, hipStream_t st) {
    Impl& I = *p_;
    I.ctx.sdf = d_sdf;
    I.ctx.gsdf = d_gsdf;
    I.ctx.grad = d_grad;
    I.launch_forward(st);  // lists and occupancy for this geometry
    std::vector<BatchVariant> out;
    hipEvent_t e0, e1;
    CMG_HIP(hipEventCreate(&e0));
    CMG_HIP(hipEventCreate(&e1));
    float best_ms = 1e30f;
    int best_v = I.vpl, best_d = I.depth;
    for (int v : {1, 2})
        for (int d : {0, 32, 64}) {
            I.vpl = v;
            I.depth = d;
            float total = 0;
            for (int rep = 0; rep < 3; ++rep) {
                reset_queue_kernel<<<1, 64, 0, st>>>(I.ctx.ctr, 7, 8);
                zero_grad_kernel<<<I.grid256, 256, 0, st>>>(I.ctx.grad, 3 * I.total_pts, I.ctx.ctr);
                CMG_HIP(hipEventRecord(e0, st));
                I.launch_vox<false>(st);
                I.launch_vox<true>(st);
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
    if (!((vpl == 1 || vpl == 2) && (depth == 0 || depth == 32 || depth == 64)))
        throw InvalidInput("batch: variant must have vpl in {1,2} and depth in {0,32,64}");
    Impl& I = *p_;
    if (I.vpl != vpl || I.depth != depth) {
        I.vpl = vpl;
        I.depth = depth;
        I.graph_flags = -1;  // recapture with the new kernels
    }
}

BatchVariant BatchEngine::variant() const { return BatchVariant{p_->vpl, p_->depth, 0.0f}; }

void BatchEngine::status(hipStream_t st, unsigned* overflow, unsigned long long* pairs) {
    BatchCounters h{};
    CMG_HIP(hipMemcpyAsync(&h, p_->ctx.ctr, sizeof h, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    if (overflow) *overflow = h.overflow;
    if (pairs) *pairs = h.pairs;
}

const uint8_t* BatchEngine::occupancy() const { return p_->ctx.occ; }

}  // namespace cmg::hip
