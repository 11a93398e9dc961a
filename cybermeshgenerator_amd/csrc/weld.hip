// weld.hip — surface input straight to device arrays: binary STL decoding and the merge
// of coincident vertices, on the GPU.
//
// The reference reads an STL file into a vertex soup (three vertices per triangle) and
// merges coincident vertices through an ordered map on the exact coordinate triple,
// numbering the merged points in order of first occurrence (src/io/stl.cpp:24-61,
// 94-113). Here:
//   stl_decode_kernel  binary records (50 B each, 2-byte aligned) -> FP64 soup; a
//                      workgroup stages 256 records (12,800 B, 4-byte aligned) in LDS
//                      with coalesced dword loads, then each lane unpacks its 9 floats
//   weld_insert_kernel open-addressing hash table of vertex numbers (2x capacity,
//                      linear probing): the first vertex of a key claims a slot with a
//                      CAS, later ones with an equal key lower it with atomicMin, so the
//                      slot ends holding the key's first occurrence; each vertex records
//                      the slot of its key. A workgroup merges its 256 vertices in LDS
//                      first, and only each key's smallest local index probes globally.
//   weld_rep_kernel    rep = the final value of the vertex's slot (its key's first
//                      occurrence); flags the vertices that are their own rep
//   (exclusive scan of the flags: the merged number of each first occurrence)
//   weld_emit_kernel   merged index of every vertex, first occurrences' coordinates
// Keys compare as doubles (-0 == +0, as the map's operator< treats them) and hash on
// bits with -0 folded to +0; the point written is the first occurrence's own value.
// A NaN coordinate makes its vertex unequal to every vertex (the reference's map has no
// defined behaviour for NaN keys): such a vertex skips the hash tables and is its own
// representative.
//
// Bound: HBM and atomic latency; the table and soup accesses of the insert and rep
// kernels are random. Work buffers come from the library's per-device scratch (reused
// across calls; calls on one device serialise on it).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>

#include <cstring>
#include <vector>

#include "device_common.hpp"
#include "voxelize_internal.hpp"
#include "weld_internal.hpp"

namespace cmg::hip {
namespace {

constexpr unsigned kEmpty = 0xffffffffu;
constexpr int kDecodeTris = 256;  // records per workgroup (12,800 B)

// 16-byte vector of dwords that needs only 4-byte alignment (a record block starts 4 bytes
// past a 16-byte boundary: 84 + 12,800 k)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4), aligned(4)));

// One workgroup decodes kDecodeTris records: the block's bytes go to LDS (full blocks:
// 16-byte loads, all in flight together), each lane converts its record's 9 floats, and
// the 9 * 256 doubles leave through LDS as coalesced 16-byte stores.
__global__ __launch_bounds__(256) void stl_decode_kernel(const uint8_t* bytes, long long ntris,
                                                         double* soup) {
    __shared__ __align__(16) double lds_d[kDecodeTris * 9];  // 18,432 B; input uses 12,800
    uint32_t* lds = reinterpret_cast<uint32_t*>(lds_d);
    const long long t0 = (long long)blockIdx.x * kDecodeTris;
    const int nt = (int)min((long long)kDecodeTris, ntris - t0);
    const uint8_t* base = bytes + 84 + 50 * t0;  // 84 + 12,800 * block: 4-byte aligned
    const int nbytes = 50 * nt;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(base);
    if (nt == kDecodeTris) {
        constexpr int kVec = kDecodeTris * 50 / 16;  // 800 16-byte words
        constexpr int kPer = (kVec + 255) / 256;     // 4 per lane
        const u32x4* wv = reinterpret_cast<const u32x4*>(w);
        u32x4 r[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int q = j * 256 + (int)threadIdx.x;
            r[j] = wv[q < kVec ? q : kVec - 1];  // unconditional: the loads issue together
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int q = j * 256 + (int)threadIdx.x;
            if (q < kVec) {
                lds[4 * q] = r[j].x;
                lds[4 * q + 1] = r[j].y;
                lds[4 * q + 2] = r[j].z;
                lds[4 * q + 3] = r[j].w;
            }
        }
    } else {
        for (int i = threadIdx.x; i < nbytes / 4; i += blockDim.x) lds[i] = w[i];
        if (threadIdx.x == 0 && (nbytes & 3))  // odd record count: a 2-byte tail
            lds[nbytes / 4] = reinterpret_cast<const uint16_t*>(base)[nbytes / 2 - 1];
    }
    __syncthreads();
    float f[9];
    if ((int)threadIdx.x < nt) {
        const uint16_t* h = reinterpret_cast<const uint16_t*>(lds);
        const int o = (50 * threadIdx.x + 12) / 2;  // vertex floats start 12 B into a record
#pragma unroll
        for (int j = 0; j < 9; ++j)
            f[j] = __uint_as_float((uint32_t)h[o + 2 * j] | ((uint32_t)h[o + 2 * j + 1] << 16));
    }
    __syncthreads();  // every lane has read its record: the buffer takes the doubles
    if ((int)threadIdx.x < nt) {
#pragma unroll
        for (int j = 0; j < 9; ++j) lds_d[9 * threadIdx.x + j] = (double)f[j];
    }
    __syncthreads();
    double* out = soup + 9 * t0;
    const int nd = 9 * nt;
    if ((reinterpret_cast<uintptr_t>(out) & 15) == 0) {
        const double2* src = reinterpret_cast<const double2*>(lds_d);
        double2* dst = reinterpret_cast<double2*>(out);
        for (int q = threadIdx.x; q < nd / 2; q += blockDim.x) dst[q] = src[q];
        if ((nd & 1) && threadIdx.x == 0) out[nd - 1] = lds_d[nd - 1];
    } else {
        for (int q = threadIdx.x; q < nd; q += blockDim.x) out[q] = lds_d[q];
    }
}

__device__ __forceinline__ unsigned long long key_bits(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b << 1) == 0 ? 0ull : b;  // -0 and +0 hash alike
}

__device__ __forceinline__ unsigned hash_key(const double* p) {
    unsigned long long h = key_bits(p[0]) * 0x9E3779B97F4A7C15ull;
    h ^= key_bits(p[1]) + 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= key_bits(p[2]) + 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    return (unsigned)h;
}

__device__ __forceinline__ bool same_point(const double* soup, unsigned a, long long b) {
    const double* p = soup + 3 * (long long)a;
    const double* q = soup + 3 * b;
    return p[0] == q[0] && p[1] == q[1] && p[2] == q[2];
}

// Global insert of vertex v: returns the slot of v's key. A slot's key never changes once
// set and its value only decreases, so a plain read that finds it set (even a stale
// value) names the key correctly: the CAS is needed only where the slot looks empty.
// At most nv keys in a table of >= 2 nv slots: an empty slot is always reached.
__device__ __forceinline__ unsigned global_insert(const double* soup, long long v, unsigned h,
                                                  unsigned* tab, unsigned mask) {
    unsigned s = h & mask;
    for (;;) {
        unsigned cur = __hip_atomic_load(&tab[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == kEmpty) {
            cur = atomicCAS(&tab[s], kEmpty, (unsigned)v);
            if (cur == kEmpty) return s;
        }
        if (same_point(soup, cur, v)) {
            // the slot only decreases: with cur < v it already holds less than v
            if ((unsigned)v < cur) atomicMin(&tab[s], (unsigned)v);
            return s;
        }
        s = (s + 1) & mask;
    }
}

// Each workgroup first merges its 256 consecutive vertices in an LDS table (an STL file
// repeats a vertex in neighbouring triangles), keeping the smallest local index of each
// key; only that vertex inserts into the global table, and the others take its slot. The
// global slot still ends at the key's first occurrence (the minimum over the workgroups'
// minima). slot[v] = the slot of v's key.
constexpr int kLocalSlots = 512;
__global__ __launch_bounds__(256) void weld_insert_kernel(const double* soup, long long nv,
                                                          unsigned* tab, unsigned mask,
                                                          unsigned* slot) {
    __shared__ unsigned s_key[kLocalSlots];  // local index of the key's first vertex
    __shared__ unsigned s_slot[256];         // global slot of each local representative
    __shared__ double s_xyz[256 * 3];
    const int tid = (int)threadIdx.x;
    for (long long base = (long long)blockIdx.x * 256; base < nv;
         base += (long long)gridDim.x * 256) {
        const long long v = base + tid;
        const bool in = v < nv;
        s_key[tid] = kEmpty;
        s_key[tid + 256] = kEmpty;
        unsigned h = 0;
        bool ok = false;  // a vertex that takes part in the merge
        if (in) {
            for (int a = 0; a < 3; ++a) s_xyz[3 * tid + a] = soup[3 * v + a];
            // a vertex with a NaN coordinate equals no vertex: it stays out of both
            // tables (where equal NaN bit patterns would all probe one cluster)
            ok = !(__builtin_isnan(s_xyz[3 * tid]) || __builtin_isnan(s_xyz[3 * tid + 1]) ||
                   __builtin_isnan(s_xyz[3 * tid + 2]));
            if (ok) h = hash_key(soup + 3 * v);
        }
        __syncthreads();
        unsigned ls = 0;
        if (ok) {
            ls = h & (kLocalSlots - 1);
            for (;;) {  // at most 256 keys in 512 slots
                const unsigned cur = atomicCAS(&s_key[ls], kEmpty, (unsigned)tid);
                if (cur == kEmpty) break;
                if (s_xyz[3 * cur] == s_xyz[3 * tid] && s_xyz[3 * cur + 1] == s_xyz[3 * tid + 1] &&
                    s_xyz[3 * cur + 2] == s_xyz[3 * tid + 2]) {
                    atomicMin(&s_key[ls], (unsigned)tid);
                    break;
                }
                ls = (ls + 1) & (kLocalSlots - 1);
            }
        }
        __syncthreads();
        const unsigned rep = ok ? s_key[ls] : 0u;
        if (ok && rep == (unsigned)tid) s_slot[tid] = global_insert(soup, v, h, tab, mask);
        __syncthreads();
        if (in) slot[v] = ok ? s_slot[rep] : kEmpty;  // kEmpty: its own representative
        __syncthreads();  // s_key, s_slot and s_xyz are rewritten by the next round
    }
}

// After every insert: each slot holds the first occurrence of its key, so a vertex's
// representative is the final value of the slot the insert recorded for it (rep holds the
// slot on entry and the representative on exit; a vertex with a NaN coordinate recorded
// no slot and is its own representative).
__global__ __launch_bounds__(256) void weld_rep_kernel(long long nv, const unsigned* tab,
                                                       unsigned* rep, unsigned* first) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < nv;
         v += (long long)gridDim.x * blockDim.x) {
        const unsigned s = rep[v];
        const unsigned r = s == kEmpty ? (unsigned)v : tab[s];
        rep[v] = r;
        first[v] = r == (unsigned)v;
    }
}

__global__ __launch_bounds__(256) void weld_emit_kernel(const double* soup, long long nv,
                                                        const unsigned* rep,
                                                        const unsigned* number, double* xyz,
                                                        int* idx) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < nv;
         v += (long long)gridDim.x * blockDim.x) {
        const unsigned r = rep[v];
        idx[v] = (int)number[r];
        if (r == (unsigned)v) {
            const long long m = number[v];
            xyz[3 * m] = soup[3 * v];
            xyz[3 * m + 1] = soup[3 * v + 1];
            xyz[3 * m + 2] = soup[3 * v + 2];
        }
    }
}

// Bump allocation of the work buffers out of one scratch block.
struct Carve {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t n) {
        T* p = reinterpret_cast<T*>(base + off);
        off += (n * sizeof(T) + 255) & ~size_t(255);
        return p;
    }
};

unsigned table_slots(long long nv) {
    unsigned cap = 1024;  // nv <= 2^30: at most 2^31 slots
    while ((long long)cap < 2 * nv) cap <<= 1;
    return cap;
}

size_t scan_tmp_bytes(long long nv, hipStream_t st) {
    size_t b = 0;
    CMG_HIP(rocprim::exclusive_scan(nullptr, b, (const unsigned*)nullptr, (unsigned*)nullptr,
                                    0u, (size_t)nv, rocprim::plus<unsigned>(), st));
    return b;
}

// Bytes of the merge's work buffers for nv vertices.
size_t weld_bytes(long long nv, hipStream_t st) {
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    return al(sizeof(unsigned) * table_slots(nv)) + 3 * al(sizeof(unsigned) * (size_t)nv) +
           al(scan_tmp_bytes(nv, st));
}

long long weld_impl(Carve& W, const double* d_soup, long long nv, double* d_xyz, int* d_idx,
                    hipStream_t st) {
    const unsigned cap = table_slots(nv);
    unsigned* tab = W.take<unsigned>(cap);
    unsigned* rep = W.take<unsigned>((size_t)nv);
    unsigned* first = W.take<unsigned>((size_t)nv);
    unsigned* number = W.take<unsigned>((size_t)nv);
    size_t tmp_bytes = scan_tmp_bytes(nv, st);
    void* tmp = W.take<char>(tmp_bytes);
    CMG_HIP(hipMemsetAsync(tab, 0xff, sizeof(unsigned) * cap, st));
    // 32 waves per CU: the random probes and atomics are latency-bound and these kernels
    // are small enough (<= 23 VGPRs) for 8 waves per SIMD
    const unsigned nb = grid_for((size_t)nv, 256, 32);
    weld_insert_kernel<<<nb, 256, 0, st>>>(d_soup, nv, tab, cap - 1, rep);
    CMG_HIP_LAUNCH();
    weld_rep_kernel<<<nb, 256, 0, st>>>(nv, tab, rep, first);
    CMG_HIP_LAUNCH();
    CMG_HIP(rocprim::exclusive_scan(tmp, tmp_bytes, first, number, 0u, (size_t)nv,
                                    rocprim::plus<unsigned>(), st));
    weld_emit_kernel<<<nb, 256, 0, st>>>(d_soup, nv, rep, number, d_xyz, d_idx);
    CMG_HIP_LAUNCH();
    unsigned last[2];
    CMG_HIP(hipMemcpyAsync(&last[0], number + nv - 1, sizeof(unsigned), hipMemcpyDeviceToHost,
                           st));
    CMG_HIP(hipMemcpyAsync(&last[1], first + nv - 1, sizeof(unsigned), hipMemcpyDeviceToHost,
                           st));
    CMG_HIP(hipStreamSynchronize(st));
    return (long long)last[0] + last[1];
}

}  // namespace

long long weld_vertices(const double* d_soup, long long nv, double* d_xyz, int* d_idx,
                        hipStream_t st) {
    if (nv < 1) throw InvalidInput("weld: no vertices");
    if (nv > (1LL << 30)) throw InvalidInput("weld: more than 2^30 vertices");
    if (!d_soup || !d_xyz || !d_idx) throw InvalidInput("weld: null pointer");
    ScratchLease lease;
    Carve W{static_cast<char*>(lease.s->get(weld_bytes(nv, st)))};
    return weld_impl(W, d_soup, nv, d_xyz, d_idx, st);
}

long long stl_decode_weld(const void* d_bytes, size_t nbytes, double* d_xyz, int* d_tris,
                          hipStream_t st) {
    if (!d_bytes || !d_xyz || !d_tris) throw InvalidInput("stl: null pointer");
    if (reinterpret_cast<uintptr_t>(d_bytes) % 4)
        throw InvalidInput("stl: the byte buffer must be 4-byte aligned");
    if (nbytes < 84) throw InvalidInput("stl: not a binary STL (shorter than its header)");
    uint32_t n = 0;
    CMG_HIP(hipMemcpyAsync(&n, static_cast<const uint8_t*>(d_bytes) + 80, 4,
                           hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    // stl.cpp:158-161: binary iff the size is exactly 84 + 50 * count
    if (nbytes != 84ull + 50ull * n)
        throw InvalidInput("stl: not a binary STL (size is not 84 + 50 * triangle count)");
    // stl.cpp:50-51
    if (n == 0) throw InvalidInput(".stl: vertex count is not a positive multiple of 3");
    if (3ull * n > (1ull << 30)) throw InvalidInput("weld: more than 2^30 vertices");
    ScratchLease lease;
    const size_t soup_bytes = (9 * (size_t)n * sizeof(double) + 255) & ~size_t(255);
    Carve W{static_cast<char*>(lease.s->get(soup_bytes + weld_bytes(3LL * n, st)))};
    double* soup = W.take<double>(9 * (size_t)n);
    stl_decode_kernel<<<(unsigned)((n + kDecodeTris - 1) / kDecodeTris), 256, 0, st>>>(
        static_cast<const uint8_t*>(d_bytes), n, soup);
    CMG_HIP_LAUNCH();
    return weld_impl(W, soup, 3LL * n, d_xyz, d_tris, st);
}

}  // namespace cmg::hip
