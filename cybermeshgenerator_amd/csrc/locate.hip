// locate.hip — batched point location in a fixed tetrahedral mesh: one visibility walk
// per query, with the rules of the Delaunay insertion's locate
// (src/delaunay/incremental.cpp:321-358).
//
// A lane owns one query. At tet t (positively oriented, nbr[k] across the face opposite
// v_k) it evaluates orient3d(q, v1, v2, v3), orient3d(v0, q, v2, v3), orient3d(v0, v1,
// q, v3), orient3d(v0, v1, v2, q) in that order with the compat predicate (the naive FP64
// determinant, exact only when it is zero) and steps through the first face whose
// orientation is negative and which has a neighbour; a tet where no step is taken is the
// answer. After 2 * ntets + 16 steps the lane falls back to brute_locate: the first tet
// with all four orientations >= 0 (else tet 0). Identical walks, hence identical answers,
// to the reference's for the same mesh and start tet.
//
// Queries are walked in Morton order of their positions (a radix sort of 30-bit codes):
// the lanes of a wavefront then walk nearly the same tets, which they load together and
// in similar numbers of steps. The order changes no walk.
//
// Bound: latency of the dependent gathers (16 B of indices, 16 B of neighbours and 96 B
// of points per step); lanes diverge in walk length.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>
#include <vector>

#include "device_common.hpp"
#include "locate_internal.hpp"
#include "occ_device.hpp"

namespace cmg::hip {
namespace {

__device__ __forceinline__ void load_tet(const double* xyz, int4 v, double (&P)[4][3]) {
    const int id[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int a = 0; a < 3; ++a) P[j][a] = xyz[3 * (long long)id[j] + a];
}

__device__ __forceinline__ bool tet_ok(int4 v, long long npts) {
    return v.x >= 0 && v.y >= 0 && v.z >= 0 && v.w >= 0 && v.x < npts && v.y < npts &&
           v.z < npts && v.w < npts;
}

// Per-block partial bounds of the query points (min and max per axis; exact).
__global__ __launch_bounds__(256) void qbounds_kernel(const double* q, long long nq,
                                                      double* part) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nq;
         i += (long long)gridDim.x * blockDim.x)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin(lo[a], q[3 * i + a]);
            hi[a] = fmax(hi[a], q[3 * i + a]);
        }
    __shared__ double s[6][256];
    for (int a = 0; a < 3; ++a) s[a][threadIdx.x] = lo[a], s[3 + a][threadIdx.x] = hi[a];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fmin(s[a][threadIdx.x], s[a][threadIdx.x + o]);
                s[3 + a][threadIdx.x] = fmax(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__device__ __forceinline__ unsigned spread10(unsigned v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// 30-bit Morton code of each query on a 1024^3 grid over the queries' bounds: sorting by
// it puts nearby queries (which walk nearly the same tets) in the same wavefront.
__global__ __launch_bounds__(256) void qmorton_kernel(const double* q, long long nq,
                                                      double lx, double ly, double lz,
                                                      double sx, double sy, double sz,
                                                      unsigned* key, unsigned* val) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nq;
         i += (long long)gridDim.x * blockDim.x) {
        const double f[3] = {(q[3 * i] - lx) * sx, (q[3 * i + 1] - ly) * sy,
                             (q[3 * i + 2] - lz) * sz};
        unsigned c[3];
        for (int a = 0; a < 3; ++a) {
            const double x = f[a];
            c[a] = x > 0.0 ? (x < 1023.0 ? (unsigned)x : 1023u) : 0u;  // NaN -> 0
        }
        key[i] = spread10(c[0]) | (spread10(c[1]) << 1) | (spread10(c[2]) << 2);
        val[i] = (unsigned)i;
    }
}

__global__ __launch_bounds__(256) void locate_kernel(const double* xyz, long long npts,
                                                     const int* tets, const int* nbrs,
                                                     long long ntets, const double* q,
                                                     long long nq, const unsigned* order,
                                                     const int* start, int start_tet,
                                                     int* out, int* steps_out,
                                                     unsigned long long* ctr) {
    const long long limit = 2 * ntets + 16;
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < nq;
         j += (long long)gridDim.x * blockDim.x) {
        const long long i = order ? (long long)order[j] : j;
        const double Q[3] = {q[3 * i], q[3 * i + 1], q[3 * i + 2]};
        int t = start ? start[i] : start_tet;
        int result = -1;
        long long s = 0;
        bool bad = false;
        for (; s < limit; ++s) {
            if (t < 0 || t >= ntets) {
                bad = true;
                break;
            }
            const int4 v = *reinterpret_cast<const int4*>(tets + 4 * (long long)t);
            const int4 n = *reinterpret_cast<const int4*>(nbrs + 4 * (long long)t);
            if (!tet_ok(v, npts)) {
                bad = true;
                break;
            }
            double P[4][3];
            load_tet(xyz, v, P);
            if (n.x >= 0 && occdev::o3_compat(Q, P[1], P[2], P[3]) < 0) {
                t = n.x;
                continue;
            }
            if (n.y >= 0 && occdev::o3_compat(P[0], Q, P[2], P[3]) < 0) {
                t = n.y;
                continue;
            }
            if (n.z >= 0 && occdev::o3_compat(P[0], P[1], Q, P[3]) < 0) {
                t = n.z;
                continue;
            }
            if (n.w >= 0 && occdev::o3_compat(P[0], P[1], P[2], Q) < 0) {
                t = n.w;
                continue;
            }
            result = t;
            break;
        }
        if (bad) {
            out[i] = -1;
            if (steps_out) steps_out[i] = 0;
            atomicAdd(&ctr[0], 1ull);
            continue;
        }
        if (result < 0) {
            // brute_locate (incremental.cpp:345-358)
            atomicAdd(&ctr[1], 1ull);
            result = 0;
            for (long long u = 0; u < ntets; ++u) {
                const int4 v = *reinterpret_cast<const int4*>(tets + 4 * u);
                if (!tet_ok(v, npts)) continue;
                double P[4][3];
                load_tet(xyz, v, P);
                if (occdev::o3_compat(Q, P[1], P[2], P[3]) >= 0 &&
                    occdev::o3_compat(P[0], Q, P[2], P[3]) >= 0 &&
                    occdev::o3_compat(P[0], P[1], Q, P[3]) >= 0 &&
                    occdev::o3_compat(P[0], P[1], P[2], Q) >= 0) {
                    result = (int)u;
                    break;
                }
            }
            s = -1;
        }
        out[i] = result;
        if (steps_out) steps_out[i] = (int)s;
    }
}

}  // namespace

LocateStats locate_points(const double* d_xyz, long long npts, const int* d_tets,
                          const int* d_nbrs, long long ntets, const double* d_q, long long nq,
                          const int* d_start, int start_tet, int* d_out, int* d_steps,
                          hipStream_t st) {
    if (nq < 0 || ntets < 0 || npts < 0) throw InvalidInput("locate: negative count");
    if (nq == 0) return {};
    if (ntets < 1) throw InvalidInput("locate: no tetrahedra");
    if (ntets >= (1LL << 31)) throw InvalidInput("locate: more than 2^31 - 1 tetrahedra");
    // the Morton order carries 32-bit query indices
    if (nq >= (1LL << 32)) throw InvalidInput("locate: n must be < 2^32 queries");
    if (!d_xyz || !d_tets || !d_nbrs || !d_q || !d_out) throw InvalidInput("locate: null pointer");
    if (reinterpret_cast<uintptr_t>(d_tets) % 16 || reinterpret_cast<uintptr_t>(d_nbrs) % 16)
        throw InvalidInput("locate: tets and nbrs must be 16-byte aligned");
    if (!d_start && (start_tet < 0 || start_tet >= ntets))
        throw InvalidInput("locate: start tet out of range");
    // Morton order of the queries (walk order only: every walk is the same as in input
    // order, so the results are too)
    const unsigned nb = grid_for((size_t)nq, 256, 4);
    size_t sort_tmp = 0;
    CMG_HIP(rocprim::radix_sort_pairs(nullptr, sort_tmp, (const unsigned*)nullptr,
                                      (unsigned*)nullptr, (const unsigned*)nullptr,
                                      (unsigned*)nullptr, (size_t)nq, 0, 30, st));
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t b_pair = al(sizeof(unsigned) * nq);
    ScratchLease lease;
    char* base = static_cast<char*>(lease.s->get(al(16) + al(sizeof(double) * 6 * nb) +
                                                 4 * b_pair + al(sort_tmp)));
    auto* ctr = reinterpret_cast<unsigned long long*>(base);
    double* part = reinterpret_cast<double*>(base + al(16));
    char* pb = base + al(16) + al(sizeof(double) * 6 * nb);
    unsigned* key = reinterpret_cast<unsigned*>(pb);
    unsigned* val = reinterpret_cast<unsigned*>(pb + b_pair);
    unsigned* skey = reinterpret_cast<unsigned*>(pb + 2 * b_pair);
    unsigned* sval = reinterpret_cast<unsigned*>(pb + 3 * b_pair);
    void* tmp = pb + 4 * b_pair;
    CMG_HIP(hipMemsetAsync(ctr, 0, 2 * sizeof(unsigned long long), st));
    qbounds_kernel<<<nb, 256, 0, st>>>(d_q, nq, part);
    CMG_HIP_LAUNCH();
    std::vector<double> hp((size_t)6 * nb);
    CMG_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost,
                           st));
    CMG_HIP(hipStreamSynchronize(st));
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (unsigned b = 0; b < nb; ++b)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::fmin(lo[a], hp[6 * b + a]);
            hi[a] = std::fmax(hi[a], hp[6 * b + 3 + a]);
        }
    double sc[3];
    for (int a = 0; a < 3; ++a) {
        const double e = hi[a] - lo[a];
        sc[a] = e > 0 && std::isfinite(e) ? 1024.0 / e : 0.0;
        if (!std::isfinite(lo[a])) lo[a] = 0.0;
    }
    qmorton_kernel<<<nb, 256, 0, st>>>(d_q, nq, lo[0], lo[1], lo[2], sc[0], sc[1], sc[2], key,
                                       val);
    CMG_HIP_LAUNCH();
    CMG_HIP(rocprim::radix_sort_pairs(tmp, sort_tmp, key, skey, val, sval, (size_t)nq, 0, 30,
                                      st));
    {
        TimedLaunch tl("locate", st);
        locate_kernel<<<grid_for((size_t)nq, 256, 8), 256, 0, st>>>(
            d_xyz, npts, d_tets, d_nbrs, ntets, d_q, nq, sval, d_start, start_tet, d_out,
            d_steps, ctr);
        CMG_HIP_LAUNCH();
    }
    unsigned long long h[2];
    CMG_HIP(hipMemcpyAsync(h, ctr, sizeof h, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    if (h[0])
        throw InvalidInput("locate: " + std::to_string(h[0]) +
                           " walks met a tet or neighbour index out of range");
    return LocateStats{(long long)h[1]};
}

}  // namespace cmg::hip
