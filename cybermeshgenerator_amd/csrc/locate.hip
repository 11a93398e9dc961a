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
// Bound: latency of the dependent gathers (16 B of indices, 16 B of neighbours and 96 B
// of points per step); lanes diverge in walk length.
#include <hip/hip_runtime.h>

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

__global__ __launch_bounds__(256) void locate_kernel(const double* xyz, long long npts,
                                                     const int* tets, const int* nbrs,
                                                     long long ntets, const double* q,
                                                     long long nq, const int* start,
                                                     int start_tet, int* out, int* steps_out,
                                                     unsigned long long* ctr) {
    const long long limit = 2 * ntets + 16;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nq;
         i += (long long)gridDim.x * blockDim.x) {
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
    if (!d_xyz || !d_tets || !d_nbrs || !d_q || !d_out) throw InvalidInput("locate: null pointer");
    if (reinterpret_cast<uintptr_t>(d_tets) % 16 || reinterpret_cast<uintptr_t>(d_nbrs) % 16)
        throw InvalidInput("locate: tets and nbrs must be 16-byte aligned");
    if (!d_start && (start_tet < 0 || start_tet >= ntets))
        throw InvalidInput("locate: start tet out of range");
    ScratchLease lease;
    auto* ctr = static_cast<unsigned long long*>(lease.s->get(2 * sizeof(unsigned long long)));
    CMG_HIP(hipMemsetAsync(ctr, 0, 2 * sizeof(unsigned long long), st));
    {
        TimedLaunch tl("locate", st);
        locate_kernel<<<grid_for((size_t)nq, 256, 8), 256, 0, st>>>(
            d_xyz, npts, d_tets, d_nbrs, ntets, d_q, nq, d_start, start_tet, d_out, d_steps,
            ctr);
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
