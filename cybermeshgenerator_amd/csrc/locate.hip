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
#include "expansion.hpp"
#include "locate_internal.hpp"
#include "occ_device.hpp"
#include "pred_naive.hpp"

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

// ---- Delaunay cavities ----------------------------------------------------------------

// Exact insphere sign for the rare tuples whose naive determinant is exactly 0. FULL =
// false tries the register-only stages (fixed point with 31-bit differences, the
// double-word bound, fixed point with 62- and 126-bit differences) and returns kUndecided
// when they decline (differences spread over more than 126 bits); FULL = true then uses
// expansions. Out of line, so that neither weighs on the common path.
constexpr int kUndecided = 3;
template <bool FULL>
__device__ __noinline__ int is_exact_slow(const double* v) {
    int s;
    if (insphere_sign_inline(v, v + 3, v + 6, v + 9, v + 12, s) ||
        insphere_sign_fixed(v, v + 3, v + 6, v + 9, v + 12, s) ||
        insphere_sign_fixed_wide(v, v + 3, v + 6, v + 9, v + 12, s))
        return s;
    if (!FULL) return kUndecided;
    s = insphere_exact_sign<40, 96>(v, v + 3, v + 6, v + 9, v + 12);
    if (s == kOverflow) s = insphere_exact_sign<200, 1024>(v, v + 3, v + 6, v + 9, v + 12);
    return s == kOverflow ? 0 : s;
}

// The sign of cmg::robust::insphere(a, b, c, d, e) as shipped: the naive FP64 determinant
// (predicates.cpp:4063-4106) when nonzero, else the exact sign (or kUndecided, FULL =
// false only).
template <bool FULL>
__device__ __forceinline__ int is_compat(const double* v) {
    const double det = is_naive(v, nullptr);
    if (det > 0.0) return 1;
    if (det < 0.0) return -1;
    return is_exact_slow<FULL>(v);
}

// The Delaunay cavity of each query point p (incremental.cpp:148-161, the flood the
// insertion performs before retriangulating): from its located tet t0, breadth first over
// neighbours k = 0..3, every tet not yet in the cavity whose circumsphere contains p
// (robust::insphere(v0, v1, v2, v3, p) > 0) joins it. The cavity of query i is written in
// that discovery order to cav[i*cap ...]; count[i] = its size, or -1 when it has more than
// cap tets (the first cap are written). ctr[0] counts queries with a tet or neighbour index
// out of range (count -1 as well); a query whose located tet is -1 gets count 0. Pass
// FULL = false leaves count -2 on the queries whose insphere tests need expansions
// (counted in ctr[1]); pass FULL = true redoes exactly those.
template <bool FULL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void cavity_kernel(
    const double* xyz, long long npts,
                                                     const int* tets, const int* nbrs,
                                                     long long ntets, const double* q,
                                                     long long nq, const int* located,
                                                     int cap, int* count, int* cav,
                                                     unsigned long long* ctr) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nq;
         i += (long long)gridDim.x * blockDim.x) {
        if (FULL && count[i] != -2) continue;
        const int t0 = located[i];
        if (t0 < 0) {
            count[i] = 0;
            continue;
        }
        int* list = cav + (size_t)i * cap;
        bool bad = t0 >= ntets, full = false, undecided = false;
        int n = 0;
        if (!bad) list[n++] = t0;
        double v[15];
        v[12] = q[3 * i];
        v[13] = q[3 * i + 1];
        v[14] = q[3 * i + 2];
        for (int h = 0; h < n && !bad && !full && !undecided; ++h) {
            const int4 nb = *reinterpret_cast<const int4*>(nbrs + 4 * (long long)list[h]);
            const int nk[4] = {nb.x, nb.y, nb.z, nb.w};
            for (int k = 0; k < 4 && !bad && !full && !undecided; ++k) {
                const int m = nk[k];
                if (m < 0) continue;
                if (m >= ntets) {
                    bad = true;
                    break;
                }
                bool in = false;
                for (int e = 0; e < n && !in; ++e) in = list[e] == m;
                if (in) continue;
                const int4 tv = *reinterpret_cast<const int4*>(tets + 4 * (long long)m);
                if (!tet_ok(tv, npts)) {
                    bad = true;
                    break;
                }
                const int id[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int c = 0; c < 3; ++c) v[3 * a + c] = xyz[3 * (long long)id[a] + c];
                const int sg = is_compat<FULL>(v);
                if (sg == kUndecided) {
                    undecided = true;
                } else if (sg > 0) {
                    if (n == cap) full = true;
                    else list[n++] = m;
                }
            }
        }
        if (bad) atomicAdd(&ctr[0], 1ull);
        if (undecided && !bad) atomicAdd(&ctr[1], 1ull);
        count[i] = bad || full ? -1 : (undecided ? -2 : n);
    }
}

}  // namespace

long long cavities(const double* d_xyz, long long npts, const int* d_tets, const int* d_nbrs,
                   long long ntets, const double* d_q, long long nq, const int* d_located,
                   int cap, int* d_count, int* d_cav, hipStream_t st) {
    if (nq < 0 || ntets < 0 || npts < 0) throw InvalidInput("cavity: negative count");
    if (cap < 1) throw InvalidInput("cavity: cap must be >= 1");
    if (nq == 0) return 0;
    if (ntets < 1) throw InvalidInput("cavity: no tetrahedra");
    if (ntets >= (1LL << 31)) throw InvalidInput("cavity: more than 2^31 - 1 tetrahedra");
    if (!d_xyz || !d_tets || !d_nbrs || !d_q || !d_located || !d_count || !d_cav)
        throw InvalidInput("cavity: null pointer");
    if (reinterpret_cast<uintptr_t>(d_tets) % 16 || reinterpret_cast<uintptr_t>(d_nbrs) % 16)
        throw InvalidInput("cavity: tets and nbrs must be 16-byte aligned");
    ScratchLease lease;
    auto* ctr = static_cast<unsigned long long*>(lease.s->get(256));
    CMG_HIP(hipMemsetAsync(ctr, 0, 2 * sizeof(unsigned long long), st));
    {
        TimedLaunch tl("cavity", st);
        cavity_kernel<false><<<grid_for((size_t)nq, 256, 8), 256, 0, st>>>(
            d_xyz, npts, d_tets, d_nbrs, ntets, d_q, nq, d_located, cap, d_count, d_cav, ctr);
        CMG_HIP_LAUNCH();
    }
    unsigned long long c[2] = {0, 0};
    CMG_HIP(hipMemcpyAsync(c, ctr, sizeof c, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    if (c[1]) {  // queries whose insphere tests need expansions: redone with them
        cavity_kernel<true><<<grid_for((size_t)nq, 256, 2), 256, 0, st>>>(
            d_xyz, npts, d_tets, d_nbrs, ntets, d_q, nq, d_located, cap, d_count, d_cav, ctr);
        CMG_HIP_LAUNCH();
        CMG_HIP(hipMemcpyAsync(c, ctr, sizeof c, hipMemcpyDeviceToHost, st));
        CMG_HIP(hipStreamSynchronize(st));
    }
    return (long long)c[0];
}

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
