// quality.hip — the mesh quality report (src/quality/metrics.cpp:81-169) as one scan over
// the tetrahedra plus a selection of the worst ones.
//
// quality_kernel gives each lane one tet at a time (grid-stride): it gathers the four
// points (4 x 24 B) and evaluates, in the reference's operation order, the volume
// (metrics.cpp:36-40), the shortest edge (:43-49), the circumradius
// (include/cmg/core/circumcenter.hpp:15-46) and radius-edge ratio, and the six dihedral
// angles (:52-77), then accumulates min / max / sum / count and the 18-bucket dihedral
// histogram (LDS atomics) per workgroup. The ratio of every tet (DBL_MAX for degenerate
// ones, as the reference's sentinel) is written for the worst-tets selection: for up to 64
// worst tets, each workgroup selects its chunk's largest and a stable descending radix
// sort orders the candidates; otherwise all (ratio, index) pairs are sorted. Either way
// equal ratios come out by ascending index.
//
// Per tet: 16 B of indices and 96 B of gathered points in, 8 B out; ~300 FP64
// operations including 6 acos, 7 sqrt and 4 divisions: FP64 VALU bound once the points
// are cached. Min, max, volumes, ratios and histogram are bit-identical to the
// reference's sequential loop; the sums are reduced per workgroup and then in workgroup
// order, so they differ from its left-to-right sums by rounding only; acos is the device
// library's (within a few ulp of the host's).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <cfloat>
#include <climits>
#include <cmath>
#include <vector>

#include "device_common.hpp"
#include "quality_internal.hpp"

namespace cmg::hip {
namespace {

constexpr double kPiQ = 3.14159265358979323846;
constexpr int kBuckets = 18;

struct Partial {
    double min_vol, max_vol, sum_vol;
    double min_re, max_re, sum_re;
    double min_dih, max_dih, sum_dih;
    unsigned long long re_count;
    unsigned hist[kBuckets];
};

struct V3 {
    double x, y, z;
};
__device__ __forceinline__ V3 vsub(const double* a, const double* b) {
    return {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
}
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double vnorm(V3 a) { return sqrt(vdot(a, a)); }

// circumcenter.hpp:15-46: false for a near-degenerate tet or a non-finite result
__device__ bool circumradius(const double* a, const double* b, const double* c,
                             const double* d, double& radius) {
    const double B[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double C[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double D[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
    const double CD[3] = {C[1] * D[2] - C[2] * D[1], C[2] * D[0] - C[0] * D[2],
                          C[0] * D[1] - C[1] * D[0]};
    const double DB[3] = {D[1] * B[2] - D[2] * B[1], D[2] * B[0] - D[0] * B[2],
                          D[0] * B[1] - D[1] * B[0]};
    const double BC[3] = {B[1] * C[2] - B[2] * C[1], B[2] * C[0] - B[0] * C[2],
                          B[0] * C[1] - B[1] * C[0]};
    const double denom = 2.0 * (B[0] * CD[0] + B[1] * CD[1] + B[2] * CD[2]);
    const double nB = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
    const double nC = C[0] * C[0] + C[1] * C[1] + C[2] * C[2];
    const double nD = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    double m = nB;  // std::max({nB, nC, nD, 1e-300})
    if (m < nC) m = nC;
    if (m < nD) m = nD;
    if (m < 1e-300) m = 1e-300;
    const double scale = sqrt(m);
    if (fabs(denom) < 1e-12 * scale * scale * scale) return false;
    double num[3];
    for (int i = 0; i < 3; i++) num[i] = (nB * CD[i] + nC * DB[i] + nD * BC[i]) / denom;
    const double cx = a[0] + num[0], cy = a[1] + num[1], cz = a[2] + num[2];
    radius = sqrt(num[0] * num[0] + num[1] * num[1] + num[2] * num[2]);
    return isfinite(cx) && isfinite(cy) && isfinite(cz) && isfinite(radius);
}

__global__ __launch_bounds__(256) void quality_kernel(const double* xyz, long long npts,
                                                      const int* tets, long long n,
                                                      double* ratio, Partial* part,
                                                      unsigned long long* bad) {
    __shared__ unsigned hist[kBuckets];
    if (threadIdx.x < kBuckets) hist[threadIdx.x] = 0;
    __syncthreads();
    double mnv = INFINITY, mxv = -INFINITY, sv = 0, mnr = INFINITY, mxr = -INFINITY, sr = 0;
    double mnd = INFINITY, mxd = -INFINITY, sd = 0;
    unsigned long long rc = 0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int4 q = *reinterpret_cast<const int4*>(tets + 4 * t);
        const int v[4] = {q.x, q.y, q.z, q.w};
        bool ok_idx = true;
        for (int j = 0; j < 4; ++j) ok_idx &= v[j] >= 0 && v[j] < npts;
        if (!ok_idx) {
            ratio[t] = -1.0;  // below every real ratio: never listed among the worst
            atomicAdd(bad, 1ull);
            continue;
        }
        double P[4][3];
        for (int j = 0; j < 4; ++j)
            for (int a = 0; a < 3; ++a) P[j][a] = xyz[3 * (long long)v[j] + a];
        // metrics.cpp:36-40
        const V3 ab = vsub(P[1], P[0]), ac = vsub(P[2], P[0]), ad = vsub(P[3], P[0]);
        const double vol = fabs(vdot(ab, vcross(ac, ad))) / 6.0;
        mnv = fmin(mnv, vol);
        mxv = fmax(mxv, vol);
        sv += vol;
        // metrics.cpp:43-49, 109-131
        double se = INFINITY;
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j) {
                const double e = vnorm(vsub(P[j], P[i]));
                se = e < se ? e : se;
            }
        double radius = 0;
        const bool okc = circumradius(P[0], P[1], P[2], P[3], radius);
        if (okc && se > 0.0 && isfinite(radius)) {
            const double re = radius / se;
            ratio[t] = re;
            mnr = fmin(mnr, re);
            mxr = fmax(mxr, re);
            sr += re;
            ++rc;
        } else {
            ratio[t] = DBL_MAX;
        }
        // metrics.cpp:52-77
        V3 nrm[4];
        for (int e = 0; e < 4; ++e) {
            const int a = (e + 1) & 3, b = (e + 2) & 3, c = (e + 3) & 3;
            V3 nv = vcross(vsub(P[b], P[a]), vsub(P[c], P[a]));
            if (vdot(nv, vsub(P[e], P[a])) > 0.0) nv = {-nv.x, -nv.y, -nv.z};
            nrm[e] = nv;
        }
        // edge (i, j) -> the two faces opposite its other vertices k < l
        constexpr int kk[6] = {2, 1, 1, 0, 0, 0};
        constexpr int ll[6] = {3, 3, 2, 3, 2, 1};
#pragma unroll
        for (int q2 = 0; q2 < 6; ++q2) {
            const V3 A = nrm[kk[q2]], Bn = nrm[ll[q2]];
            const double na = vnorm(A), nb = vnorm(Bn);
            double ang;
            if (na == 0.0 || nb == 0.0) {
                ang = 0.0;
            } else {
                double cd = vdot(A, Bn) / (na * nb);
                cd = cd < -1.0 ? -1.0 : (1.0 < cd ? 1.0 : cd);
                ang = (kPiQ - acos(cd)) * 180.0 / kPiQ;
            }
            mnd = fmin(mnd, ang);
            mxd = fmax(mxd, ang);
            sd += ang;
            int bk = (int)floor(ang / 10.0);
            bk = bk < 0 ? 0 : (bk > kBuckets - 1 ? kBuckets - 1 : bk);
            atomicAdd(&hist[bk], 1u);
        }
    }
    // workgroup reduction (tree order)
    __shared__ double red[9][256];
    __shared__ unsigned long long redc[256];
    const double mine[9] = {mnv, mxv, sv, mnr, mxr, sr, mnd, mxd, sd};
    for (int k = 0; k < 9; ++k) red[k][threadIdx.x] = mine[k];
    redc[threadIdx.x] = rc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const int u = threadIdx.x, w = threadIdx.x + o;
            for (int k = 0; k < 9; k += 3) {
                red[k][u] = fmin(red[k][u], red[k][w]);
                red[k + 1][u] = fmax(red[k + 1][u], red[k + 1][w]);
                red[k + 2][u] = red[k + 2][u] + red[k + 2][w];
            }
            redc[u] += redc[w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Partial& p = part[blockIdx.x];
        p.min_vol = red[0][0], p.max_vol = red[1][0], p.sum_vol = red[2][0];
        p.min_re = red[3][0], p.max_re = red[4][0], p.sum_re = red[5][0];
        p.min_dih = red[6][0], p.max_dih = red[7][0], p.sum_dih = red[8][0];
        p.re_count = redc[0];
    }
    if (threadIdx.x < kBuckets) part[blockIdx.x].hist[threadIdx.x] = hist[threadIdx.x];
}

// Per-workgroup selection of the k largest ratios of a chunk of kChunk tets, largest first,
// equal ratios by ascending index (k rounds of a workgroup arg-max with the winner taken
// out). Every tet of the global k largest is among its chunk's k largest under this
// order, and chunks are in index order, so a stable descending sort of the candidates
// yields the global list.
constexpr int kTopPer = 32;               // ratios per lane
constexpr int kChunk = 256 * kTopPer;     // tets per workgroup
constexpr int kMaxTopK = 64;

// (ra, ia) comes before (rb, ib) in the worst list: larger ratio, then smaller index
__device__ __forceinline__ bool worse(double ra, long long ia, double rb, long long ib) {
    return ra > rb || (ra == rb && ia < ib);
}

__global__ __launch_bounds__(256) void topk_kernel(const double* ratio, long long n, int k,
                                                   double* cand_r, int* cand_i) {
    const long long c0 = (long long)blockIdx.x * kChunk;
    const int tid = (int)threadIdx.x;
    double r[kTopPer];
#pragma unroll
    for (int q = 0; q < kTopPer; ++q) {
        const long long t = c0 + (long long)q * 256 + tid;
        r[q] = t < n ? ratio[t] : -INFINITY;
    }
    // each lane's best of its own values, rescanned only when that best leaves the chunk
    double br;
    long long bi;
    auto scan = [&]() {
        br = -INFINITY;
        bi = LLONG_MAX;
#pragma unroll
        for (int q = 0; q < kTopPer; ++q) {
            const long long t = c0 + (long long)q * 256 + tid;
            if (worse(r[q], t, br, bi)) br = r[q], bi = t;
        }
    };
    scan();
    __shared__ double sr[256];
    __shared__ long long si[256];
    __shared__ double s_wr;
    __shared__ long long s_wi;
    for (int round = 0; round < k; ++round) {
        // a tree over the lanes' bests, lane t taking lane t + o's when it is worse, for
        // o = 128, 64 through LDS and 32 .. 1 by shuffles inside the first wavefront
        sr[tid] = br;
        si[tid] = bi;
        __syncthreads();
        double xr = br;
        long long xi = bi;
        if (tid < 128) {
            const double orr = sr[tid + 128];
            const long long oi = si[tid + 128];
            if (worse(orr, oi, xr, xi)) xr = orr, xi = oi;
            sr[tid] = xr;
            si[tid] = xi;
        }
        __syncthreads();
        if (tid < 64) {
            const double orr = sr[tid + 64];
            const long long oi = si[tid + 64];
            if (worse(orr, oi, xr, xi)) xr = orr, xi = oi;
            for (int o = 32; o > 0; o >>= 1) {
                const double dr = __shfl_down(xr, o);
                const long long di = __shfl_down(xi, o);
                if (tid < o && worse(dr, di, xr, xi)) xr = dr, xi = di;
            }
            if (tid == 0) {
                s_wr = xr;
                s_wi = xi;
            }
        }
        __syncthreads();
        const double wr = s_wr;
        const long long wi = s_wi;
        if (tid == 0) {
            cand_r[(long long)blockIdx.x * k + round] = wr;
            cand_i[(long long)blockIdx.x * k + round] = wi < n ? (int)wi : -1;
        }
        // the winner leaves the chunk; its lane rescans
        if (wi < n && (int)((wi - c0) % 256) == tid) {
            const int q = (int)((wi - c0) / 256);
#pragma unroll
            for (int u = 0; u < kTopPer; ++u)
                if (u == q) r[u] = -INFINITY;
            scan();
        }
        __syncthreads();  // sr, si and s_wr / s_wi are rewritten next round
    }
}

__global__ void iota_kernel(int* v, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        v[i] = (int)i;
}

}  // namespace

QualityReport quality_report(const double* d_xyz, long long npts, const int* d_tets,
                             long long n, int worst_count, std::vector<int>* worst,
                             hipStream_t st) {
    QualityReport r;
    r.num_tets = n;
    if (n < 0 || npts < 0) throw InvalidInput("quality: negative count");
    if (n >= (1LL << 31)) throw InvalidInput("quality: more than 2^31 - 1 tets");
    if (n == 0) {
        if (worst) worst->clear();
        return r;
    }
    if (!d_xyz || !d_tets) throw InvalidInput("quality: null pointer");
    if (reinterpret_cast<uintptr_t>(d_tets) % 16)
        throw InvalidInput("quality: tets must be 16-byte aligned");
    // 3 workgroups of 256 per CU: as many waves as the scan kernel's registers let run
    const unsigned nb = grid_for((size_t)n, 256, 12);
    const bool want_worst = worst && worst_count != 0;
    // small lists: per-workgroup selection, then a sort of the candidates only
    const bool select = want_worst && worst_count > 0 && worst_count <= kMaxTopK;
    const long long nchunk = (n + kChunk - 1) / kChunk;
    const long long nsort = select ? nchunk * worst_count : n;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t sort_tmp = 0;
    if (want_worst)
        CMG_HIP(rocprim::radix_sort_pairs_desc(nullptr, sort_tmp, (const double*)nullptr,
                                               (double*)nullptr, (const int*)nullptr,
                                               (int*)nullptr, (size_t)nsort, 0, 64, st));
    const size_t bytes = al(sizeof(double) * n) + al(sizeof(Partial) * nb) + al(8) +
                         (want_worst ? 2 * al(sizeof(double) * nsort) +
                                           2 * al(sizeof(int) * nsort) + al(sort_tmp)
                                     : 0);
    ScratchLease lease;
    char* base = static_cast<char*>(lease.s->get(bytes));
    size_t off = 0;
    auto take = [&](size_t b) {
        char* p = base + off;
        off += al(b);
        return p;
    };
    double* ratio = reinterpret_cast<double*>(take(sizeof(double) * n));
    Partial* part = reinterpret_cast<Partial*>(take(sizeof(Partial) * nb));
    auto* bad = reinterpret_cast<unsigned long long*>(take(8));
    CMG_HIP(hipMemsetAsync(bad, 0, 8, st));
    {
        TimedLaunch tl("quality", st);
        quality_kernel<<<nb, 256, 0, st>>>(d_xyz, npts, d_tets, n, ratio, part, bad);
        CMG_HIP_LAUNCH();
    }
    std::vector<Partial> hp(nb);
    unsigned long long hbad = 0;
    CMG_HIP(hipMemcpyAsync(hp.data(), part, sizeof(Partial) * nb, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipMemcpyAsync(&hbad, bad, 8, hipMemcpyDeviceToHost, st));
    long long k = 0;
    std::vector<int> top;
    if (want_worst) {
        k = worst_count < 0 ? n : std::min<long long>(worst_count, n);
        double* keys_in = select ? reinterpret_cast<double*>(take(sizeof(double) * nsort))
                                 : ratio;
        double* keys_out = reinterpret_cast<double*>(take(sizeof(double) * nsort));
        int* idx_in = reinterpret_cast<int*>(take(sizeof(int) * nsort));
        int* idx_out = reinterpret_cast<int*>(take(sizeof(int) * nsort));
        void* tmp = take(sort_tmp);
        if (select) {
            topk_kernel<<<(unsigned)nchunk, 256, 0, st>>>(ratio, n, worst_count, keys_in,
                                                          idx_in);
        } else {
            iota_kernel<<<grid_for((size_t)n, 256, 4), 256, 0, st>>>(idx_in, n);
        }
        CMG_HIP_LAUNCH();
        CMG_HIP(rocprim::radix_sort_pairs_desc(tmp, sort_tmp, keys_in, keys_out, idx_in,
                                               idx_out, (size_t)nsort, 0, 64, st));
        top.resize((size_t)k);
        CMG_HIP(hipMemcpyAsync(top.data(), idx_out, sizeof(int) * k, hipMemcpyDeviceToHost,
                               st));
    }
    CMG_HIP(hipStreamSynchronize(st));
    if (hbad)
        throw InvalidInput("quality: " + std::to_string(hbad) +
                           " tets have a vertex index outside the point array");
    // workgroup partials in workgroup order (metrics.cpp:133-155 semantics)
    double mnv = INFINITY, mxv = -INFINITY, sv = 0, mnr = INFINITY, mxr = -INFINITY, sr = 0;
    double mnd = INFINITY, mxd = -INFINITY, sd = 0;
    unsigned long long rc = 0;
    for (const Partial& p : hp) {
        mnv = std::fmin(mnv, p.min_vol), mxv = std::fmax(mxv, p.max_vol), sv += p.sum_vol;
        mnr = std::fmin(mnr, p.min_re), mxr = std::fmax(mxr, p.max_re), sr += p.sum_re;
        mnd = std::fmin(mnd, p.min_dih), mxd = std::fmax(mxd, p.max_dih), sd += p.sum_dih;
        rc += p.re_count;
        for (int b = 0; b < kBuckets; ++b) r.dihedral_histogram[b] += (int)p.hist[b];
    }
    r.min_volume = mnv, r.max_volume = mxv, r.total_volume = sv;
    if (rc) {
        r.min_radius_edge = mnr, r.max_radius_edge = mxr;
        r.mean_radius_edge = sr / (double)rc;
    }
    r.min_dihedral = mnd, r.max_dihedral = mxd;
    r.mean_dihedral = sd / (double)(6 * n);
    if (worst) *worst = std::move(top);
    return r;
}

}  // namespace cmg::hip
