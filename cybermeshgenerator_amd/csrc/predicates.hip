// predicates.hip — batched orient3d / insphere signs on gfx950.
//
// Pass 1 (filter kernel, HBM-bound): each thread streams two tuples per iteration with
// 16-byte non-temporal loads from the structure-of-arrays planes and decides a sign
// through a staged filter:
//   exact mode  : FP32 filter with a forward error bound (input rounding + Shewchuk's
//                 o3derrboundA/isperrboundA evaluation bound at u = 2^-24), then the
//                 FP64 naive determinant with the FP64 semi-static bound
//                 (predicates.cpp:2235-2245, :4121-4161);
//   compat mode : the naive FP64 determinant in the reference's operation order, whose
//                 sign is the answer whenever it is nonzero (what cmg::robust returns
//                 with its zero static filter, predicates.cpp:2230-2231, :4113).
// Lanes still undecided are compacted with a wavefront ballot and a popcount prefix into
// the wave's own queue segment (no atomics); each wave records its segment length.
// Pass 2 (exact kernel): one 64-lane block per segment (lengths are read on the device,
// so the host never waits), exact expansion arithmetic per lane; a lane whose
// expansion outgrows the small private arrays is re-queued for a wide-capacity pass.
//
// This TU is compiled with -ffp-contract=off: the naive determinants must round exactly
// as the reference's (no FMA contraction), and the error bounds assume it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cmg_hip_internal.hpp"
#include "device_common.hpp"
#include "expansion.hpp"
#include "pred_naive.hpp"

namespace cmg::hip {

namespace {

constexpr int kUndecided = 3;

typedef double d2v __attribute__((ext_vector_type(2)));

// ---- FP32 stage ------------------------------------------------------------------
//
// Inputs are rounded to float (|f - x| <= u|x|, u = 2^-24) and the naive formula is
// evaluated in FP32. With M_a = max |rounded coordinate| on axis a (orient3d) or over all
// coordinates (insphere), the true determinant D satisfies
//   orient3d : |D32 - D| <= (7u + 56u^2) P32 + 145 u Mx My Mz
//   insphere : |D32 - D| <= (16u + 224u^2) P32 + 11521 u M^5
// where P32 is the FP32 permanent (Shewchuk's evaluation bound at FP32 precision) and the
// second term bounds how far rounding the inputs moves the exact determinant (6 resp.
// 72 monomials of degree 3 resp. 5 in differences bounded by 2M, each difference moved
// by at most 2uM). The coefficients below are rounded up so that computing the bound in
// FP32 cannot undercut it. The range checks keep every product normal and finite, so
// the relative error model holds; outside them the lane goes to the FP64 stage.

__device__ __forceinline__ int o3_fp32(const double* v) {
    float f[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) f[q] = (float)v[q];
    float mx = fmaxf(fmaxf(__builtin_fabsf(f[0]), __builtin_fabsf(f[3])),
                     fmaxf(__builtin_fabsf(f[6]), __builtin_fabsf(f[9])));
    float my = fmaxf(fmaxf(__builtin_fabsf(f[1]), __builtin_fabsf(f[4])),
                     fmaxf(__builtin_fabsf(f[7]), __builtin_fabsf(f[10])));
    float mz = fmaxf(fmaxf(__builtin_fabsf(f[2]), __builtin_fabsf(f[5])),
                     fmaxf(__builtin_fabsf(f[8]), __builtin_fabsf(f[11])));
    const float lo = 0x1p-30f, hi = 0x1p30f;
    if (!(mx >= lo && mx <= hi && my >= lo && my <= hi && mz >= lo && mz <= hi))
        return kUndecided;
    float adx = f[0] - f[9], ady = f[1] - f[10], adz = f[2] - f[11];
    float bdx = f[3] - f[9], bdy = f[4] - f[10], bdz = f[5] - f[11];
    float cdx = f[6] - f[9], cdy = f[7] - f[10], cdz = f[8] - f[11];
    float bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    float cdxady = cdx * ady, adxcdy = adx * cdy;
    float adxbdy = adx * bdy, bdxady = bdx * ady;
    float det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    float perm = (__builtin_fabsf(bdxcdy) + __builtin_fabsf(cdxbdy)) * __builtin_fabsf(adz) +
                 (__builtin_fabsf(cdxady) + __builtin_fabsf(adxcdy)) * __builtin_fabsf(bdz) +
                 (__builtin_fabsf(adxbdy) + __builtin_fabsf(bdxady)) * __builtin_fabsf(cdz);
    float bound = 0x1p-21f * perm + (146.0f * 0x1p-24f) * (mx * my * mz);
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return kUndecided;
}

__device__ __forceinline__ int is_fp32(const double* v) {
    float f[15];
    float m = 0.0f;
#pragma unroll
    for (int q = 0; q < 15; ++q) {
        f[q] = (float)v[q];
        m = fmaxf(m, __builtin_fabsf(f[q]));
    }
    if (!(m >= 0x1p-20f && m <= 0x1p20f)) return kUndecided;
    float aex = f[0] - f[12], bex = f[3] - f[12], cex = f[6] - f[12], dex = f[9] - f[12];
    float aey = f[1] - f[13], bey = f[4] - f[13], cey = f[7] - f[13], dey = f[10] - f[13];
    float aez = f[2] - f[14], bez = f[5] - f[14], cez = f[8] - f[14], dez = f[11] - f[14];
    float aexbey = aex * bey, bexaey = bex * aey, ab = aexbey - bexaey;
    float bexcey = bex * cey, cexbey = cex * bey, bc = bexcey - cexbey;
    float cexdey = cex * dey, dexcey = dex * cey, cd = cexdey - dexcey;
    float dexaey = dex * aey, aexdey = aex * dey, da = dexaey - aexdey;
    float aexcey = aex * cey, cexaey = cex * aey, ac = aexcey - cexaey;
    float bexdey = bex * dey, dexbey = dex * bey, bd = bexdey - dexbey;
    float abc = aez * bc - bez * ac + cez * ab;
    float bcd = bez * cd - cez * bd + dez * bc;
    float cda = cez * da + dez * ac + aez * cd;
    float dab = dez * ab + aez * bd + bez * da;
    float alift = aex * aex + aey * aey + aez * aez;
    float blift = bex * bex + bey * bey + bez * bez;
    float clift = cex * cex + cey * cey + cez * cez;
    float dlift = dex * dex + dey * dey + dez * dez;
    float det = (dlift * abc - clift * dab) + (blift * cda - alift * bcd);
    float aezp = __builtin_fabsf(aez), bezp = __builtin_fabsf(bez), cezp = __builtin_fabsf(cez),
          dezp = __builtin_fabsf(dez);
    float p_ab = __builtin_fabsf(aexbey) + __builtin_fabsf(bexaey);
    float p_bc = __builtin_fabsf(bexcey) + __builtin_fabsf(cexbey);
    float p_cd = __builtin_fabsf(cexdey) + __builtin_fabsf(dexcey);
    float p_da = __builtin_fabsf(dexaey) + __builtin_fabsf(aexdey);
    float p_ac = __builtin_fabsf(aexcey) + __builtin_fabsf(cexaey);
    float p_bd = __builtin_fabsf(bexdey) + __builtin_fabsf(dexbey);
    float perm = (p_cd * bezp + p_bd * cezp + p_bc * dezp) * alift +
                 (p_da * cezp + p_ac * dezp + p_cd * aezp) * blift +
                 (p_ab * dezp + p_bd * aezp + p_da * bezp) * clift +
                 (p_bc * aezp + p_ac * bezp + p_ab * cezp) * dlift;
    float m2 = m * m;
    float bound = (17.0f * 0x1p-24f) * perm + (11600.0f * 0x1p-24f) * (m2 * m2 * m);
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return kUndecided;
}

// ---- staged decision per tuple -----------------------------------------------------

__device__ __forceinline__ int sgn(double x) { return x > 0.0 ? 1 : (x < 0.0 ? -1 : 0); }

constexpr double kEps = 0x1p-53;
constexpr double kO3ErrA = (7.0 + 56.0 * kEps) * kEps;     // predicates.cpp:619
constexpr double kIspErrA = (16.0 + 224.0 * kEps) * kEps;  // predicates.cpp:625

// The FP64 semi-static stage of exact mode: the naive determinant and its error bound.
template <int NPTS>
__device__ __forceinline__ int decide_fp64(const double* v) {
    double perm;
    const double det = NPTS == 4 ? o3_naive(v, &perm) : is_naive(v, &perm);
    const double eb = (NPTS == 4 ? kO3ErrA : kIspErrA) * perm;
    if (det > eb || -det > eb) return sgn(det);
    return kUndecided;
}

// Returns the sign, or kUndecided; `stage` records which stage decided (1 fp32, 2 fp64).
// FP32_ONLY (exact mode) stops after the FP32 stage: the caller runs the FP64 stage later,
// on full wavefronts of the lanes the FP32 stage left.
template <int NPTS, int MODE, bool FP32_ONLY = false>
__device__ __forceinline__ int decide(const double* v, int& stage) {
    if (MODE == kModeExact) {
        int s = NPTS == 4 ? o3_fp32(v) : is_fp32(v);
        if (s != kUndecided) {
            stage = 1;
            return s;
        }
        if (FP32_ONLY) return kUndecided;
        double perm;
        double det = NPTS == 4 ? o3_naive(v, &perm) : is_naive(v, &perm);
        double eb = (NPTS == 4 ? kO3ErrA : kIspErrA) * perm;
        stage = 2;
        if (det > eb || -det > eb) return sgn(det);
        return kUndecided;
    } else {
        double det = NPTS == 4 ? o3_naive(v, nullptr) : is_naive(v, nullptr);
        stage = 2;
        if (det > 0.0 || det < -0.0) return sgn(det);
        return kUndecided;
    }
}

// Register-only exact stage run inside the streaming kernels for the lanes the filters
// leave undecided. orient3d: fixed point when the coordinate differences are exact and
// span at most 31 bits (exact zeros and small determinants on a common grid; 12 integer
// products, where the exact sum of the 24 FMA-split triple-product terms took several
// error-free passes), else double-word arithmetic with a rigorous bound (nonzero
// determinants); the lanes neither decides (in practice exact zeros with inexact or wider
// differences) are queued for the exact pass, whose 62- and 126-bit fixed-point stages
// resolve them. insphere: fixed point with 62-bit differences, else the double-word
// bound; the rest is queued for the 126-bit stage. (For insphere a cheaper 31-bit first
// stage does not pay: a wavefront runs every stage any of its lanes needs, the drained
// lanes mix narrow and wide tuples, and the extra live registers spill.)
template <int NPTS>
__device__ __forceinline__ bool decide_inline_exact(const double* v, int& s) {
    if constexpr (NPTS == 4) {
        if (orient3d_sign_fixed31(v, v + 3, v + 6, v + 9, s)) return true;
        return orient3d_sign_dw(v, v + 3, v + 6, v + 9, s);
    } else {
        return insphere_sign_fixed(v, v + 3, v + 6, v + 9, v + 12, s) ||
               insphere_sign_dw(v, v + 3, v + 6, v + 9, v + 12, s);
    }
}

// ---- tuple sources -------------------------------------------------------------------

template <int NPTS>
struct SoaSource {
    const double* soa;
    size_t ld;
    __device__ __forceinline__ void load(size_t i, double* v) const {
#pragma unroll
        for (int q = 0; q < NPTS * 3; ++q) v[q] = soa[(size_t)q * ld + i];
    }
};

template <int NPTS>
struct IndexedSource {  // NPTS-point tuples as index records into AoS points
    const double* xyz;
    const int* idx;
    __device__ __forceinline__ void load(size_t i, double* v) const {
#pragma unroll
        for (int s = 0; s < NPTS; ++s) {
            const size_t id = (size_t)idx[(size_t)NPTS * i + s];
#pragma unroll
            for (int a = 0; a < 3; ++a) v[3 * s + a] = xyz[3 * id + a];
        }
    }
};

// ---- wave helpers ----------------------------------------------------------------------

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Per-wave compaction: every wave owns queue segment [wave*seg, (wave+1)*seg) and a
// wave-uniform fill count `off`; flagged lanes write their value at off + (number of
// flagged lanes below them). No atomics: the exact pass reads each segment's count.
__device__ __forceinline__ void wave_push(bool flag, uint32_t value, uint32_t* seg_base,
                                          unsigned& off) {
    unsigned long long mask = __ballot(flag);
    if (flag) {
        unsigned long long below = mask & ((1ull << lane_id()) - 1ull);
        seg_base[off + __popcll(below)] = value;
    }
    off += (unsigned)__popcll(mask);
}

__device__ __forceinline__ unsigned wave_sum(unsigned v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ---- kernels -------------------------------------------------------------------------
//
// Grid-stride kernels run the same number of iterations in every lane (`iters`), so the
// ballots and the per-wave fill count stay wave-uniform.

// Intra-wave ordering of LDS traffic: the staging ring below is written by some lanes
// and read by others of the same wave, with no workgroup barrier (waves progress
// independently).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-wave LDS ring of tuples the filters left undecided (orient3d): they are pushed as
// the stream goes by and handed to the register-only exact stage 64 at a time, so that
// stage runs on full wavefronts instead of on the ~10 % of lanes that need it.
// Ring entries: orient3d drains once per iteration (>= 63 pending + 128 pushed); insphere,
// whose 15 coordinates take more LDS, drains after each of its two pushes (>= 63 + 64).
template <int NPL>
constexpr unsigned ring_size() { return NPL == 12 ? 192 : 128; }

template <int NPL>
struct StageRing {
    static constexpr unsigned kRing = ring_size<NPL>();
    double* v;      // [NPL][kRing] for this wave
    uint32_t* idx;  // [kRing]
    unsigned head = 0, count = 0;  // wave-uniform

    __device__ __forceinline__ void push(bool flag, const double* t, uint32_t i) {
        const unsigned long long mask = __ballot(flag);
        if (flag) {
            unsigned pos = head + count + (unsigned)__popcll(mask & ((1ull << __lane_id()) - 1ull));
            if (pos >= kRing) pos -= kRing;
#pragma unroll
            for (int q = 0; q < NPL; ++q) v[q * kRing + pos] = t[q];
            idx[pos] = i;
        }
        count += (unsigned)__popcll(mask);
    }

    // Pops min(count, 64) entries, one per lane; returns whether this lane got one.
    __device__ __forceinline__ bool pop(double* t, uint32_t& i) {
        wave_sync();
        const unsigned take = count < 64 ? count : 64;
        const bool has = __lane_id() < take;
        if (has) {
            unsigned pos = head + __lane_id();
            if (pos >= kRing) pos -= kRing;
#pragma unroll
            for (int q = 0; q < NPL; ++q) t[q] = v[q * kRing + pos];
            i = idx[pos];
        }
        wave_sync();
        head += take;
        if (head >= kRing) head -= kRing;
        count -= take;
        return has;
    }
};

// Structure-of-arrays filter: two tuples per thread per iteration, 16-byte
// non-temporal loads (requires 16-byte aligned planes and an even leading dimension).
// Every output byte is written exactly once: by this kernel when a filter or the inline
// exact stage decides it, otherwise by the expansion pass.
template <int NPTS, int MODE, bool VEC>
__global__ __launch_bounds__(256) void filter_soa_kernel(const double* __restrict__ soa,
                                                         size_t n, size_t ld,
                                                         int8_t* __restrict__ out,
                                                         PredWorkspace ws, unsigned iters) {
    constexpr int NPL = NPTS * 3;
    constexpr bool STAGE = true;
    // exact mode: about 10 % of the lanes leave the FP32 stage undecided, so nearly every
    // wavefront iteration would run the FP64 stage for a handful of lanes, and on the
    // workloads it decides almost none of them. The streaming loop stops after the FP32
    // stage instead, and the ring drain runs the FP64 stage on full wavefronts of
    // undecided tuples before the exact stages.
    constexpr bool FP64_IN_DRAIN = MODE == kModeExact;
    // the global counters are only touched by the passes after this one
    if (blockIdx.x == 0 && threadIdx.x == 0) *ws.ctr = PredCounters{};
    constexpr unsigned kRing = ring_size<NPL>();
    __shared__ double s_ring[STAGE ? 4 : 1][STAGE ? NPL * kRing : 1];
    __shared__ uint32_t s_idx[STAGE ? 4 : 1][STAGE ? kRing : 1];
    unsigned n32 = 0, n64 = 0, nx = 0;
    const size_t npairs = (n + 1) / 2;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned wave = (unsigned)(tid / 64);
    uint32_t* seg = ws.queue + (size_t)wave * ws.seg;
    unsigned off = 0;
    StageRing<NPL> ring;
    ring.v = &s_ring[STAGE ? threadIdx.x / 64 : 0][0];
    ring.idx = &s_idx[STAGE ? threadIdx.x / 64 : 0][0];
    // the inline exact stage over up to 64 ring entries
    auto drain = [&]() {
        double t[NPL];
        uint32_t i = 0;
        const bool has = ring.pop(t, i);
        int s = kUndecided;
        bool dec64 = false;
        if (FP64_IN_DRAIN && has) {
            s = decide_fp64<NPTS>(t);
            dec64 = s != kUndecided;
        }
        const bool dec = has && !dec64 && decide_inline_exact<NPTS>(t, s);
        if (dec || dec64) out[i] = (int8_t)s;
        nx += dec;
        n64 += dec64;
        wave_push(has && !dec && !dec64, i, seg, off);
    };
    // orient3d: the next iteration's planes are loaded before this iteration's tuples are
    // decided and the ring drained, so the wave's loads stay in flight while it computes
    // (48 more VGPRs; insphere's kernel has no room for its 60)
    constexpr bool PREFETCH = VEC && NPTS == 4;
    auto full_pair = [&](size_t q) { return q < npairs && 2 * q + 1 < n; };
    d2v nxt[PREFETCH ? NPL : 1] = {};  // lanes without a full pair never read it
    if (PREFETCH && iters > 0 && full_pair(tid)) {
#pragma unroll
        for (int q = 0; q < NPL; ++q)
            nxt[q] = __builtin_nontemporal_load(
                reinterpret_cast<const d2v*>(soa + (size_t)q * ld + 2 * tid));
    }
    for (unsigned it = 0; it < iters; ++it) {
        const size_t p = tid + (size_t)it * stride;
        const bool valid = p < npairs;
        const size_t i0 = 2 * p;
        const bool two = valid && i0 + 1 < n;
        bool u0 = false, u1 = false;
        double v0[NPL], v1[NPL];
        d2v cur[PREFETCH ? NPL : 1];
        if (PREFETCH) {
#pragma unroll
            for (int q = 0; q < NPL; ++q) cur[q] = nxt[q];
            const size_t pn = p + stride;
            if (it + 1 < iters && full_pair(pn)) {
#pragma unroll
                for (int q = 0; q < NPL; ++q)
                    nxt[q] = __builtin_nontemporal_load(
                        reinterpret_cast<const d2v*>(soa + (size_t)q * ld + 2 * pn));
            }
        }
        if (valid) {
            if (PREFETCH && two) {
#pragma unroll
                for (int q = 0; q < NPL; ++q) {
                    v0[q] = cur[q].x;
                    v1[q] = cur[q].y;
                }
            } else if (VEC && two) {
#pragma unroll
                for (int q = 0; q < NPL; ++q) {
                    d2v x = __builtin_nontemporal_load(
                        reinterpret_cast<const d2v*>(soa + (size_t)q * ld + i0));
                    v0[q] = x.x;
                    v1[q] = x.y;
                }
            } else {
#pragma unroll
                for (int q = 0; q < NPL; ++q) {
                    v0[q] = soa[(size_t)q * ld + i0];
                    v1[q] = two ? soa[(size_t)q * ld + i0 + 1] : v0[q];
                }
            }
            int st0 = 0, st1 = 0;
            const int s0 = decide<NPTS, MODE, FP64_IN_DRAIN>(v0, st0);
            const int s1 = decide<NPTS, MODE, FP64_IN_DRAIN>(v1, st1);
            u0 = s0 == kUndecided;
            u1 = two && s1 == kUndecided;
            const int8_t o0 = (int8_t)s0, o1 = (int8_t)s1;
            if (two && !u0 && !u1 && ((reinterpret_cast<uintptr_t>(out + i0) & 1) == 0)) {
                uint16_t packed = (uint16_t)(uint8_t)o0 | ((uint16_t)(uint8_t)o1 << 8);
                __builtin_nontemporal_store(packed, reinterpret_cast<uint16_t*>(out + i0));
            } else {
                if (!u0) out[i0] = o0;
                if (two && !u1) out[i0 + 1] = o1;
            }
            n32 += (!u0 && st0 == 1) + (two && !u1 && st1 == 1);
            n64 += (!u0 && st0 == 2) + (two && !u1 && st1 == 2);
        }
        if (STAGE && NPTS == 4) {
            ring.push(u0, v0, (uint32_t)i0);
            ring.push(u1, v1, (uint32_t)(i0 + 1));
            while (ring.count >= 64) drain();
        } else if (STAGE) {
#pragma unroll 1
            for (int h = 0; h < 2; ++h) {
                ring.push(h ? u1 : u0, h ? v1 : v0, (uint32_t)(i0 + h));
                if (ring.count >= 64) drain();
            }
        } else {
            wave_push(u0, (uint32_t)i0, seg, off);
            wave_push(u1, (uint32_t)(i0 + 1), seg, off);
        }
    }
    if (STAGE)
        while (ring.count > 0) drain();
    n32 = wave_sum(n32);
    n64 = wave_sum(n64);
    nx = wave_sum(nx);
    if (lane_id() == 0) {
        ws.wave_q[wave] = off;
        ws.wave_32[wave] = n32;
        ws.wave_64[wave] = n64;
        ws.wave_x[wave] = nx;
    }
}

// Generic filter over any source (indexed form), one tuple per thread.
template <int NPTS, int MODE, class Src>
__global__ __launch_bounds__(256) void filter_kernel(Src src, size_t n,
                                                     int8_t* __restrict__ out,
                                                     PredWorkspace ws, unsigned iters) {
    unsigned n32 = 0, n64 = 0, nx = 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned wave = (unsigned)(tid / 64);
    uint32_t* seg = ws.queue + (size_t)wave * ws.seg;
    unsigned off = 0;
    for (unsigned it = 0; it < iters; ++it) {
        const size_t i = tid + (size_t)it * stride;
        bool u = false;
        if (i < n) {
            double v[NPTS * 3];
            src.load(i, v);
            int st = 0;
            int s = decide<NPTS, MODE>(v, st);
            if (s == kUndecided && decide_inline_exact<NPTS>(v, s)) st = 3;
            u = s == kUndecided;
            out[i] = u ? 0 : (int8_t)s;
            n32 += !u && st == 1;
            n64 += !u && st == 2;
            nx += !u && st == 3;
        }
        wave_push(u, (uint32_t)i, seg, off);
    }
    n32 = wave_sum(n32);
    n64 = wave_sum(n64);
    nx = wave_sum(nx);
    if (lane_id() == 0) {
        ws.wave_q[wave] = off;
        ws.wave_32[wave] = n32;
        ws.wave_64[wave] = n64;
        ws.wave_x[wave] = nx;
    }
}

// Exclusive prefix of the per-wave queue counts (one workgroup of 1024), so that the
// exact pass can give every queued tuple its own lane wherever its wave put it.
__global__ __launch_bounds__(1024) void queue_scan_kernel(PredWorkspace ws) {
    __shared__ uint32_t part[1024];
    const unsigned per = (ws.nwaves + 1023) / 1024;
    const unsigned b = threadIdx.x * per, e = min(b + per, ws.nwaves);
    uint32_t sum = 0;
    for (unsigned w = b; w < e; ++w) sum += ws.wave_q[w];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (unsigned o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan
        const uint32_t v = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (unsigned w = b; w < e; ++w) {
        ws.qpref[w] = run;
        run += ws.wave_q[w];
    }
    if (threadIdx.x == 1023) ws.qpref[ws.nwaves] = part[1023];
}

// Queued tuple k (0 <= k < qpref[nwaves]) -> its index in the input.
__device__ __forceinline__ uint32_t queued(const PredWorkspace& ws, uint32_t k) {
    unsigned lo = 0, hi = ws.nwaves;  // qpref[lo] <= k < qpref[hi]
    while (hi - lo > 1) {
        const unsigned mid = (lo + hi) >> 1;
        if (ws.qpref[mid] <= k) lo = mid;
        else hi = mid;
    }
    return ws.queue[(size_t)lo * ws.seg + (k - ws.qpref[lo])];
}

// Register-only exact stage over the insphere queue (the lanes the streaming kernel's
// fixed-point and double-word stages left), one lane per queued tuple, no private arrays:
// fixed point with 126-bit differences. What is left goes to the rest list (one atomic
// per wavefront) for the expansion pass.
template <int NPTS, class Src>
__global__ __launch_bounds__(256) void fixed_kernel(Src src, int8_t* __restrict__ out,
                                                    PredWorkspace ws) {
    const uint32_t total = ws.qpref[ws.nwaves];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (total + stride - 1) / stride;  // the same in every lane
    unsigned long long done = 0;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x + r * stride;
        bool left = false;
        uint32_t i = 0;
        if (k < total) {
            i = queued(ws, k);
            double v[NPTS * 3];
            src.load(i, v);
            int s;
            const bool ok = insphere_sign_fixed_wide(v, v + 3, v + 6, v + 9, v + 12, s);
            if (ok) {
                out[i] = (int8_t)s;
                ++done;
            } else {
                left = true;
            }
        }
        const unsigned long long m = __ballot(left);
        if (m) {
            uint32_t base = 0;
            if (lane_id() == 0) base = atomicAdd(&ws.ctr->rest_n, (unsigned)__popcll(m));
            base = __shfl(base, 0);
            if (left) ws.rest[base + __popcll(m & ((1ull << lane_id()) - 1ull))] = i;
        }
    }
    for (int o = 32; o > 0; o >>= 1) done += __shfl_xor(done, o);
    if (lane_id() == 0 && done) atomicAdd(&ws.ctr->fixed, done);
}

// Exact pass. Pass 1 (WIDE = 0): one lane per queued tuple (orient3d: the queue, through
// the prefix of the per-wave counts; insphere: the rest list of the register-only pass):
// fixed-point stages first, then expansions in small private arrays; lanes that outgrow
// them are appended to the wide queue (rare, so one atomic each). Pass 2 (WIDE = 1): the wide queue with wide
// arrays; a lane that still overflows is counted in ctr->failed and left at 0.
template <int NPTS, bool WIDE, class Src>
__global__ __launch_bounds__(64) void exact_kernel(Src src, int8_t* __restrict__ out,
                                                   PredWorkspace ws) {
    if (!WIDE) {
        const uint32_t total = NPTS == 5 ? ws.ctr->rest_n : ws.qpref[ws.nwaves];
        const uint32_t stride = gridDim.x * blockDim.x;
        for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < total; k += stride) {
            const uint32_t i = NPTS == 5 ? ws.rest[k] : queued(ws, k);
            {
                double v[NPTS * 3];
                src.load(i, v);
                int s = NPTS == 4 ? orient3d_exact_sign<40, false>(v, v + 3, v + 6, v + 9)
                                  : insphere_exact_sign<40, 96>(v, v + 3, v + 6, v + 9, v + 12);
                if (s == kOverflow) {
                    unsigned slot = atomicAdd(&ws.ctr->wide_n, 1u);
                    ws.wide[slot] = i;
                } else {
                    out[i] = (int8_t)s;
                }
            }
        }
    } else {
        const unsigned count = ws.ctr->wide_n;
        const unsigned stride = gridDim.x * blockDim.x;
        for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
            uint32_t i = ws.wide[k];
            double v[NPTS * 3];
            src.load(i, v);
            int s = NPTS == 4 ? orient3d_exact_sign<320, false>(v, v + 3, v + 6, v + 9)
                              : insphere_exact_sign<200, 1024>(v, v + 3, v + 6, v + 9, v + 12);
            if (s == kOverflow) {
                atomicAdd(&ws.ctr->failed, 1u);
                s = 0;
            }
            out[i] = (int8_t)s;
        }
    }
}

// Predicate values (cmg::robust::orient3d_batch, cmg::robust::insphere): the naive FP64
// determinant when it is nonzero, else the exact determinant's leading expansion
// component (same sign).
template <int NPTS, class Src>
__global__ __launch_bounds__(256) void values_kernel(Src src, size_t n,
                                                     double* __restrict__ out,
                                                     PredWorkspace ws, unsigned iters) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned wave = (unsigned)(tid / 64);
    uint32_t* seg = ws.queue + (size_t)wave * ws.seg;
    unsigned off = 0;
    for (unsigned it = 0; it < iters; ++it) {
        const size_t i = tid + (size_t)it * stride;
        bool u = false;
        if (i < n) {
            double v[NPTS * 3];
            src.load(i, v);
            double det = NPTS == 4 ? o3_naive(v, nullptr) : is_naive(v, nullptr);
            u = !(det > 0.0 || det < -0.0);
            out[i] = det;
        }
        wave_push(u, (uint32_t)i, seg, off);
    }
    if (lane_id() == 0) {
        ws.wave_q[wave] = off;
        ws.wave_32[wave] = 0;
        ws.wave_64[wave] = 0;
        ws.wave_x[wave] = 0;
    }
}

template <int NPTS, class Src>
__global__ __launch_bounds__(64) void values_exact_kernel(Src src,
                                                          double* __restrict__ out,
                                                          PredWorkspace ws) {
    const uint32_t total = ws.qpref[ws.nwaves];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < total; k += stride) {
        {
            const uint32_t i = queued(ws, k);
            double v[NPTS * 3];
            src.load(i, v);
            double top = 0.0;
            const bool ok =
                NPTS == 4 ? orient3d_exact<320>(v, v + 3, v + 6, v + 9, top)
                          : insphere_exact<200, 1024>(v, v + 3, v + 6, v + 9, v + 12, top);
            if (!ok) {
                atomicAdd(&ws.ctr->failed, 1u);
                top = 0.0;
            }
            out[i] = top;
        }
    }
}

// Sums the per-wave statistics into ctr (only when the caller asks for them).
__global__ __launch_bounds__(256) void reduce_stats_kernel(PredWorkspace ws) {
    unsigned long long q = 0, a = 0, b = 0, x = 0;
    for (unsigned w = threadIdx.x; w < ws.nwaves; w += blockDim.x) {
        q += ws.wave_q[w];
        a += ws.wave_32[w];
        b += ws.wave_64[w];
        x += ws.wave_x[w];
    }
    __shared__ unsigned long long sq[256], sa[256], sb[256], sx[256];
    sq[threadIdx.x] = q;
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    sx[threadIdx.x] = x;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sq[threadIdx.x] += sq[threadIdx.x + o];
            sa[threadIdx.x] += sa[threadIdx.x + o];
            sb[threadIdx.x] += sb[threadIdx.x + o];
            sx[threadIdx.x] += sx[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ws.ctr->exact = sq[0];
        ws.ctr->fp32 = sa[0];
        ws.ctr->fp64 = sb[0];
        ws.ctr->inline_ = sx[0];
    }
}

__global__ void zero_counters_kernel(PredCounters* c) {
    if (threadIdx.x == 0) *c = PredCounters{};
}

template <int NPTS, class Src>
void launch_exact(Src src, int8_t* out, const PredWorkspace& ws, hipStream_t stream) {
    queue_scan_kernel<<<1, 1024, 0, stream>>>(ws);
    CMG_HIP_LAUNCH();
    if constexpr (NPTS == 5) {
        // grid-stride; the queue is usually short, and a large grid of these waves costs
        // tens of microseconds even when it is empty
        fixed_kernel<NPTS, Src><<<(unsigned)cu_count(), 256, 0, stream>>>(src, out, ws);
        CMG_HIP_LAUNCH();
    }
    // few lanes reach this pass: a small grid keeps the scratch set-up of its waves cheap
    const unsigned g = (unsigned)cu_count();
    exact_kernel<NPTS, false, Src><<<g, 64, 0, stream>>>(src, out, ws);
    CMG_HIP_LAUNCH();
    unsigned gw = cu_count() / 4 > 0 ? cu_count() / 4 : 1;
    exact_kernel<NPTS, true, Src><<<gw, 64, 0, stream>>>(src, out, ws);
    CMG_HIP_LAUNCH();
}

}  // namespace

PredPlan pred_plan(size_t items, unsigned per_thread) {
    PredPlan p;
    p.block = 256;
    size_t units = (items + per_thread - 1) / per_thread;
    p.grid = grid_for(units ? units : 1, p.block, 16);
    size_t stride = (size_t)p.grid * p.block;
    p.iters = (unsigned)((units + stride - 1) / stride);
    p.nwaves = p.grid * (p.block / 64);
    p.seg = 64 * per_thread * p.iters;
    return p;
}

size_t pred_workspace_bytes(const PredPlan& p, size_t n) {
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    return 256 + 4 * al((size_t)p.nwaves * 4) + al(((size_t)p.nwaves + 1) * 4) +
           al((size_t)p.nwaves * p.seg * 4) + 2 * al(n * 4);
}

PredWorkspace pred_workspace(void* base, const PredPlan& p, size_t n) {
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    PredWorkspace ws;
    char* c = static_cast<char*>(base);
    ws.ctr = reinterpret_cast<PredCounters*>(c);
    c += 256;
    ws.wave_q = reinterpret_cast<uint32_t*>(c);
    c += al((size_t)p.nwaves * 4);
    ws.wave_32 = reinterpret_cast<uint32_t*>(c);
    c += al((size_t)p.nwaves * 4);
    ws.wave_64 = reinterpret_cast<uint32_t*>(c);
    c += al((size_t)p.nwaves * 4);
    ws.wave_x = reinterpret_cast<uint32_t*>(c);
    c += al((size_t)p.nwaves * 4);
    ws.qpref = reinterpret_cast<uint32_t*>(c);
    c += al(((size_t)p.nwaves + 1) * 4);
    ws.queue = reinterpret_cast<uint32_t*>(c);
    c += al((size_t)p.nwaves * p.seg * 4);
    ws.wide = reinterpret_cast<uint32_t*>(c);
    c += al(n * 4);
    ws.rest = reinterpret_cast<uint32_t*>(c);
    ws.seg = p.seg;
    ws.nwaves = p.nwaves;
    return ws;
}

void launch_pred_stats(const PredWorkspace& ws, hipStream_t stream) {
    reduce_stats_kernel<<<1, 256, 0, stream>>>(ws);
    CMG_HIP_LAUNCH();
}

void launch_signs_soa(int npts, int mode, const double* soa, size_t n, size_t ld,
                      int8_t* out, const PredPlan& p, const PredWorkspace& ws,
                      hipStream_t stream) {
    if (n == 0) {  // (otherwise the filter kernel clears the counters)
        zero_counters_kernel<<<1, 64, 0, stream>>>(ws.ctr);
        CMG_HIP_LAUNCH();
        return;
    }
    const bool vec = (reinterpret_cast<uintptr_t>(soa) % 16 == 0) && (ld % 2 == 0);
#define CMG_FILTER(NP, MD)                                                                \
    do {                                                                                  \
        {                                                                                 \
            TimedLaunch timed(NP == 4 ? "orient3d_filter" : "insphere_filter", stream);   \
            if (vec)                                                                      \
                filter_soa_kernel<NP, MD, true>                                           \
                    <<<p.grid, p.block, 0, stream>>>(soa, n, ld, out, ws, p.iters);       \
            else                                                                          \
                filter_soa_kernel<NP, MD, false>                                          \
                    <<<p.grid, p.block, 0, stream>>>(soa, n, ld, out, ws, p.iters);       \
            CMG_HIP_LAUNCH();                                                             \
        }                                                                                 \
        launch_exact<NP>(SoaSource<NP>{soa, ld}, out, ws, stream);                        \
    } while (0)
    if (npts == 4) {
        if (mode == kModeExact) CMG_FILTER(4, kModeExact);
        else CMG_FILTER(4, kModeCompat);
    } else {
        if (mode == kModeExact) CMG_FILTER(5, kModeExact);
        else CMG_FILTER(5, kModeCompat);
    }
#undef CMG_FILTER
}

void launch_signs_indexed(int npts, int mode, const double* xyz, const int* idx, size_t n,
                          int8_t* out, const PredPlan& p, const PredWorkspace& ws,
                          hipStream_t stream) {
    zero_counters_kernel<<<1, 64, 0, stream>>>(ws.ctr);
    CMG_HIP_LAUNCH();
    if (n == 0) return;
#define CMG_FILTER_IDX(NP, MD)                                                            \
    do {                                                                                  \
        IndexedSource<NP> src{xyz, idx};                                                  \
        filter_kernel<NP, MD, IndexedSource<NP>>                                          \
            <<<p.grid, p.block, 0, stream>>>(src, n, out, ws, p.iters);                   \
        CMG_HIP_LAUNCH();                                                                 \
        launch_exact<NP>(src, out, ws, stream);                                           \
    } while (0)
    if (npts == 4) {
        if (mode == kModeExact) CMG_FILTER_IDX(4, kModeExact);
        else CMG_FILTER_IDX(4, kModeCompat);
    } else {
        if (mode == kModeExact) CMG_FILTER_IDX(5, kModeExact);
        else CMG_FILTER_IDX(5, kModeCompat);
    }
#undef CMG_FILTER_IDX
}

void launch_values_indexed(int npts, const double* xyz, const int* idx, size_t n,
                           double* out, const PredPlan& p, const PredWorkspace& ws,
                           hipStream_t stream) {
    zero_counters_kernel<<<1, 64, 0, stream>>>(ws.ctr);
    CMG_HIP_LAUNCH();
    if (n == 0) return;
    const unsigned g = (unsigned)cu_count() * 8;
    if (npts == 4) {
        IndexedSource<4> src{xyz, idx};
        values_kernel<4, IndexedSource<4>><<<p.grid, p.block, 0, stream>>>(src, n, out, ws,
                                                                            p.iters);
        CMG_HIP_LAUNCH();
        queue_scan_kernel<<<1, 1024, 0, stream>>>(ws);
        CMG_HIP_LAUNCH();
        values_exact_kernel<4, IndexedSource<4>><<<g, 64, 0, stream>>>(src, out, ws);
        CMG_HIP_LAUNCH();
    } else {
        IndexedSource<5> src{xyz, idx};
        values_kernel<5, IndexedSource<5>><<<p.grid, p.block, 0, stream>>>(src, n, out, ws,
                                                                            p.iters);
        CMG_HIP_LAUNCH();
        queue_scan_kernel<<<1, 1024, 0, stream>>>(ws);
        CMG_HIP_LAUNCH();
        // the wide-capacity insphere expansion spills heavily: a small grid
        const unsigned gw = cu_count() / 4 > 0 ? cu_count() / 4 : 1;
        values_exact_kernel<5, IndexedSource<5>><<<gw, 64, 0, stream>>>(src, out, ws);
        CMG_HIP_LAUNCH();
    }
}

}  // namespace cmg::hip
