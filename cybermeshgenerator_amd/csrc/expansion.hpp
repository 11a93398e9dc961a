// expansion.hpp — exact FP64 sign evaluation on gfx950 for the lanes the filters leave
// uncertain.
//
// Expansion arithmetic after Shewchuk, "Adaptive Precision Floating-Point Arithmetic and
// Fast Robust Geometric Predicates" (CMU-CS-96-140): a value is an array of FP64
// components, nonoverlapping, increasing in magnitude, zeros eliminated; its sign is the
// sign of the last component. Two-Product uses the hardware FMA (p = a*b,
// e = fma(a, b, -p) is exact), which replaces Dekker's split. Sums merge the two
// inputs by magnitude and run one Two-Sum chain (linear time, never reading past either
// input), and Compress keeps every intermediate short.
//
// Both determinants are evaluated exactly in the same difference form as the naive FP64
// formula of src/predicates/predicates.cpp:2201-2222 (orient3d) and :4063-4106
// (insphere): each difference is first made exact as a 1- or 2-component expansion.
// Intermediates live in fixed-capacity private arrays; a lane whose expansion would
// exceed a capacity reports kOverflow and is re-run by a wide-capacity kernel.
#pragma once

#include <hip/hip_runtime.h>

namespace cmg::hip {

constexpr int kOverflow = 2;  // sign slot value meaning "capacity exceeded"

__host__ __device__ __forceinline__ void two_sum(double a, double b, double& x, double& y) {
    double s = a + b;
    double bv = s - a;
    double av = s - bv;
    x = s;
    y = (a - av) + (b - bv);
}

__host__ __device__ __forceinline__ void fast_two_sum(double a, double b, double& x, double& y) {
    double s = a + b;
    x = s;
    y = b - (s - a);
}

__host__ __device__ __forceinline__ void two_prod(double a, double b, double& x, double& y) {
    double p = a * b;
    x = p;
    y = __builtin_fma(a, b, -p);
}

// h = e + f. h must hold en + fn; h may not alias e or f.
__host__ __device__ inline int exp_sum(const double* e, int en, const double* f, int fn, double* h) {
    int i = 0, j = 0, k = 0;
    double q;
    // smallest-magnitude component first
    if (j >= fn || (i < en && __builtin_fabs(e[i]) <= __builtin_fabs(f[j]))) q = e[i++];
    else q = f[j++];
    while (i < en || j < fn) {
        double x;
        if (j >= fn || (i < en && __builtin_fabs(e[i]) <= __builtin_fabs(f[j]))) x = e[i++];
        else x = f[j++];
        double s, t;
        two_sum(q, x, s, t);
        if (t != 0.0) h[k++] = t;
        q = s;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

// h = b * e. h must hold 2*en; h may not alias e.
__host__ __device__ inline int exp_scale(const double* e, int en, double b, double* h) {
    int k = 0;
    double q, t;
    two_prod(e[0], b, q, t);
    if (t != 0.0) h[k++] = t;
    for (int i = 1; i < en; ++i) {
        double p1, p0, s;
        two_prod(e[i], b, p1, p0);
        two_sum(q, p0, s, t);
        if (t != 0.0) h[k++] = t;
        fast_two_sum(p1, s, q, t);
        if (t != 0.0) h[k++] = t;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

// In-place Compress: same value, nonadjacent components, usually very few.
__host__ __device__ inline int exp_compress(double* e, int n) {
    if (n <= 1) return n;
    int bottom = n - 1;
    double q = e[bottom];
    for (int i = n - 2; i >= 0; --i) {
        double qn, r;
        fast_two_sum(q, e[i], qn, r);
        if (r != 0.0) {
            e[bottom--] = qn;
            q = r;
        } else {
            q = qn;
        }
    }
    int top = 0;
    for (int i = bottom + 1; i < n; ++i) {
        double qn, r;
        fast_two_sum(e[i], q, qn, r);
        if (r != 0.0) e[top++] = r;
        q = qn;
    }
    e[top++] = q;
    return top;
}

__host__ __device__ __forceinline__ void exp_negate(double* e, int n) {
    for (int i = 0; i < n; ++i) e[i] = -e[i];
}

__host__ __device__ __forceinline__ int exp_sign(const double* e, int n) {
    double t = e[n - 1];
    return t > 0.0 ? 1 : (t < 0.0 ? -1 : 0);
}

// exact a - b as a 1- or 2-component expansion
__host__ __device__ __forceinline__ int exp_diff(double a, double b, double* h) {
    double x, y;
    two_sum(a, -b, x, y);
    if (y != 0.0) {
        h[0] = y;
        h[1] = x;
        return 2;
    }
    h[0] = x;
    return 1;
}

// h = e * f, e and f short (<= 2 components each): at most 8 components.
__host__ __device__ inline int exp_mul_small(const double* e, int en, const double* f, int fn,
                                    double* h) {
    double a[4], b[4];
    int an = exp_scale(e, en, f[0], a);
    if (fn == 1) {
        for (int i = 0; i < an; ++i) h[i] = a[i];
        return an;
    }
    int bn = exp_scale(e, en, f[1], b);
    return exp_sum(a, an, b, bn, h);
}

// h = e * f where f is short (fn components) and e has en <= ECAP components; the
// result is compressed after every partial sum. Returns -1 if HCAP would be exceeded.
template <int HCAP>
__host__ __device__ inline int exp_mul(const double* e, int en, const double* f, int fn, double* h,
                              double* tmp /* HCAP */, double* part /* 2*en */) {
    if (2 * en > HCAP) return -1;
    int hn = exp_scale(e, en, f[0], h);
    hn = exp_compress(h, hn);
    for (int j = 1; j < fn; ++j) {
        int pn = exp_scale(e, en, f[j], part);
        if (hn + pn > HCAP) return -1;
        int tn = exp_sum(h, hn, part, pn, tmp);
        tn = exp_compress(tmp, tn);
        for (int i = 0; i < tn; ++i) h[i] = tmp[i];
        hn = tn;
    }
    return hn;
}

// 2x2 minor p*q - r*s of 1-2 component entries (<= 16 components, compressed).
__host__ __device__ inline int exp_minor(const double* p, int pn, const double* q, int qn,
                                const double* r, int rn, const double* s, int sn,
                                double* out /* 16 */) {
    double a[8], b[8];
    int an = exp_mul_small(p, pn, q, qn, a);
    int bn = exp_mul_small(r, rn, s, sn, b);
    exp_negate(b, bn);
    int n = exp_sum(a, an, b, bn, out);
    return exp_compress(out, n);
}

// acc += sgn * z * m  (z: 1-2 components, m: <= 16). Returns false on overflow.
template <int CAP>
__host__ __device__ inline bool exp_accumulate_zm(double* acc, int& an, const double* z, int zn,
                                         const double* m, int mn, int sgn) {
    double t[64];
    double tmp[64];
    double part[32];
    int tn = exp_mul<64>(m, mn, z, zn, t, tmp, part);
    if (tn < 0) return false;
    if (sgn < 0) exp_negate(t, tn);
    if (an + tn > CAP) return false;
    double sum[CAP];
    int sn = exp_sum(acc, an, t, tn, sum);
    sn = exp_compress(sum, sn);
    for (int i = 0; i < sn; ++i) acc[i] = sum[i];
    an = sn;
    return true;
}

// Exact orient3d(pa, pb, pc, pd) as an expansion; `top` receives its largest
// component (same sign as the determinant, within a relative 2^-52 of its value).
// Returns false if CAP is too small.
template <int CAP>
__host__ __device__ inline bool orient3d_exact(const double* pa, const double* pb, const double* pc,
                                      const double* pd, double& top) {
    double adx[2], ady[2], adz[2], bdx[2], bdy[2], bdz[2], cdx[2], cdy[2], cdz[2];
    int nadx = exp_diff(pa[0], pd[0], adx), nady = exp_diff(pa[1], pd[1], ady),
        nadz = exp_diff(pa[2], pd[2], adz);
    int nbdx = exp_diff(pb[0], pd[0], bdx), nbdy = exp_diff(pb[1], pd[1], bdy),
        nbdz = exp_diff(pb[2], pd[2], bdz);
    int ncdx = exp_diff(pc[0], pd[0], cdx), ncdy = exp_diff(pc[1], pd[1], cdy),
        ncdz = exp_diff(pc[2], pd[2], cdz);
    double acc[CAP];
    int an = 0;
    double m[16];
    int mn;
    // adz * (bdx*cdy - cdx*bdy)
    mn = exp_minor(bdx, nbdx, cdy, ncdy, cdx, ncdx, bdy, nbdy, m);
    if (!exp_accumulate_zm<CAP>(acc, an, adz, nadz, m, mn, 1)) return false;
    // + bdz * (cdx*ady - adx*cdy)
    mn = exp_minor(cdx, ncdx, ady, nady, adx, nadx, cdy, ncdy, m);
    if (!exp_accumulate_zm<CAP>(acc, an, bdz, nbdz, m, mn, 1)) return false;
    // + cdz * (adx*bdy - bdx*ady)
    mn = exp_minor(adx, nadx, bdy, nbdy, bdx, nbdx, ady, nady, m);
    if (!exp_accumulate_zm<CAP>(acc, an, cdz, ncdz, m, mn, 1)) return false;
    top = acc[an - 1];
    return true;
}

// Sign of the exact sum of N doubles held in registers (static indices only), by
// repeated error-free VecSum passes (Ogita, Rump, Oishi, "Accurate sum and dot product",
// SISC 2005): each pass replaces v by its rounded sum in v[N-1] and the rounding errors
// below it, without changing the exact sum. The sign is certain once |v[N-1]| exceeds
// the sum of the other magnitudes (with a margin for rounding that sum), and the sum is 0
// once every entry is 0. Returns false if neither happened within MAXPASS passes.
template <int N, int MAXPASS>
__host__ __device__ inline bool sum_sign(double (&v)[N], int& sign) {
#pragma unroll 1
    for (int pass = 0; pass < MAXPASS; ++pass) {
#pragma unroll
        for (int i = 1; i < N; ++i) {
            double s, e;
            two_sum(v[i], v[i - 1], s, e);
            v[i] = s;
            v[i - 1] = e;
        }
        double tail = 0.0;
#pragma unroll
        for (int i = 0; i < N - 1; ++i) tail += __builtin_fabs(v[i]);
        const double top = v[N - 1];
        if (__builtin_fabs(top) > tail * (1.0 + 0x1p-40)) {
            sign = top > 0.0 ? 1 : -1;
            return true;
        }
        if (top == 0.0 && tail == 0.0) {
            sign = 0;
            return true;
        }
    }
    return false;
}

// orient3d sign when every difference a - d etc. is exact in FP64: the determinant is the
// exact sum of six triple products of those differences, each split into four doubles
// with FMA (x*y = p + e exactly, then p*z and e*z exactly). False if undecided.
__host__ __device__ inline bool orient3d_sign_exact_diffs(const double* pa, const double* pb,
                                                 const double* pc, const double* pd,
                                                 int& sign) {
    double d[9];
    for (int s = 0; s < 3; ++s) {
        const double* p = s == 0 ? pa : (s == 1 ? pb : pc);
        for (int a = 0; a < 3; ++a) {
            double x, y;
            two_sum(p[a], -pd[a], x, y);
            if (y != 0.0) return false;
            // keep every triple product and its FMA error terms normal and finite
            const double ax = __builtin_fabs(x);
            if (ax != 0.0 && !(ax >= 0x1p-300 && ax <= 0x1p300)) return false;
            d[3 * s + a] = x;
        }
    }
    const double adx = d[0], ady = d[1], adz = d[2], bdx = d[3], bdy = d[4], bdz = d[5],
                 cdx = d[6], cdy = d[7], cdz = d[8];
    // adz*bdx*cdy - adz*cdx*bdy + bdz*cdx*ady - bdz*adx*cdy + cdz*adx*bdy - cdz*bdx*ady
    const double t[6][3] = {{bdx, cdy, adz},  {-cdx, bdy, adz}, {cdx, ady, bdz},
                            {-adx, cdy, bdz}, {adx, bdy, cdz},  {-bdx, ady, cdz}};
    double v[24];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double p, e;
        two_prod(t[k][0], t[k][1], p, e);
        two_prod(p, t[k][2], v[4 * k], v[4 * k + 1]);
        two_prod(e, t[k][2], v[4 * k + 2], v[4 * k + 3]);
    }
    return sum_sign<24, 6>(v, sign);
}

// Double-word (hi + lo) arithmetic with the FMA algorithms of Joldes, Muller and Popescu,
// "Tight and rigorous error bounds for basic building blocks of double-word arithmetic"
// (ACM TOMS 2017): DWTimesDW3 (relative error <= 4u^2) and AccurateDWPlusDW
// (<= 3u^2 + 13u^3), u = 2^-53.
struct DW {
    double hi, lo;
};

__host__ __device__ __forceinline__ DW dw_mul(DW x, DW y) {
    double ch, cl1;
    two_prod(x.hi, y.hi, ch, cl1);
    const double tl0 = x.lo * y.lo;
    const double tl1 = __builtin_fma(x.hi, y.lo, tl0);
    const double cl2 = __builtin_fma(x.lo, y.hi, tl1);
    DW z;
    fast_two_sum(ch, cl1 + cl2, z.hi, z.lo);
    return z;
}

__host__ __device__ __forceinline__ DW dw_add(DW x, DW y) {
    double sh, sl, th, tl;
    two_sum(x.hi, y.hi, sh, sl);
    two_sum(x.lo, y.lo, th, tl);
    sl += th;
    double vh, vl;
    fast_two_sum(sh, sl, vh, vl);
    vl += tl;
    DW z;
    fast_two_sum(vh, vl, z.hi, z.lo);
    return z;
}

__host__ __device__ __forceinline__ DW dw_neg(DW x) { return DW{-x.hi, -x.lo}; }

// orient3d sign from exact double-word differences evaluated in double-word arithmetic.
// Each of the 15 operations adds at most ~5u^2 of the magnitudes it combines, all bounded
// by the permanent P, so |det_dw - det| <= 2^-96 P with a wide margin; lanes inside that
// band (in practice only exact zeros) are left to the expansion arithmetic. Inputs must
// keep every product normal (same range check as the exact-difference path).
__host__ __device__ inline bool orient3d_sign_dw(const double* pa, const double* pb, const double* pc,
                                        const double* pd, int& sign) {
    DW d[9];
    double perm_in[9];
    for (int s = 0; s < 3; ++s) {
        const double* p = s == 0 ? pa : (s == 1 ? pb : pc);
        for (int a = 0; a < 3; ++a) {
            DW x;
            two_sum(p[a], -pd[a], x.hi, x.lo);
            const double ax = __builtin_fabs(x.hi);
            if (ax != 0.0 && !(ax >= 0x1p-250 && ax <= 0x1p250)) return false;
            d[3 * s + a] = x;
            perm_in[3 * s + a] = ax;
        }
    }
    const DW adx = d[0], ady = d[1], adz = d[2], bdx = d[3], bdy = d[4], bdz = d[5],
             cdx = d[6], cdy = d[7], cdz = d[8];
    const DW m1 = dw_add(dw_mul(bdx, cdy), dw_neg(dw_mul(cdx, bdy)));
    const DW m2 = dw_add(dw_mul(cdx, ady), dw_neg(dw_mul(adx, cdy)));
    const DW m3 = dw_add(dw_mul(adx, bdy), dw_neg(dw_mul(bdx, ady)));
    const DW det = dw_add(dw_add(dw_mul(adz, m1), dw_mul(bdz, m2)), dw_mul(cdz, m3));
    const double* A = perm_in;  // |adx| |ady| |adz| |bdx| ... (hi parts)
    const double perm = (A[3] * A[7] + A[6] * A[4]) * A[2] + (A[6] * A[1] + A[0] * A[7]) * A[5] +
                        (A[0] * A[4] + A[3] * A[1]) * A[8];
    const double bound = 0x1p-96 * perm * (1.0 + 0x1p-40);
    const double v = det.hi + det.lo;
    if (v > bound) {
        sign = 1;
        return true;
    }
    if (-v > bound) {
        sign = -1;
        return true;
    }
    return false;
}


// ---- fixed-point exact insphere ---------------------------------------------------
//
// Each difference p - pe is exactly x + y (Two-Sum). When all twelve are integer
// multiples of one power of two 2^s with magnitudes below 2^(s+62), the determinant is an
// integer polynomial in 62-bit integers D = (x + y) / 2^s and is evaluated exactly in two's
// complement multi-limb integers held in registers: lifts (unsigned, < 2^126), xy minors
// (signed 128-bit), cofactors (192-bit), lifted products and their sum (320-bit).

typedef unsigned __int128 u128;

struct I192 { unsigned long long w[3]; };  // two's complement
struct I320 { unsigned long long w[5]; };

// w = -w when neg (two's complement, without branches)
template <int N>
__host__ __device__ __forceinline__ void limbs_neg(unsigned long long (&w)[N], bool neg = true) {
    const unsigned long long m = 0ull - (unsigned long long)neg;
    unsigned long long c = (unsigned long long)neg;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const u128 t = (u128)(w[i] ^ m) + c;
        w[i] = (unsigned long long)t;
        c = (unsigned long long)(t >> 64);
    }
}

// a * b for |a|, |b| < 2^63 as a signed 128-bit integer (magnitudes, then the sign)
__host__ __device__ __forceinline__ __int128 smul(long long a, long long b) {
    const unsigned long long ua = a < 0 ? 0ull - (unsigned long long)a : (unsigned long long)a;
    const unsigned long long ub = b < 0 ? 0ull - (unsigned long long)b : (unsigned long long)b;
    const u128 p = (u128)ua * ub;
    const u128 m = (u128)0 - (u128)((a < 0) != (b < 0));
    return (__int128)((p ^ m) - m);
}

template <int N>
__host__ __device__ __forceinline__ void limbs_add(unsigned long long (&a)[N],
                                          const unsigned long long (&b)[N]) {
    unsigned long long c = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const u128 t = (u128)a[i] + b[i] + c;
        a[i] = (unsigned long long)t;
        c = (unsigned long long)(t >> 64);
    }
}

// z * m (z < 2^62 in magnitude, m a signed 128-bit minor) as a 192-bit integer
__host__ __device__ __forceinline__ I192 mul_z_minor(long long z, __int128 m) {
    const bool neg = (z < 0) != (m < 0);
    const unsigned long long az = z < 0 ? (unsigned long long)(-z) : (unsigned long long)z;
    const u128 am = m < 0 ? (u128)(-m) : (u128)m;
    const u128 p0 = (u128)az * (unsigned long long)am;
    const u128 p1 = (u128)az * (unsigned long long)(am >> 64);
    I192 r;
    r.w[0] = (unsigned long long)p0;
    const u128 mid = (p0 >> 64) + (unsigned long long)p1;
    r.w[1] = (unsigned long long)mid;
    r.w[2] = (unsigned long long)(p1 >> 64) + (unsigned long long)(mid >> 64);
    limbs_neg(r.w, neg);
    return r;
}

// lift * c (lift < 2^126 unsigned, c a signed 192-bit cofactor) as a 320-bit integer
__host__ __device__ __forceinline__ I320 mul_lift_cof(u128 lift, I192 c) {
    const bool neg = (long long)c.w[2] < 0;
    limbs_neg(c.w, neg);
    const unsigned long long l[2] = {(unsigned long long)lift, (unsigned long long)(lift >> 64)};
    I320 r{{0, 0, 0, 0, 0}};
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        unsigned long long carry = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const u128 t = (u128)l[i] * c.w[j] + r.w[i + j] + carry;
            r.w[i + j] = (unsigned long long)t;
            carry = (unsigned long long)(t >> 64);
        }
        r.w[i + 3] += carry;  // i + 3 <= 4; no carry out: |product| < 2^316
    }
    limbs_neg(r.w, neg);
    return r;
}

// Exponent range of a nonzero finite normal double: |x| = mant * 2^e, lowest set bit at
// 2^low, |x| < 2^high. False for subnormals, infinities and NaNs.
__host__ __device__ __forceinline__ bool bit_range(double x, int& low, int& high) {
    const unsigned long long bits = __builtin_bit_cast(unsigned long long, x);
    const int ef = (int)((bits >> 52) & 0x7ff);
    if (ef == 0 || ef == 0x7ff) return false;
    const unsigned long long mant = (bits & 0xfffffffffffffull) | (1ull << 52);
    const int e = ef - 1075;
    low = e + __builtin_ctzll(mant);
    high = e + 53;
    return true;
}

// x / 2^s as an integer, for a double x whose lowest set bit is at or above 2^s
__host__ __device__ __forceinline__ long long scaled_int(double x, int s) {
    if (x == 0.0) return 0;
    const unsigned long long bits = __builtin_bit_cast(unsigned long long, x);
    const unsigned long long mant = (bits & 0xfffffffffffffull) | (1ull << 52);
    const int sh = (int)((bits >> 52) & 0x7ff) - 1075 - s;
    const long long q = (long long)(sh >= 0 ? mant << sh : mant >> -sh);
    return (bits >> 63) ? -q : q;
}

// The twelve differences p - pe (vertex v = a, b, c, d; axis a at 3v + a) as Two-Sum
// pairs x + y, and the bit range of their nonzero parts: lowest set bit 2^lo, magnitudes
// below 2^hi (hi < lo when every difference is zero). False when a part is subnormal.
struct InsDiffs {
    double x[12], y[12];
    int lo, hi;
};

__host__ __device__ __forceinline__ bool insphere_diffs(const double* pa, const double* pb,
                                                        const double* pc, const double* pd,
                                                        const double* pe, InsDiffs& d) {
    const double* P[4] = {pa, pb, pc, pd};
    d.lo = 1 << 20;
    d.hi = -(1 << 20);
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double x, y;
            two_sum(P[v][a], -pe[a], x, y);
            d.x[3 * v + a] = x;
            d.y[3 * v + a] = y;
            const double xy[2] = {x, y};
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                if (xy[q] == 0.0) continue;
                int l, h;
                if (!bit_range(xy[q], l, h)) return false;
                d.lo = l < d.lo ? l : d.lo;
                d.hi = h > d.hi ? h : d.hi;
            }
        }
    return true;
}

// Whether the fixed-point stages may evaluate d with differences of at most `bits` bits.
// Besides the width, stay where every product of five differences and its rounding-error
// terms is a normal double (components within [2^-150, 2^150]): there the expansion
// arithmetic is exact too, so every stage gives the same sign; outside, leave the lane to
// the expansions.
__host__ __device__ __forceinline__ bool insphere_fixed_fits(const InsDiffs& d, int bits) {
    return d.hi - d.lo <= bits && d.hi <= 150 && d.lo >= -150;
}

// Exact insphere sign from differences of at most 62 bits: the determinant is an integer
// polynomial in the 62-bit integers D = (x + y) / 2^lo, evaluated in two's complement
// multi-limb integers: lifts (unsigned, < 2^126), xy minors (signed 128-bit), cofactors
// (192-bit), lifted products and their sum (320-bit).
__host__ __device__ inline int insphere_fixed62_sign(const InsDiffs& d) {
    long long D[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) D[i] = scaled_int(d.x[i], d.lo) + scaled_int(d.y[i], d.lo);
    // D[3v + a]: vertex v (a, b, c, d) minus e, axis a
    u128 lift[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const long long x = D[3 * v], y = D[3 * v + 1], z = D[3 * v + 2];
        const unsigned long long ax = x < 0 ? -x : x, ay = y < 0 ? -y : y, az = z < 0 ? -z : z;
        lift[v] = (u128)ax * ax + (u128)ay * ay + (u128)az * az;
    }
    auto minor = [&](int u, int v) -> __int128 {  // uex * vey - vex * uey
        return smul(D[3 * u], D[3 * v + 1]) - smul(D[3 * v], D[3 * u + 1]);
    };
    constexpr int A = 0, B = 1, C = 2, Dv = 3;
    const __int128 ab = minor(A, B), bc = minor(B, C), cd = minor(C, Dv), da = minor(Dv, A),
                   ac = minor(A, C), bd = minor(B, Dv);
    auto z = [&](int v) { return D[3 * v + 2]; };
    auto cof = [&](int v0, __int128 m0, int v1, __int128 m1, int v2, __int128 m2) {
        I192 r = mul_z_minor(z(v0), m0);
        const I192 t1 = mul_z_minor(z(v1), m1), t2 = mul_z_minor(z(v2), m2);
        limbs_add(r.w, t1.w);
        limbs_add(r.w, t2.w);
        return r;
    };
    // abc = aez*bc - bez*ac + cez*ab      bcd = bez*cd - cez*bd + dez*bc
    // cda = cez*da + dez*ac + aez*cd      dab = dez*ab + aez*bd + bez*da
    const I192 abc = cof(A, bc, B, -ac, C, ab);
    const I192 bcd = cof(B, cd, C, -bd, Dv, bc);
    const I192 cda = cof(C, da, Dv, ac, A, cd);
    const I192 dab = cof(Dv, ab, A, bd, B, da);
    // det = (dlift*abc - clift*dab) + (blift*cda - alift*bcd)
    I320 det = mul_lift_cof(lift[Dv], abc);
    I320 t = mul_lift_cof(lift[C], dab);
    limbs_neg(t.w);
    limbs_add(det.w, t.w);
    t = mul_lift_cof(lift[B], cda);
    limbs_add(det.w, t.w);
    t = mul_lift_cof(lift[A], bcd);
    limbs_neg(t.w);
    limbs_add(det.w, t.w);
    if ((det.w[0] | det.w[1] | det.w[2] | det.w[3] | det.w[4]) == 0) return 0;
    return (long long)det.w[4] < 0 ? -1 : 1;
}

// |a| * b for |a| < 2^31 and b < 2^64 as an unsigned 128-bit integer, from two 32 x 32-bit
// products
__host__ __device__ __forceinline__ u128 mul_u32_u64(unsigned a, unsigned long long b) {
    const unsigned long long p0 = (unsigned long long)a * (unsigned)b;
    const unsigned long long p1 = (unsigned long long)a * (unsigned)(b >> 32);
    return ((u128)p1 << 32) + p0;
}

// Exact insphere sign from differences of at most 31 bits (D = (x + y) / 2^lo, |D| < 2^31):
// lifts fit 64 bits unsigned (3 squares below 2^62), xy minors 64 bits signed (two
// products below 2^62), cofactors 128 bits signed (below 2^96), and the lifted products
// and their sum 192 bits; every product is built from 32 x 32-bit ones (84 in all, against
// about 290 for the 62-bit evaluation).
__host__ __device__ inline int insphere_fixed31_sign(const InsDiffs& d) {
    int D[12];
#pragma unroll
    for (int i = 0; i < 12; ++i)
        D[i] = (int)(scaled_int(d.x[i], d.lo) + scaled_int(d.y[i], d.lo));
    unsigned long long lift[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        unsigned long long l = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const unsigned m = (unsigned)(D[3 * v + a] < 0 ? -D[3 * v + a] : D[3 * v + a]);
            l += (unsigned long long)m * m;
        }
        lift[v] = l;
    }
    auto minor = [&](int u, int v) -> long long {  // uex * vey - vex * uey
        return (long long)D[3 * u] * D[3 * v + 1] - (long long)D[3 * v] * D[3 * u + 1];
    };
    constexpr int A = 0, B = 1, C = 2, Dv = 3;
    const long long ab = minor(A, B), bc = minor(B, C), cd = minor(C, Dv), da = minor(Dv, A),
                    ac = minor(A, C), bd = minor(B, Dv);
    auto zm = [&](int v, long long m) -> __int128 {  // z_v * m, |z| < 2^31, |m| < 2^63
        const int zv = D[3 * v + 2];
        const unsigned az = (unsigned)(zv < 0 ? -zv : zv);
        const unsigned long long am = m < 0 ? 0ull - (unsigned long long)m : (unsigned long long)m;
        const __int128 p = (__int128)mul_u32_u64(az, am);
        return ((zv < 0) != (m < 0)) ? -p : p;
    };
    const __int128 abc = zm(A, bc) - zm(B, ac) + zm(C, ab);
    const __int128 bcd = zm(B, cd) - zm(C, bd) + zm(Dv, bc);
    const __int128 cda = zm(C, da) + zm(Dv, ac) + zm(A, cd);
    const __int128 dab = zm(Dv, ab) + zm(A, bd) + zm(B, da);
    // det = (dlift*abc - clift*dab) + (blift*cda - alift*bcd), each lifted product below
    // 2^160 in magnitude: |cof| (< 2^96) times the 64-bit lift in 192-bit limbs
    unsigned long long det[3] = {0, 0, 0};
    auto add_lifted = [&](unsigned long long lift_v, __int128 c, bool negate) {
        const bool neg = (c < 0) != negate;
        const u128 ac = c < 0 ? (u128)0 - (u128)c : (u128)c;
        const unsigned long long c0 = (unsigned long long)ac, c1 = (unsigned long long)(ac >> 64);
        const u128 p0 = (u128)lift_v * c0;                // < 2^128
        const u128 p1 = mul_u32_u64((unsigned)c1, lift_v);  // c1 < 2^32
        unsigned long long w[3];
        w[0] = (unsigned long long)p0;
        const u128 mid = (p0 >> 64) + (unsigned long long)p1;
        w[1] = (unsigned long long)mid;
        w[2] = (unsigned long long)(p1 >> 64) + (unsigned long long)(mid >> 64);
        limbs_neg(w, neg);
        limbs_add(det, w);
    };
    add_lifted(lift[Dv], abc, false);
    add_lifted(lift[C], dab, true);
    add_lifted(lift[B], cda, false);
    add_lifted(lift[A], bcd, true);
    if ((det[0] | det[1] | det[2]) == 0) return 0;
    return (long long)det[2] < 0 ? -1 : 1;
}

// Exact insphere sign in fixed point with differences of at most 62 bits; false when a
// difference is subnormal or they span more (the caller then uses expansion arithmetic).
__host__ __device__ inline bool insphere_sign_fixed(const double* pa, const double* pb, const double* pc,
                                           const double* pd, const double* pe, int& sign) {
    InsDiffs d;
    if (!insphere_diffs(pa, pb, pc, pd, pe, d)) return false;
    if (d.hi < d.lo) {  // every difference is zero
        sign = 0;
        return true;
    }
    if (!insphere_fixed_fits(d, 62)) return false;
    sign = insphere_fixed62_sign(d);
    return true;
}

// The same with differences of at most 31 bits (the cheaper evaluation).
__host__ __device__ inline bool insphere_sign_fixed31(const double* pa, const double* pb,
                                                      const double* pc, const double* pd,
                                                      const double* pe, int& sign) {
    InsDiffs d;
    if (!insphere_diffs(pa, pb, pc, pd, pe, d)) return false;
    if (d.hi < d.lo) {
        sign = 0;
        return true;
    }
    if (!insphere_fixed_fits(d, 31)) return false;
    sign = insphere_fixed31_sign(d);
    return true;
}

// Exact orient3d sign in fixed point: each difference p - pd is exactly x + y (Two-Sum);
// when all nine are integer multiples of one power of two with magnitudes within 62
// bits, det = adz*(bdx*cdy - cdx*bdy) + bdz*(cdx*ady - adx*cdy) + cdz*(adx*bdy - bdx*ady)
// is evaluated exactly (minors 128-bit, products and sum 192-bit). Decides the exact
// zeros whose differences are inexact in FP64, which the exact-difference and
// double-word stages cannot. Same range limits as the insphere stage.
__host__ __device__ inline bool orient3d_sign_fixed(const double* pa, const double* pb,
                                                    const double* pc, const double* pd,
                                                    int& sign) {
    const double* P[3] = {pa, pb, pc};
    double dx[9], dy[9];
    int lo = 1 << 20, hi = -(1 << 20);
    for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a) {
            double x, y;
            two_sum(P[v][a], -pd[a], x, y);
            dx[3 * v + a] = x;
            dy[3 * v + a] = y;
            const double xy[2] = {x, y};
            for (int q = 0; q < 2; ++q) {
                if (xy[q] == 0.0) continue;
                int l, h;
                if (!bit_range(xy[q], l, h)) return false;
                lo = l < lo ? l : lo;
                hi = h > hi ? h : hi;
            }
        }
    if (hi < lo) {
        sign = 0;
        return true;
    }
    if (hi - lo > 62 || hi > 300 || lo < -300) return false;
    long long D[9];
    for (int i = 0; i < 9; ++i) D[i] = scaled_int(dx[i], lo) + scaled_int(dy[i], lo);
    const long long adx = D[0], ady = D[1], adz = D[2], bdx = D[3], bdy = D[4], bdz = D[5],
                    cdx = D[6], cdy = D[7], cdz = D[8];
    I192 det = mul_z_minor(adz, smul(bdx, cdy) - smul(cdx, bdy));
    const I192 t1 = mul_z_minor(bdz, smul(cdx, ady) - smul(adx, cdy));
    const I192 t2 = mul_z_minor(cdz, smul(adx, bdy) - smul(bdx, ady));
    limbs_add(det.w, t1.w);
    limbs_add(det.w, t2.w);
    sign = (det.w[0] | det.w[1] | det.w[2]) == 0 ? 0 : ((long long)det.w[2] < 0 ? -1 : 1);
    return true;
}

// Exact orient3d sign when the nine differences p - pd are exact in FP64 and integer
// multiples of 2^(E-31), where 2^(E-1) <= max |difference| < 2^E <= 2^31: then
// D = difference * 2^(31-E) is an integer below 2^31 in magnitude (the scaling is exact:
// by 2^(31-E) >= 1, no overflow), the minors bdx*cdy - cdx*bdy (two products below 2^62)
// fit 64 bits, and det = adz*m1 + bdz*m2 + cdz*m3 (three products below 2^94) fits 128
// bits, each product built from 32 x 32-bit ones (12 in all). Decides the exact zeros
// and the small determinants of tuples on a common grid, whatever their sign, in a few
// hundred instructions. False when a difference is inexact, the differences span more
// than 31 bits, or max |difference| lies outside [2^-260, 2^31) (the lower limit keeps
// every nonzero difference above 2^-291, inside the range of the other stages).
__host__ __device__ inline bool orient3d_sign_fixed31(const double* pa, const double* pb,
                                                      const double* pc, const double* pd,
                                                      int& sign) {
    double d[9];
    double amax = 0.0;
    bool exact = true;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const double* p = v == 0 ? pa : (v == 1 ? pb : pc);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double x, y;
            two_sum(p[a], -pd[a], x, y);
            exact = exact && y == 0.0;
            d[3 * v + a] = x;
            amax = __builtin_fmax(amax, __builtin_fabs(x));
        }
    }
    if (!exact) return false;
    if (amax == 0.0) {  // four equal points
        sign = 0;
        return true;
    }
    if (!(amax >= 0x1p-260 && amax < 0x1p31)) return false;
    int E = 0;
    (void)__builtin_frexp(amax, &E);  // amax in [2^(E-1), 2^E), E <= 31
    const double scale = __builtin_ldexp(1.0, 31 - E);
    int D[9];
    bool ints = true;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const double t = d[i] * scale;  // exact, |t| < 2^31
        ints = ints && t == __builtin_rint(t);
        D[i] = (int)t;
    }
    if (!ints) return false;
    const int adx = D[0], ady = D[1], adz = D[2], bdx = D[3], bdy = D[4], bdz = D[5],
              cdx = D[6], cdy = D[7], cdz = D[8];
    const long long m1 = (long long)bdx * cdy - (long long)cdx * bdy;
    const long long m2 = (long long)cdx * ady - (long long)adx * cdy;
    const long long m3 = (long long)adx * bdy - (long long)bdx * ady;
    auto zm = [](int z, long long m) -> __int128 {  // z * m, |z| < 2^31, |m| < 2^63
        const unsigned az = (unsigned)(z < 0 ? -z : z);
        const unsigned long long am = m < 0 ? 0ull - (unsigned long long)m : (unsigned long long)m;
        const __int128 p = (__int128)mul_u32_u64(az, am);
        return ((z < 0) != (m < 0)) ? -p : p;
    };
    const __int128 det = zm(adz, m1) + zm(bdz, m2) + zm(cdz, m3);
    sign = det == 0 ? 0 : (det < 0 ? -1 : 1);
    return true;
}

// ---- wide fixed point: the same evaluation for differences spanning up to 126 bits ----

// r = a * b (unsigned, little-endian 64-bit limbs)
template <int NA, int NB>
__host__ __device__ __forceinline__ void limbs_mul(const unsigned long long (&a)[NA],
                                                   const unsigned long long (&b)[NB],
                                                   unsigned long long (&r)[NA + NB]) {
#pragma unroll
    for (int i = 0; i < NA + NB; ++i) r[i] = 0;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        unsigned long long carry = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const u128 t = (u128)a[i] * b[j] + r[i + j] + carry;
            r[i + j] = (unsigned long long)t;
            carry = (unsigned long long)(t >> 64);
        }
        r[i + NB] = carry;
    }
}

// w = |w| of a two's complement value; returns whether it was negative
template <int N>
__host__ __device__ __forceinline__ bool limbs_abs(unsigned long long (&w)[N]) {
    const bool neg = (long long)w[N - 1] < 0;
    limbs_neg(w, neg);
    return neg;
}

// signed a * b, both two's complement, into NA + NB limbs (the magnitudes must leave the
// top bit of the product clear; a and b must not be the same array)
template <int NA, int NB>
__host__ __device__ __forceinline__ void limbs_smul(unsigned long long (&a)[NA],
                                                    unsigned long long (&b)[NB],
                                                    unsigned long long (&r)[NA + NB]) {
    const bool na = limbs_abs(a), nb = limbs_abs(b);
    limbs_mul(a, b, r);
    limbs_neg(r, na != nb);
    limbs_neg(a, na);  // restore the operands
    limbs_neg(b, nb);
}

// x / 2^s as a 128-bit integer, for a double x whose lowest set bit is at or above 2^s and
// whose magnitude is below 2^(s+126)
__host__ __device__ __forceinline__ __int128 scaled_int128(double x, int s) {
    if (x == 0.0) return 0;
    const unsigned long long bits = __builtin_bit_cast(unsigned long long, x);
    const u128 mant = (bits & 0xfffffffffffffull) | (1ull << 52);
    const int sh = (int)((bits >> 52) & 0x7ff) - 1075 - s;
    const u128 q = sh >= 0 ? mant << sh : mant >> -sh;
    return (bits >> 63) ? -(__int128)q : (__int128)q;
}

// Exact insphere sign in fixed point with 126-bit differences: lifts and minors in 256
// bits, cofactors in 384, lifted products and their sum in 640. False when a component
// is outside [2^-150, 2^150] or the differences span more than 126 bits.
__host__ __device__ inline bool insphere_sign_fixed_wide(const double* pa, const double* pb,
                                                         const double* pc, const double* pd,
                                                         const double* pe, int& sign) {
    const double* P[4] = {pa, pb, pc, pd};
    double dx[12], dy[12];
    int lo = 1 << 20, hi = -(1 << 20);
    for (int v = 0; v < 4; ++v)
        for (int a = 0; a < 3; ++a) {
            double x, y;
            two_sum(P[v][a], -pe[a], x, y);
            dx[3 * v + a] = x;
            dy[3 * v + a] = y;
            const double xy[2] = {x, y};
            for (int q = 0; q < 2; ++q) {
                if (xy[q] == 0.0) continue;
                int l, h;
                if (!bit_range(xy[q], l, h)) return false;
                lo = l < lo ? l : lo;
                hi = h > hi ? h : hi;
            }
        }
    if (hi < lo) {
        sign = 0;
        return true;
    }
    if (hi - lo > 126 || hi > 150 || lo < -150) return false;
    unsigned long long D[12][2];
    for (int i = 0; i < 12; ++i) {
        const __int128 q = scaled_int128(dx[i], lo) + scaled_int128(dy[i], lo);
        D[i][0] = (unsigned long long)q;
        D[i][1] = (unsigned long long)((u128)q >> 64);
    }
    typedef unsigned long long L4[4];
    typedef unsigned long long L6[6];
    typedef unsigned long long L10[10];
    L4 lift[4];
    for (int v = 0; v < 4; ++v) {
        for (int i = 0; i < 4; ++i) lift[v][i] = 0;
        for (int a = 0; a < 3; ++a) {
            unsigned long long x[2] = {D[3 * v + a][0], D[3 * v + a][1]};
            limbs_abs(x);
            L4 sq;
            limbs_mul(x, x, sq);
            limbs_add(lift[v], sq);
        }
    }
    auto minor = [&](int u, int v, L4& m) {  // uex * vey - vex * uey
        L4 t;
        limbs_smul(D[3 * u], D[3 * v + 1], m);
        limbs_smul(D[3 * v], D[3 * u + 1], t);
        limbs_neg(t);
        limbs_add(m, t);
    };
    constexpr int A = 0, B = 1, C = 2, Dv = 3;
    L4 ab, bc, cd, da, ac, bd;
    minor(A, B, ab);
    minor(B, C, bc);
    minor(C, Dv, cd);
    minor(Dv, A, da);
    minor(A, C, ac);
    minor(B, Dv, bd);
    auto cof = [&](L6& r, int v0, L4& m0, bool n0, int v1, L4& m1, bool n1, int v2, L4& m2,
                   bool n2) {
        L6 t;
        limbs_smul(D[3 * v0 + 2], m0, r);
        limbs_neg(r, n0);
        limbs_smul(D[3 * v1 + 2], m1, t);
        limbs_neg(t, n1);
        limbs_add(r, t);
        limbs_smul(D[3 * v2 + 2], m2, t);
        limbs_neg(t, n2);
        limbs_add(r, t);
    };
    // abc = aez*bc - bez*ac + cez*ab      bcd = bez*cd - cez*bd + dez*bc
    // cda = cez*da + dez*ac + aez*cd      dab = dez*ab + aez*bd + bez*da
    L6 abc, bcd, cda, dab;
    cof(abc, A, bc, false, B, ac, true, C, ab, false);
    cof(bcd, B, cd, false, C, bd, true, Dv, bc, false);
    cof(cda, C, da, false, Dv, ac, false, A, cd, false);
    cof(dab, Dv, ab, false, A, bd, false, B, da, false);
    // det = (dlift*abc - clift*dab) + (blift*cda - alift*bcd)
    L10 det, t;
    limbs_smul(lift[Dv], abc, det);
    limbs_smul(lift[C], dab, t);
    limbs_neg(t);
    limbs_add(det, t);
    limbs_smul(lift[B], cda, t);
    limbs_add(det, t);
    limbs_smul(lift[A], bcd, t);
    limbs_neg(t);
    limbs_add(det, t);
    unsigned long long any = 0;
    for (int i = 0; i < 10; ++i) any |= det[i];
    sign = any == 0 ? 0 : ((long long)det[9] < 0 ? -1 : 1);
    return true;
}

// orient3d in fixed point with 126-bit differences (minors 256-bit, products and sum
// 384-bit), for the exact zeros whose differences span more than 62 bits.
__host__ __device__ inline bool orient3d_sign_fixed_wide(const double* pa, const double* pb,
                                                         const double* pc, const double* pd,
                                                         int& sign) {
    const double* P[3] = {pa, pb, pc};
    double dx[9], dy[9];
    int lo = 1 << 20, hi = -(1 << 20);
    for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a) {
            double x, y;
            two_sum(P[v][a], -pd[a], x, y);
            dx[3 * v + a] = x;
            dy[3 * v + a] = y;
            const double xy[2] = {x, y};
            for (int q = 0; q < 2; ++q) {
                if (xy[q] == 0.0) continue;
                int l, h;
                if (!bit_range(xy[q], l, h)) return false;
                lo = l < lo ? l : lo;
                hi = h > hi ? h : hi;
            }
        }
    if (hi < lo) {
        sign = 0;
        return true;
    }
    if (hi - lo > 126 || hi > 300 || lo < -300) return false;
    unsigned long long D[9][2];
    for (int i = 0; i < 9; ++i) {
        const __int128 q = scaled_int128(dx[i], lo) + scaled_int128(dy[i], lo);
        D[i][0] = (unsigned long long)q;
        D[i][1] = (unsigned long long)((u128)q >> 64);
    }
    typedef unsigned long long L4[4];
    typedef unsigned long long L6[6];
    auto minor = [&](int u, int v, L4& m) {  // D[u]x * D[v]y - D[v]x * D[u]y (rows a, b, c)
        L4 t;
        limbs_smul(D[3 * u], D[3 * v + 1], m);
        limbs_smul(D[3 * v], D[3 * u + 1], t);
        limbs_neg(t);
        limbs_add(m, t);
    };
    // adz*(bdx*cdy - cdx*bdy) + bdz*(cdx*ady - adx*cdy) + cdz*(adx*bdy - bdx*ady)
    L4 mbc, mca, mab;
    minor(1, 2, mbc);
    minor(2, 0, mca);
    minor(0, 1, mab);
    L6 det, t;
    limbs_smul(D[2], mbc, det);
    limbs_smul(D[5], mca, t);
    limbs_add(det, t);
    limbs_smul(D[8], mab, t);
    limbs_add(det, t);
    unsigned long long any = 0;
    for (int i = 0; i < 6; ++i) any |= det[i];
    sign = any == 0 ? 0 : ((long long)det[5] < 0 ? -1 : 1);
    return true;
}

// insphere sign from the exact double-word differences p - pe, evaluated in double-word
// arithmetic (dw_mul relative error <= 4u^2, dw_add <= 3u^2 + 13u^3, u = 2^-53). Following
// the error through minors (<= 7.1u^2 of their permanents), cofactors (<= 17.4u^2), lifts
// (<= 10.2u^2), lifted products (<= 31.6u^2) and the final sums gives
// |det_dw - det| <= 41u^2 P < 2^-100 P, P the permanent (every product taken in absolute
// value); the test uses 2^-96 P. Components must lie in [2^-150, 2^150] so that no product
// or error term leaves the normal range. False when |det_dw| is inside the bound (in
// practice only exact zeros, which the fixed-point stage or the expansions resolve).
__host__ __device__ inline bool insphere_sign_dw(const double* pa, const double* pb,
                                                 const double* pc, const double* pd,
                                                 const double* pe, int& sign) {
    const double* P[4] = {pa, pb, pc, pd};
    DW d[12];
    double h[12];  // |hi| of each difference
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            DW x;
            two_sum(P[v][a], -pe[a], x.hi, x.lo);
            const double ax = x.hi < 0.0 ? -x.hi : x.hi;
            if (ax != 0.0 && !(ax >= 0x1p-150 && ax <= 0x1p150)) return false;
            const double al = x.lo < 0.0 ? -x.lo : x.lo;
            if (al != 0.0 && !(al >= 0x1p-150)) return false;
            d[3 * v + a] = x;
            h[3 * v + a] = ax;
        }
    auto X = [&](int v) { return d[3 * v]; };
    auto Y = [&](int v) { return d[3 * v + 1]; };
    auto Z = [&](int v) { return d[3 * v + 2]; };
    auto minor = [&](int u, int v) { return dw_add(dw_mul(X(u), Y(v)), dw_neg(dw_mul(X(v), Y(u)))); };
    auto mperm = [&](int u, int v) { return h[3 * u] * h[3 * v + 1] + h[3 * v] * h[3 * u + 1]; };
    constexpr int A = 0, B = 1, C = 2, Dv = 3;
    const DW ab = minor(A, B), bc = minor(B, C), cd = minor(C, Dv), da = minor(Dv, A),
             ac = minor(A, C), bd = minor(B, Dv);
    const double pab = mperm(A, B), pbc = mperm(B, C), pcd = mperm(C, Dv), pda = mperm(Dv, A),
                 pac = mperm(A, C), pbd = mperm(B, Dv);
    auto zm = [&](int v, DW m) { return dw_mul(Z(v), m); };
    // abc = aez*bc - bez*ac + cez*ab      bcd = bez*cd - cez*bd + dez*bc
    // cda = cez*da + dez*ac + aez*cd      dab = dez*ab + aez*bd + bez*da
    const DW abc = dw_add(dw_add(zm(A, bc), dw_neg(zm(B, ac))), zm(C, ab));
    const DW bcd = dw_add(dw_add(zm(B, cd), dw_neg(zm(C, bd))), zm(Dv, bc));
    const DW cda = dw_add(dw_add(zm(C, da), zm(Dv, ac)), zm(A, cd));
    const DW dab = dw_add(dw_add(zm(Dv, ab), zm(A, bd)), zm(B, da));
    const double zh[4] = {h[2], h[5], h[8], h[11]};
    const double qabc = zh[A] * pbc + zh[B] * pac + zh[C] * pab;
    const double qbcd = zh[B] * pcd + zh[C] * pbd + zh[Dv] * pbc;
    const double qcda = zh[C] * pda + zh[Dv] * pac + zh[A] * pcd;
    const double qdab = zh[Dv] * pab + zh[A] * pbd + zh[B] * pda;
    DW lift[4];
    double lh[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        lift[v] = dw_add(dw_add(dw_mul(X(v), X(v)), dw_mul(Y(v), Y(v))), dw_mul(Z(v), Z(v)));
        lh[v] = h[3 * v] * h[3 * v] + h[3 * v + 1] * h[3 * v + 1] + h[3 * v + 2] * h[3 * v + 2];
    }
    // det = (dlift*abc - clift*dab) + (blift*cda - alift*bcd)
    const DW det = dw_add(dw_add(dw_mul(lift[Dv], abc), dw_neg(dw_mul(lift[C], dab))),
                          dw_add(dw_mul(lift[B], cda), dw_neg(dw_mul(lift[A], bcd))));
    const double perm = lh[Dv] * qabc + lh[C] * qdab + lh[B] * qcda + lh[A] * qbcd;
    // the hi parts understate the double-word magnitudes by at most a factor 1 + 2u per
    // difference; (1 + 2^-40) covers that and the rounding of perm
    const double bound = 0x1p-96 * perm * (1.0 + 0x1p-40);
    const double v = det.hi + det.lo;
    if (v > bound) {
        sign = 1;
        return true;
    }
    if (-v > bound) {
        sign = -1;
        return true;
    }
    return false;
}

// The two cheapest register-only exact insphere stages for one lane: fixed point with
// 31-bit differences (exact zeros and small determinants of moderately spread tuples),
// then the double-word bound (nonzero determinants of any spread within range). False
// when both decline; the caller then tries fixed point with 62- and 126-bit differences.
__host__ __device__ inline bool insphere_sign_inline(const double* pa, const double* pb,
                                                     const double* pc, const double* pd,
                                                     const double* pe, int& sign) {
    {
        InsDiffs d;
        if (insphere_diffs(pa, pb, pc, pd, pe, d)) {
            if (d.hi < d.lo) {
                sign = 0;
                return true;
            }
            if (insphere_fixed_fits(d, 31)) {
                sign = insphere_fixed31_sign(d);
                return true;
            }
        }
    }
    return insphere_sign_dw(pa, pb, pc, pd, pe, sign);
}

// Exact sign of orient3d(pa, pb, pc, pd); kOverflow if CAP is too small. FAST = false
// skips the register-only stages (for lanes that already failed them).
template <int CAP, bool FAST = true>
__host__ __device__ inline int orient3d_exact_sign(const double* pa, const double* pb,
                                          const double* pc, const double* pd) {
    int fast;
    if (FAST && orient3d_sign_exact_diffs(pa, pb, pc, pd, fast)) return fast;
    if (FAST && orient3d_sign_dw(pa, pb, pc, pd, fast)) return fast;
    if (orient3d_sign_fixed(pa, pb, pc, pd, fast)) return fast;
    if (orient3d_sign_fixed_wide(pa, pb, pc, pd, fast)) return fast;
    double top;
    if (!orient3d_exact<CAP>(pa, pb, pc, pd, top)) return kOverflow;
    return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

// Exact insphere(pa, pb, pc, pd, pe) as an expansion; `top` receives its largest
// component (same sign as the determinant, within a relative 2^-52 of its value).
// Returns false if a capacity is too small: CCAP bounds a 3x3 cofactor, LCAP a lifted
// product.
template <int CCAP, int LCAP>
__host__ __device__ inline bool insphere_exact(const double* pa, const double* pb, const double* pc,
                                      const double* pd, const double* pe, double& top) {
    double ex[4][2], ey[4][2], ez[4][2];
    int nx[4], ny[4], nz[4];
    const double* P[4] = {pa, pb, pc, pd};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        nx[v] = exp_diff(P[v][0], pe[0], ex[v]);
        ny[v] = exp_diff(P[v][1], pe[1], ey[v]);
        nz[v] = exp_diff(P[v][2], pe[2], ez[v]);
    }
    // The four cofactors of the naive formula (predicates.cpp:4072-4102), as
    //   abc = aez*bc - bez*ac + cez*ab      bcd = bez*cd - cez*bd + dez*bc
    //   cda = cez*da + dez*ac + aez*cd      dab = dez*ab + aez*bd + bez*da
    // with xy minors uv = uex*vey - vex*uey, and
    //   det = (dlift*abc - clift*dab) + (blift*cda - alift*bcd).
    // Table: for each term, (lift vertex, sign, then three (z vertex, minor u, minor v,
    // sign) entries).
    constexpr int A = 0, B = 1, C = 2, D = 3;
    struct Entry { int z, u, v, s; };
    struct Term { int lift, sgn; Entry e[3]; };
    const Term terms[4] = {
        {D, 1, {{A, B, C, 1}, {B, A, C, -1}, {C, A, B, 1}}},   // dlift * abc
        {C, -1, {{D, A, B, 1}, {A, B, D, 1}, {B, D, A, 1}}},   // - clift * dab
        {B, 1, {{C, D, A, 1}, {D, A, C, 1}, {A, C, D, 1}}},    // + blift * cda
        {A, -1, {{B, C, D, 1}, {C, B, D, -1}, {D, B, C, 1}}},  // - alift * bcd
    };
    double acc[LCAP];
    int an = 0;
    for (int ti = 0; ti < 4; ++ti) {
        const Term& T = terms[ti];
        double cof[CCAP];
        int cn = 0;
        for (int k = 0; k < 3; ++k) {
            const Entry& E = T.e[k];
            double m[16];
            int mn = exp_minor(ex[E.u], nx[E.u], ey[E.v], ny[E.v], ex[E.v], nx[E.v],
                               ey[E.u], ny[E.u], m);
            if (!exp_accumulate_zm<CCAP>(cof, cn, ez[E.z], nz[E.z], m, mn, E.s))
                return false;
        }
        // lift = x^2 + y^2 + z^2 of the term's vertex
        double lift[24];
        int ln;
        {
            double a[8], b[8], t[16];
            int an2 = exp_mul_small(ex[T.lift], nx[T.lift], ex[T.lift], nx[T.lift], a);
            int bn2 = exp_mul_small(ey[T.lift], ny[T.lift], ey[T.lift], ny[T.lift], b);
            int tn = exp_sum(a, an2, b, bn2, t);
            int cn2 = exp_mul_small(ez[T.lift], nz[T.lift], ez[T.lift], nz[T.lift], a);
            ln = exp_sum(t, tn, a, cn2, lift);
            ln = exp_compress(lift, ln);
        }
        // acc += sgn * lift * cof, one lift component at a time
        for (int j = 0; j < ln; ++j) {
            if (2 * cn > LCAP) return false;
            double part[2 * CCAP];
            int pn = exp_scale(cof, cn, lift[j], part);
            if (T.sgn < 0) exp_negate(part, pn);
            if (an + pn > LCAP) return false;
            double sum[LCAP];
            int sn = exp_sum(acc, an, part, pn, sum);
            sn = exp_compress(sum, sn);
            for (int i = 0; i < sn; ++i) acc[i] = sum[i];
            an = sn;
        }
    }
    top = acc[an - 1];
    return true;
}

// Exact sign of insphere(pa, pb, pc, pd, pe); kOverflow if a capacity is too small.
template <int CCAP, int LCAP>
__host__ __device__ inline int insphere_exact_sign(const double* pa, const double* pb,
                                          const double* pc, const double* pd,
                                          const double* pe) {
    int fixed;  // differences within 126 bits: exact in registers
    if (insphere_sign_fixed_wide(pa, pb, pc, pd, pe, fixed)) return fixed;
    double top;
    if (!insphere_exact<CCAP, LCAP>(pa, pb, pc, pd, pe, top)) return kOverflow;
    return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

}  // namespace cmg::hip
