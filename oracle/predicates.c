/* predicates.c — CPU oracle for orient3d / insphere signs.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Compiled with -ffp-contract=off.
 *
 * Naive determinants restate src/predicates/predicates.cpp:2194-2222 (orient3d) and
 * :4050-4106 (insphere) operation for operation. Exact signs use expansion arithmetic
 * after Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast Robust
 * Geometric Predicates" (CMU-CS-96-140): error-free Two-Sum and Dekker/Veltkamp
 * Two-Product, expansions grown one component at a time (the paper's
 * Grow-Expansion, Theorem 10), and Compress. The determinant is evaluated exactly in
 * the same difference form as the naive one: each difference a_k - e_k is first
 * made exact as a two-component expansion, then every product and sum is exact.
 */
#include <math.h>
#include <omp.h>
#include <string.h>

#include "oracle.h"

#define EPS 1.1102230246251565e-16 /* 2^-53, what exactinit derives (predicates.cpp:597-610) */
#define SPLITTER 134217729.0       /* 2^27 + 1 */

/* ---- error-free transformations -------------------------------------------------- */

static inline void two_sum(double a, double b, double* x, double* y) {
    double s = a + b;
    double bv = s - a;
    double av = s - bv;
    *x = s;
    *y = (a - av) + (b - bv);
}

static inline void fast_two_sum(double a, double b, double* x, double* y) {
    double s = a + b;
    *x = s;
    *y = b - (s - a);
}

static inline void split(double a, double* hi, double* lo) {
    double c = SPLITTER * a;
    double big = c - a;
    *hi = c - big;
    *lo = a - *hi;
}

static inline void two_prod(double a, double b, double* x, double* y) {
    double p = a * b;
    double ah, al, bh, bl;
    split(a, &ah, &al);
    split(b, &bh, &bl);
    double e1 = p - ah * bh;
    double e2 = e1 - al * bh;
    double e3 = e2 - ah * bl;
    *x = p;
    *y = al * bl - e3;
}

/* ---- expansions: components nonoverlapping, increasing magnitude, no zeros ------- */

/* h = h + b, in place (Grow-Expansion with zero elimination). Returns new length. */
static int grow(double* h, int n, double b) {
    double q = b;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        double s, t;
        two_sum(q, h[i], &s, &t);
        if (t != 0.0) h[m++] = t;
        q = s;
    }
    if (q != 0.0 || m == 0) h[m++] = q;
    return m;
}

/* h (len n) += f (len fn), in place; h needs room for n + fn. */
static int add_into(double* h, int n, const double* f, int fn) {
    for (int j = 0; j < fn; ++j) n = grow(h, n, f[j]);
    return n;
}

/* Compress (Shewchuk): same value, usually far fewer components. */
static int compress(double* e, int n) {
    if (n <= 1) return n;
    double q = e[n - 1];
    int bottom = n - 1;
    for (int i = n - 2; i >= 0; --i) {
        double qn, r;
        fast_two_sum(q, e[i], &qn, &r);
        if (r != 0.0) {
            e[bottom--] = qn;
            q = r;
        } else {
            q = qn;
        }
    }
    e[bottom] = q;
    int top = 0;
    q = e[bottom];
    for (int i = bottom + 1; i < n; ++i) {
        double qn, r;
        fast_two_sum(e[i], q, &qn, &r);
        if (r != 0.0) e[top++] = r;
        q = qn;
    }
    e[top++] = q;
    return top;
}

/* h = e * f exactly; h needs room for 2*en*fn. */
static int mul(const double* e, int en, const double* f, int fn, double* h) {
    int n = 0;
    for (int j = 0; j < fn; ++j)
        for (int i = 0; i < en; ++i) {
            double p, q;
            two_prod(e[i], f[j], &p, &q);
            if (q != 0.0) n = grow(h, n, q);
            n = grow(h, n, p);
        }
    if (n == 0) h[n++] = 0.0;
    return compress(h, n);
}

static void negate(double* e, int n) {
    for (int i = 0; i < n; ++i) e[i] = -e[i];
}

static int sign_of_expansion(const double* e, int n) {
    double top = e[n - 1];
    return top > 0 ? 1 : (top < 0 ? -1 : 0);
}

/* exact difference a - b as a 1- or 2-component expansion */
static int diff2(double a, double b, double* h) {
    double x, y;
    two_sum(a, -b, &x, &y);
    if (y != 0.0) {
        h[0] = y;
        h[1] = x;
        return 2;
    }
    h[0] = x;
    return 1;
}

/* minor = p*q - r*s over 2-component entries; out needs 32 */
static int minor2(const double* p, int pn, const double* q, int qn, const double* r,
                  int rn, const double* s, int sn, double* out) {
    double t[8];
    int n = mul(p, pn, q, qn, out);
    int tn = mul(r, rn, s, sn, t);
    negate(t, tn);
    n = add_into(out, n, t, tn);
    return compress(out, n);
}

/* ---- naive determinants ---------------------------------------------------------- */

double orc_orient3d_naive(const double* pa, const double* pb, const double* pc,
                          const double* pd) {
    /* predicates.cpp:2201-2222 */
    double adx = pa[0] - pd[0], ady = pa[1] - pd[1], adz = pa[2] - pd[2];
    double bdx = pb[0] - pd[0], bdy = pb[1] - pd[1], bdz = pb[2] - pd[2];
    double cdx = pc[0] - pd[0], cdy = pc[1] - pd[1], cdz = pc[2] - pd[2];
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    double cdxady = cdx * ady, adxcdy = adx * cdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    return adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
}

static double orient3d_permanent(const double* pa, const double* pb, const double* pc,
                                 const double* pd) {
    /* predicates.cpp:2235-2237 */
    double adx = pa[0] - pd[0], ady = pa[1] - pd[1], adz = pa[2] - pd[2];
    double bdx = pb[0] - pd[0], bdy = pb[1] - pd[1], bdz = pb[2] - pd[2];
    double cdx = pc[0] - pd[0], cdy = pc[1] - pd[1], cdz = pc[2] - pd[2];
    return (fabs(bdx * cdy) + fabs(cdx * bdy)) * fabs(adz) +
           (fabs(cdx * ady) + fabs(adx * cdy)) * fabs(bdz) +
           (fabs(adx * bdy) + fabs(bdx * ady)) * fabs(cdz);
}

double orc_insphere_naive(const double* pa, const double* pb, const double* pc,
                          const double* pd, const double* pe) {
    /* predicates.cpp:4063-4106 */
    double aex = pa[0] - pe[0], bex = pb[0] - pe[0], cex = pc[0] - pe[0],
           dex = pd[0] - pe[0];
    double aey = pa[1] - pe[1], bey = pb[1] - pe[1], cey = pc[1] - pe[1],
           dey = pd[1] - pe[1];
    double aez = pa[2] - pe[2], bez = pb[2] - pe[2], cez = pc[2] - pe[2],
           dez = pd[2] - pe[2];
    double aexbey = aex * bey, bexaey = bex * aey;
    double ab = aexbey - bexaey;
    double bexcey = bex * cey, cexbey = cex * bey;
    double bc = bexcey - cexbey;
    double cexdey = cex * dey, dexcey = dex * cey;
    double cd = cexdey - dexcey;
    double dexaey = dex * aey, aexdey = aex * dey;
    double da = dexaey - aexdey;
    double aexcey = aex * cey, cexaey = cex * aey;
    double ac = aexcey - cexaey;
    double bexdey = bex * dey, dexbey = dex * bey;
    double bd = bexdey - dexbey;
    double abc = aez * bc - bez * ac + cez * ab;
    double bcd = bez * cd - cez * bd + dez * bc;
    double cda = cez * da + dez * ac + aez * cd;
    double dab = dez * ab + aez * bd + bez * da;
    double alift = aex * aex + aey * aey + aez * aez;
    double blift = bex * bex + bey * bey + bez * bez;
    double clift = cex * cex + cey * cey + cez * cez;
    double dlift = dex * dex + dey * dey + dez * dez;
    return (dlift * abc - clift * dab) + (blift * cda - alift * bcd);
}

static double insphere_permanent(const double* pa, const double* pb, const double* pc,
                                 const double* pd, const double* pe) {
    /* predicates.cpp:4122-4156 */
    double aex = pa[0] - pe[0], bex = pb[0] - pe[0], cex = pc[0] - pe[0],
           dex = pd[0] - pe[0];
    double aey = pa[1] - pe[1], bey = pb[1] - pe[1], cey = pc[1] - pe[1],
           dey = pd[1] - pe[1];
    double aez = pa[2] - pe[2], bez = pb[2] - pe[2], cez = pc[2] - pe[2],
           dez = pd[2] - pe[2];
    double aezp = fabs(aez), bezp = fabs(bez), cezp = fabs(cez), dezp = fabs(dez);
    double aexbeyp = fabs(aex * bey), bexaeyp = fabs(bex * aey);
    double bexceyp = fabs(bex * cey), cexbeyp = fabs(cex * bey);
    double cexdeyp = fabs(cex * dey), dexceyp = fabs(dex * cey);
    double dexaeyp = fabs(dex * aey), aexdeyp = fabs(aex * dey);
    double aexceyp = fabs(aex * cey), cexaeyp = fabs(cex * aey);
    double bexdeyp = fabs(bex * dey), dexbeyp = fabs(dex * bey);
    double alift = aex * aex + aey * aey + aez * aez;
    double blift = bex * bex + bey * bey + bez * bez;
    double clift = cex * cex + cey * cey + cez * cez;
    double dlift = dex * dex + dey * dey + dez * dez;
    return ((cexdeyp + dexceyp) * bezp + (dexbeyp + bexdeyp) * cezp +
            (bexceyp + cexbeyp) * dezp) * alift +
           ((dexaeyp + aexdeyp) * cezp + (aexceyp + cexaeyp) * dezp +
            (cexdeyp + dexceyp) * aezp) * blift +
           ((aexbeyp + bexaeyp) * dezp + (bexdeyp + dexbeyp) * aezp +
            (dexaeyp + aexdeyp) * bezp) * clift +
           ((bexceyp + cexbeyp) * aezp + (cexaeyp + aexceyp) * bezp +
            (aexbeyp + bexaeyp) * cezp) * dlift;
}

/* ---- exact signs ----------------------------------------------------------------- */

static int orient3d_exact_expansion(const double* pa, const double* pb, const double* pc,
                                    const double* pd) {
    double adx[2], ady[2], adz[2], bdx[2], bdy[2], bdz[2], cdx[2], cdy[2], cdz[2];
    int nadx = diff2(pa[0], pd[0], adx), nady = diff2(pa[1], pd[1], ady),
        nadz = diff2(pa[2], pd[2], adz);
    int nbdx = diff2(pb[0], pd[0], bdx), nbdy = diff2(pb[1], pd[1], bdy),
        nbdz = diff2(pb[2], pd[2], bdz);
    int ncdx = diff2(pc[0], pd[0], cdx), ncdy = diff2(pc[1], pd[1], cdy),
        ncdz = diff2(pc[2], pd[2], cdz);
    double m[32], t[128], acc[400];
    int na = 0;
    /* adz * (bdx*cdy - cdx*bdy) */
    int nm = minor2(bdx, nbdx, cdy, ncdy, cdx, ncdx, bdy, nbdy, m);
    int nt = mul(adz, nadz, m, nm, t);
    na = add_into(acc, na, t, nt);
    /* + bdz * (cdx*ady - adx*cdy) */
    nm = minor2(cdx, ncdx, ady, nady, adx, nadx, cdy, ncdy, m);
    nt = mul(bdz, nbdz, m, nm, t);
    na = add_into(acc, na, t, nt);
    /* + cdz * (adx*bdy - bdx*ady) */
    nm = minor2(adx, nadx, bdy, nbdy, bdx, nbdx, ady, nady, m);
    nt = mul(cdz, ncdz, m, nm, t);
    na = add_into(acc, na, t, nt);
    na = compress(acc, na);
    return sign_of_expansion(acc, na);
}

static int lift2(const double* x, int xn, const double* y, int yn, const double* z,
                 int zn, double* out) {
    double t[8];
    int n = mul(x, xn, x, xn, out);
    int tn = mul(y, yn, y, yn, t);
    n = add_into(out, n, t, tn);
    tn = mul(z, zn, z, zn, t);
    n = add_into(out, n, t, tn);
    return compress(out, n);
}

/* c = sa*za*m1 + sb*zb*m2 + sc*zc*m3 for 3x3 cofactors, signs +-1 */
static int cofactor3(const double* z1, int n1, const double* m1, int k1, int s1,
                     const double* z2, int n2, const double* m2, int k2, int s2,
                     const double* z3, int n3, const double* m3, int k3, int s3,
                     double* out) {
    double t[128];
    int n = 0;
    int tn = mul(z1, n1, m1, k1, t);
    if (s1 < 0) negate(t, tn);
    n = add_into(out, n, t, tn);
    tn = mul(z2, n2, m2, k2, t);
    if (s2 < 0) negate(t, tn);
    n = add_into(out, n, t, tn);
    tn = mul(z3, n3, m3, k3, t);
    if (s3 < 0) negate(t, tn);
    n = add_into(out, n, t, tn);
    return compress(out, n);
}

static int insphere_exact_expansion(const double* pa, const double* pb, const double* pc,
                                    const double* pd, const double* pe) {
    double aex[2], bex[2], cex[2], dex[2], aey[2], bey[2], cey[2], dey[2];
    double aez[2], bez[2], cez[2], dez[2];
    int naex = diff2(pa[0], pe[0], aex), nbex = diff2(pb[0], pe[0], bex),
        ncex = diff2(pc[0], pe[0], cex), ndex = diff2(pd[0], pe[0], dex);
    int naey = diff2(pa[1], pe[1], aey), nbey = diff2(pb[1], pe[1], bey),
        ncey = diff2(pc[1], pe[1], cey), ndey = diff2(pd[1], pe[1], dey);
    int naez = diff2(pa[2], pe[2], aez), nbez = diff2(pb[2], pe[2], bez),
        ncez = diff2(pc[2], pe[2], cez), ndez = diff2(pd[2], pe[2], dez);
    double ab[32], bc[32], cd[32], da[32], ac[32], bd[32];
    int nab = minor2(aex, naex, bey, nbey, bex, nbex, aey, naey, ab);
    int nbc = minor2(bex, nbex, cey, ncey, cex, ncex, bey, nbey, bc);
    int ncd = minor2(cex, ncex, dey, ndey, dex, ndex, cey, ncey, cd);
    int nda = minor2(dex, ndex, aey, naey, aex, naex, dey, ndey, da);
    int nac = minor2(aex, naex, cey, ncey, cex, ncex, aey, naey, ac);
    int nbd = minor2(bex, nbex, dey, ndey, dex, ndex, bey, nbey, bd);
    double abc[400], bcd[400], cda[400], dab[400];
    /* abc = aez*bc - bez*ac + cez*ab */
    int nabc = cofactor3(aez, naez, bc, nbc, 1, bez, nbez, ac, nac, -1, cez, ncez, ab, nab,
                         1, abc);
    /* bcd = bez*cd - cez*bd + dez*bc */
    int nbcd = cofactor3(bez, nbez, cd, ncd, 1, cez, ncez, bd, nbd, -1, dez, ndez, bc, nbc,
                         1, bcd);
    /* cda = cez*da + dez*ac + aez*cd */
    int ncda = cofactor3(cez, ncez, da, nda, 1, dez, ndez, ac, nac, 1, aez, naez, cd, ncd,
                         1, cda);
    /* dab = dez*ab + aez*bd + bez*da */
    int ndab = cofactor3(dez, ndez, ab, nab, 1, aez, naez, bd, nbd, 1, bez, nbez, da, nda,
                         1, dab);
    double al[32], bl[32], cl[32], dl[32];
    int nal = lift2(aex, naex, aey, naey, aez, naez, al);
    int nbl = lift2(bex, nbex, bey, nbey, bez, nbez, bl);
    int ncl = lift2(cex, ncex, cey, ncey, cez, ncez, cl);
    int ndl = lift2(dex, ndex, dey, ndey, dez, ndez, dl);
    /* det = (dlift*abc - clift*dab) + (blift*cda - alift*bcd); sizes after compress are
       small, so the products are sized from the actual lengths. */
    int cap = 2 * (ndl * nabc + ncl * ndab + nbl * ncda + nal * nbcd) + 8;
    double acc[cap];
    double tmp[cap];
    int n = 0, tn;
    tn = mul(dl, ndl, abc, nabc, tmp);
    n = add_into(acc, n, tmp, tn);
    tn = mul(cl, ncl, dab, ndab, tmp);
    negate(tmp, tn);
    n = add_into(acc, n, tmp, tn);
    tn = mul(bl, nbl, cda, ncda, tmp);
    n = add_into(acc, n, tmp, tn);
    tn = mul(al, nal, bcd, nbcd, tmp);
    negate(tmp, tn);
    n = add_into(acc, n, tmp, tn);
    n = compress(acc, n);
    return sign_of_expansion(acc, n);
}

static int sgn(double v) { return v > 0 ? 1 : (v < 0 ? -1 : 0); }

int orc_orient3d_exact(const double* pa, const double* pb, const double* pc,
                       const double* pd) {
    /* Shewchuk's first filter (o3derrboundA, predicates.cpp:619 and :2235-2245) is a
       proven bound on the naive value, so it may decide the sign before the exact
       stage. */
    double det = orc_orient3d_naive(pa, pb, pc, pd);
    double errbound = (7.0 + 56.0 * EPS) * EPS * orient3d_permanent(pa, pb, pc, pd);
    if (det > errbound || -det > errbound) return sgn(det);
    return orient3d_exact_expansion(pa, pb, pc, pd);
}

int orc_insphere_exact(const double* pa, const double* pb, const double* pc,
                       const double* pd, const double* pe) {
    double det = orc_insphere_naive(pa, pb, pc, pd, pe);
    double errbound =
        (16.0 + 224.0 * EPS) * EPS * insphere_permanent(pa, pb, pc, pd, pe);
    if (det > errbound || -det > errbound) return sgn(det);
    return insphere_exact_expansion(pa, pb, pc, pd, pe);
}

int orc_orient3d_sign(const double* pa, const double* pb, const double* pc,
                      const double* pd, int mode) {
    if (mode == ORC_COMPAT) {
        /* zero static filter: any nonzero naive value is returned (predicates.cpp:
           2230-2231); 0.0 falls through to the adaptive, exact-signed stages. */
        double det = orc_orient3d_naive(pa, pb, pc, pd);
        if (det > 0.0 || det < -0.0) return sgn(det);
        return orient3d_exact_expansion(pa, pb, pc, pd);
    }
    return orc_orient3d_exact(pa, pb, pc, pd);
}

int orc_insphere_sign(const double* pa, const double* pb, const double* pc,
                      const double* pd, const double* pe, int mode) {
    if (mode == ORC_COMPAT) {
        double det = orc_insphere_naive(pa, pb, pc, pd, pe);
        if (fabs(det) > 0.0) return sgn(det); /* predicates.cpp:4113 */
        return insphere_exact_expansion(pa, pb, pc, pd, pe);
    }
    return orc_insphere_exact(pa, pb, pc, pd, pe);
}

void orc_orient3d_batch(const double* soa, size_t n, size_t ld, int mode, int8_t* out,
                        int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (size_t i = 0; i < n; ++i) {
        double p[4][3];
        for (int s = 0; s < 4; ++s)
            for (int a = 0; a < 3; ++a) p[s][a] = soa[(size_t)(s * 3 + a) * ld + i];
        out[i] = (int8_t)orc_orient3d_sign(p[0], p[1], p[2], p[3], mode);
    }
}

void orc_insphere_batch(const double* soa, size_t n, size_t ld, int mode, int8_t* out,
                        int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (size_t i = 0; i < n; ++i) {
        double p[5][3];
        for (int s = 0; s < 5; ++s)
            for (int a = 0; a < 3; ++a) p[s][a] = soa[(size_t)(s * 3 + a) * ld + i];
        out[i] = (int8_t)orc_insphere_sign(p[0], p[1], p[2], p[3], p[4], mode);
    }
}

void orc_orient3d_indexed(const double* xyz, const int* quads, size_t n, int mode,
                          int8_t* out, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (size_t i = 0; i < n; ++i) {
        const int* q = quads + 4 * i;
        out[i] = (int8_t)orc_orient3d_sign(xyz + 3 * (size_t)q[0], xyz + 3 * (size_t)q[1],
                                           xyz + 3 * (size_t)q[2], xyz + 3 * (size_t)q[3],
                                           mode);
    }
}
