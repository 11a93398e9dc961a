/* quality.c — CPU oracle for the mesh quality report.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Compiled with -ffp-contract=off.
 *
 * Restates src/quality/metrics.cpp:1-172 (volume, shortest edge, radius-edge ratio via
 * include/cmg/core/circumcenter.hpp:15-46, the six dihedral angles, the 18-bucket
 * histogram, sequential sums in tet order) operation for operation. The worst-tets list
 * is ordered by descending ratio with ties by ascending index (the reference's std::sort
 * leaves the order of equal ratios unspecified).
 *
 * Also the point-location walk of src/delaunay/incremental.cpp:321-358 over a fixed
 * tetrahedral mesh with neighbours.
 */
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const double kPiQ = 3.14159265358979323846;

typedef struct {
    double x, y, z;
} v3;

static inline v3 vsub(const double* a, const double* b) {
    v3 r = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return r;
}
static inline v3 vcross(v3 a, v3 b) {
    v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double vnorm(v3 a) { return sqrt(vdot(a, a)); }

/* std::min / std::max argument order */
static inline double smin(double a, double b) { return b < a ? b : a; }
static inline double smax(double a, double b) { return a < b ? b : a; }

/* circumcenter.hpp:15-46 */
static int circumradius(const double* a, const double* b, const double* c, const double* d,
                        double* radius) {
    const double B[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double C[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double D[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
    double CD[3] = {C[1] * D[2] - C[2] * D[1], C[2] * D[0] - C[0] * D[2],
                    C[0] * D[1] - C[1] * D[0]};
    double DB[3] = {D[1] * B[2] - D[2] * B[1], D[2] * B[0] - D[0] * B[2],
                    D[0] * B[1] - D[1] * B[0]};
    double BC[3] = {B[1] * C[2] - B[2] * C[1], B[2] * C[0] - B[0] * C[2],
                    B[0] * C[1] - B[1] * C[0]};
    const double denom = 2.0 * (B[0] * CD[0] + B[1] * CD[1] + B[2] * CD[2]);
    const double nB = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
    const double nC = C[0] * C[0] + C[1] * C[1] + C[2] * C[2];
    const double nD = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
    /* std::max({nB, nC, nD, 1e-300}): the first of the largest */
    double m = nB;
    if (m < nC) m = nC;
    if (m < nD) m = nD;
    if (m < 1e-300) m = 1e-300;
    const double scale = sqrt(m);
    if (fabs(denom) < 1e-12 * scale * scale * scale) return 0;
    double num[3];
    for (int i = 0; i < 3; i++) num[i] = (nB * CD[i] + nC * DB[i] + nD * BC[i]) / denom;
    const double cx = a[0] + num[0], cy = a[1] + num[1], cz = a[2] + num[2];
    *radius = sqrt(num[0] * num[0] + num[1] * num[1] + num[2] * num[2]);
    return isfinite(cx) && isfinite(cy) && isfinite(cz) && isfinite(*radius);
}

/* metrics.cpp:52-77: the six dihedral angles in degrees */
void orc_tet_dihedrals(const double* P0, const double* P1, const double* P2,
                       const double* P3, double* out) {
    const double* P[4] = {P0, P1, P2, P3};
    v3 nrm[4];
    for (int e = 0; e < 4; ++e) {
        const int a = (e + 1) % 4, b = (e + 2) % 4, c = (e + 3) % 4;
        v3 nv = vcross(vsub(P[b], P[a]), vsub(P[c], P[a]));
        if (vdot(nv, vsub(P[e], P[a])) > 0.0) {
            nv.x = -nv.x;
            nv.y = -nv.y;
            nv.z = -nv.z;
        }
        nrm[e] = nv;
    }
    int idx = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) {
            int k = -1, l = -1;
            for (int mm = 0; mm < 4; ++mm)
                if (mm != i && mm != j) {
                    if (k < 0)
                        k = mm;
                    else
                        l = mm;
                }
            const double na = vnorm(nrm[k]), nb = vnorm(nrm[l]);
            double ang;
            if (na == 0.0 || nb == 0.0) {
                ang = 0.0;
            } else {
                double cd = vdot(nrm[k], nrm[l]) / (na * nb);
                cd = cd < -1.0 ? -1.0 : (1.0 < cd ? 1.0 : cd); /* std::clamp */
                ang = (kPiQ - acos(cd)) * 180.0 / kPiQ;
            }
            out[idx++] = ang;
        }
}

/* metrics.cpp:36-48 */
double orc_tet_volume(const double* a, const double* b, const double* c, const double* d) {
    const v3 ab = vsub(b, a), ac = vsub(c, a), ad = vsub(d, a);
    return fabs(vdot(ab, vcross(ac, ad))) / 6.0;
}

/* per-tet values: ratio (DBL_MAX when degenerate), valid flag, volume, 6 dihedrals */
void orc_tet_quality(const double* pts, const int* tet, double* ratio, int* valid,
                     double* vol, double* dih) {
    const double* P[4] = {pts + 3 * (size_t)tet[0], pts + 3 * (size_t)tet[1],
                          pts + 3 * (size_t)tet[2], pts + 3 * (size_t)tet[3]};
    *vol = orc_tet_volume(P[0], P[1], P[2], P[3]);
    double se = INFINITY;
    for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) se = smin(se, vnorm(vsub(P[j], P[i])));
    double radius = 0;
    int ok = circumradius(P[0], P[1], P[2], P[3], &radius);
    if (ok && se > 0.0 && isfinite(radius)) {
        *ratio = radius / se;
        *valid = 1;
    } else {
        *ratio = DBL_MAX;
        *valid = 0;
    }
    orc_tet_dihedrals(P[0], P[1], P[2], P[3], dih);
}

static const double* g_ratio;
static int cmp_worst(const void* x, const void* y) {
    const int a = *(const int*)x, b = *(const int*)y;
    if (g_ratio[a] != g_ratio[b]) return g_ratio[a] > g_ratio[b] ? -1 : 1;
    return a < b ? -1 : (a > b);
}

/* metrics.cpp:81-169 */
void orc_quality_report(const double* pts, const int* tets, size_t n, int worst_count,
                        orc_quality* r, int* worst) {
    memset(r, 0, sizeof *r);
    r->num_tets = (int64_t)n;
    double* ratio = (double*)malloc(sizeof(double) * (n ? n : 1));
    double min_re = 0, max_re = 0, sum_re = 0, min_dih = 0, max_dih = 0, sum_dih = 0;
    double min_vol = 0, max_vol = 0, total_vol = 0;
    size_t re_count = 0, dih_count = 0;
    for (size_t t = 0; t < n; ++t) {
        double vol, re, dih[6];
        int valid;
        orc_tet_quality(pts, tets + 4 * t, &re, &valid, &vol, dih);
        total_vol += vol;
        if (t == 0) {
            min_vol = max_vol = vol;
        } else {
            min_vol = smin(min_vol, vol);
            max_vol = smax(max_vol, vol);
        }
        ratio[t] = re;
        if (valid) {
            if (re_count == 0) {
                min_re = max_re = re;
            } else {
                min_re = smin(min_re, re);
                max_re = smax(max_re, re);
            }
            sum_re += re;
            ++re_count;
        }
        for (int q = 0; q < 6; ++q) {
            const double a = dih[q];
            if (dih_count == 0) {
                min_dih = max_dih = a;
            } else {
                min_dih = smin(min_dih, a);
                max_dih = smax(max_dih, a);
            }
            sum_dih += a;
            ++dih_count;
            int b = (int)floor(a / 10.0);
            b = b < 0 ? 0 : (b > 17 ? 17 : b);
            ++r->dihedral_histogram[b];
        }
    }
    r->min_radius_edge = min_re;
    r->max_radius_edge = max_re;
    r->mean_radius_edge = re_count ? sum_re / (double)re_count : 0.0;
    r->min_dihedral = min_dih;
    r->max_dihedral = max_dih;
    r->mean_dihedral = dih_count ? sum_dih / (double)dih_count : 0.0;
    r->min_volume = min_vol;
    r->max_volume = max_vol;
    r->total_volume = total_vol;
    /* every ratio is finite (DBL_MAX marks degenerate tets): all tets are candidates */
    int* order = (int*)malloc(sizeof(int) * (n ? n : 1));
    for (size_t t = 0; t < n; ++t) order[t] = (int)t;
    g_ratio = ratio;
    qsort(order, n, sizeof(int), cmp_worst);
    size_t k = n;
    if (worst_count >= 0 && k > (size_t)worst_count) k = (size_t)worst_count;
    r->num_worst = (int64_t)k;
    if (worst) memcpy(worst, order, sizeof(int) * k);
    free(order);
    free(ratio);
}

/* The six dihedral angles of each of n tets (6n doubles). */
void orc_dihedrals_batch(const double* pts, const int* tets, size_t n, double* out) {
    for (size_t t = 0; t < n; ++t) {
        const int* q = tets + 4 * t;
        orc_tet_dihedrals(pts + 3 * (size_t)q[0], pts + 3 * (size_t)q[1],
                          pts + 3 * (size_t)q[2], pts + 3 * (size_t)q[3], out + 6 * t);
    }
}

/* ---- point location (src/delaunay/incremental.cpp:321-358) ------------------------- */

static double o3(const double* a, const double* b, const double* c, const double* d) {
    return (double)orc_orient3d_sign(a, b, c, d, ORC_COMPAT);
}

/* brute_locate (:345-358): the first tet with all four orientations >= 0, else 0 */
static int locate_brute(const double* pts, const int* tets, size_t ntets, const double* q) {
    for (size_t i = 0; i < ntets; ++i) {
        const int* v = tets + 4 * i;
        const double* P0 = pts + 3 * (size_t)v[0];
        const double* P1 = pts + 3 * (size_t)v[1];
        const double* P2 = pts + 3 * (size_t)v[2];
        const double* P3 = pts + 3 * (size_t)v[3];
        if (o3(q, P1, P2, P3) >= 0 && o3(P0, q, P2, P3) >= 0 && o3(P0, P1, q, P3) >= 0 &&
            o3(P0, P1, P2, q) >= 0)
            return (int)i;
    }
    return 0;
}

/* locate (:321-337) from tet `start`; *steps (optional) receives the walk length, -1
 * when the walk gave up and brute_locate answered */
int orc_locate(const double* pts, const int* tets, const int* nbrs, size_t ntets,
               const double* q, int start, int* steps) {
    int t = start;
    const long long limit = 2 * (long long)ntets + 16;
    for (long long s = 0; s < limit; ++s) {
        const int* v = tets + 4 * (size_t)t;
        const int* n = nbrs + 4 * (size_t)t;
        const double* P0 = pts + 3 * (size_t)v[0];
        const double* P1 = pts + 3 * (size_t)v[1];
        const double* P2 = pts + 3 * (size_t)v[2];
        const double* P3 = pts + 3 * (size_t)v[3];
        if (o3(q, P1, P2, P3) < 0 && n[0] >= 0) { t = n[0]; continue; }
        if (o3(P0, q, P2, P3) < 0 && n[1] >= 0) { t = n[1]; continue; }
        if (o3(P0, P1, q, P3) < 0 && n[2] >= 0) { t = n[2]; continue; }
        if (o3(P0, P1, P2, q) < 0 && n[3] >= 0) { t = n[3]; continue; }
        if (steps) *steps = (int)s;
        return t;
    }
    if (steps) *steps = -1;
    return locate_brute(pts, tets, ntets, q);
}

void orc_locate_batch(const double* pts, const int* tets, const int* nbrs, size_t ntets,
                      const double* q, size_t nq, const int* start, int start_tet,
                      int* out, int* steps, int nthreads) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (size_t i = 0; i < nq; ++i)
        out[i] = orc_locate(pts, tets, nbrs, ntets, q + 3 * i, start ? start[i] : start_tet,
                            steps ? steps + i : 0);
}

/* Delaunay cavity of point q (src/delaunay/incremental.cpp:148-161, the flood before the
 * retriangulation, unweighted): from tet t0, breadth first, each neighbour m (k = 0..3 in
 * order) not yet in the cavity joins it when robust::insphere(v0, v1, v2, v3, q) > 0 (the
 * compat sign). Writes the cavity in discovery order to out[0..cap) and returns its size,
 * or -1 when it has more than cap tets. */
int orc_cavity(const double* pts, const int* tets, const int* nbrs, const double* q, int t0,
               int cap, int* out) {
    if (t0 < 0) return 0;
    int n = 0;
    out[n++] = t0;
    for (int h = 0; h < n; ++h) {
        const int* nb = nbrs + 4 * (size_t)out[h];
        for (int k = 0; k < 4; ++k) {
            const int m = nb[k];
            if (m < 0) continue;
            int in = 0;
            for (int e = 0; e < n && !in; ++e) in = out[e] == m;
            if (in) continue;
            const int* v = tets + 4 * (size_t)m;
            if (orc_insphere_sign(pts + 3 * (size_t)v[0], pts + 3 * (size_t)v[1],
                                  pts + 3 * (size_t)v[2], pts + 3 * (size_t)v[3], q,
                                  ORC_COMPAT) > 0) {
                if (n == cap) return -1;
                out[n++] = m;
            }
        }
    }
    return n;
}

void orc_cavity_batch(const double* pts, const int* tets, const int* nbrs, const double* q,
                      size_t nq, const int* located, int cap, int* count, int* out,
                      int nthreads) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (size_t i = 0; i < nq; ++i)
        count[i] = orc_cavity(pts, tets, nbrs, q + 3 * i, located[i], cap, out + i * (size_t)cap);
}
