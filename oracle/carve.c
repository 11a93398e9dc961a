/* carve.c — CPU oracle for the point-in-domain classification of the PLC carve.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Compiled with -ffp-contract=off so every
 * FP64 expression rounds exactly as the reference's x86-64 build does.
 *
 * Two classifiers:
 *  - rays: src/constrained/tetrahedralize_plc.cpp:39-168 (CarveGrid): three fixed
 *    generic ray directions, Moller-Trumbore hits with the reference's epsilons, one
 *    projection grid per direction whose cell lists decide which triangles a query
 *    tests, majority of the three parities; tet centroids as at :170-174.
 *  - parity: the voxelizer's exact column parity (src/voxelize/voxelize.cpp:78-195)
 *    evaluated at an arbitrary point instead of a cell centre: the column through
 *    (x, y) with the xy bucket of the cell that contains it, kNudge retries, and the
 *    count of crossing heights strictly above z.
 */
#include <limits.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NT(n) ((n) > 0 ? (n) : omp_get_max_threads())

/* tetrahedralize_plc.cpp:63-65 */
static const double kRayDirs[3][3] = {{0.5773269, 0.5774341, 0.5771914},
                                      {0.3251097, -0.7211021, 0.6119274},
                                      {-0.8017412, 0.2673021, 0.5346203}};

/* static_cast<int>(double) as the reference's x86-64 build executes it (cvttsd2si):
 * truncation toward zero, and INT_MIN for NaN and for values outside the int range. */
int orc_x86_trunc(double v) {
    if (!(v > -2147483649.0 && v < 2147483648.0)) return INT_MIN;
    return (int)v;
}

/* tetrahedralize_plc.cpp:39-58 */
int orc_ray_hits(const double* o, const double* d, const double* a, const double* b,
                 const double* c) {
    double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2],
                   d[0] * e2[1] - d[1] * e2[0]};
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (fabs(det) < 1e-12) return 0;
    double inv = 1.0 / det;
    double t0[3] = {o[0] - a[0], o[1] - a[1], o[2] - a[2]};
    double u = (t0[0] * p[0] + t0[1] * p[1] + t0[2] * p[2]) * inv;
    if (u < 0 || u > 1) return 0;
    double q[3] = {t0[1] * e1[2] - t0[2] * e1[1], t0[2] * e1[0] - t0[0] * e1[2],
                   t0[0] * e1[1] - t0[1] * e1[0]};
    double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    if (v < 0 || u + v > 1) return 0;
    double t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    return t > 1e-9;
}

typedef struct {
    double u[3], v[3];
    double min_u, min_v, inv_u, inv_v;
    int* start; /* r*r + 1 */
    int* items;
} carve_dir;

struct orc_carve {
    const double* pts;
    const int* tris;
    size_t ntris;
    int r;
    carve_dir dir[3];
};

/* tetrahedralize_plc.cpp:100-119 */
static void carve_basis(const double* d, double* u, double* v) {
    double n[3] = {d[0], d[1], d[2]};
    double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int i = 0; i < 3; ++i) n[i] /= len;
    double h[3] = {1, 0, 0};
    if (fabs(n[0]) > 0.9) {
        h[0] = 0;
        h[1] = 1;
    }
    u[0] = n[1] * h[2] - n[2] * h[1];
    u[1] = n[2] * h[0] - n[0] * h[2];
    u[2] = n[0] * h[1] - n[1] * h[0];
    double ul = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int i = 0; i < 3; ++i) u[i] /= ul;
    v[0] = n[1] * u[2] - n[2] * u[1];
    v[1] = n[2] * u[0] - n[0] * u[2];
    v[2] = n[0] * u[1] - n[1] * u[0];
}

static inline double carve_proj(const double* a, const double* p) {
    return a[0] * p[0] + a[1] * p[1] + a[2] * p[2];
}

static inline int carve_clampi(int i, int r) { return i < 0 ? 0 : (i >= r ? r - 1 : i); }

/* tetrahedralize_plc.cpp:123-128 */
static inline int carve_iu(const carve_dir* g, const double* p, int r) {
    return carve_clampi(orc_x86_trunc((carve_proj(g->u, p) - g->min_u) * g->inv_u), r);
}
static inline int carve_iv(const carve_dir* g, const double* p, int r) {
    return carve_clampi(orc_x86_trunc((carve_proj(g->v, p) - g->min_v) * g->inv_v), r);
}

int orc_carve_resolution(size_t ntris) {
    /* tetrahedralize_plc.cpp:73-75 */
    int r = 1;
    while ((long long)r * r < (long long)ntris) ++r;
    return r < 256 ? r : 256;
}

orc_carve* orc_carve_create(const double* pts, size_t npts, const int* tris,
                            size_t ntris) {
    orc_carve* c = (orc_carve*)calloc(1, sizeof(orc_carve));
    c->pts = pts;
    c->tris = tris;
    c->ntris = ntris;
    const int r = orc_carve_resolution(ntris);
    c->r = r;
    const size_t ncell = (size_t)r * r;
    for (int k = 0; k < 3; ++k) {
        /* tetrahedralize_plc.cpp:131-158 */
        carve_dir* g = &c->dir[k];
        carve_basis(kRayDirs[k], g->u, g->v);
        double lou = 1e300, lov = 1e300, hiu = -1e300, hiv = -1e300;
        for (size_t i = 0; i < npts; ++i) {
            const double* p = pts + 3 * i;
            double pu = carve_proj(g->u, p), pv = carve_proj(g->v, p);
            /* std::min(a, b) = b < a ? b : a; std::max(a, b) = a < b ? b : a */
            lou = pu < lou ? pu : lou;
            hiu = hiu < pu ? pu : hiu;
            lov = pv < lov ? pv : lov;
            hiv = hiv < pv ? pv : hiv;
        }
        double eu = hiu - lou, ev = hiv - lov;
        g->min_u = lou;
        g->min_v = lov;
        g->inv_u = r / (eu > 0 ? eu * 1.0000001 : 1.0);
        g->inv_v = r / (ev > 0 ? ev * 1.0000001 : 1.0);
        g->start = (int*)calloc(ncell + 1, sizeof(int));
        int* rect = (int*)malloc(sizeof(int) * 4 * (ntris ? ntris : 1));
        for (size_t t = 0; t < ntris; ++t) {
            int u0 = r, u1 = 0, v0 = r, v1 = 0;
            for (int j = 0; j < 3; ++j) {
                const double* p = pts + 3 * (size_t)tris[3 * t + j];
                int iu = carve_iu(g, p, r), iv = carve_iv(g, p, r);
                u0 = iu < u0 ? iu : u0;
                u1 = iu > u1 ? iu : u1;
                v0 = iv < v0 ? iv : v0;
                v1 = iv > v1 ? iv : v1;
            }
            rect[4 * t] = u0, rect[4 * t + 1] = u1, rect[4 * t + 2] = v0, rect[4 * t + 3] = v1;
            for (int iu = u0; iu <= u1; ++iu)
                for (int iv = v0; iv <= v1; ++iv) g->start[(size_t)iu * r + iv + 1]++;
        }
        for (size_t q = 0; q < ncell; ++q) g->start[q + 1] += g->start[q];
        int* fill = (int*)malloc(sizeof(int) * ncell);
        memcpy(fill, g->start, sizeof(int) * ncell);
        g->items = (int*)malloc(sizeof(int) * ((size_t)g->start[ncell] + 1));
        for (size_t t = 0; t < ntris; ++t)
            for (int iu = rect[4 * t]; iu <= rect[4 * t + 1]; ++iu)
                for (int iv = rect[4 * t + 2]; iv <= rect[4 * t + 3]; ++iv)
                    g->items[fill[(size_t)iu * r + iv]++] = (int)t;
        free(fill);
        free(rect);
    }
    return c;
}

void orc_carve_destroy(orc_carve* c) {
    if (!c) return;
    for (int k = 0; k < 3; ++k) {
        free(c->dir[k].start);
        free(c->dir[k].items);
    }
    free(c);
}

int orc_carve_cell_entries(const orc_carve* c, int k) {
    return c->dir[k].start[(size_t)c->r * c->r];
}

/* tetrahedralize_plc.cpp:79-93 */
int orc_carve_inside(const orc_carve* c, const double* q) {
    int votes = 0;
    for (int k = 0; k < 3; ++k) {
        const carve_dir* g = &c->dir[k];
        const size_t cell = (size_t)carve_iu(g, q, c->r) * c->r + carve_iv(g, q, c->r);
        int crossings = 0;
        for (int e = g->start[cell]; e < g->start[cell + 1]; ++e) {
            const int t = g->items[e];
            const double* a = c->pts + 3 * (size_t)c->tris[3 * t];
            const double* b = c->pts + 3 * (size_t)c->tris[3 * t + 1];
            const double* cc = c->pts + 3 * (size_t)c->tris[3 * t + 2];
            if (orc_ray_hits(q, kRayDirs[k], a, b, cc)) ++crossings;
        }
        if (crossings & 1) ++votes;
    }
    return votes >= 2;
}

/* The same vote with every triangle tested (no projection grid). The reference's
 * comment (tetrahedralize_plc.cpp:67-72) states that the grid changes nothing. */
int orc_carve_inside_brute(const double* pts, const int* tris, size_t ntris,
                           const double* q) {
    int votes = 0;
    for (int k = 0; k < 3; ++k) {
        int crossings = 0;
        for (size_t t = 0; t < ntris; ++t)
            crossings += orc_ray_hits(q, kRayDirs[k], pts + 3 * (size_t)tris[3 * t],
                                      pts + 3 * (size_t)tris[3 * t + 1],
                                      pts + 3 * (size_t)tris[3 * t + 2]);
        if (crossings & 1) ++votes;
    }
    return votes >= 2;
}

/* tetrahedralize_plc.cpp:170-174 */
void orc_tet_centroid(const double* pts, const int* tet, double* c) {
    const double* a = pts + 3 * (size_t)tet[0];
    const double* b = pts + 3 * (size_t)tet[1];
    const double* cc = pts + 3 * (size_t)tet[2];
    const double* d = pts + 3 * (size_t)tet[3];
    for (int i = 0; i < 3; ++i) c[i] = (a[i] + b[i] + cc[i] + d[i]) / 4;
}

void orc_carve_points(const orc_carve* c, const double* q, size_t nq, uint8_t* out,
                      int nthreads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(NT(nthreads))
    for (size_t i = 0; i < nq; ++i) out[i] = (uint8_t)orc_carve_inside(c, q + 3 * i);
}

/* tetrahedralize_plc.cpp:258-268: keep[i] = inside(centroid of tet i) */
void orc_carve_tets(const orc_carve* c, const int* tets, size_t ntets, uint8_t* keep,
                    int nthreads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(NT(nthreads))
    for (size_t i = 0; i < ntets; ++i) {
        double q[3];
        orc_tet_centroid(c->pts, tets + 4 * i, q);
        keep[i] = (uint8_t)orc_carve_inside(c, q);
    }
}

/* ---- parity classifier --------------------------------------------------------- */

static inline int parity_cell(double v, double corner, double spacing, int n) {
    /* voxelize.cpp:82-89 */
    double f = floor((v - corner) / spacing);
    if (!(f >= 0.0)) return 0;
    if (f >= (double)n) return n - 1;
    return (int)f;
}

static inline int sgn3(const double* P, const double* a, const double* b, const double* c) {
    return orc_orient3d_sign(P, a, b, c, ORC_COMPAT);
}

/* voxelize.cpp:107-114 */
static double parity_plane_z(const double* a, const double* b, const double* c, double x,
                             double y) {
    double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    double vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    double nx = uy * vz - uz * vy;
    double ny = uz * vx - ux * vz;
    double nz = ux * vy - uy * vx;
    return a[2] - (nx * (x - a[0]) + ny * (y - a[1])) / nz;
}

/* voxelize.cpp:121-143 with the heights folded into the count above z: returns 0 when
 * a sign is exactly zero; *above then holds the count of the crossings found before
 * the grazing triangle (the partial list the reference keeps). */
static int parity_column(double x, double y, double zmin, double zmax, double z,
                         const double* pts, const int* tris, const int* bucket, int nb,
                         int* above) {
    *above = 0;
    const double P[3] = {x, y, zmin};
    const double Q[3] = {x, y, zmax};
    for (int k = 0; k < nb; ++k) {
        int t = bucket[k];
        const double* a = pts + 3 * (size_t)tris[3 * t];
        const double* b = pts + 3 * (size_t)tris[3 * t + 1];
        const double* c = pts + 3 * (size_t)tris[3 * t + 2];
        int sa = sgn3(P, a, b, c), sb = sgn3(Q, a, b, c);
        if (sa == 0 || sb == 0) return 0;
        if (sa == sb) continue;
        int s1 = sgn3(P, Q, a, b), s2 = sgn3(P, Q, b, c), s3 = sgn3(P, Q, c, a);
        if (s1 == 0 || s2 == 0 || s3 == 0) return 0;
        if (s1 != s2 || s2 != s3) continue;
        *above += parity_plane_z(a, b, c, x, y) > z;
    }
    return 1;
}

static const double kNudgeP[6][2] = {
    {1.0e-6, 3.0e-7},   {-7.0e-7, 5.0e-7}, {4.0e-7, -9.0e-7},
    {-2.0e-6, -1.0e-6}, {8.0e-7, 6.0e-7},  {-5.0e-7, -4.0e-6},
};

void orc_carve_parity_points(const double* pts, const int* tris, size_t ntris,
                             const orc_grid* g, const double* q, size_t nq, uint8_t* out,
                             int nthreads) {
    const int nx = g->nx, ny = g->ny;
    const size_t ncol = (size_t)nx * ny;
    /* xy buckets (voxelize.cpp:78-103), ascending triangle order */
    int* count = (int*)calloc(ncol + 1, sizeof(int));
    int* rng = (int*)malloc(sizeof(int) * 4 * (ntris ? ntris : 1));
    for (size_t t = 0; t < ntris; ++t) {
        int i0 = nx, i1 = 0, j0 = ny, j1 = 0;
        for (int v = 0; v < 3; ++v) {
            const double* p = pts + 3 * (size_t)tris[3 * t + v];
            int i = parity_cell(p[0], g->corner[0], g->spacing, nx);
            int j = parity_cell(p[1], g->corner[1], g->spacing, ny);
            i0 = i < i0 ? i : i0;
            i1 = i > i1 ? i : i1;
            j0 = j < j0 ? j : j0;
            j1 = j > j1 ? j : j1;
        }
        rng[4 * t] = i0, rng[4 * t + 1] = i1, rng[4 * t + 2] = j0, rng[4 * t + 3] = j1;
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) count[(size_t)j * nx + i + 1]++;
    }
    for (size_t c = 0; c < ncol; ++c) count[c + 1] += count[c];
    int* fill = (int*)malloc(sizeof(int) * (ncol ? ncol : 1));
    memcpy(fill, count, sizeof(int) * ncol);
    int* items = (int*)malloc(sizeof(int) * ((size_t)count[ncol] + 1));
    for (size_t t = 0; t < ntris; ++t)
        for (int j = rng[4 * t + 2]; j <= rng[4 * t + 3]; ++j)
            for (int i = rng[4 * t]; i <= rng[4 * t + 1]; ++i)
                items[fill[(size_t)j * nx + i]++] = (int)t;
    free(fill);
    free(rng);

    const double zmin = g->corner[2] - g->spacing;
    const double zmax = g->corner[2] + (g->nz + 1) * g->spacing;
#pragma omp parallel for schedule(dynamic, 256) num_threads(NT(nthreads))
    for (size_t n = 0; n < nq; ++n) {
        const double x = q[3 * n], y = q[3 * n + 1], z = q[3 * n + 2];
        const int i = parity_cell(x, g->corner[0], g->spacing, nx);
        const int j = parity_cell(y, g->corner[1], g->spacing, ny);
        const size_t col = (size_t)j * nx + i;
        const int* bucket = items + count[col];
        const int nb = count[col + 1] - count[col];
        int above = 0;
        /* voxelize.cpp:154-165 */
        if (!parity_column(x, y, zmin, zmax, z, pts, tris, bucket, nb, &above)) {
            int ok = 0;
            for (int d = 0; d < 6 && !ok; ++d)
                ok = parity_column(x + kNudgeP[d][0] * g->spacing,
                                   y + kNudgeP[d][1] * g->spacing, zmin, zmax, z, pts,
                                   tris, bucket, nb, &above);
        }
        out[n] = (uint8_t)(above & 1);
    }
    free(items);
    free(count);
}
