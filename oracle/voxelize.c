/* voxelize.c — CPU oracle for occupancy and signed-distance voxelization.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Compiled with -ffp-contract=off so every
 * FP64 expression rounds exactly as the reference's x86-64 build does.
 *
 * Restates src/voxelize/voxelize.cpp operation for operation where the result depends
 * on rounding (grid set-up, bucket membership, the crossing tests, plane heights, cell
 * centres, point-triangle distance). The SDF has three equivalent searches: the
 * reference's own shell search (for timing and for pinning), brute force, and an exact
 * BVH search (fast enough for 256^3 grids); all three return the FP64 minimum of the
 * same per-pair value, so they agree bit for bit.
 */
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NT(n) ((n) > 0 ? (n) : omp_get_max_threads())

size_t orc_fan_triangulate(const int* off, size_t npoly, const int* idx, int* tris) {
    /* voxelize.cpp:29-38: (v0, v[i-1], v[i]) for i >= 2 */
    size_t m = 0;
    for (size_t p = 0; p < npoly; ++p) {
        const int* v = idx + off[p];
        int n = off[p + 1] - off[p];
        for (int i = 2; i < n; ++i) {
            tris[3 * m] = v[0];
            tris[3 * m + 1] = v[i - 1];
            tris[3 * m + 2] = v[i];
            ++m;
        }
    }
    return m;
}

int orc_make_grid(const double* pts, size_t npts, int resolution, int pad, orc_grid* g) {
    /* voxelize.cpp:52-74 */
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = pts[a];
    for (size_t i = 0; i < npts; ++i)
        for (int a = 0; a < 3; ++a) {
            double v = pts[3 * i + a];
            if (v < lo[a]) lo[a] = v;
            if (v > hi[a]) hi[a] = v;
        }
    double ext = 0;
    for (int a = 0; a < 3; ++a) {
        double e = hi[a] - lo[a];
        if (e > ext) ext = e; /* std::max(ext, e) keeps ext unless e > ext */
    }
    memset(g, 0, sizeof(*g));
    g->spacing = ext / resolution;
    if (!(g->spacing > 0)) return -1;
    int* dims[3] = {&g->nx, &g->ny, &g->nz};
    for (int a = 0; a < 3; ++a) {
        double e = hi[a] - lo[a];
        int core = (int)ceil(e / g->spacing);
        if (core < 1) core = 1;
        double covered = core * g->spacing;
        g->corner[a] = lo[a] - (covered - e) / 2.0 - pad * g->spacing;
        *dims[a] = core + 2 * pad;
    }
    return 0;
}

static inline int clamp_cell(double v, double corner, double spacing, int n) {
    int i = (int)floor((v - corner) / spacing);
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

static inline int sgn(double v) { return v > 0 ? 1 : (v < 0 ? -1 : 0); }

/* voxelize.cpp:107-114 */
static double plane_z(const double* a, const double* b, const double* c, double x,
                      double y) {
    double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    double vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    double nx = uy * vz - uz * vy;
    double ny = uz * vx - ux * vz;
    double nz = ux * vy - uy * vx;
    return a[2] - (nx * (x - a[0]) + ny * (y - a[1])) / nz;
}

typedef struct {
    double* h;
    int n, cap;
} heights_t;

static void push_height(heights_t* hs, double z) {
    if (hs->n == hs->cap) {
        hs->cap = hs->cap ? 2 * hs->cap : 16;
        hs->h = (double*)realloc(hs->h, sizeof(double) * (size_t)hs->cap);
    }
    hs->h[hs->n++] = z;
}

/* voxelize.cpp:121-143: 0 if some orient3d is exactly zero (the line grazes) */
static int column_crossings(double x, double y, double zmin, double zmax,
                            const double* pts, const int* tris, const int* bucket, int nb,
                            heights_t* hs) {
    hs->n = 0;
    const double P[3] = {x, y, zmin};
    const double Q[3] = {x, y, zmax};
    for (int k = 0; k < nb; ++k) {
        int t = bucket[k];
        const double* a = pts + 3 * (size_t)tris[3 * t];
        const double* b = pts + 3 * (size_t)tris[3 * t + 1];
        const double* c = pts + 3 * (size_t)tris[3 * t + 2];
        int sa = orc_orient3d_sign(P, a, b, c, ORC_COMPAT);
        int sb = orc_orient3d_sign(Q, a, b, c, ORC_COMPAT);
        if (sa == 0 || sb == 0) return 0;
        if (sa == sb) continue;
        int s1 = orc_orient3d_sign(P, Q, a, b, ORC_COMPAT);
        int s2 = orc_orient3d_sign(P, Q, b, c, ORC_COMPAT);
        int s3 = orc_orient3d_sign(P, Q, c, a, ORC_COMPAT);
        if (s1 == 0 || s2 == 0 || s3 == 0) return 0;
        if (s1 != s2 || s2 != s3) continue;
        push_height(hs, plane_z(a, b, c, x, y));
    }
    return 1;
}

/* voxelize.cpp:148-151 */
static const double kNudge[6][2] = {
    {1.0e-6, 3.0e-7},   {-7.0e-7, 5.0e-7}, {4.0e-7, -9.0e-7},
    {-2.0e-6, -1.0e-6}, {8.0e-7, 6.0e-7},  {-5.0e-7, -4.0e-6},
};

void orc_occupancy(const double* pts, const int* tris, size_t ntris, const orc_grid* g,
                   uint8_t* occ, int64_t* stats, int nthreads) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t ncol = (size_t)nx * ny;
    /* xy buckets (voxelize.cpp:78-103) as CSR, triangles in ascending order */
    int* count = (int*)calloc(ncol + 1, sizeof(int));
    int* rng = (int*)malloc(sizeof(int) * 4 * (ntris ? ntris : 1));
    for (size_t t = 0; t < ntris; ++t) {
        int i0 = nx, i1 = 0, j0 = ny, j1 = 0;
        for (int v = 0; v < 3; ++v) {
            const double* p = pts + 3 * (size_t)tris[3 * t + v];
            int i = clamp_cell(p[0], g->corner[0], g->spacing, nx);
            int j = clamp_cell(p[1], g->corner[1], g->spacing, ny);
            if (i < i0) i0 = i;
            if (i > i1) i1 = i;
            if (j < j0) j0 = j;
            if (j > j1) j1 = j;
        }
        rng[4 * t] = i0, rng[4 * t + 1] = i1, rng[4 * t + 2] = j0, rng[4 * t + 3] = j1;
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) count[(size_t)j * nx + i + 1]++;
    }
    for (size_t c = 0; c < ncol; ++c) count[c + 1] += count[c];
    int* fill = (int*)malloc(sizeof(int) * (ncol ? ncol : 1));
    for (size_t c = 0; c < ncol; ++c) fill[c] = count[c];
    int* items = (int*)malloc(sizeof(int) * ((size_t)count[ncol] + 1));
    for (size_t t = 0; t < ntris; ++t)
        for (int j = rng[4 * t + 2]; j <= rng[4 * t + 3]; ++j)
            for (int i = rng[4 * t]; i <= rng[4 * t + 1]; ++i)
                items[fill[(size_t)j * nx + i]++] = (int)t;
    free(fill);
    free(rng);

    /* voxelize.cpp:169-195 */
    memset(occ, 0, (size_t)nx * ny * nz);
    const double zmin = g->corner[2] - g->spacing;
    const double zmax = g->corner[2] + (nz + 1) * g->spacing;
    int64_t nudged = 0, failed = 0;
#pragma omp parallel num_threads(NT(nthreads)) reduction(+ : nudged, failed)
    {
        heights_t hs = {0, 0, 0};
#pragma omp for schedule(dynamic, 16)
        for (size_t col = 0; col < ncol; ++col) {
            int j = (int)(col / nx), i = (int)(col % nx);
            double cx = g->corner[0] + (i + 0.5) * g->spacing;
            double cy = g->corner[1] + (j + 0.5) * g->spacing;
            const int* bucket = items + count[col];
            int nb = count[col + 1] - count[col];
            /* voxelize.cpp:154-165 */
            if (!column_crossings(cx, cy, zmin, zmax, pts, tris, bucket, nb, &hs)) {
                ++nudged;
                int ok = 0;
                for (int d = 0; d < 6 && !ok; ++d)
                    ok = column_crossings(cx + kNudge[d][0] * g->spacing,
                                          cy + kNudge[d][1] * g->spacing, zmin, zmax, pts,
                                          tris, bucket, nb, &hs);
                if (!ok) ++failed; /* keeps the last attempt's partial list */
            }
            for (int k = 0; k < nz; ++k) {
                double zc = g->corner[2] + (k + 0.5) * g->spacing;
                int above = 0;
                for (int h = 0; h < hs.n; ++h) above += hs.h[h] > zc;
                if (above & 1) occ[((size_t)k * ny + j) * nx + i] = 1;
            }
        }
        free(hs.h);
    }
    if (stats) {
        stats[0] = nudged;
        stats[1] = failed;
    }
    free(items);
    free(count);
}

static inline double dot3(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static inline double dist2_to(const double* p, const double* q) {
    double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
    return dx * dx + dy * dy + dz * dz;
}

double orc_point_tri_dist2(const double* p, const double* a, const double* b,
                           const double* c, double* bary) {
    /* voxelize.cpp:202-247 (Ericson, Real-Time Collision Detection 5.1.5) */
    double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0 && d2 <= 0) {
        if (bary) bary[0] = 1, bary[1] = 0, bary[2] = 0;
        return dist2_to(p, a);
    }
    double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0 && d4 <= d3) {
        if (bary) bary[0] = 0, bary[1] = 1, bary[2] = 0;
        return dist2_to(p, b);
    }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        double v = d1 / (d1 - d3);
        double q[3] = {a[0] + v * ab[0], a[1] + v * ab[1], a[2] + v * ab[2]};
        if (bary) bary[0] = 1 - v, bary[1] = v, bary[2] = 0;
        return dist2_to(p, q);
    }
    double cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0 && d5 <= d6) {
        if (bary) bary[0] = 0, bary[1] = 0, bary[2] = 1;
        return dist2_to(p, c);
    }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        double w = d2 / (d2 - d6);
        double q[3] = {a[0] + w * ac[0], a[1] + w * ac[1], a[2] + w * ac[2]};
        if (bary) bary[0] = 1 - w, bary[1] = 0, bary[2] = w;
        return dist2_to(p, q);
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        double q[3] = {b[0] + w * (c[0] - b[0]), b[1] + w * (c[1] - b[1]),
                       b[2] + w * (c[2] - b[2])};
        if (bary) bary[0] = 0, bary[1] = 1 - w, bary[2] = w;
        return dist2_to(p, q);
    }
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom, w = vc * denom;
    double q[3] = {a[0] + ab[0] * v + ac[0] * w, a[1] + ab[1] * v + ac[1] * w,
                   a[2] + ab[2] * v + ac[2] * w};
    if (bary) bary[0] = 1 - v - w, bary[1] = v, bary[2] = w;
    return dist2_to(p, q);
}

/* std::min(best, x) keeps best unless x < best (a NaN x never replaces it) */
static inline double keep_min(double best, double x) { return x < best ? x : best; }

static inline void cell_center(const orc_grid* g, int i, int j, int k, double* p) {
    p[0] = g->corner[0] + (i + 0.5) * g->spacing;
    p[1] = g->corner[1] + (j + 0.5) * g->spacing;
    p[2] = g->corner[2] + (k + 0.5) * g->spacing;
}

static inline float signed_dist(double best, uint8_t inside) {
    double d = sqrt(best);
    return (float)(inside ? -d : d); /* voxelize.cpp:354-355 */
}

/* ---- the reference's shell search (voxelize.cpp:251-326) ------------------------- */

typedef struct {
    const double* pts;
    const int* tris;
    const orc_grid* g;
    const int* cnt;
    const int* items;
    int* seen;
    int stamp;
    double best;
    int any;
    double p[3];
} shell_ctx;

/* the lambda body of nearest_tri_dist2 (voxelize.cpp:311-321) */
static void shell_visit(shell_ctx* c, int i, int j, int k) {
    const orc_grid* g = c->g;
    if (i < 0 || i >= g->nx || j < 0 || j >= g->ny || k < 0 || k >= g->nz) return;
    c->any = 1;
    size_t cell = ((size_t)k * g->ny + j) * g->nx + i;
    for (int q = c->cnt[cell]; q < c->cnt[cell + 1]; ++q) {
        int t = c->items[q];
        if (c->seen[t] == c->stamp) continue;
        c->seen[t] = c->stamp;
        c->best = keep_min(c->best, orc_point_tri_dist2(c->p, c->pts + 3 * (size_t)c->tris[3 * t],
                                                         c->pts + 3 * (size_t)c->tris[3 * t + 1],
                                                         c->pts + 3 * (size_t)c->tris[3 * t + 2],
                                                         NULL));
    }
}

void orc_sdf_shell(const double* pts, const int* tris, size_t ntris, const orc_grid* g,
                   const uint8_t* occ, float* dist, size_t first, size_t count,
                   size_t stride) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t ncell = (size_t)nx * ny * nz;
    if (stride == 0) stride = 1;
    if (count == 0) count = (ncell - first + stride - 1) / stride;
    int* cnt = (int*)calloc(ncell + 1, sizeof(int));
    int* rng = (int*)malloc(sizeof(int) * 6 * (ntris ? ntris : 1));
    for (size_t t = 0; t < ntris; ++t) {
        int i0 = nx, i1 = 0, j0 = ny, j1 = 0, k0 = nz, k1 = 0;
        for (int v = 0; v < 3; ++v) {
            const double* p = pts + 3 * (size_t)tris[3 * t + v];
            int i = clamp_cell(p[0], g->corner[0], g->spacing, nx);
            int j = clamp_cell(p[1], g->corner[1], g->spacing, ny);
            int k = clamp_cell(p[2], g->corner[2], g->spacing, nz);
            if (i < i0) i0 = i;
            if (i > i1) i1 = i;
            if (j < j0) j0 = j;
            if (j > j1) j1 = j;
            if (k < k0) k0 = k;
            if (k > k1) k1 = k;
        }
        int* r = rng + 6 * t;
        r[0] = i0, r[1] = i1, r[2] = j0, r[3] = j1, r[4] = k0, r[5] = k1;
        for (int k = k0; k <= k1; ++k)
            for (int j = j0; j <= j1; ++j)
                for (int i = i0; i <= i1; ++i) cnt[((size_t)k * ny + j) * nx + i + 1]++;
    }
    for (size_t c = 0; c < ncell; ++c) cnt[c + 1] += cnt[c];
    int* fill = (int*)malloc(sizeof(int) * ncell);
    memcpy(fill, cnt, sizeof(int) * ncell);
    int* items = (int*)malloc(sizeof(int) * ((size_t)cnt[ncell] + 1));
    for (size_t t = 0; t < ntris; ++t) {
        int* r = rng + 6 * t;
        for (int k = r[4]; k <= r[5]; ++k)
            for (int j = r[2]; j <= r[3]; ++j)
                for (int i = r[0]; i <= r[1]; ++i)
                    items[fill[((size_t)k * ny + j) * nx + i]++] = (int)t;
    }
    free(fill);
    free(rng);
    shell_ctx sc = {pts, tris, g, cnt, items, NULL, 0, 0.0, 0, {0, 0, 0}};
    sc.seen = (int*)malloc(sizeof(int) * (ntris ? ntris : 1));
    for (size_t t = 0; t < ntris; ++t) sc.seen[t] = -1;
    const int max_r = nx + ny + nz;
    for (size_t c = 0; c < count; ++c) {
        const size_t idx = first + c * stride;
        if (idx >= ncell) break;
        int k = (int)(idx / ((size_t)nx * ny));
        int j = (int)((idx / nx) % ny);
        int i = (int)(idx % nx);
        cell_center(g, i, j, k, sc.p);
        sc.best = DBL_MAX;
        /* nearest_tri_dist2 (voxelize.cpp:299-326) */
        for (int r = 0; r <= max_r; ++r) {
            double ring_min = (r - 1) * g->spacing;
            if (r > 0 && sc.best != DBL_MAX && ring_min * ring_min >= sc.best) break;
            sc.any = 0;
            /* for_shell (voxelize.cpp:281-293) */
            if (r == 0) {
                shell_visit(&sc, i, j, k);
            } else {
                for (int s = -r; s <= r; s += 2 * r)
                    for (int dj = -r; dj <= r; ++dj)
                        for (int di = -r; di <= r; ++di) shell_visit(&sc, i + di, j + dj, k + s);
                for (int s = -r; s <= r; s += 2 * r)
                    for (int dk = -r + 1; dk <= r - 1; ++dk)
                        for (int di = -r; di <= r; ++di) shell_visit(&sc, i + di, j + s, k + dk);
                for (int s = -r; s <= r; s += 2 * r)
                    for (int dk = -r + 1; dk <= r - 1; ++dk)
                        for (int dj = -r + 1; dj <= r - 1; ++dj)
                            shell_visit(&sc, i + s, j + dj, k + dk);
            }
            if (!sc.any && sc.best != DBL_MAX) break;
        }
        ++sc.stamp;
        dist[idx] = signed_dist(sc.best, occ[idx]);
    }
    free(sc.seen);
    free(items);
    free(cnt);
}

void orc_sdf_brute(const double* pts, const int* tris, size_t ntris, const orc_grid* g,
                   const uint8_t* occ, float* dist, int nthreads) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t ncell = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(dynamic, 64) num_threads(NT(nthreads))
    for (size_t idx = 0; idx < ncell; ++idx) {
        int k = (int)(idx / ((size_t)nx * ny));
        int j = (int)((idx / nx) % ny);
        int i = (int)(idx % nx);
        double p[3];
        cell_center(g, i, j, k, p);
        double best = DBL_MAX;
        for (size_t t = 0; t < ntris; ++t)
            best = keep_min(best, orc_point_tri_dist2(p, pts + 3 * (size_t)tris[3 * t],
                                                      pts + 3 * (size_t)tris[3 * t + 1],
                                                      pts + 3 * (size_t)tris[3 * t + 2],
                                                      NULL));
        dist[idx] = signed_dist(best, occ[idx]);
    }
}

/* ---- exact BVH nearest-triangle search ----------------------------------------- */

typedef struct {
    double lo[3], hi[3];
    int left, right; /* children, or -1 for a leaf */
    int first, count;
} bvh_node;

typedef struct {
    bvh_node* nodes;
    int nnodes;
    int* order; /* triangle ids, leaves index ranges of it */
} bvh_t;

typedef struct {
    const double* pts;
    const int* tris;
    double* cen;
} bvh_ctx;

static void tri_box(const double* pts, const int* tris, int t, double* lo, double* hi) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = DBL_MAX;
        hi[a] = -DBL_MAX;
    }
    for (int v = 0; v < 3; ++v) {
        const double* p = pts + 3 * (size_t)tris[3 * t + v];
        for (int a = 0; a < 3; ++a) {
            if (p[a] < lo[a]) lo[a] = p[a];
            if (p[a] > hi[a]) hi[a] = p[a];
        }
    }
}

static const double* g_sort_cen;
static int g_sort_axis;
static int cmp_cen(const void* x, const void* y) {
    double a = g_sort_cen[3 * (size_t)(*(const int*)x) + g_sort_axis];
    double b = g_sort_cen[3 * (size_t)(*(const int*)y) + g_sort_axis];
    return a < b ? -1 : (a > b ? 1 : 0);
}

static int bvh_build_rec(bvh_t* b, bvh_ctx* c, int first, int count) {
    int id = b->nnodes++;
    bvh_node* n = &b->nodes[id];
    for (int a = 0; a < 3; ++a) {
        n->lo[a] = DBL_MAX;
        n->hi[a] = -DBL_MAX;
    }
    double clo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, chi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int q = first; q < first + count; ++q) {
        int t = b->order[q];
        double lo[3], hi[3];
        tri_box(c->pts, c->tris, t, lo, hi);
        for (int a = 0; a < 3; ++a) {
            if (lo[a] < n->lo[a]) n->lo[a] = lo[a];
            if (hi[a] > n->hi[a]) n->hi[a] = hi[a];
            double ce = c->cen[3 * (size_t)t + a];
            if (ce < clo[a]) clo[a] = ce;
            if (ce > chi[a]) chi[a] = ce;
        }
    }
    if (count <= 4) {
        n->left = n->right = -1;
        n->first = first;
        n->count = count;
        return id;
    }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
        if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
    g_sort_cen = c->cen;
    g_sort_axis = axis;
    qsort(b->order + first, (size_t)count, sizeof(int), cmp_cen);
    int half = count / 2;
    int l = bvh_build_rec(b, c, first, half);
    int r = bvh_build_rec(b, c, first + half, count - half);
    n = &b->nodes[id];
    n->left = l;
    n->right = r;
    n->first = first;
    n->count = count;
    return id;
}

static bvh_t bvh_build(const double* pts, const int* tris, size_t ntris) {
    bvh_t b;
    b.nodes = (bvh_node*)malloc(sizeof(bvh_node) * (2 * ntris + 1));
    b.nnodes = 0;
    b.order = (int*)malloc(sizeof(int) * (ntris ? ntris : 1));
    double* cen = (double*)malloc(sizeof(double) * 3 * (ntris ? ntris : 1));
    for (size_t t = 0; t < ntris; ++t) {
        b.order[t] = (int)t;
        for (int a = 0; a < 3; ++a)
            cen[3 * t + a] = (pts[3 * (size_t)tris[3 * t] + a] +
                              pts[3 * (size_t)tris[3 * t + 1] + a] +
                              pts[3 * (size_t)tris[3 * t + 2] + a]) /
                             3.0;
    }
    bvh_ctx c = {pts, tris, cen};
    if (ntris) bvh_build_rec(&b, &c, 0, (int)ntris);
    free(cen);
    return b;
}

static void bvh_free(bvh_t* b) {
    free(b->nodes);
    free(b->order);
}

static inline double box_d2(const bvh_node* n, const double* p) {
    double s = 0;
    for (int a = 0; a < 3; ++a) {
        double d = 0;
        if (p[a] < n->lo[a]) d = n->lo[a] - p[a];
        else if (p[a] > n->hi[a]) d = p[a] - n->hi[a];
        s += d * d;
    }
    return s;
}

/* Minimum FP64 squared distance over all triangles (excluding those with excl[t] set,
 * if excl is given). A subtree is skipped only when its box lies farther than the best
 * value found by a relative margin of 1e-6, far above the FP64 rounding of either
 * quantity, so the triangle whose computed value is smallest is always evaluated. */
static double bvh_nearest(const bvh_t* b, const double* pts, const int* tris,
                          const double* p, const uint8_t* excl, int* arg) {
    double best = DBL_MAX;
    int besti = -1;
    int stack[128];
    int sp = 0;
    if (b->nnodes == 0) return best;
    stack[sp++] = 0;
    while (sp) {
        const bvh_node* n = &b->nodes[stack[--sp]];
        if (best != DBL_MAX && box_d2(n, p) > best * (1.0 + 1e-6)) continue;
        if (n->left < 0) {
            for (int q = n->first; q < n->first + n->count; ++q) {
                int t = b->order[q];
                if (excl && excl[t]) continue;
                double d = orc_point_tri_dist2(p, pts + 3 * (size_t)tris[3 * t],
                                               pts + 3 * (size_t)tris[3 * t + 1],
                                               pts + 3 * (size_t)tris[3 * t + 2], NULL);
                if (d < best) {
                    best = d;
                    besti = t;
                }
            }
            continue;
        }
        const bvh_node* L = &b->nodes[n->left];
        const bvh_node* R = &b->nodes[n->right];
        double dl = box_d2(L, p), dr = box_d2(R, p);
        if (dl <= dr) {
            stack[sp++] = n->right;
            stack[sp++] = n->left;
        } else {
            stack[sp++] = n->left;
            stack[sp++] = n->right;
        }
    }
    if (arg) *arg = besti;
    return best;
}

void orc_sdf_bvh(const double* pts, const int* tris, size_t ntris, const orc_grid* g,
                 const uint8_t* occ, float* dist, double* best, int32_t* arg,
                 size_t first, size_t count, int nthreads) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t ncell = (size_t)nx * ny * nz;
    if (count == 0) count = ncell - first;
    bvh_t b = bvh_build(pts, tris, ntris);
#pragma omp parallel for schedule(dynamic, 256) num_threads(NT(nthreads))
    for (size_t idx = first; idx < first + count; ++idx) {
        int k = (int)(idx / ((size_t)nx * ny));
        int j = (int)((idx / nx) % ny);
        int i = (int)(idx % nx);
        double p[3];
        cell_center(g, i, j, k, p);
        int t = -1;
        double d2 = bvh_nearest(&b, pts, tris, p, NULL, &t);
        if (dist) dist[idx] = signed_dist(d2, occ[idx]);
        if (best) best[idx] = d2;
        if (arg) arg[idx] = t;
    }
    bvh_free(&b);
}

void orc_sdf_grad(const double* pts, size_t npts, const int* tris, size_t ntris,
                  const orc_grid* g, const uint8_t* occ, const float* gout,
                  const int32_t* arg, double* grad, int nthreads) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t ncell = (size_t)nx * ny * nz;
    /* cells within 2^-40 of the grid extent of the surface are on it up to rounding */
    const int mdim = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
    const double dzero = ldexp(mdim * g->spacing, -40);
    bvh_t b = bvh_build(pts, tris, ntris);
    memset(grad, 0, sizeof(double) * 3 * npts);
    int nt = NT(nthreads);
    double* part = (double*)calloc((size_t)nt * 3 * npts, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        double* mine = part + (size_t)omp_get_thread_num() * 3 * npts;
#pragma omp for schedule(dynamic, 256)
        for (size_t idx = 0; idx < ncell; ++idx) {
            if (gout[idx] == 0.0f) continue;
            int k = (int)(idx / ((size_t)nx * ny));
            int j = (int)((idx / nx) % ny);
            int i = (int)(idx % nx);
            double p[3];
            cell_center(g, i, j, k, p);
            int t = -1;
            if (arg) t = arg[idx];
            else bvh_nearest(&b, pts, tris, p, NULL, &t);
            if (t < 0) continue;
            const int* tv = tris + 3 * (size_t)t;
            double w[3];
            double d2 = orc_point_tri_dist2(p, pts + 3 * (size_t)tv[0],
                                            pts + 3 * (size_t)tv[1], pts + 3 * (size_t)tv[2], w);
            if (!(sqrt(d2) > dzero)) continue;
            double q[3];
            for (int a = 0; a < 3; ++a)
                q[a] = w[0] * pts[3 * (size_t)tv[0] + a] + w[1] * pts[3 * (size_t)tv[1] + a] +
                       w[2] * pts[3 * (size_t)tv[2] + a];
            double d = sqrt(d2);
            double s = occ[idx] ? -1.0 : 1.0;
            double coef = (double)gout[idx] * s / d;
            /* d = |p - q|, dd/dq = -(p - q)/d, dq/dv_i = w_i (envelope theorem) */
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 3; ++a)
                    mine[3 * (size_t)tv[v] + a] += -coef * w[v] * (p[a] - q[a]);
        }
    }
    for (int r = 0; r < nt; ++r)
        for (size_t q = 0; q < 3 * npts; ++q) grad[q] += part[(size_t)r * 3 * npts + q];
    free(part);
    bvh_free(&b);
}

double orc_sdf_fd(const double* pts, size_t npts, const int* tris, size_t ntris,
                  const orc_grid* g, const uint8_t* occ, const float* gout,
                  const double* base_best, const int32_t* base_arg, int vert, int axis,
                  double h, double* tie_bound, int nthreads) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t ncell = (size_t)nx * ny * nz;
    bvh_t b = bvh_build(pts, tris, ntris);
    uint8_t* inc = (uint8_t*)calloc(ntris ? ntris : 1, 1);
    int* inc_list = (int*)malloc(sizeof(int) * (ntris ? ntris : 1));
    int ninc = 0;
    for (size_t t = 0; t < ntris; ++t)
        for (int v = 0; v < 3; ++v)
            if (tris[3 * t + v] == vert) {
                inc[t] = 1;
                inc_list[ninc++] = (int)t;
                break;
            }
    /* union box of the incident triangles, widened by h on every side */
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int q = 0; q < ninc; ++q) {
        double tl[3], th[3];
        tri_box(pts, tris, inc_list[q], tl, th);
        for (int a = 0; a < 3; ++a) {
            if (tl[a] - h < lo[a]) lo[a] = tl[a] - h;
            if (th[a] + h > hi[a]) hi[a] = th[a] + h;
        }
    }
    bvh_node box;
    memcpy(box.lo, lo, sizeof lo);
    memcpy(box.hi, hi, sizeof hi);
    double total = 0, ties = 0;
#pragma omp parallel for schedule(dynamic, 256) num_threads(NT(nthreads)) reduction(+ : total, ties)
    for (size_t idx = 0; idx < ncell; ++idx) {
        if (gout[idx] == 0.0f) continue;
        int k = (int)(idx / ((size_t)nx * ny));
        int j = (int)((idx / nx) % ny);
        int i = (int)(idx % nx);
        double p[3];
        cell_center(g, i, j, k, p);
        int t0 = -1;
        double base;
        if (base_best) {
            base = base_best[idx];
            t0 = base_arg[idx];
        } else {
            base = bvh_nearest(&b, pts, tris, p, NULL, &t0);
        }
        if (ninc == 0) continue;
        if (box_d2(&box, p) > base * (1.0 + 1e-6) && !(t0 >= 0 && inc[t0])) continue;
        double others = bvh_nearest(&b, pts, tris, p, inc, NULL);
        const double ro = sqrt(others);
        /* incident distances at +h, -h and 0, and each incident triangle's derivative
         * -w_vert (p - q)_axis / d at 0 (a non-incident triangle's is 0) */
        double dist[3][64], der[64];
        int nc = 0;
        for (int q = 0; q < ninc && q < 64; ++q) nc = q + 1;
        double dpm[2];
        for (int s = 0; s < 3; ++s) {
            double moved[3];
            for (int a = 0; a < 3; ++a) moved[a] = pts[3 * (size_t)vert + a];
            moved[axis] += s == 0 ? h : (s == 1 ? -h : 0.0);
            double m = others;
            for (int q = 0; q < nc; ++q) {
                const int* tv = tris + 3 * (size_t)inc_list[q];
                const double* v3[3];
                for (int v = 0; v < 3; ++v)
                    v3[v] = tv[v] == vert ? moved : pts + 3 * (size_t)tv[v];
                double w[3];
                const double d2 = orc_point_tri_dist2(p, v3[0], v3[1], v3[2], w);
                m = keep_min(m, d2);
                dist[s][q] = sqrt(d2);
                if (s == 2) {
                    double wv = 0.0, qa = 0.0;
                    for (int v = 0; v < 3; ++v) {
                        if (tv[v] == vert) wv = w[v];
                        qa += w[v] * v3[v][axis];
                    }
                    der[q] = dist[2][q] > 0.0 ? -wv * (p[axis] - qa) / dist[2][q] : 0.0;
                }
            }
            if (s < 2) dpm[s] = sqrt(m);
        }
        double sgnv = occ[idx] ? -1.0 : 1.0;
        total += (double)gout[idx] * sgnv * (dpm[0] - dpm[1]);
        /* Tie allowance: the candidates that are nearest at -h, 0 or +h. When more than
         * one is (the nearest triangle changes within the step), the difference quotient
         * mixes their derivatives while the adjoint takes the derivative of one of them:
         * the two differ by at most |g| times the spread of those derivatives (equal
         * derivatives, e.g. of triangles sharing the closest point, add nothing). */
        double lo_d = DBL_MAX, hi_d = -DBL_MAX;
        for (int s = 0; s < 3; ++s) {
            double mn = ro;
            for (int q = 0; q < nc; ++q) mn = dist[s][q] < mn ? dist[s][q] : mn;
            const double cut = mn * (1.0 + 1e-12);
            if (ro <= cut) {
                lo_d = lo_d < 0.0 ? lo_d : 0.0;
                hi_d = hi_d > 0.0 ? hi_d : 0.0;
            }
            for (int q = 0; q < nc; ++q)
                if (dist[s][q] <= cut) {
                    if (der[q] < lo_d) lo_d = der[q];
                    if (der[q] > hi_d) hi_d = der[q];
                }
        }
        if (hi_d > lo_d) ties += fabs((double)gout[idx]) * (hi_d - lo_d);
    }
    free(inc);
    free(inc_list);
    bvh_free(&b);
    (void)npts;
    if (tie_bound) *tie_bound = ties;
    return total / (2.0 * h);
}

void orc_sdf_dist2_to(const double* pts, const int* tris, const orc_grid* g,
                      const int32_t* arg, double* out, int nthreads) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t ncell = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static) num_threads(NT(nthreads))
    for (size_t idx = 0; idx < ncell; ++idx) {
        const int t = arg[idx];
        if (t < 0) {
            out[idx] = -1.0;
            continue;
        }
        int k = (int)(idx / ((size_t)nx * ny));
        int j = (int)((idx / nx) % ny);
        int i = (int)(idx % nx);
        double p[3];
        cell_center(g, i, j, k, p);
        const int* tv = tris + 3 * (size_t)t;
        out[idx] = orc_point_tri_dist2(p, pts + 3 * (size_t)tv[0], pts + 3 * (size_t)tv[1],
                                       pts + 3 * (size_t)tv[2], NULL);
    }
}
