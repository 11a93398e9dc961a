/* oracle.h — CPU restatement of CyberMeshGenerator's predicate and voxelization path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the GPU path is compared
 * against; nothing in the shipped package links, loads or calls it. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Each function cites the reference file:line whose behaviour it restates
 * (paths are relative to the CyberMeshGenerator source tree).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off: no FMA contraction, so every
 * FP64 expression rounds exactly as the reference's x86-64 build does).
 */
#ifndef CMG_ORACLE_H
#define CMG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- predicates ---------------------------------------------------- */

/* Predicate modes: which oracle a sign is taken from.
 *  ORC_EXACT  — the exact sign of the determinant (Shewchuk's orient3dexact /
 *               insphereexact semantics, src/predicates/predicates.cpp:1609,3237).
 *  ORC_COMPAT — what cmg::robust::orient3d / insphere return as shipped: exactinit is
 *               called with zero bounding-box extents (src/predicates/robust.cpp:28-34),
 *               which zeroes the static filters (predicates.cpp:645-646), so the naive
 *               FP64 determinant's sign is returned whenever it is nonzero
 *               (predicates.cpp:2230-2231, 4113) and the exact sign only when it is 0. */
enum { ORC_EXACT = 0, ORC_COMPAT = 1 };

/* Naive FP64 determinants in the reference's operation order
 * (orient3d: predicates.cpp:2201-2222; insphere: predicates.cpp:4063-4106). */
double orc_orient3d_naive(const double* pa, const double* pb, const double* pc,
                          const double* pd);
double orc_insphere_naive(const double* pa, const double* pb, const double* pc,
                          const double* pd, const double* pe);

/* Exact signs (-1, 0, +1) by expansion arithmetic (Shewchuk 1997, CMU-CS-96-140). */
int orc_orient3d_exact(const double* pa, const double* pb, const double* pc,
                       const double* pd);
int orc_insphere_exact(const double* pa, const double* pb, const double* pc,
                       const double* pd, const double* pe);

/* Signs in the given mode. */
int orc_orient3d_sign(const double* pa, const double* pb, const double* pc,
                      const double* pd, int mode);
int orc_insphere_sign(const double* pa, const double* pb, const double* pc,
                      const double* pd, const double* pe, int mode);

/* Batched signs over structure-of-arrays tuples: plane q = slot*3 + axis holds
 * coordinate `axis` of point `slot` for all n tuples at soa[q*ld + i]
 * (orient3d: 12 planes, insphere: 15 planes). nthreads <= 0 uses OpenMP's default. */
void orc_orient3d_batch(const double* soa, size_t n, size_t ld, int mode, int8_t* out,
                        int nthreads);
void orc_insphere_batch(const double* soa, size_t n, size_t ld, int mode, int8_t* out,
                        int nthreads);

/* Indexed form: xyz is AoS [3*npts], quads [4*n] (the shape of
 * cmg::robust::orient3d_signs, src/predicates/robust.cpp:68-105). */
void orc_orient3d_indexed(const double* xyz, const int* quads, size_t n, int mode,
                          int8_t* out, int nthreads);

/* ---------------- voxelization -------------------------------------------------- */

typedef struct {
    double corner[3]; /* world coordinate of the MIN corner of cell (0,0,0) */
    double spacing;
    int nx, ny, nz;
} orc_grid;

/* Fan-triangulate polygons (src/voxelize/voxelize.cpp:29-38). Polygon p has vertices
 * idx[off[p] .. off[p+1]-1]. Writes 3 ints per triangle; returns the count. */
size_t orc_fan_triangulate(const int* off, size_t npoly, const int* idx, int* tris);

/* Bounding-box grid (voxelize.cpp:52-74). Returns 0, or -1 for zero extent. */
int orc_make_grid(const double* pts, size_t npts, int resolution, int pad, orc_grid* g);

/* Occupancy by exact vertical ray parity with the compat predicate, xy buckets, kNudge
 * retries and the strict "height > zc" count (voxelize.cpp:78-195). occ has
 * nx*ny*nz bytes, idx = (k*ny + j)*nx + i. stats (optional, 2 entries): columns that
 * needed at least one nudge, columns where every nudge grazed. */
void orc_occupancy(const double* pts, const int* tris, size_t ntris, const orc_grid* g,
                   uint8_t* occ, int64_t* stats, int nthreads);

/* Ericson squared point-triangle distance, FP64, reference operation order
 * (voxelize.cpp:197-247). bary (optional) receives the closest point's barycentric
 * weights on a, b, c. */
double orc_point_tri_dist2(const double* p, const double* a, const double* b,
                           const double* c, double* bary);

/* SDF by the reference's own 3-D bucket + Chebyshev-shell search
 * (voxelize.cpp:251-358). Single-threaded like the reference. Cells first,
 * first+stride, ... (count of them; count = 0 means to the end of the grid) are
 * computed in that order; other cells are left untouched. */
void orc_sdf_shell(const double* pts, const int* tris, size_t ntris, const orc_grid* g,
                   const uint8_t* occ, float* dist, size_t first, size_t count,
                   size_t stride);

/* SDF by testing every triangle (the brute-force check of tests/test_voxelize.cpp:
 * 194-275). Same output as orc_sdf_shell. */
void orc_sdf_brute(const double* pts, const int* tris, size_t ntris, const orc_grid* g,
                   const uint8_t* occ, float* dist, int nthreads);

/* SDF by an exact BVH nearest-triangle search: the FP64 minimum over all triangles,
 * identical to orc_sdf_brute, fast enough for 256^3 grids. Optional outputs: best
 * (FP64 squared distance) and arg (index of a minimising triangle). Any of dist,
 * best, arg may be NULL. Cells [first, first+count) only (count = 0: all). */
void orc_sdf_bvh(const double* pts, const int* tris, size_t ntris, const orc_grid* g,
                 const uint8_t* occ, float* dist, double* best, int32_t* arg,
                 size_t first, size_t count, int nthreads);

/* FP64 SDF-magnitude adjoint: for L = sum_v g[v] * s[v] * d[v] with the sign s
 * (+1 outside, -1 inside, from occ) held constant, accumulates dL/dxyz into
 * grad[3*npts] (zeroed first). The closest feature is the closest point of the
 * minimising triangle, or of triangle arg[cell] when arg is given (-1: no
 * contribution); a cell within 2^-40 of the grid extent of the surface (on it up to
 * rounding, where the unsigned distance has a kink) carries zero gradient. */
void orc_sdf_grad(const double* pts, size_t npts, const int* tris, size_t ntris,
                  const orc_grid* g, const uint8_t* occ, const float* gout,
                  const int32_t* arg, double* grad, int nthreads);

/* FP64 squared distance from every cell centre to triangle arg[cell] (-1 where
 * arg[cell] < 0). */
void orc_sdf_dist2_to(const double* pts, const int* tris, const orc_grid* g,
                      const int32_t* arg, double* out, int nthreads);

/* FP64 central finite difference of the same L with respect to coordinate `axis` of
 * vertex `vert`, step h, signs held at their unperturbed values. base_best/base_arg
 * (optional, from orc_sdf_bvh) are the unperturbed squared distances and minimising
 * triangles, so that many differences can share one search. tie_bound (optional)
 * receives, over the cells whose nearest triangle differs between the samples -h, 0
 * and +h, the sum of |g| times the spread of those triangles' derivatives: how far the
 * difference quotient and the adjoint may legitimately differ. */
double orc_sdf_fd(const double* pts, size_t npts, const int* tris, size_t ntris,
                  const orc_grid* g, const uint8_t* occ, const float* gout,
                  const double* base_best, const int32_t* base_arg, int vert, int axis,
                  double h, double* tie_bound, int nthreads);

/* ---------------- PLC carve: point-in-domain classification --------------------- */

/* static_cast<int>(double) as the reference's x86-64 build executes it (cvttsd2si):
 * truncation toward zero; INT_MIN for NaN and for values outside the int range. */
int orc_x86_trunc(double v);

/* Moller-Trumbore ray/triangle hit, forward only (tetrahedralize_plc.cpp:39-58). */
int orc_ray_hits(const double* o, const double* d, const double* a, const double* b,
                 const double* c);

/* CarveGrid (tetrahedralize_plc.cpp:60-168): the projection grids of the three ray
 * directions over pts (npts points) and tris (ntris triangles, 3 ints each); both
 * arrays must outlive the object. */
typedef struct orc_carve orc_carve;
int orc_carve_resolution(size_t ntris); /* cells per axis, :73-75 */
orc_carve* orc_carve_create(const double* pts, size_t npts, const int* tris, size_t ntris);
void orc_carve_destroy(orc_carve* c);
int orc_carve_cell_entries(const orc_carve* c, int k); /* list entries of direction k */
int orc_carve_inside(const orc_carve* c, const double* q); /* :79-93 */
int orc_carve_inside_brute(const double* pts, const int* tris, size_t ntris,
                           const double* q);
void orc_tet_centroid(const double* pts, const int* tet, double* c); /* :170-174 */
void orc_carve_points(const orc_carve* c, const double* q, size_t nq, uint8_t* out,
                      int nthreads);
/* keep[i] = inside(centroid of tets[4i..4i+3]) (tetrahedralize_plc.cpp:258-268) */
void orc_carve_tets(const orc_carve* c, const int* tets, size_t ntets, uint8_t* keep,
                    int nthreads);

/* Exact column parity (voxelize.cpp:78-195) at arbitrary points q (3 doubles each):
 * out[n] = 1 iff an odd number of the crossings of the vertical line through
 * (q.x, q.y), bucketed by the grid cell containing it and nudged as at
 * voxelize.cpp:154-165, lie strictly above q.z. */
void orc_carve_parity_points(const double* pts, const int* tris, size_t ntris,
                             const orc_grid* g, const double* q, size_t nq, uint8_t* out,
                             int nthreads);

/* ---------------- surface file readers ------------------------------------------ */

/* Merge of coincident vertices on the exact coordinate triple, first-occurrence order
 * (src/io/stl.cpp:24-61): soup holds nv vertices (3 doubles each, no NaN); pts receives
 * the merged points (up to nv), idx the merged index of every vertex. Returns the
 * number of points. */
int64_t orc_weld(const double* soup, int64_t nv, double* pts, int32_t* idx);

/* read_stl (stl.cpp:153-165): binary iff the size is 84 + 50 * count, else ASCII
 * "vertex" records; merged as orc_weld. Returns the point count and mallocs *pts
 * (3 per point) and *tris (3 per triangle), or -1 with err filled (the reference's
 * messages). *binary tells which form was read. Free with orc_free. */
int64_t orc_read_stl(const char* path, double** pts, int32_t** tris, int64_t* ntris,
                     int* binary, char* err, size_t len);
/* The STL file's vertex soup before merging (3 doubles per vertex). */
int64_t orc_stl_soup(const char* path, double** soup, int* binary, char* err, size_t len);

/* read_obj (src/io/obj.cpp:33-72): points, and one polygon per "f" statement with
 * vertex indices idx[off[f] .. off[f+1]) (0-based, unchecked as in the reference).
 * Returns 0, or -1 with err filled. Free with orc_free. */
int orc_read_obj(const char* path, double** pts, int64_t* npts, int32_t** off,
                 int32_t** idx, int64_t* nfacets, char* err, size_t len);
void orc_free(void* p);

/* ---------------- mesh quality report ------------------------------------------- */

/* src/quality/metrics.hpp:14-40 (QualityReport), with the worst list's length */
typedef struct {
    int64_t num_tets;
    double min_radius_edge, max_radius_edge, mean_radius_edge;
    double min_dihedral, max_dihedral, mean_dihedral;
    double min_volume, max_volume, total_volume;
    int32_t dihedral_histogram[18];
    int64_t num_worst;
} orc_quality;

/* metrics.cpp:36-48, 52-77 */
double orc_tet_volume(const double* a, const double* b, const double* c, const double* d);
void orc_tet_dihedrals(const double* P0, const double* P1, const double* P2,
                       const double* P3, double* out);
/* One tet's radius-edge ratio (DBL_MAX and *valid = 0 when degenerate, metrics.cpp:
 * 109-131), volume and six dihedral angles. */
void orc_tet_quality(const double* pts, const int* tet, double* ratio, int* valid,
                     double* vol, double* dih);
/* quality::report (metrics.cpp:81-169) over n tets (4 ints each); worst (optional)
 * receives the num_worst indices, descending ratio, ties by ascending index. */
void orc_quality_report(const double* pts, const int* tets, size_t n, int worst_count,
                        orc_quality* r, int* worst);
/* The six dihedral angles of each of n tets (6n doubles, metrics.cpp:52-77). */
void orc_dihedrals_batch(const double* pts, const int* tets, size_t n, double* out);

/* ---------------- point location ------------------------------------------------- */

/* Triangulation::locate (src/delaunay/incremental.cpp:321-358) in a fixed mesh: tets
 * (4 ints each, orient3d(v0,v1,v2,v3) > 0), nbrs (4 ints each: the tet across the face
 * opposite v_k, or -1), walking from tet `start` with the compat orient3d; after
 * 2 * ntets + 16 steps, brute_locate. *steps (optional): the walk length, -1 when brute
 * force answered. */
int orc_locate(const double* pts, const int* tets, const int* nbrs, size_t ntets,
               const double* q, int start, int* steps);
/* One walk per query (start[i], or start_tet when start is NULL). */
void orc_locate_batch(const double* pts, const int* tets, const int* nbrs, size_t ntets,
                      const double* q, size_t nq, const int* start, int start_tet,
                      int* out, int* steps, int nthreads);

/* Delaunay cavity of q from its located tet t0 (src/delaunay/incremental.cpp:148-161):
 * breadth first through neighbours k = 0..3, tets whose circumsphere contains q
 * (compat insphere > 0) join, in discovery order; returns the size (-1 past cap; 0 for
 * t0 < 0). */
int orc_cavity(const double* pts, const int* tets, const int* nbrs, const double* q, int t0,
               int cap, int* out);
void orc_cavity_batch(const double* pts, const int* tets, const int* nbrs, const double* q,
                      size_t nq, const int* located, int cap, int* count, int* out,
                      int nthreads);

#ifdef __cplusplus
}
#endif
#endif
