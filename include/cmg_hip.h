/* cmg_hip.h — C ABI of libcmg_hip: CyberMeshGenerator's predicate and voxelization
 * path on AMD Instinct MI355X (gfx950).
 *
 * Plain C types only (pointers, sizes, status codes). Groups of entry points:
 *
 *  1. The stable C ABI subset the reference exports for voxelization, byte-compatible
 *     with bindings/c/include/cmg_c/cmg_c.h (declarations cited per function). A caller
 *     linked against libcmg_c.so for these symbols can link libcmg_hip.so instead.
 *  2. Additive batched-predicate and SDF-adjoint entry points (host pointers), proposed
 *     as C-ABI additions for this path; they have no counterpart in cmg_c.h.
 *  3. The device layer that replaces src/backend/cuda (include/cmg/backend/
 *     cuda_kernels.hpp:11-21): device-pointer, stream-ordered entry points.
 *  4. Batched voxelization of many meshes (device pointers).
 *  5. The PLC carve's point-in-domain classification (device and host pointers).
 *  6. Surface files (STL, OBJ) to PLCs and to device arrays.
 *  7. The mesh quality report (device pointers).
 *  8. Batched point location and Delaunay cavities in a tetrahedral mesh (device
 *     pointers).
 *
 * Every entry point runs on the GPU. There is no CPU fallback: with no usable HIP
 * device a call returns CMG_ERR_INTERNAL and fills errbuf. Inputs must be finite.
 * Status codes, errbuf conventions and handle ownership follow cmg_c.h:18-33 and
 * bindings/c/src/cmg_c.cpp:34-52: errbuf (optional) receives a NUL-terminated message,
 * no C++ exception crosses the boundary, HIP failures map to CMG_ERR_INTERNAL.
 */
#ifndef CMG_HIP_H
#define CMG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. the reference's C ABI subset ---------------------------------------------- */

/* cmg_c.h:18-27 */
typedef enum cmg_status {
    CMG_OK = 0,
    CMG_ERR_INVALID_INPUT = 1,
    CMG_ERR_SELF_INTERSECTION = 2,
    CMG_ERR_UNSUPPORTED_OPTION = 3,
    CMG_ERR_QUALITY_NOT_MET = 4,
    CMG_ERR_NOT_IMPLEMENTED = 5,
    CMG_ERR_INTERNAL = 6,
    CMG_ERR_PARSE = 7
} cmg_status;

/* cmg_c.h:30-33 (opaque handles) */
typedef struct cmg_plc cmg_plc;
typedef struct cmg_voxels cmg_voxels;

/* cmg_c.h:36-37 */
cmg_plc* cmg_plc_create(void);
void cmg_plc_destroy(cmg_plc*);

/* cmg_c.h:44 — replaces the point set from a flat [x0,y0,z0, ...] array */
cmg_status cmg_plc_set_points(cmg_plc*, const double* xyz, size_t count);

/* cmg_c.h:51-52 — appends a one-polygon facet (count >= 3 indices) */
cmg_status cmg_plc_add_facet(cmg_plc*, const int* poly_vertex_indices, size_t count,
                             int marker);

/* cmg_c.h:137-138 — mode 1 = signed distance, anything else = occupancy; grid as in
 * src/voxelize/voxelize.cpp:52-74. On success *out is a handle to destroy with
 * cmg_voxels_destroy; on failure *out is NULL and errbuf is filled. */
cmg_status cmg_plc_voxelize(const cmg_plc*, int resolution, int mode, int pad,
                            cmg_voxels** out, char* errbuf, size_t errbuf_len);

/* cmg_c.h:141-153 */
void cmg_voxels_destroy(cmg_voxels*);
void cmg_voxels_dims(const cmg_voxels*, int* nx, int* ny, int* nz);
void cmg_voxels_origin(const cmg_voxels*, double* x, double* y, double* z);
double cmg_voxels_spacing(const cmg_voxels*);
const unsigned char* cmg_voxels_occupancy(const cmg_voxels*);
const float* cmg_voxels_distance(const cmg_voxels*);

/* cmg_c.h:156 */
const char* cmg_version(void);

/* ---- 2. additive host-pointer entry points ---------------------------------------- */

/* Predicate modes. CMG_PRED_EXACT: exact signs (Shewchuk orient3dexact/insphereexact,
 * src/predicates/predicates.cpp:1609,3237). CMG_PRED_COMPAT: the signs
 * cmg::robust::orient3d/insphere return as shipped (naive FP64 determinant when nonzero,
 * exact sign when it is 0; src/predicates/robust.cpp:28-50, predicates.cpp:2230,4113). */
enum { CMG_PRED_EXACT = 0, CMG_PRED_COMPAT = 1 };

/* The batched predicate entry points below take n < 2^32 tuples per call (their
 * compaction queue holds 32-bit tuple indices); larger n returns CMG_ERR_INVALID_INPUT. */

/* Batched orient3d signs over index quadruples into AoS points: the batch of
 * cmg::robust::orient3d_signs (src/predicates/robust.cpp:68-105), int8 output. */
cmg_status cmg_orient3d_signs(const double* xyz, size_t npts, const int* quads, size_t n,
                              int mode, signed char* out, char* errbuf, size_t errbuf_len);

/* Batched orient3d values: cmg::robust::orient3d_batch (robust.cpp:52-66). The naive
 * FP64 determinant (bit-identical) when nonzero; when it is exactly 0 the exact
 * determinant's leading expansion component (exact sign). */
cmg_status cmg_orient3d_values(const double* xyz, size_t npts, const int* quads, size_t n,
                               double* out, char* errbuf, size_t errbuf_len);

/* Batched insphere values: cmg::robust::insphere (robust.cpp:44-50) over index
 * quintuples (a,b,c,d,e). The naive FP64 lifted determinant in the reference's operation
 * order (predicates.cpp:4063-4106), bit-identical, when nonzero; when it is exactly 0 the
 * exact determinant's leading expansion component (exact sign). */
cmg_status cmg_insphere_values(const double* xyz, size_t npts, const int* quints, size_t n,
                               double* out, char* errbuf, size_t errbuf_len);

/* Batched insphere signs over index quintuples (a,b,c,d,e) into AoS points; the batched
 * form of cmg::robust::insphere (robust.cpp:44-50). */
cmg_status cmg_insphere_signs(const double* xyz, size_t npts, const int* quints, size_t n,
                              int mode, signed char* out, char* errbuf, size_t errbuf_len);

/* SDF adjoint for a grid produced by cmg_plc_voxelize(mode=1) from `plc`: given
 * grad_sdf (dL/dSDF per cell, flat idx = (k*ny + j)*nx + i), writes dL/dxyz
 * (3 floats per PLC point, anchors included) to grad_xyz. The sign is held constant;
 * cells on the surface (distance at most 2^-40 of the grid extent) carry no gradient. */
cmg_status cmg_voxels_sdf_backward(const cmg_plc* plc, const cmg_voxels* grid,
                                   const float* grad_sdf, float* grad_xyz, char* errbuf,
                                   size_t errbuf_len);

/* For the same grid: the index (into the PLC's fan-triangulated facet list) of the
 * triangle whose closest point the adjoint uses for each cell, i.e. the FP32 nearest
 * triangle (-1 where there is none). No reference counterpart (the reference's
 * fill_distance, voxelize.cpp:331-358, keeps only the distance); useful for attribute
 * transfer and for checking the adjoint. */
cmg_status cmg_voxels_sdf_nearest(const cmg_plc* plc, const cmg_voxels* grid,
                                  int32_t* tri_out, char* errbuf, size_t errbuf_len);

/* Name of the backend that served the calling thread's last predicate or voxelize
 * call ("hip", or "none" before any call). Mirrors cmg::backend::last_backend
 * (src/backend/dispatch.cpp:65). */
const char* cmg_backend_last(void);

/* ---- 3. device layer (device pointers, stream-ordered) ----------------------------- */

/* cuda_kernels.hpp:14 — 1 when a usable HIP device is present. */
int cmg_hip_available(void);

/* cuda_kernels.hpp:16-21, same arguments: host AoS points xyz[3*npts], host quadruples
 * tets[4*n], host out_sign[n]. The signs are final compat signs (the reference's CPU
 * escalation of zeros, robust.cpp:97-98, becomes a no-op). Blocking. Returns 0 on
 * success, nonzero on a HIP failure. */
int cmg_hip_orient3d_signs(const double* xyz, int npts, const int* tets, int n,
                           int* out_sign);

/* Per-stage lane counts of the last batched-sign call on a workspace. */
typedef struct cmg_pred_stats {
    uint64_t lanes;        /* tuples in the batch */
    uint64_t fp32;         /* decided by the FP32 filter (exact mode only) */
    uint64_t fp64;         /* decided by the FP64 stage */
    uint64_t exact;        /* left undecided by the filters, resolved exactly */
    uint64_t wide;         /* of those, re-run with wide expansion capacity */
    uint64_t failed;       /* exceeded even the wide capacity (sign left 0) */
    uint64_t exact_inline; /* of `exact`, decided by register-only exact stages: inside
                              the streaming kernel (orient3d: exact difference sums or a
                              double-word bound) or by fixed-point integer evaluation in
                              the exact pass (insphere); the rest went through expansion
                              arithmetic */
} cmg_pred_stats;

/* Workspace bytes a batched-sign call over n tuples needs (npts 4 = orient3d,
 * 5 = insphere). */
size_t cmg_hip_pred_workspace_size(size_t n, int npts);

/* Batched signs over device-resident structure-of-arrays tuples: plane q = slot*3 + axis
 * (orient3d: 12 planes, insphere: 15) holds that coordinate of every tuple at
 * soa[q*ld + i]; out[i] receives -1/0/+1. `workspace` is a device buffer of at least
 * cmg_hip_pred_workspace_size bytes, or NULL to use the library's own (then the call
 * synchronises `stream` before returning). `stream` is a hipStream_t (NULL = default).
 * With a caller workspace the call is stream-ordered and returns at once. */
cmg_status cmg_hip_orient3d_signs_soa(const double* soa, size_t n, size_t ld, int mode,
                                      signed char* out, void* workspace,
                                      size_t workspace_bytes, void* stream, char* errbuf,
                                      size_t errbuf_len);
cmg_status cmg_hip_insphere_signs_soa(const double* soa, size_t n, size_t ld, int mode,
                                      signed char* out, void* workspace,
                                      size_t workspace_bytes, void* stream, char* errbuf,
                                      size_t errbuf_len);

/* No counterpart in the reference: cmg_hip_orient3d_signs_soa for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The same kernels and results; the workspace is a fixed 16 MiB
 * device array inside the library, one per GPU, held until the process ends, so the call
 * performs no allocation of any kind. n <= 2^20 tuples (larger n: CMG_ERR_INVALID_INPUT).
 * Stream-ordered, returns at once. Every call on one GPU shares that workspace: calls
 * that overlap on one GPU (on different streams, or from different threads) race. */
cmg_status cmg_hip_orient3d_signs_soa_noalloc(const double* soa, size_t n, size_t ld,
                                              int mode, signed char* out, void* stream,
                                              char* errbuf, size_t errbuf_len);

/* Reads the per-stage counts of the last call that used `workspace` (synchronises
 * `stream`). */
cmg_status cmg_hip_pred_stats(void* workspace, size_t n, int npts, void* stream,
                              cmg_pred_stats* out, char* errbuf, size_t errbuf_len);

/* ---- 4. batched voxelization of many meshes (no counterpart in the reference: the
 * proposed cmg_voxelize_batch addition) ------------------------------------------------ */

typedef struct cmg_batch cmg_batch;

enum { CMG_BATCH_FORWARD = 1, CMG_BATCH_BACKWARD = 2, CMG_BATCH_GRAPH = 4 };

/* Plans a batch of nmesh meshes stored back to back in device memory: mesh m owns
 * points [point_off[m], point_off[m+1]) of xyz (3 doubles each) and triangles
 * [tri_off[m], tri_off[m+1]) of tris (3 ints each, indices local to the mesh's points).
 * Grid m is make_grid(points of m, resolution[m], pad[m]) exactly as cmg_plc_voxelize
 * builds it (voxelize.cpp:52-74). point_off/tri_off/resolution/pad are host arrays.
 * Planning synchronises `stream` (it orders each mesh's triangles for the nearest-
 * triangle search and sizes the column-tile pool); xyz and tris must stay valid and at
 * the same address for the batch's lifetime (the point values may change between steps:
 * every step rebuilds the search bounds from them). */
cmg_status cmg_voxelize_batch_create(int nmesh, const int64_t* point_off,
                                     const int64_t* tri_off, const int* resolution,
                                     const int* pad, const double* xyz, const int* tris,
                                     void* stream, cmg_batch** out, char* errbuf,
                                     size_t errbuf_len);
void cmg_voxelize_batch_destroy(cmg_batch*);

/* Cells over all grids (the length of the concatenated SDF), and mesh m's grid: dims,
 * origin (centre of cell (0,0,0)), spacing and the offset of its first cell. */
int64_t cmg_voxelize_batch_cells(const cmg_batch*);
cmg_status cmg_voxelize_batch_grid(const cmg_batch*, int mesh, int* dims, double* origin,
                                   double* spacing, int64_t* cell_offset);

/* One step on `stream`, no host synchronisation. CMG_BATCH_FORWARD writes every mesh's
 * signed distance to sdf (device, cells floats, each grid flat idx = (k*ny + j)*nx + i at
 * its cell offset). CMG_BATCH_BACKWARD (after a forward on the same points) writes
 * dL/dxyz to grad_xyz (device, 3 floats per point) for the upstream grad_sdf (device,
 * cells floats), signs held constant. CMG_BATCH_GRAPH captures the step into a hipGraph
 * on first use for this buffer set and replays it afterwards. */
cmg_status cmg_voxelize_batch_run(cmg_batch*, int flags, const double* xyz, float* sdf,
                                  const float* grad_sdf, float* grad_xyz, void* stream,
                                  char* errbuf, size_t errbuf_len);

/* Synchronises `stream`; overflow != 0 means the column-tile pool was too small for the
 * last step (results incomplete: create the batch again), pairs counts the FP32
 * point-triangle evaluations (per wavefront of 64 lanes) of the last passes. */
cmg_status cmg_voxelize_batch_status(cmg_batch*, void* stream, unsigned* overflow,
                                     uint64_t* pairs, char* errbuf, size_t errbuf_len);

/* Autotuner: times every voxel-pass variant on this batch — voxels per lane (1: 4x4x4
 * voxels per wavefront, 2: 4x4x8, 4: 4x4x16) x LDS tile depth of a search leaf's
 * triangle records (0: read with scalar loads, 8: staged through LDS as one tile) — with
 * HIP events on `stream` (forward + backward voxel passes, buffers as in
 * cmg_voxelize_batch_run) and keeps the fastest. ms (optional, 6 entries) receives the
 * timings in the order (1,0) (1,8) (2,0) (2,8) (4,0) (4,8). */
cmg_status cmg_voxelize_batch_autotune(cmg_batch*, float* sdf, const float* grad_sdf,
                                       float* grad_xyz, void* stream, int* vpl, int* depth,
                                       float* ms, char* errbuf, size_t errbuf_len);

/* Selects a voxel-pass variant explicitly (vpl 1, 2 or 4, depth 0 or 8). */
cmg_status cmg_voxelize_batch_set_variant(cmg_batch*, int vpl, int depth);

/* Device pointer to the occupancy bytes of every grid (valid after a forward). */
const unsigned char* cmg_voxelize_batch_occupancy(const cmg_batch*);

/* After a forward: for every cell (device, cells ints, same layout as sdf) the
 * mesh-local index of the triangle whose closest point the adjoint uses, -1 where there
 * is none. Stream-ordered. */
cmg_status cmg_voxelize_batch_nearest(cmg_batch*, int32_t* tri, void* stream,
                                      char* errbuf, size_t errbuf_len);

/* Synchronises `stream` and reads the work counters of the last step: per wavefront of
 * the voxel passes, out[0] bounding-box tests of search-tree nodes, out[1] bounding-box
 * tests of triangles, out[2] FP32 point-triangle distance evaluations; out[3] overflow
 * flags; out[4] column-tile list entries of the occupancy pass. */
cmg_status cmg_voxelize_batch_counters(cmg_batch*, void* stream, uint64_t* out,
                                       char* errbuf, size_t errbuf_len);

/* Profiling hook (no counterpart in the reference): enable or disable HIP-event timing
 * of the launches of one kernel, by name ("orient3d_filter", "insphere_filter",
 * "batch_occupancy", "batch_sdf", "batch_sdf_backward", "carve_query", "quality",
 * "locate").
 * Events are recorded on the stream the kernel is launched on; launches recorded into a
 * graph are not timed. */
void cmg_hip_kernel_timer(const char* name, int enable);

/* Total milliseconds and launch count of the timed launches of `name` since the last
 * read (synchronises on the recorded events, then clears them). */
cmg_status cmg_hip_kernel_time(const char* name, double* total_ms, uint64_t* count,
                               char* errbuf, size_t errbuf_len);

/* ---- 5. PLC carve: point-in-domain classification ---------------------------------- */

/* The classification step of the PLC tetrahedralization (no counterpart in cmg_c.h): the
 * loop of src/constrained/tetrahedralize_plc.cpp:258-268 that keeps the Delaunay
 * tetrahedra whose centroid (:170-174) lies inside the boundary, as batched GPU calls.
 *   CMG_CARVE_RAYS   the reference's classifier (CarveGrid, :60-168): majority of three
 *                    Moller-Trumbore ray parities over its projection grids; results are
 *                    bit-identical to it.
 *   CMG_CARVE_PARITY the voxelizer's exact column parity (src/voxelize/voxelize.cpp:
 *                    78-195) at the query point: compat orient3d crossing tests, kNudge
 *                    retries, crossings strictly above the point, on the grid
 *                    make_grid(points, resolution, 1). */
enum { CMG_CARVE_RAYS = 0, CMG_CARVE_PARITY = 1 };

typedef struct cmg_carve cmg_carve;

/* Builds the classifier over a closed boundary surface given as device arrays: xyz
 * (npts points, 3 doubles each) and tris (ntris triangles, 3 ints each), which must stay
 * valid and unchanged for the handle's lifetime. resolution: parity grid cells per axis
 * (<= 0: the rays grid's, ceil(sqrt(ntris)) capped at 256; ignored by CMG_CARVE_RAYS).
 * Synchronises `stream`. ntris == 0 fails with the reference's message "PLC has facets but
 * no boundary triangles" (tetrahedralize_plc.cpp:250-253); a triangle vertex index outside
 * [0, npts) fails with CMG_ERR_INVALID_INPUT (checked on the device before any use). */
cmg_status cmg_carve_create(const double* xyz, size_t npts, const int* tris, size_t ntris,
                            int mode, int resolution, void* stream, cmg_carve** out,
                            char* errbuf, size_t errbuf_len);
void cmg_carve_destroy(cmg_carve*);

/* inside[i] = 1 iff point q[3i..3i+2] is inside (device arrays; stream-ordered, no host
 * synchronisation). */
cmg_status cmg_carve_points(cmg_carve*, const double* q, size_t nq, uint8_t* inside,
                            void* stream, char* errbuf, size_t errbuf_len);

/* keep[i] = 1 iff the centroid of tets[4i..4i+3] (indices into xyz; device array, 16-byte
 * aligned) is inside. A tet with an index outside [0, npts) gets 0 and is counted:
 * cmg_carve_status reports the count (stream-ordered, no host synchronisation). */
cmg_status cmg_carve_tets(cmg_carve*, const int* tets, size_t ntets, uint8_t* keep,
                          void* stream, char* errbuf, size_t errbuf_len);

/* Synchronises `stream`; *bad_tets = tets rejected for out-of-range indices since the
 * last call. */
cmg_status cmg_carve_status(cmg_carve*, void* stream, uint64_t* bad_tets, char* errbuf,
                            size_t errbuf_len);

/* Host-memory form of tetrahedralize_plc.cpp:256-268 over a PLC: the boundary triangles
 * are the fan triangulation of its facets (:26-37), the points its points; keep[i] for
 * each of the ntets tets (4 ints each, indices into the PLC's points, checked).
 * Synchronous. */
cmg_status cmg_plc_carve(const cmg_plc*, const int* tets, size_t ntets, int mode,
                         uint8_t* keep, char* errbuf, size_t errbuf_len);

/* ---- 6. surface files -------------------------------------------------------------- */

/* cmg_c.h:104-106: read a PLC boundary from a surface file, format from the extension.
 * This library reads .stl (ASCII or binary, coincident vertices merged in first-
 * occurrence order on the GPU; src/io/stl.cpp) and .obj (src/io/obj.cpp) with the
 * reference's grammar and messages; .poly/.smesh/.off/.ply give CMG_ERR_NOT_IMPLEMENTED,
 * other extensions CMG_ERR_INVALID_INPUT ("unsupported file format for <path>",
 * src/io/dispatch.cpp:33-36). On CMG_OK *out is a plc to free with cmg_plc_destroy. */
cmg_status cmg_read_plc(const char* path, cmg_plc** out, char* errbuf, size_t errbuf_len);

/* cmg_c.h:115-121: a PLC's geometry. Points as 3 * num_points doubles; facets
 * fan-triangulated to 3 * num_triangles ints. Pointers are valid until the PLC is
 * modified or destroyed. */
size_t cmg_plc_num_points(const cmg_plc*);
const double* cmg_plc_points(cmg_plc*);
size_t cmg_plc_num_triangles(cmg_plc*);
const int* cmg_plc_triangles(cmg_plc*);

/* Device layer of the STL reader (no counterpart in the reference: it replaces the host
 * path of stl.cpp:94-113 and 24-61 for callers that keep meshes on the GPU). A whole
 * binary STL file in device memory (bytes, 4-byte aligned, nbytes = 84 + 50 * count) ->
 * merged points xyz (room for 9 * count doubles) and triangles tris (3 * count ints),
 * the points numbered in first-occurrence order as the reference numbers them.
 * Synchronises `stream`; *npoints = number of points. */
cmg_status cmg_hip_stl_decode(const void* bytes, size_t nbytes, double* xyz, int* tris,
                              int64_t* npoints, void* stream, char* errbuf,
                              size_t errbuf_len);

/* The merge alone (stl.cpp:24-61) on a device vertex soup of nverts vertices (3 doubles
 * each, nverts <= 2^30): xyz receives the merged points (room for 3 * nverts doubles),
 * idx the merged index of every vertex. Synchronises `stream`. */
cmg_status cmg_hip_weld(const double* soup, size_t nverts, double* xyz, int* idx,
                        int64_t* npoints, void* stream, char* errbuf, size_t errbuf_len);

/* ---- 7. mesh quality report -------------------------------------------------------- */

/* cmg::quality::QualityReport (include/cmg/quality/metrics.hpp:14-40); the worst tets are
 * returned separately (num_worst of them). */
typedef struct cmg_quality_report {
    uint64_t num_tets;
    double min_radius_edge, max_radius_edge, mean_radius_edge;
    double min_dihedral, max_dihedral, mean_dihedral; /* degrees */
    double min_volume, max_volume, total_volume;
    int dihedral_histogram[18]; /* [0,10), [10,20), ... degrees; the last bucket is >= 170 */
    uint64_t num_worst;
} cmg_quality_report;

/* cmg::quality::report (src/quality/metrics.cpp:81-169), the accelerable QualityScan
 * phase (include/cmg/backend/dispatch.hpp:25), over device arrays: xyz (npts points, 3
 * doubles each) and tets (ntets tets, 4 ints each, 16-byte aligned). worst (host array,
 * optional; room for worst_count entries, or ntets when worst_count < 0) receives the
 * tets of largest radius-edge ratio, largest first, equal ratios by ascending index.
 * Min, max, volumes, ratios and the worst list equal the reference's; the sums (total
 * volume, means) are summed in a different order; dihedral angles use the device acos.
 * A tet index outside [0, npts) fails with CMG_ERR_INVALID_INPUT. Synchronises
 * `stream`. */
cmg_status cmg_hip_quality_report(const double* xyz, size_t npts, const int* tets,
                                  size_t ntets, int worst_count, cmg_quality_report* out,
                                  int* worst, void* stream, char* errbuf, size_t errbuf_len);

/* ---- 8. point location ------------------------------------------------------------- */

/* Batched form of the Delaunay insertion's point location (Triangulation::locate,
 * src/delaunay/incremental.cpp:321-358; the PointLocation phase the reference marks as
 * accelerable at :424-427): one visibility walk per query point q[3i..3i+2] through a
 * fixed mesh of ntets tets (device arrays): tets (4 ints each, positively oriented,
 * orient3d(v0,v1,v2,v3) > 0) and nbrs (4 ints each: the tet across the face opposite
 * v_k, -1 on the hull), both 16-byte aligned. The walk starts at start[i] (device array)
 * or, when start is NULL, at start_tet; it tests the faces in the reference's order with
 * the compat orient3d and steps through the first negative one that has a neighbour;
 * after 2 * ntets + 16 steps it falls back to the reference's brute_locate. out[i]
 * (device) = the tet found, identical to the reference's walk from the same start;
 * steps[i] (device, optional) = the walk length, -1 where brute force answered;
 * *brute (optional) = how many walks did. Synchronises `stream`; a walk that meets an
 * index out of range fails the call with CMG_ERR_INVALID_INPUT, and so does nq >= 2^32
 * (queries are ordered by 32-bit index). */
cmg_status cmg_hip_locate(const double* xyz, size_t npts, const int* tets, const int* nbrs,
                          size_t ntets, const double* q, size_t nq, const int* start,
                          int start_tet, int* out, int* steps, uint64_t* brute,
                          void* stream, char* errbuf, size_t errbuf_len);

/* Delaunay cavities of query points in a fixed mesh: the flood Triangulation::insert
 * performs before it retriangulates (src/delaunay/incremental.cpp:148-161; conflict at
 * :132-141), as one batched call. For query i located in tet located[i] (device, e.g. from
 * cmg_hip_locate), breadth first from that tet through neighbours k = 0..3, every tet not
 * yet in the cavity whose circumsphere contains q[i] (the compat insphere,
 * robust::insphere(v0, v1, v2, v3, q) > 0) joins it. cavity[i*cap ...] (device, nq * cap
 * ints) receives the cavity in that discovery order, count[i] (device) its size: -1 when
 * it has more than cap tets (the first cap are written) or meets an index out of range,
 * 0 when located[i] < 0. tets (positively oriented) and nbrs as for cmg_hip_locate, both
 * 16-byte aligned. *bad (optional) = queries that met an index out of range. Synchronises
 * `stream`. */
cmg_status cmg_hip_cavity(const double* xyz, size_t npts, const int* tets, const int* nbrs,
                          size_t ntets, const double* q, size_t nq, const int* located,
                          int cap, int* count, int* cavity, uint64_t* bad, void* stream,
                          char* errbuf, size_t errbuf_len);

#ifdef __cplusplus
} /* extern "C" */
#endif
#endif /* CMG_HIP_H */
