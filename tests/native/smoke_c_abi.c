/* C-ABI smoke test, pure C: the voxel entry points of bindings/c/include/cmg_c/cmg_c.h
 * on the 4x4-subdivided unit cube (tests/test_voxelize.cpp:140-163 known answer), plus the
 * device layer of include/cmg/backend/cuda_kernels.hpp:16-21. Exit 0 on success. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cmg_hip.h"

#define CHECK(c)                                                                      \
    do {                                                                              \
        if (!(c)) {                                                                   \
            fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);              \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

static int vid(int n, int x, int y, int z) { return (z * (n + 1) + y) * (n + 1) + x; }

int main(void) {
    const int n = 4;
    double xyz[3 * 125];
    for (int z = 0; z <= n; ++z)
        for (int y = 0; y <= n; ++y)
            for (int x = 0; x <= n; ++x) {
                int v = vid(n, x, y, z);
                xyz[3 * v] = (double)x / n;
                xyz[3 * v + 1] = (double)y / n;
                xyz[3 * v + 2] = (double)z / n;
            }
    cmg_plc* p = cmg_plc_create();
    CHECK(p != NULL);
    CHECK(cmg_plc_set_points(p, xyz, 125) == CMG_OK);
    /* the six faces as quads split into two triangles (outward winding) */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            int q[6][4] = {
                {vid(n, i, j, 0), vid(n, i, j + 1, 0), vid(n, i + 1, j + 1, 0), vid(n, i + 1, j, 0)},
                {vid(n, i, j, n), vid(n, i + 1, j, n), vid(n, i + 1, j + 1, n), vid(n, i, j + 1, n)},
                {vid(n, i, 0, j), vid(n, i + 1, 0, j), vid(n, i + 1, 0, j + 1), vid(n, i, 0, j + 1)},
                {vid(n, i, n, j), vid(n, i, n, j + 1), vid(n, i + 1, n, j + 1), vid(n, i + 1, n, j)},
                {vid(n, 0, i, j), vid(n, 0, i, j + 1), vid(n, 0, i + 1, j + 1), vid(n, 0, i + 1, j)},
                {vid(n, n, i, j), vid(n, n, i + 1, j), vid(n, n, i + 1, j + 1), vid(n, n, i, j + 1)}};
            for (int f = 0; f < 6; ++f) {
                int t1[3] = {q[f][0], q[f][1], q[f][2]}, t2[3] = {q[f][0], q[f][2], q[f][3]};
                CHECK(cmg_plc_add_facet(p, t1, 3, 0) == CMG_OK);
                CHECK(cmg_plc_add_facet(p, t2, 3, 0) == CMG_OK);
            }
        }
    int bad[2] = {0, 1};
    CHECK(cmg_plc_add_facet(p, bad, 2, 0) == CMG_ERR_INVALID_INPUT);
    char err[256];
    cmg_voxels* v = (cmg_voxels*)1;
    CHECK(cmg_plc_voxelize(p, 0, 0, 1, &v, err, sizeof err) == CMG_ERR_INVALID_INPUT);
    CHECK(v == NULL && strstr(err, "resolution") != NULL);
    CHECK(cmg_plc_voxelize(p, 16, 0, 1, &v, err, sizeof err) == CMG_OK);
    int nx, ny, nz;
    cmg_voxels_dims(v, &nx, &ny, &nz);
    CHECK(nx == 18 && ny == 18 && nz == 18);
    double ox, oy, oz;
    cmg_voxels_origin(v, &ox, &oy, &oz);
    CHECK(ox == -1.0 / 32 && cmg_voxels_spacing(v) == 1.0 / 16);
    const unsigned char* occ = cmg_voxels_occupancy(v);
    CHECK(occ != NULL && cmg_voxels_distance(v) == NULL);
    size_t inside = 0, stray = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int blk = i >= 1 && i <= 16 && j >= 1 && j <= 16 && k >= 1 && k <= 16;
                unsigned char o = occ[((size_t)k * ny + j) * nx + i];
                inside += o;
                stray += o && !blk;
            }
    CHECK(inside == 16 * 16 * 16 && stray == 0);
    cmg_voxels_destroy(v);
    CHECK(cmg_plc_voxelize(p, 16, 1, 2, &v, err, sizeof err) == CMG_OK);
    const float* d = cmg_voxels_distance(v);
    CHECK(d != NULL && cmg_voxels_occupancy(v) == NULL);
    cmg_voxels_dims(v, &nx, &ny, &nz);
    CHECK(d[((size_t)10 * ny + 10) * nx + 10] < 0.0f && d[0] > 0.0f);
    cmg_voxels_destroy(v);
    cmg_plc_destroy(p);
    /* device layer: cuda::orient3d_signs' argument list */
    double pts[15] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, -1};
    int tets[8] = {0, 1, 2, 3, 0, 1, 2, 4};
    int sign[2] = {9, 9};
    CHECK(cmg_hip_available() == 1);
    CHECK(cmg_hip_orient3d_signs(pts, 5, tets, 2, sign) == 0);
    CHECK(sign[0] == -1 && sign[1] == 1);
    CHECK(strcmp(cmg_backend_last(), "hip") == 0);
    printf("C ABI smoke: OK\n");
    return 0;
}
