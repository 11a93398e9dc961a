// weld_internal.hpp — surface files to device arrays (weld.hip, meshio.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

namespace cmg::hip {

// Merges coincident vertices of a device vertex soup (nv vertices, 3 doubles each) in
// first-occurrence order (src/io/stl.cpp:24-61): d_xyz receives the merged points (room
// for nv), d_idx the merged index of every vertex. Synchronises st; returns the number
// of points.
long long weld_vertices(const double* d_soup, long long nv, double* d_xyz, int* d_idx,
                        hipStream_t st);

// A whole binary STL file in device memory (4-byte aligned) -> merged points (room for
// 3 * count) and triangles (3 ints each; count = its header's). Synchronises st;
// returns the number of points.
long long stl_decode_weld(const void* d_bytes, size_t nbytes, double* d_xyz, int* d_tris,
                          hipStream_t st);

// A surface read from a file: points (3 doubles each) and one polygon per facet.
struct SurfaceFile {
    std::vector<double> xyz;
    std::vector<std::vector<int>> polys;
};

// read_stl (src/io/stl.cpp:153-165) and read_obj (src/io/obj.cpp:33-72) with the
// reference's grammar and messages; the STL merge runs on the GPU. Throw InvalidInput
// with the reference's message on malformed input.
SurfaceFile read_stl_file(const std::string& path);
SurfaceFile read_obj_file(const std::string& path);

}  // namespace cmg::hip
