// quality_internal.hpp — the mesh quality report (quality.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace cmg::hip {

// src/quality/metrics.hpp:14-40 (QualityReport without the worst list)
struct QualityReport {
    long long num_tets = 0;
    double min_radius_edge = 0, max_radius_edge = 0, mean_radius_edge = 0;
    double min_dihedral = 0, max_dihedral = 0, mean_dihedral = 0;
    double min_volume = 0, max_volume = 0, total_volume = 0;
    int dihedral_histogram[18] = {};
};

// quality::report (metrics.cpp:81-169) over device arrays: d_xyz (npts points, 3 doubles
// each), d_tets (n tets, 4 ints each, 16-byte aligned). worst (optional) receives the
// worst_count (< 0: all) tets of largest radius-edge ratio, largest first, equal ratios
// by ascending index. Synchronises st. Throws InvalidInput for out-of-range indices.
QualityReport quality_report(const double* d_xyz, long long npts, const int* d_tets,
                             long long n, int worst_count, std::vector<int>* worst,
                             hipStream_t st);

}  // namespace cmg::hip
