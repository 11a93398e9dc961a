// locate_internal.hpp — batched point location (locate.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace cmg::hip {

struct LocateStats {
    long long brute = 0;  // walks that hit the step limit and fell back to brute force
};

// One walk per query (src/delaunay/incremental.cpp:321-358) through a fixed mesh of
// ntets positively oriented tets (d_tets, 4 ints each) with neighbours (d_nbrs, 4 ints
// each, -1 on the hull), both 16-byte aligned; start tet d_start[i] or start_tet.
// d_out[i] = the located tet; d_steps (optional) = walk length, -1 for brute force.
// Synchronises st; throws InvalidInput when a walk meets an index out of range.
LocateStats locate_points(const double* d_xyz, long long npts, const int* d_tets,
                          const int* d_nbrs, long long ntets, const double* d_q, long long nq,
                          const int* d_start, int start_tet, int* d_out, int* d_steps,
                          hipStream_t st);

// Delaunay cavities of query points against a fixed mesh (incremental.cpp:148-161): for
// query i located in d_located[i], the tets reached breadth first from it through
// neighbours whose circumsphere contains the point (compat insphere > 0), in discovery
// order, into d_cav[i*cap ...]; d_count[i] = the cavity size, -1 when it exceeds cap or an
// index is out of range, 0 when d_located[i] < 0. Returns the number of queries that met
// an index out of range. Synchronises st.
long long cavities(const double* d_xyz, long long npts, const int* d_tets, const int* d_nbrs,
                   long long ntets, const double* d_q, long long nq, const int* d_located,
                   int cap, int* d_count, int* d_cav, hipStream_t st);

}  // namespace cmg::hip
