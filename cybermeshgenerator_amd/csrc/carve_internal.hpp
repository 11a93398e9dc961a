// carve_internal.hpp — point-in-domain classification for the PLC carve (carve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>

namespace cmg::hip {

enum CarveMode : int {
    // src/constrained/tetrahedralize_plc.cpp:39-168: majority of three Moller-Trumbore
    // ray parities, each over the triangle list of the query's cell in a projection grid
    kCarveRays = 0,
    // src/voxelize/voxelize.cpp:78-195 at an arbitrary point: exact column parity over
    // the xy bucket of make_grid(points, resolution, 1), kNudge retries
    kCarveParity = 1,
};

class CarveEngine {
  public:
    // d_xyz: npts points (3 doubles each), d_tris: ntris triangles (3 ints each), both
    // device arrays that must stay valid (and unchanged) for the engine's lifetime.
    // resolution (parity mode; <= 0: the rays grid's cells per axis). Builds the cell
    // lists on `st` and synchronises it.
    CarveEngine(const double* d_xyz, long long npts, const int* d_tris, long long ntris,
                int mode, int resolution, hipStream_t st);
    ~CarveEngine();
    CarveEngine(const CarveEngine&) = delete;
    CarveEngine& operator=(const CarveEngine&) = delete;

    // inside[i] = 1 iff point q[3i..3i+2] is inside (device arrays, asynchronous on st).
    void points(const double* d_q, long long nq, uint8_t* d_inside, hipStream_t st);
    // keep[i] = 1 iff the centroid of tets[4i..4i+3] is inside (tetrahedralize_plc.cpp:
    // 258-268); tets with an index outside [0, npts) get 0 and are counted (bad_tets).
    void tets(const int* d_tets, long long nt, uint8_t* d_keep, hipStream_t st);
    // Synchronises st; tets rejected for out-of-range indices since the last call.
    unsigned long long bad_tets(hipStream_t st);

    int mode() const;
    int cells_per_axis() const;      // rays: r; parity: nx of the grid
    long long list_entries() const;  // total entries of the cell lists

    struct Impl;  // defined in carve.hip

  private:
    std::unique_ptr<Impl> p_;
};

}  // namespace cmg::hip
