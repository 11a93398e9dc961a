// voxelize.hip — occupancy and signed-distance voxelization on gfx950 (in progress).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "cmg_hip_internal.hpp"
#include "device_common.hpp"

namespace cmg::hip {

bool make_grid(const double* xyz, size_t npts, int resolution, int pad, GridSpec& g) {
    // voxelize.cpp:52-74
    double lo[3] = {xyz[0], xyz[1], xyz[2]}, hi[3] = {xyz[0], xyz[1], xyz[2]};
    for (size_t i = 0; i < npts; ++i)
        for (int a = 0; a < 3; ++a) {
            double v = xyz[3 * i + a];
            if (v < lo[a]) lo[a] = v;
            if (v > hi[a]) hi[a] = v;
        }
    double ext = 0;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, hi[a] - lo[a]);
    g = GridSpec{};
    g.spacing = ext / resolution;
    if (!(g.spacing > 0)) return false;
    int* dims[3] = {&g.nx, &g.ny, &g.nz};
    for (int a = 0; a < 3; ++a) {
        double e = hi[a] - lo[a];
        int core = std::max(1, static_cast<int>(std::ceil(e / g.spacing)));
        double covered = core * g.spacing;
        g.corner[a] = lo[a] - (covered - e) / 2.0 - pad * g.spacing;
        *dims[a] = core + 2 * pad;
    }
    return true;
}

void voxelize_host(const double*, size_t, const int*, size_t, const GridSpec&, bool,
                   uint8_t*, float*) {
    throw HipError("voxelize: not implemented yet");
}

void sdf_backward_host(const double*, size_t, const int*, size_t, const GridSpec&,
                       const uint8_t*, const float*, float*) {
    throw HipError("sdf backward: not implemented yet");
}

}  // namespace cmg::hip
