// capi.cpp — the C ABI of libcmg_hip (include/cmg_hip.h).
//
// Handle semantics, status codes and errbuf behaviour follow the reference's stable C
// ABI (bindings/c/src/cmg_c.cpp:23-52, 79-106, 242-291): handles are created with new
// and released by *_destroy, *out is cleared first, errbuf receives a NUL-terminated
// message, and no C++ exception crosses the boundary. All computation runs on the GPU
// through the launchers in cmg_hip_internal.hpp; without a HIP device every compute
// entry point fails with CMG_ERR_INTERNAL.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cmg_hip.h"
#include "batch_internal.hpp"
#include "carve_internal.hpp"
#include "cmg_hip_internal.hpp"
#include "device_common.hpp"
#include "locate_internal.hpp"
#include "quality_internal.hpp"
#include "weld_internal.hpp"

struct cmg_plc {
    std::vector<double> xyz;            // 3 per point
    std::vector<std::vector<int>> polys;  // one polygon per facet (cmg_c.cpp:88-100)
    std::vector<int> markers;
    std::vector<int> tri_cache;           // cmg_plc_triangles (cmg_c.cpp:392-415)
};

struct cmg_batch {
    std::unique_ptr<cmg::hip::BatchEngine> engine;
};

struct cmg_carve {
    std::unique_ptr<cmg::hip::CarveEngine> engine;
};

struct cmg_voxels {
    cmg::hip::GridSpec grid;
    int mode = 0;                        // 0 occupancy, 1 signed distance
    std::vector<unsigned char> occupancy;  // exposed in occupancy mode only
    std::vector<float> distance;           // exposed in SDF mode only
    std::vector<unsigned char> inside;     // SDF mode: the sign bits, kept for backward
};

namespace {

thread_local const char* g_last_backend = "none";

void set_err(char* buf, size_t len, const std::string& msg) {
    if (buf && len) {
        std::strncpy(buf, msg.c_str(), len - 1);
        buf[len - 1] = '\0';
    }
}

template <class F>
cmg_status guarded(char* errbuf, size_t len, F&& f) {
    try {
        f();
        return CMG_OK;
    } catch (const cmg::hip::InvalidInput& e) {
        set_err(errbuf, len, e.what());
        return CMG_ERR_INVALID_INPUT;
    } catch (const std::exception& e) {
        set_err(errbuf, len, e.what());
        return CMG_ERR_INTERNAL;
    } catch (...) {
        set_err(errbuf, len, "unknown error");
        return CMG_ERR_INTERNAL;
    }
}

void require_device() {
    if (!cmg::hip::device_available())
        throw cmg::hip::HipError("no usable HIP device: libcmg_hip has no CPU path");
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// Fan-triangulate every polygon: (v0, v[i-1], v[i]) (src/voxelize/voxelize.cpp:29-38).
std::vector<int> fan_triangles(const cmg_plc& p) {
    std::vector<int> t;
    for (const auto& v : p.polys)
        for (size_t i = 2; i < v.size(); ++i) {
            t.push_back(v[0]);
            t.push_back(v[i - 1]);
            t.push_back(v[i]);
        }
    return t;
}

// Host-pointer batched signs: stage inputs in the library scratch, run, copy back.
void host_signs(int npts_per, int mode, const double* xyz, size_t npts, const int* idx,
                size_t n, int8_t* out) {
    require_device();
    if (mode != cmg::hip::kModeExact && mode != cmg::hip::kModeCompat)
        throw cmg::hip::InvalidInput("mode must be CMG_PRED_EXACT (0) or CMG_PRED_COMPAT (1)");
    if (n && (!xyz || !idx || !out)) throw cmg::hip::InvalidInput("null pointer");
    // the compaction queue holds 32-bit tuple indices
    if (n > 0xFFFFFFFFull) throw cmg::hip::InvalidInput("n must be < 2^32");
    for (size_t i = 0; i < n * (size_t)npts_per; ++i)
        if (idx[i] < 0 || (size_t)idx[i] >= npts)
            throw cmg::hip::InvalidInput("point index out of range");
    g_last_backend = "hip";
    if (n == 0) return;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    cmg::hip::PredPlan plan = cmg::hip::pred_plan(n, 1);
    size_t wsb = cmg::hip::pred_workspace_bytes(plan, n);
    size_t b_xyz = al(npts * 3 * sizeof(double)), b_idx = al(n * npts_per * sizeof(int)),
           b_out = al(n);
    cmg::hip::ScratchLease lease;
    char* base = static_cast<char*>(lease.s->get(b_xyz + b_idx + b_out + wsb));
    double* d_xyz = reinterpret_cast<double*>(base);
    int* d_idx = reinterpret_cast<int*>(base + b_xyz);
    int8_t* d_out = reinterpret_cast<int8_t*>(base + b_xyz + b_idx);
    void* d_ws = base + b_xyz + b_idx + b_out;
    hipStream_t st = nullptr;
    CMG_HIP(hipMemcpyAsync(d_xyz, xyz, npts * 3 * sizeof(double), hipMemcpyHostToDevice, st));
    CMG_HIP(hipMemcpyAsync(d_idx, idx, n * npts_per * sizeof(int), hipMemcpyHostToDevice, st));
    cmg::hip::PredWorkspace ws = cmg::hip::pred_workspace(d_ws, plan, n);
    cmg::hip::launch_signs_indexed(npts_per, mode, d_xyz, d_idx, n, d_out, plan, ws, st);
    CMG_HIP(hipMemcpyAsync(out, d_out, n, hipMemcpyDeviceToHost, st));
    cmg::hip::PredCounters c{};
    CMG_HIP(hipMemcpyAsync(&c, ws.ctr, sizeof c, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    if (c.failed)
        throw cmg::hip::HipError("exact predicate: expansion capacity exceeded on " +
                                 std::to_string(c.failed) + " tuples");
}

// Host-pointer batched values (npts_per 4 = orient3d, 5 = insphere).
void host_values(int npts_per, const double* xyz, size_t npts, const int* idx, size_t n,
                 double* out) {
    require_device();
    if (n && (!xyz || !idx || !out)) throw cmg::hip::InvalidInput("null pointer");
    if (n > 0xFFFFFFFFull) throw cmg::hip::InvalidInput("n must be < 2^32");
    for (size_t i = 0; i < n * (size_t)npts_per; ++i)
        if (idx[i] < 0 || (size_t)idx[i] >= npts)
            throw cmg::hip::InvalidInput("point index out of range");
    g_last_backend = "hip";
    if (n == 0) return;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    cmg::hip::PredPlan plan = cmg::hip::pred_plan(n, 1);
    size_t wsb = cmg::hip::pred_workspace_bytes(plan, n);
    size_t b_xyz = al(npts * 3 * sizeof(double)), b_idx = al(n * npts_per * sizeof(int)),
           b_out = al(n * sizeof(double));
    cmg::hip::ScratchLease lease;
    char* base = static_cast<char*>(lease.s->get(b_xyz + b_idx + b_out + wsb));
    double* d_xyz = reinterpret_cast<double*>(base);
    int* d_idx = reinterpret_cast<int*>(base + b_xyz);
    double* d_out = reinterpret_cast<double*>(base + b_xyz + b_idx);
    cmg::hip::PredWorkspace ws = cmg::hip::pred_workspace(base + b_xyz + b_idx + b_out, plan, n);
    hipStream_t st = nullptr;
    CMG_HIP(hipMemcpyAsync(d_xyz, xyz, npts * 3 * sizeof(double), hipMemcpyHostToDevice, st));
    CMG_HIP(hipMemcpyAsync(d_idx, idx, n * npts_per * sizeof(int), hipMemcpyHostToDevice, st));
    cmg::hip::launch_values_indexed(npts_per, d_xyz, d_idx, n, d_out, plan, ws, st);
    CMG_HIP(hipMemcpyAsync(out, d_out, n * sizeof(double), hipMemcpyDeviceToHost, st));
    cmg::hip::PredCounters c{};
    CMG_HIP(hipMemcpyAsync(&c, ws.ctr, sizeof c, hipMemcpyDeviceToHost, st));
    CMG_HIP(hipStreamSynchronize(st));
    if (c.failed)
        throw cmg::hip::HipError("exact predicate: expansion capacity exceeded on " +
                                 std::to_string(c.failed) + " tuples");
}

cmg_status soa_signs(int npts, const double* soa, size_t n, size_t ld, int mode,
                     signed char* out, void* workspace, size_t workspace_bytes,
                     void* stream, char* errbuf, size_t len) {
    return guarded(errbuf, len, [&] {
        require_device();
        if (mode != cmg::hip::kModeExact && mode != cmg::hip::kModeCompat)
            throw cmg::hip::InvalidInput("mode must be CMG_PRED_EXACT or CMG_PRED_COMPAT");
        if (n && (!soa || !out)) throw cmg::hip::InvalidInput("null pointer");
        if (n && ld < n) throw cmg::hip::InvalidInput("ld must be >= n");
        if (n > 0xFFFFFFFFull) throw cmg::hip::InvalidInput("n must be < 2^32");
        g_last_backend = "hip";
        cmg::hip::PredPlan plan = cmg::hip::pred_plan(n, 2);
        size_t need = cmg::hip::pred_workspace_bytes(plan, n);
        hipStream_t st = as_stream(stream);
        if (workspace) {
            if (workspace_bytes < need)
                throw cmg::hip::InvalidInput("workspace smaller than cmg_hip_pred_workspace_size");
            cmg::hip::PredWorkspace ws = cmg::hip::pred_workspace(workspace, plan, n);
            cmg::hip::launch_signs_soa(npts, mode, soa, n, ld, reinterpret_cast<int8_t*>(out),
                                       plan, ws, st);
        } else {
            cmg::hip::ScratchLease lease;
            cmg::hip::PredWorkspace ws = cmg::hip::pred_workspace(lease.s->get(need), plan, n);
            cmg::hip::launch_signs_soa(npts, mode, soa, n, ld, reinterpret_cast<int8_t*>(out),
                                       plan, ws, st);
            CMG_HIP(hipStreamSynchronize(st));
        }
    });
}

// ---- fixed workspace of the no-allocation orient3d entry point ------------------------
//
// cmg_hip_orient3d_signs_soa_noalloc runs the same kernels as cmg_hip_orient3d_signs_soa
// with its workspace in this file-scope device array instead of a caller or library
// allocation: every GPU the library runs on holds one copy (16 MiB) until the process
// ends, and every call on that GPU shares it. In host code `scratch` only names the
// symbol; its device address on each GPU comes from hipGetSymbolAddress, taken once per
// device with that device current.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
constexpr size_t kNoAllocMaxTuples = size_t(1) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

// This GPU's address of `scratch` (looked up on the first call on each device).
void* noalloc_scratch() {
    int dev = 0;
    CMG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw cmg::hip::HipError("device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw cmg::hip::HipError(std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

// ---- 1. the reference's C ABI subset ---------------------------------------------------

cmg_plc* cmg_plc_create(void) { return new cmg_plc{}; }
void cmg_plc_destroy(cmg_plc* p) { delete p; }

cmg_status cmg_plc_set_points(cmg_plc* plc, const double* xyz, size_t count) {
    if (!plc || (!xyz && count)) return CMG_ERR_INVALID_INPUT;
    plc->xyz.assign(xyz, xyz + 3 * count);
    return CMG_OK;
}

cmg_status cmg_plc_add_facet(cmg_plc* plc, const int* idx, size_t count, int marker) {
    if (!plc || (!idx && count)) return CMG_ERR_INVALID_INPUT;
    if (count < 3) return CMG_ERR_INVALID_INPUT;
    plc->polys.emplace_back(idx, idx + count);
    plc->markers.push_back(marker);
    return CMG_OK;
}

cmg_status cmg_plc_voxelize(const cmg_plc* in, int resolution, int mode, int pad,
                            cmg_voxels** out, char* errbuf, size_t len) {
    if (out) *out = nullptr;
    if (!in || !out) return CMG_ERR_INVALID_INPUT;
    if (resolution < 1) {  // cmg_c.cpp:246-249
        set_err(errbuf, len, "resolution must be >= 1");
        return CMG_ERR_INVALID_INPUT;
    }
    return guarded(errbuf, len, [&] {
        // voxelize.cpp:363-379
        const size_t npts = in->xyz.size() / 3;
        if (npts == 0) throw cmg::hip::InvalidInput("voxelize: PLC has no points");
        std::vector<int> tris = fan_triangles(*in);
        if (tris.empty())
            throw cmg::hip::InvalidInput("voxelize: PLC has no boundary triangles");
        for (int v : tris)
            if (v < 0 || (size_t)v >= npts)
                throw cmg::hip::InvalidInput("voxelize: facet index out of range");
        cmg::hip::GridSpec g;
        if (!cmg::hip::make_grid(in->xyz.data(), npts, resolution, pad, g))
            throw cmg::hip::InvalidInput("voxelize: degenerate (zero-extent) PLC");
        if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0)
            throw cmg::hip::InvalidInput("voxelize: grid has no cells (negative pad)");
        require_device();
        g_last_backend = "hip";
        auto* v = new cmg_voxels{};
        v->grid = g;
        v->mode = mode == 1 ? 1 : 0;
        const size_t ncell = (size_t)g.nx * g.ny * g.nz;
        std::vector<unsigned char> occ(ncell);
        if (v->mode == 1) v->distance.resize(ncell);
        try {
            cmg::hip::voxelize_host(in->xyz.data(), npts, tris.data(), tris.size() / 3, g,
                                    v->mode == 1, occ.data(),
                                    v->mode == 1 ? v->distance.data() : nullptr);
        } catch (...) {
            delete v;
            throw;
        }
        if (v->mode == 1) v->inside = std::move(occ);
        else v->occupancy = std::move(occ);
        *out = v;
    });
}

void cmg_voxels_destroy(cmg_voxels* v) { delete v; }

void cmg_voxels_dims(const cmg_voxels* v, int* nx, int* ny, int* nz) {
    if (nx) *nx = v ? v->grid.nx : 0;
    if (ny) *ny = v ? v->grid.ny : 0;
    if (nz) *nz = v ? v->grid.nz : 0;
}

void cmg_voxels_origin(const cmg_voxels* v, double* x, double* y, double* z) {
    // origin = corner + 0.5*spacing, the centre of cell (0,0,0) (voxelize.cpp:389-391)
    if (x) *x = v ? v->grid.corner[0] + 0.5 * v->grid.spacing : 0.0;
    if (y) *y = v ? v->grid.corner[1] + 0.5 * v->grid.spacing : 0.0;
    if (z) *z = v ? v->grid.corner[2] + 0.5 * v->grid.spacing : 0.0;
}

double cmg_voxels_spacing(const cmg_voxels* v) { return v ? v->grid.spacing : 0.0; }

const unsigned char* cmg_voxels_occupancy(const cmg_voxels* v) {
    return (v && !v->occupancy.empty()) ? v->occupancy.data() : nullptr;
}

const float* cmg_voxels_distance(const cmg_voxels* v) {
    return (v && !v->distance.empty()) ? v->distance.data() : nullptr;
}

const char* cmg_version(void) { return "0.5.0-hip"; }

// ---- 2. additive host-pointer entry points ---------------------------------------------

cmg_status cmg_orient3d_signs(const double* xyz, size_t npts, const int* quads, size_t n,
                              int mode, signed char* out, char* errbuf, size_t len) {
    return guarded(errbuf, len, [&] {
        host_signs(4, mode, xyz, npts, quads, n, reinterpret_cast<int8_t*>(out));
    });
}

cmg_status cmg_insphere_signs(const double* xyz, size_t npts, const int* quints, size_t n,
                              int mode, signed char* out, char* errbuf, size_t len) {
    return guarded(errbuf, len, [&] {
        host_signs(5, mode, xyz, npts, quints, n, reinterpret_cast<int8_t*>(out));
    });
}

cmg_status cmg_orient3d_values(const double* xyz, size_t npts, const int* quads, size_t n,
                               double* out, char* errbuf, size_t len) {
    return guarded(errbuf, len, [&] { host_values(4, xyz, npts, quads, n, out); });
}

cmg_status cmg_insphere_values(const double* xyz, size_t npts, const int* quints, size_t n,
                               double* out, char* errbuf, size_t len) {
    return guarded(errbuf, len, [&] { host_values(5, xyz, npts, quints, n, out); });
}

cmg_status cmg_voxels_sdf_backward(const cmg_plc* plc, const cmg_voxels* grid,
                                   const float* grad_sdf, float* grad_xyz, char* errbuf,
                                   size_t len) {
    if (!plc || !grid || !grad_sdf || !grad_xyz) {
        set_err(errbuf, len, "null argument");
        return CMG_ERR_INVALID_INPUT;
    }
    return guarded(errbuf, len, [&] {
        if (grid->mode != 1)
            throw cmg::hip::InvalidInput("sdf backward needs a grid voxelized with mode 1");
        const size_t npts = plc->xyz.size() / 3;
        std::vector<int> tris = fan_triangles(*plc);
        for (int v : tris)
            if (v < 0 || (size_t)v >= npts)
                throw cmg::hip::InvalidInput("facet index out of range");
        require_device();
        g_last_backend = "hip";
        cmg::hip::sdf_backward_host(plc->xyz.data(), npts, tris.data(), tris.size() / 3,
                                    grid->grid, grid->inside.data(), grad_sdf, grad_xyz,
                                    nullptr);
    });
}

cmg_status cmg_voxels_sdf_nearest(const cmg_plc* plc, const cmg_voxels* grid,
                                  int32_t* tri_out, char* errbuf, size_t len) {
    if (!plc || !grid || !tri_out) {
        set_err(errbuf, len, "null argument");
        return CMG_ERR_INVALID_INPUT;
    }
    return guarded(errbuf, len, [&] {
        if (grid->mode != 1)
            throw cmg::hip::InvalidInput("sdf nearest needs a grid voxelized with mode 1");
        const size_t npts = plc->xyz.size() / 3;
        std::vector<int> tris = fan_triangles(*plc);
        for (int v : tris)
            if (v < 0 || (size_t)v >= npts)
                throw cmg::hip::InvalidInput("facet index out of range");
        require_device();
        g_last_backend = "hip";
        cmg::hip::sdf_backward_host(plc->xyz.data(), npts, tris.data(), tris.size() / 3,
                                    grid->grid, grid->inside.data(), nullptr, nullptr,
                                    tri_out);
    });
}

const char* cmg_backend_last(void) { return g_last_backend; }

// ---- 3. device layer -------------------------------------------------------------------

int cmg_hip_available(void) { return cmg::hip::device_available() ? 1 : 0; }

int cmg_hip_orient3d_signs(const double* xyz, int npts, const int* tets, int n,
                           int* out_sign) {
    if (n < 0 || npts < 0) return 1;
    std::vector<signed char> s((size_t)n);
    cmg_status st = cmg_orient3d_signs(xyz, (size_t)npts, tets, (size_t)n,
                                       CMG_PRED_COMPAT, s.data(), nullptr, 0);
    if (st != CMG_OK) return (int)st;
    for (int i = 0; i < n; ++i) out_sign[i] = s[(size_t)i];
    return 0;
}

size_t cmg_hip_pred_workspace_size(size_t n, int npts) {
    (void)npts;
    return cmg::hip::pred_workspace_bytes(cmg::hip::pred_plan(n, 2), n);
}

cmg_status cmg_hip_orient3d_signs_soa(const double* soa, size_t n, size_t ld, int mode,
                                      signed char* out, void* workspace,
                                      size_t workspace_bytes, void* stream, char* errbuf,
                                      size_t len) {
    return soa_signs(4, soa, n, ld, mode, out, workspace, workspace_bytes, stream, errbuf,
                     len);
}

cmg_status cmg_hip_orient3d_signs_soa_noalloc(const double* soa, size_t n, size_t ld,
                                              int mode, signed char* out, void* stream,
                                              char* errbuf, size_t len) {
    return guarded(errbuf, len, [&] {
        require_device();
        if (mode != cmg::hip::kModeExact && mode != cmg::hip::kModeCompat)
            throw cmg::hip::InvalidInput("mode must be CMG_PRED_EXACT or CMG_PRED_COMPAT");
        if (n && (!soa || !out)) throw cmg::hip::InvalidInput("null pointer");
        if (n && ld < n) throw cmg::hip::InvalidInput("ld must be >= n");
        if (n > kNoAllocMaxTuples)
            throw cmg::hip::InvalidInput(
                "n must be <= 2^20 tuples (the size of the library's fixed workspace)");
        const cmg::hip::PredPlan plan = cmg::hip::pred_plan(n, 2);
        if (cmg::hip::pred_workspace_bytes(plan, n) > kNoAllocScratchBytes)
            throw cmg::hip::InvalidInput("batch needs more than the 16 MiB fixed workspace");
        g_last_backend = "hip";
        const cmg::hip::PredWorkspace ws = cmg::hip::pred_workspace(noalloc_scratch(), plan, n);
        cmg::hip::launch_signs_soa(4, mode, soa, n, ld, reinterpret_cast<int8_t*>(out), plan,
                                   ws, as_stream(stream));
    });
}

cmg_status cmg_hip_insphere_signs_soa(const double* soa, size_t n, size_t ld, int mode,
                                      signed char* out, void* workspace,
                                      size_t workspace_bytes, void* stream, char* errbuf,
                                      size_t len) {
    return soa_signs(5, soa, n, ld, mode, out, workspace, workspace_bytes, stream, errbuf,
                     len);
}

cmg_status cmg_hip_pred_stats(void* workspace, size_t n, int npts, void* stream,
                              cmg_pred_stats* out, char* errbuf, size_t len) {
    (void)npts;
    if (!out) return CMG_ERR_INVALID_INPUT;
    return guarded(errbuf, len, [&] {
        require_device();
        cmg::hip::PredPlan plan = cmg::hip::pred_plan(n, 2);
        hipStream_t st = as_stream(stream);
        cmg::hip::PredCounters c{};
        auto read = [&](void* base) {
            cmg::hip::PredWorkspace ws = cmg::hip::pred_workspace(base, plan, n);
            cmg::hip::launch_pred_stats(ws, st);
            CMG_HIP(hipMemcpyAsync(&c, ws.ctr, sizeof c, hipMemcpyDeviceToHost, st));
            CMG_HIP(hipStreamSynchronize(st));
        };
        if (workspace) {
            read(workspace);
        } else {
            cmg::hip::ScratchLease lease;
            read(lease.s->get(cmg::hip::pred_workspace_bytes(plan, n)));
        }
        out->lanes = n;
        out->fp32 = c.fp32;
        out->fp64 = c.fp64;
        out->exact = c.exact + c.inline_;
        out->exact_inline = c.inline_ + c.fixed;
        out->wide = c.wide_n;
        out->failed = c.failed;
    });
}

// ---- 4. batched voxelization -----------------------------------------------------------

cmg_status cmg_voxelize_batch_create(int nmesh, const int64_t* point_off,
                                     const int64_t* tri_off, const int* resolution,
                                     const int* pad, const double* xyz, const int* tris,
                                     void* stream, cmg_batch** out, char* errbuf,
                                     size_t len) {
    if (out) *out = nullptr;
    if (!out || nmesh < 1 || !point_off || !tri_off || !resolution || !pad || !xyz || !tris) {
        set_err(errbuf, len, "batch: invalid arguments");
        return CMG_ERR_INVALID_INPUT;
    }
    return guarded(errbuf, len, [&] {
        require_device();
        g_last_backend = "hip";
        std::vector<long long> ps(point_off, point_off + nmesh + 1),
            ts(tri_off, tri_off + nmesh + 1);
        auto b = std::make_unique<cmg_batch>();
        b->engine = std::make_unique<cmg::hip::BatchEngine>(nmesh, ps.data(), ts.data(),
                                                            resolution, pad, xyz, tris,
                                                            as_stream(stream));
        *out = b.release();
    });
}

void cmg_voxelize_batch_destroy(cmg_batch* b) { delete b; }

int64_t cmg_voxelize_batch_cells(const cmg_batch* b) {
    return b ? b->engine->total_cells() : 0;
}

cmg_status cmg_voxelize_batch_grid(const cmg_batch* b, int mesh, int* dims, double* origin,
                                   double* spacing, int64_t* cell_offset) {
    if (!b || mesh < 0 || mesh >= b->engine->meshes()) return CMG_ERR_INVALID_INPUT;
    const cmg::hip::MeshDesc& M = b->engine->mesh(mesh);
    if (dims) dims[0] = M.g.nx, dims[1] = M.g.ny, dims[2] = M.g.nz;
    if (origin)
        for (int a = 0; a < 3; ++a) origin[a] = M.g.corner[a] + 0.5 * M.g.spacing;
    if (spacing) *spacing = M.g.spacing;
    if (cell_offset) *cell_offset = M.cell_base;
    return CMG_OK;
}

cmg_status cmg_voxelize_batch_run(cmg_batch* b, int flags, const double* xyz, float* sdf,
                                  const float* grad_sdf, float* grad_xyz, void* stream,
                                  char* errbuf, size_t len) {
    if (!b) return CMG_ERR_INVALID_INPUT;
    return guarded(errbuf, len, [&] {
        g_last_backend = "hip";
        b->engine->run(flags, xyz, sdf, grad_sdf, grad_xyz, as_stream(stream));
    });
}

cmg_status cmg_voxelize_batch_status(cmg_batch* b, void* stream, unsigned* overflow,
                                     uint64_t* pairs, char* errbuf, size_t len) {
    if (!b) return CMG_ERR_INVALID_INPUT;
    return guarded(errbuf, len, [&] {
        unsigned long long pr = 0;
        b->engine->status(as_stream(stream), overflow, &pr);
        if (pairs) *pairs = pr;
    });
}

cmg_status cmg_voxelize_batch_autotune(cmg_batch* b, float* sdf, const float* grad_sdf,
                                       float* grad_xyz, void* stream, int* vpl, int* depth,
                                       float* ms, char* errbuf, size_t len) {
    if (!b || !sdf || !grad_sdf || !grad_xyz) return CMG_ERR_INVALID_INPUT;
    return guarded(errbuf, len, [&] {
        auto all = b->engine->autotune(grad_sdf, sdf, grad_xyz, as_stream(stream));
        if (ms)
            for (size_t i = 0; i < all.size() && i < 6; ++i) ms[i] = all[i].ms;
        cmg::hip::BatchVariant v = b->engine->variant();
        if (vpl) *vpl = v.vpl;
        if (depth) *depth = v.depth;
    });
}

cmg_status cmg_voxelize_batch_set_variant(cmg_batch* b, int vpl, int depth) {
    if (!b) return CMG_ERR_INVALID_INPUT;
    return guarded(nullptr, 0, [&] { b->engine->set_variant(vpl, depth); });
}

const unsigned char* cmg_voxelize_batch_occupancy(const cmg_batch* b) {
    return b ? b->engine->occupancy() : nullptr;
}

cmg_status cmg_voxelize_batch_nearest(cmg_batch* b, int32_t* tri, void* stream,
                                      char* errbuf, size_t len) {
    if (!b || !tri) return CMG_ERR_INVALID_INPUT;
    return guarded(errbuf, len, [&] { b->engine->nearest(tri, as_stream(stream)); });
}

cmg_status cmg_voxelize_batch_counters(cmg_batch* b, void* stream, uint64_t* out,
                                       char* errbuf, size_t len) {
    if (!b || !out) return CMG_ERR_INVALID_INPUT;
    return guarded(errbuf, len, [&] {
        const cmg::hip::BatchCounters c = b->engine->counters(as_stream(stream));
        out[0] = c.node_tests;
        out[1] = c.tri_tests;
        out[2] = c.tri_evals;
        out[3] = c.overflow;
        out[4] = c.tile_used;
    });
}

void cmg_hip_kernel_timer(const char* name, int enable) {
    if (name) cmg::hip::timer_enable(name, enable != 0);
}

cmg_status cmg_hip_kernel_time(const char* name, double* total_ms, uint64_t* count,
                               char* errbuf, size_t len) {
    if (!name) return CMG_ERR_INVALID_INPUT;
    return guarded(errbuf, len, [&] { cmg::hip::timer_read(name, total_ms, count); });
}

// ---- 5. PLC carve ------------------------------------------------------------------------

cmg_status cmg_carve_create(const double* xyz, size_t npts, const int* tris, size_t ntris,
                            int mode, int resolution, void* stream, cmg_carve** out,
                            char* errbuf, size_t len) {
    if (out) *out = nullptr;
    if (!out) {
        set_err(errbuf, len, "carve: null output handle");
        return CMG_ERR_INVALID_INPUT;
    }
    return guarded(errbuf, len, [&] {
        require_device();
        g_last_backend = "hip";
        auto c = std::make_unique<cmg_carve>();
        c->engine = std::make_unique<cmg::hip::CarveEngine>(
            xyz, (long long)npts, tris, (long long)ntris, mode, resolution, as_stream(stream));
        *out = c.release();
    });
}

void cmg_carve_destroy(cmg_carve* c) { delete c; }

cmg_status cmg_carve_points(cmg_carve* c, const double* q, size_t nq, uint8_t* inside,
                            void* stream, char* errbuf, size_t len) {
    if (!c) return CMG_ERR_INVALID_INPUT;
    return guarded(errbuf, len, [&] {
        g_last_backend = "hip";
        c->engine->points(q, (long long)nq, inside, as_stream(stream));
    });
}

cmg_status cmg_carve_tets(cmg_carve* c, const int* tets, size_t ntets, uint8_t* keep,
                          void* stream, char* errbuf, size_t len) {
    if (!c) return CMG_ERR_INVALID_INPUT;
    return guarded(errbuf, len, [&] {
        g_last_backend = "hip";
        c->engine->tets(tets, (long long)ntets, keep, as_stream(stream));
    });
}

cmg_status cmg_carve_status(cmg_carve* c, void* stream, uint64_t* bad_tets, char* errbuf,
                            size_t len) {
    if (!c) return CMG_ERR_INVALID_INPUT;
    return guarded(errbuf, len, [&] {
        const unsigned long long b = c->engine->bad_tets(as_stream(stream));
        if (bad_tets) *bad_tets = b;
    });
}

cmg_status cmg_plc_carve(const cmg_plc* plc, const int* tets, size_t ntets, int mode,
                         uint8_t* keep, char* errbuf, size_t len) {
    if (!plc || (ntets && (!tets || !keep))) {
        set_err(errbuf, len, "carve: null pointer");
        return CMG_ERR_INVALID_INPUT;
    }
    return guarded(errbuf, len, [&] {
        require_device();
        g_last_backend = "hip";
        const size_t npts = plc->xyz.size() / 3;
        if (npts == 0) throw cmg::hip::InvalidInput("carve: PLC has no points");
        std::vector<int> tris = fan_triangles(*plc);
        if (tris.empty())
            throw cmg::hip::InvalidInput("PLC has facets but no boundary triangles");
        for (int v : tris)
            if (v < 0 || (size_t)v >= npts)
                throw cmg::hip::InvalidInput("carve: facet vertex index out of range");
        for (size_t i = 0; i < 4 * ntets; ++i)
            if (tets[i] < 0 || (size_t)tets[i] >= npts)
                throw cmg::hip::InvalidInput("carve: tet vertex index out of range");
        hipStream_t st = nullptr;
        auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
        const size_t b_xyz = al(plc->xyz.size() * sizeof(double)),
                     b_tri = al(tris.size() * sizeof(int)), b_tet = al(4 * ntets * sizeof(int)),
                     b_out = al(ntets);
        cmg::hip::ScratchLease lease;
        char* base = static_cast<char*>(lease.s->get(b_xyz + b_tri + b_tet + b_out));
        double* d_xyz = reinterpret_cast<double*>(base);
        int* d_tri = reinterpret_cast<int*>(base + b_xyz);
        int* d_tet = reinterpret_cast<int*>(base + b_xyz + b_tri);
        uint8_t* d_out = reinterpret_cast<uint8_t*>(base + b_xyz + b_tri + b_tet);
        CMG_HIP(hipMemcpyAsync(d_xyz, plc->xyz.data(), plc->xyz.size() * sizeof(double),
                               hipMemcpyHostToDevice, st));
        CMG_HIP(hipMemcpyAsync(d_tri, tris.data(), tris.size() * sizeof(int),
                               hipMemcpyHostToDevice, st));
        if (ntets)
            CMG_HIP(hipMemcpyAsync(d_tet, tets, 4 * ntets * sizeof(int), hipMemcpyHostToDevice,
                                   st));
        cmg::hip::CarveEngine eng(d_xyz, (long long)npts, d_tri, (long long)(tris.size() / 3),
                                  mode, 0, st);
        eng.tets(d_tet, (long long)ntets, d_out, st);
        if (ntets) CMG_HIP(hipMemcpyAsync(keep, d_out, ntets, hipMemcpyDeviceToHost, st));
        CMG_HIP(hipStreamSynchronize(st));
    });
}

// ---- 6. surface files -------------------------------------------------------------------

cmg_status cmg_read_plc(const char* path, cmg_plc** out, char* errbuf, size_t len) {
    if (out) *out = nullptr;
    if (!path || !out) return CMG_ERR_INVALID_INPUT;
    // extension dispatch (src/io/dispatch.cpp:20-27, 41-55, 63-73)
    std::string p(path), ext;
    const size_t dot = p.rfind('.');
    if (dot != std::string::npos)
        for (char c : p.substr(dot + 1)) ext.push_back((char)std::tolower((unsigned char)c));
    if (ext != "stl" && ext != "obj") {
        if (ext == "poly" || ext == "smesh" || ext == "off" || ext == "ply") {
            set_err(errbuf, len, "libcmg_hip reads .stl and .obj surfaces only: " + p);
            return CMG_ERR_NOT_IMPLEMENTED;
        }
        set_err(errbuf, len, "unsupported file format for " + p);
        return CMG_ERR_INVALID_INPUT;
    }
    return guarded(errbuf, len, [&] {
        cmg::hip::SurfaceFile f;
        if (ext == "stl") {
            require_device();
            g_last_backend = "hip";
            f = cmg::hip::read_stl_file(p);
        } else {
            f = cmg::hip::read_obj_file(p);
        }
        auto plc = std::make_unique<cmg_plc>();
        plc->xyz = std::move(f.xyz);
        plc->polys = std::move(f.polys);
        plc->markers.assign(plc->polys.size(), 0);
        *out = plc.release();
    });
}

size_t cmg_plc_num_points(const cmg_plc* p) { return p ? p->xyz.size() / 3 : 0; }

const double* cmg_plc_points(cmg_plc* p) { return p ? p->xyz.data() : nullptr; }

size_t cmg_plc_num_triangles(cmg_plc* p) {
    if (!p) return 0;
    p->tri_cache = fan_triangles(*p);
    return p->tri_cache.size() / 3;
}

const int* cmg_plc_triangles(cmg_plc* p) {
    if (!p) return nullptr;
    p->tri_cache = fan_triangles(*p);
    return p->tri_cache.data();
}

cmg_status cmg_hip_stl_decode(const void* bytes, size_t nbytes, double* xyz, int* tris,
                              int64_t* npoints, void* stream, char* errbuf, size_t len) {
    if (npoints) *npoints = 0;
    return guarded(errbuf, len, [&] {
        require_device();
        g_last_backend = "hip";
        const long long n = cmg::hip::stl_decode_weld(bytes, nbytes, xyz, tris, as_stream(stream));
        if (npoints) *npoints = n;
    });
}

cmg_status cmg_hip_weld(const double* soup, size_t nverts, double* xyz, int* idx,
                        int64_t* npoints, void* stream, char* errbuf, size_t len) {
    if (npoints) *npoints = 0;
    return guarded(errbuf, len, [&] {
        require_device();
        g_last_backend = "hip";
        const long long n = cmg::hip::weld_vertices(soup, (long long)nverts, xyz, idx,
                                                   as_stream(stream));
        if (npoints) *npoints = n;
    });
}

// ---- 7. mesh quality report --------------------------------------------------------------

cmg_status cmg_hip_quality_report(const double* xyz, size_t npts, const int* tets,
                                  size_t ntets, int worst_count, cmg_quality_report* out,
                                  int* worst, void* stream, char* errbuf, size_t len) {
    if (!out) {
        set_err(errbuf, len, "quality: null report");
        return CMG_ERR_INVALID_INPUT;
    }
    *out = cmg_quality_report{};
    return guarded(errbuf, len, [&] {
        require_device();
        g_last_backend = "hip";
        std::vector<int> w;
        const cmg::hip::QualityReport r = cmg::hip::quality_report(
            xyz, (long long)npts, tets, (long long)ntets, worst_count, worst ? &w : nullptr,
            as_stream(stream));
        out->num_tets = (uint64_t)r.num_tets;
        out->min_radius_edge = r.min_radius_edge;
        out->max_radius_edge = r.max_radius_edge;
        out->mean_radius_edge = r.mean_radius_edge;
        out->min_dihedral = r.min_dihedral;
        out->max_dihedral = r.max_dihedral;
        out->mean_dihedral = r.mean_dihedral;
        out->min_volume = r.min_volume;
        out->max_volume = r.max_volume;
        out->total_volume = r.total_volume;
        for (int b = 0; b < 18; ++b) out->dihedral_histogram[b] = r.dihedral_histogram[b];
        out->num_worst = (uint64_t)w.size();
        if (worst && !w.empty()) std::memcpy(worst, w.data(), sizeof(int) * w.size());
    });
}

// ---- 8. point location -------------------------------------------------------------------

cmg_status cmg_hip_locate(const double* xyz, size_t npts, const int* tets, const int* nbrs,
                          size_t ntets, const double* q, size_t nq, const int* start,
                          int start_tet, int* out, int* steps, uint64_t* brute,
                          void* stream, char* errbuf, size_t len) {
    if (brute) *brute = 0;
    return guarded(errbuf, len, [&] {
        require_device();
        g_last_backend = "hip";
        const cmg::hip::LocateStats s = cmg::hip::locate_points(
            xyz, (long long)npts, tets, nbrs, (long long)ntets, q, (long long)nq, start,
            start_tet, out, steps, as_stream(stream));
        if (brute) *brute = (uint64_t)s.brute;
    });
}

cmg_status cmg_hip_cavity(const double* xyz, size_t npts, const int* tets, const int* nbrs,
                          size_t ntets, const double* q, size_t nq, const int* located,
                          int cap, int* count, int* cavity, uint64_t* bad, void* stream,
                          char* errbuf, size_t len) {
    if (bad) *bad = 0;
    return guarded(errbuf, len, [&] {
        require_device();
        g_last_backend = "hip";
        const long long b = cmg::hip::cavities(xyz, (long long)npts, tets, nbrs,
                                               (long long)ntets, q, (long long)nq, located,
                                               cap, count, cavity, as_stream(stream));
        if (bad) *bad = (uint64_t)b;
    });
}

}  // extern "C"
