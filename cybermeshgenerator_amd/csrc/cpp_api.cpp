// cpp_api.cpp — the C++20 API (include/cmg/cmg_hip.hpp) over the C ABI of this library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cmg/cmg_hip.hpp"
#include "../../include/cmg_hip.h"
#include "device_common.hpp"
#include "weld_internal.hpp"

namespace cmg {

namespace {

thread_local backend::Backend g_last = backend::Backend::Cpu;

// PLC -> C handle (points + one polygon per facet polygon)
struct PlcHandle {
    cmg_plc* h;
    explicit PlcHandle(const PLC& plc) : h(cmg_plc_create()) {
        std::vector<double> xyz;
        xyz.reserve(plc.points.size() * 3);
        for (const Point3& p : plc.points) {
            xyz.push_back(p.x);
            xyz.push_back(p.y);
            xyz.push_back(p.z);
        }
        cmg_plc_set_points(h, xyz.data(), plc.points.size());
        for (const Facet& f : plc.facets)
            for (const Polygon& poly : f.polygons)
                if (poly.vertices.size() >= 3)
                    cmg_plc_add_facet(h, poly.vertices.data(), poly.vertices.size(), f.marker);
    }
    ~PlcHandle() { cmg_plc_destroy(h); }
};

MeshErrorCode code_of(cmg_status s) {
    switch (s) {
        case CMG_ERR_INVALID_INPUT: return MeshErrorCode::InvalidInput;
        case CMG_ERR_SELF_INTERSECTION: return MeshErrorCode::SelfIntersection;
        case CMG_ERR_UNSUPPORTED_OPTION: return MeshErrorCode::UnsupportedOption;
        case CMG_ERR_QUALITY_NOT_MET: return MeshErrorCode::QualityNotMet;
        case CMG_ERR_NOT_IMPLEMENTED: return MeshErrorCode::NotImplemented;
        default: return MeshErrorCode::Internal;
    }
}

void flatten(std::span<const Point3> pts, std::vector<double>& xyz) {
    xyz.resize(pts.size() * 3);
    for (std::size_t i = 0; i < pts.size(); ++i) {
        xyz[3 * i] = pts[i].x;
        xyz[3 * i + 1] = pts[i].y;
        xyz[3 * i + 2] = pts[i].z;
    }
}

}  // namespace

namespace voxelize {

expected<VoxelGrid, MeshError> voxelize(const PLC& plc, const VoxelOptions& opts) {
    PlcHandle h(plc);
    cmg_voxels* v = nullptr;
    char err[512] = {0};
    cmg_status st = cmg_plc_voxelize(h.h, opts.resolution,
                                     opts.mode == VoxelMode::SignedDistance ? 1 : 0, opts.pad,
                                     &v, err, sizeof err);
    if (st == CMG_ERR_INVALID_INPUT)
        return unexpected(MeshError{MeshErrorCode::InvalidInput, err});
    if (st != CMG_OK) throw std::runtime_error(err);
    g_last = backend::Backend::Hip;
    VoxelGrid g;
    cmg_voxels_dims(v, &g.nx, &g.ny, &g.nz);
    cmg_voxels_origin(v, &g.origin.x, &g.origin.y, &g.origin.z);
    g.spacing = cmg_voxels_spacing(v);
    const std::size_t n = g.cell_count();
    if (const unsigned char* o = cmg_voxels_occupancy(v)) g.occupancy.assign(o, o + n);
    if (const float* d = cmg_voxels_distance(v)) g.distance.assign(d, d + n);
    cmg_voxels_destroy(v);
    return g;
}

expected<std::vector<Point3>, MeshError> sdf_backward(const PLC& plc, const VoxelOptions& opts,
                                                      std::span<const float> grad_sdf) {
    PlcHandle h(plc);
    cmg_voxels* v = nullptr;
    char err[512] = {0};
    cmg_status st = cmg_plc_voxelize(h.h, opts.resolution, 1, opts.pad, &v, err, sizeof err);
    if (st == CMG_ERR_INVALID_INPUT)
        return unexpected(MeshError{MeshErrorCode::InvalidInput, err});
    if (st != CMG_OK) throw std::runtime_error(err);
    int nx, ny, nz;
    cmg_voxels_dims(v, &nx, &ny, &nz);
    if (grad_sdf.size() != static_cast<std::size_t>(nx) * ny * nz) {
        cmg_voxels_destroy(v);
        return unexpected(MeshError{MeshErrorCode::InvalidInput,
                                    "sdf_backward: one gradient value per cell expected"});
    }
    std::vector<float> g(plc.points.size() * 3);
    st = cmg_voxels_sdf_backward(h.h, v, grad_sdf.data(), g.data(), err, sizeof err);
    cmg_voxels_destroy(v);
    if (st != CMG_OK) {
        if (st == CMG_ERR_INVALID_INPUT) return unexpected(MeshError{code_of(st), err});
        throw std::runtime_error(err);
    }
    g_last = backend::Backend::Hip;
    std::vector<Point3> out(plc.points.size());
    for (std::size_t i = 0; i < out.size(); ++i) out[i] = {g[3 * i], g[3 * i + 1], g[3 * i + 2]};
    return out;
}

}  // namespace voxelize

namespace robust {

void ensure_initialized() {}

double orient3d(const Point3& a, const Point3& b, const Point3& c, const Point3& d) {
    const Point3 p[4] = {a, b, c, d};
    const Tetrahedron t{0, 1, 2, 3};
    return orient3d_batch(p, std::span<const Tetrahedron>(&t, 1))[0];
}

double insphere(const Point3& a, const Point3& b, const Point3& c, const Point3& d,
                const Point3& e) {
    const double xyz[15] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z,
                            d.x, d.y, d.z, e.x, e.y, e.z};
    const int q[5] = {0, 1, 2, 3, 4};
    double v = 0.0;
    char err[512] = {0};
    if (cmg_insphere_values(xyz, 5, q, 1, &v, err, sizeof err) != CMG_OK)
        throw std::runtime_error(err);
    g_last = backend::Backend::Hip;
    return v;
}

std::vector<double> orient3d_batch(std::span<const Point3> pts,
                                   std::span<const Tetrahedron> candidates) {
    std::vector<double> xyz;
    flatten(pts, xyz);
    std::vector<double> out(candidates.size());
    char err[512] = {0};
    if (cmg_orient3d_values(xyz.data(), pts.size(),
                            reinterpret_cast<const int*>(candidates.data()), candidates.size(),
                            out.data(), err, sizeof err) != CMG_OK)
        throw std::runtime_error(err);
    g_last = backend::Backend::Hip;
    return out;
}

std::vector<int> orient3d_signs(std::span<const Point3> pts,
                                std::span<const Tetrahedron> candidates) {
    std::vector<double> xyz;
    flatten(pts, xyz);
    std::vector<signed char> s(candidates.size());
    char err[512] = {0};
    if (cmg_orient3d_signs(xyz.data(), pts.size(),
                           reinterpret_cast<const int*>(candidates.data()), candidates.size(),
                           CMG_PRED_COMPAT, s.data(), err, sizeof err) != CMG_OK)
        throw std::runtime_error(err);
    g_last = backend::Backend::Hip;
    return std::vector<int>(s.begin(), s.end());
}

}  // namespace robust

namespace io {

namespace {

expected<PLC, MeshError> to_plc(cmg_plc* h, cmg_status st, const char* err) {
    if (st != CMG_OK) return unexpected(MeshError{code_of(st), err});
    PLC out;
    const std::size_t n = cmg_plc_num_points(h);
    const double* xyz = cmg_plc_points(h);
    out.points.resize(n);
    for (std::size_t i = 0; i < n; ++i) out.points[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    const std::size_t m = cmg_plc_num_triangles(h);
    const int* t = cmg_plc_triangles(h);
    out.facets.resize(m);
    for (std::size_t f = 0; f < m; ++f)
        out.facets[f] = Facet{{Polygon{{t[3 * f], t[3 * f + 1], t[3 * f + 2]}}}, 0};
    cmg_plc_destroy(h);
    return out;
}

template <class F>
expected<PLC, MeshError> read_with(F&& f) {
    hip::SurfaceFile sf;
    try {
        sf = f();
    } catch (const hip::InvalidInput& e) {
        return unexpected(MeshError{MeshErrorCode::InvalidInput, e.what()});
    }
    PLC out;
    out.points.resize(sf.xyz.size() / 3);
    for (std::size_t i = 0; i < out.points.size(); ++i)
        out.points[i] = {sf.xyz[3 * i], sf.xyz[3 * i + 1], sf.xyz[3 * i + 2]};
    out.facets.reserve(sf.polys.size());
    for (auto& poly : sf.polys) out.facets.push_back(Facet{{Polygon{std::move(poly)}}, 0});
    return out;
}

}  // namespace

expected<PLC, MeshError> read_stl(const std::string& path) {
    auto r = read_with([&] { return hip::read_stl_file(path); });
    if (r) g_last = backend::Backend::Hip;
    return r;
}

expected<PLC, MeshError> read_obj(const std::string& path) {
    return read_with([&] { return hip::read_obj_file(path); });
}

expected<PLC, MeshError> read_plc(const std::string& path) {
    cmg_plc* h = nullptr;
    char err[512] = {0};
    const cmg_status st = cmg_read_plc(path.c_str(), &h, err, sizeof err);
    return to_plc(h, st, err);
}

}  // namespace io

namespace quality {

QualityReport report(const Mesh& mesh, int worst_count) {
    QualityReport r;
    const std::size_t n = mesh.tetrahedra.size();
    r.num_tets = n;
    if (n == 0) return r;
    std::vector<double> xyz;
    flatten(mesh.points, xyz);
    std::vector<int> tets(4 * n);
    for (std::size_t i = 0; i < n; ++i)
        for (int k = 0; k < 4; ++k) tets[4 * i + k] = mesh.tetrahedra[i][k];
    double* d_xyz = nullptr;
    int* d_tets = nullptr;
    auto chk = [](hipError_t e) {
        if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
    };
    chk(hipMalloc(&d_xyz, sizeof(double) * (xyz.size() ? xyz.size() : 1)));
    chk(hipMalloc(&d_tets, sizeof(int) * tets.size()));
    struct Free {
        void* a;
        void* b;
        ~Free() {
            (void)hipFree(a);
            (void)hipFree(b);
        }
    } release{d_xyz, d_tets};
    if (!xyz.empty())
        chk(hipMemcpy(d_xyz, xyz.data(), sizeof(double) * xyz.size(), hipMemcpyHostToDevice));
    chk(hipMemcpy(d_tets, tets.data(), sizeof(int) * tets.size(), hipMemcpyHostToDevice));
    cmg_quality_report q;
    std::vector<int> worst(worst_count < 0 ? n : std::min<std::size_t>(worst_count, n));
    char err[512] = {0};
    const cmg_status st = cmg_hip_quality_report(d_xyz, mesh.points.size(), d_tets, n,
                                                 worst_count, &q, worst.data(), nullptr, err,
                                                 sizeof err);
    if (st != CMG_OK) throw std::runtime_error(err);
    g_last = backend::Backend::Hip;
    r.min_radius_edge = q.min_radius_edge;
    r.max_radius_edge = q.max_radius_edge;
    r.mean_radius_edge = q.mean_radius_edge;
    r.min_dihedral = q.min_dihedral;
    r.max_dihedral = q.max_dihedral;
    r.mean_dihedral = q.mean_dihedral;
    r.min_volume = q.min_volume;
    r.max_volume = q.max_volume;
    r.total_volume = q.total_volume;
    for (int b = 0; b < kHistogramBuckets; ++b) r.dihedral_histogram[b] = q.dihedral_histogram[b];
    worst.resize(q.num_worst);
    r.worst_tets = std::move(worst);
    return r;
}

}  // namespace quality

namespace backend {

std::size_t threshold_for(Op op) noexcept {
    switch (op) {  // dispatch.cpp:23-30
        case Op::OrientFilter: return 1u << 16;
        case Op::HilbertSort: return 1u << 15;
        case Op::PointLocation: return 1u << 16;
        case Op::QualityScan: return 1u << 17;
        case Op::Voxelize: return 1;
    }
    return static_cast<std::size_t>(-1);
}

// dispatch.cpp:39-63: below the op's threshold the CPU serves the call; at or above it,
// the device when one is present. The decision is recorded for last_backend().
bool should_offload(Op op, std::size_t problem_size) noexcept {
    if (problem_size < threshold_for(op)) {
        g_last = Backend::Cpu;
        return false;
    }
    if (cmg_hip_available() != 0) {
        g_last = Backend::Hip;
        return true;
    }
    g_last = Backend::Cpu;
    return false;
}

Backend last_backend() noexcept { return g_last; }

std::string_view name(Backend b) noexcept {
    switch (b) {
        case Backend::Cpu: return "cpu";
        case Backend::Cuda: return "cuda";
        case Backend::OpenCL: return "opencl";
        case Backend::Metal: return "metal";
        case Backend::Hip: return "hip";
    }
    return "cpu";
}

}  // namespace backend

}  // namespace cmg
