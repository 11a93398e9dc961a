// cmg_hip.hpp — C++20 API of libcmg_hip, the MI355X implementation of
// CyberMeshGenerator's predicate and voxelization path.
//
// Same names, argument meaning and error behaviour as the reference's
//   include/cmg/voxelize/voxelize.hpp:25-63   (VoxelMode, VoxelOptions, VoxelGrid, voxelize)
//   include/cmg/predicates/robust.hpp:22-46   (ensure_initialized, orient3d, insphere,
//                                              orient3d_batch, orient3d_signs)
//   include/cmg/backend/dispatch.hpp:17-39    (Backend, Op, should_offload, last_backend,
//                                              name)
//   include/cmg/io/formats.hpp:41,58, io.hpp:50 (read_stl, read_obj, read_plc)
//   include/cmg/quality/metrics.hpp:14-44     (QualityReport, report)
// with minimal value types (Point3, Tetrahedron, PLC, MeshError, expected) standing in
// for include/cmg/core. Every computation runs on the GPU; recoverable input errors come
// back as unexpected(MeshError{InvalidInput, ...}) exactly where the reference returns
// them, and device failures throw std::runtime_error.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <span>
#include <string>
#include <string_view>
#include <utility>
#include <variant>
#include <vector>

namespace cmg {

using Real = double;

struct Point3 {  // include/cmg/core/geometry.hpp:12-24
    Real x = 0, y = 0, z = 0;
    constexpr Real operator[](std::size_t i) const noexcept { return (&x)[i]; }
    friend constexpr bool operator==(const Point3&, const Point3&) = default;
};

using Index = int;
using Tetrahedron = std::array<Index, 4>;  // geometry.hpp:29
using Triangle = std::array<Index, 3>;

struct Polygon {  // include/cmg/core/plc.hpp:14-47 (the parts voxelize reads)
    std::vector<Index> vertices;
};
struct Facet {
    std::vector<Polygon> polygons;
    int marker = 0;
};
struct PLC {
    std::vector<Point3> points;
    std::vector<Facet> facets;
};

struct Mesh {  // include/cmg/core/mesh.hpp (the parts the quality report reads)
    std::vector<Point3> points;
    std::vector<Tetrahedron> tetrahedra;
};

enum class MeshErrorCode {  // include/cmg/core/error.hpp:40-48
    InvalidInput,
    SelfIntersection,
    UnsupportedOption,
    QualityNotMet,
    NotImplemented,
    Internal,
};

struct MeshError {
    MeshErrorCode code;
    std::string message;
};

template <class E>
struct unexpected {
    E error;
    explicit unexpected(E e) : error(std::move(e)) {}
};

// A small stand-in for std::expected (include/cmg/core/expected.hpp).
template <class T, class E>
class expected {
  public:
    expected(T v) : v_(std::in_place_index<0>, std::move(v)) {}
    expected(unexpected<E> u) : v_(std::in_place_index<1>, std::move(u.error)) {}
    explicit operator bool() const noexcept { return v_.index() == 0; }
    bool has_value() const noexcept { return v_.index() == 0; }
    T& value() & { return std::get<0>(v_); }
    const T& value() const& { return std::get<0>(v_); }
    T& operator*() & { return std::get<0>(v_); }
    const T& operator*() const& { return std::get<0>(v_); }
    T* operator->() { return &std::get<0>(v_); }
    const T* operator->() const { return &std::get<0>(v_); }
    const E& error() const& { return std::get<1>(v_); }

  private:
    std::variant<T, E> v_;
};

namespace voxelize {

enum class VoxelMode { Occupancy, SignedDistance };

struct VoxelOptions {
    int resolution = 64;
    int pad = 1;
    VoxelMode mode = VoxelMode::Occupancy;
};

struct VoxelGrid {
    Point3 origin;  // centre of cell (0,0,0)
    double spacing = 0;
    int nx = 0, ny = 0, nz = 0;
    std::vector<std::uint8_t> occupancy;
    std::vector<float> distance;
    std::size_t index(int i, int j, int k) const {
        return (static_cast<std::size_t>(k) * ny + j) * nx + i;
    }
    std::size_t cell_count() const { return static_cast<std::size_t>(nx) * ny * nz; }
};

// voxelize.hpp:63
expected<VoxelGrid, MeshError> voxelize(const PLC& plc, const VoxelOptions& opts = {});

// Additive: dL/dpoints of L = sum grad_sdf * sdf for a grid voxelize() produced in
// SignedDistance mode from the same PLC (signs held constant).
expected<std::vector<Point3>, MeshError> sdf_backward(const PLC& plc, const VoxelOptions& opts,
                                                      std::span<const float> grad_sdf);

}  // namespace voxelize

namespace robust {

// robust.hpp:22-46. Values and signs follow the shipped predicates (zero static
// filter): orient3d, insphere and orient3d_batch return the naive FP64 determinant,
// bit-identical, when it is nonzero, and otherwise the exact determinant's leading
// expansion component (the exact sign, 0 on a true zero).
void ensure_initialized();
double orient3d(const Point3& a, const Point3& b, const Point3& c, const Point3& d);
double insphere(const Point3& a, const Point3& b, const Point3& c, const Point3& d,
                const Point3& e);
std::vector<double> orient3d_batch(std::span<const Point3> pts,
                                   std::span<const Tetrahedron> candidates);
std::vector<int> orient3d_signs(std::span<const Point3> pts,
                                std::span<const Tetrahedron> candidates);

}  // namespace robust

namespace io {

// formats.hpp:41 / 58 and io.hpp:50: the reference's readers (grammar and messages of
// src/io/stl.cpp, src/io/obj.cpp); STL vertices are merged on the GPU. read_plc picks the
// reader from the extension (.stl, .obj; other formats: NotImplemented).
expected<PLC, MeshError> read_stl(const std::string& path);
expected<PLC, MeshError> read_obj(const std::string& path);
expected<PLC, MeshError> read_plc(const std::string& path);

}  // namespace io

namespace quality {

inline constexpr int kHistogramBuckets = 18;

struct QualityReport {  // metrics.hpp:14-40
    std::size_t num_tets = 0;
    double min_radius_edge = 0, max_radius_edge = 0, mean_radius_edge = 0;
    double min_dihedral = 0, max_dihedral = 0, mean_dihedral = 0;
    double min_volume = 0, max_volume = 0, total_volume = 0;
    std::array<int, kHistogramBuckets> dihedral_histogram{};
    std::vector<int> worst_tets;
};

// metrics.hpp:44, on the GPU. Equal radius-edge ratios are listed by ascending index;
// the sums are accumulated in a different order than the reference's loop, and the
// dihedral angles use the device acos. Throws std::runtime_error on a tet index out of
// range or a device failure.
QualityReport report(const Mesh& mesh, int worst_count = 10);

}  // namespace quality

namespace backend {

// dispatch.hpp:17-39 with the HIP backend added.
enum class Backend { Cpu, Cuda, OpenCL, Metal, Hip };
enum class Op { OrientFilter, HilbertSort, PointLocation, QualityScan, Voxelize };

// dispatch.cpp:39-63: false below threshold_for(op) (dispatch.cpp:23-30) and when no HIP
// device is present, true otherwise; records the decision for last_backend(). The entry
// points of this library run on the GPU at every size (it has no CPU kernels); the
// decision is for callers that dispatch between a CPU path and this one.
bool should_offload(Op op, std::size_t problem_size) noexcept;
std::size_t threshold_for(Op op) noexcept;
Backend last_backend() noexcept;
std::string_view name(Backend b) noexcept;

}  // namespace backend

}  // namespace cmg
