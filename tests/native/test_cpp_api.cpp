// C++20 API test: the reference's own test cases for this path (tests/test_voxelize.cpp,
// tests/test_backend.cpp, tests/test_foundation.cpp, tests/io/test_stl.cpp,
// tests/io/test_obj.cpp, tests/test_metrics.cpp), run against cmg::voxelize /
// cmg::robust / cmg::backend / cmg::io / cmg::quality of libcmg_hip. Exit 0 on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "cmg/cmg_hip.hpp"

using namespace cmg;

static int failures = 0;
#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);    \
            ++failures;                                                          \
        }                                                                        \
    } while (0)

static PLC subdivided_cube(int n) {
    PLC p;
    std::map<std::array<int, 3>, int> vid;
    auto vert = [&](int x, int y, int z) {
        std::array<int, 3> k{x, y, z};
        auto it = vid.find(k);
        if (it != vid.end()) return it->second;
        int id = static_cast<int>(p.points.size());
        p.points.push_back({double(x) / n, double(y) / n, double(z) / n});
        vid[k] = id;
        return id;
    };
    auto quad = [&](int a, int b, int c, int d) {
        p.facets.push_back(Facet{{Polygon{{a, b, c}}}, 0});
        p.facets.push_back(Facet{{Polygon{{a, c, d}}}, 0});
    };
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            quad(vert(i, j, 0), vert(i + 1, j, 0), vert(i + 1, j + 1, 0), vert(i, j + 1, 0));
            quad(vert(i, j, n), vert(i, j + 1, n), vert(i + 1, j + 1, n), vert(i + 1, j, n));
            quad(vert(i, 0, j), vert(i, 0, j + 1), vert(i + 1, 0, j + 1), vert(i + 1, 0, j));
            quad(vert(i, n, j), vert(i + 1, n, j), vert(i + 1, n, j + 1), vert(i, n, j + 1));
            quad(vert(0, i, j), vert(0, i + 1, j), vert(0, i + 1, j + 1), vert(0, i, j + 1));
            quad(vert(n, i, j), vert(n, i, j + 1), vert(n, i + 1, j + 1), vert(n, i + 1, j));
        }
    return p;
}

static double occupied_volume(const voxelize::VoxelGrid& g) {
    std::size_t n = 0;
    for (std::uint8_t c : g.occupancy) n += c;
    return n * g.spacing * g.spacing * g.spacing;
}

int main() {
    // test_voxelize.cpp:100-114 — cube volume converges
    PLC cube = subdivided_cube(4);
    auto coarse = voxelize::voxelize(cube, {.resolution = 12});
    auto fine = voxelize::voxelize(cube, {.resolution = 48});
    CHECK(bool(coarse) && bool(fine));
    double ec = std::fabs(occupied_volume(*coarse) - 1.0),
           ef = std::fabs(occupied_volume(*fine) - 1.0);
    CHECK(ec < 0.03 && ef < 0.03 && ef <= ec + 1e-9);
    // :128-138 — determinism
    auto a = voxelize::voxelize(cube, {.resolution = 20});
    auto b = voxelize::voxelize(cube, {.resolution = 20});
    CHECK(a->occupancy == b->occupancy && a->origin == b->origin);
    // :140-163 — known answer
    auto g = voxelize::voxelize(cube, {.resolution = 16, .pad = 1});
    std::size_t occupied = 0, stray = 0, missing = 0;
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                bool in = i >= 1 && i < 17 && j >= 1 && j < 17 && k >= 1 && k < 17;
                auto v = g->occupancy[g->index(i, j, k)];
                occupied += v;
                stray += v && !in;
                missing += !v && in;
            }
    CHECK(stray == 0 && missing == 0 && occupied == 4096);
    // :165-182 — SDF signs
    auto s = voxelize::voxelize(cube, {.resolution = 16, .pad = 2,
                                       .mode = voxelize::VoxelMode::SignedDistance});
    CHECK(bool(s) && s->distance.size() == s->cell_count());
    float c = s->distance[s->index(10, 10, 10)];
    CHECK(c < 0.0f && std::fabs(c) > 0.3f && std::fabs(c) < 0.6f);
    CHECK(s->distance[s->index(0, 0, 0)] > 0.0f);
    // :184-189 — invalid input
    CHECK(!voxelize::voxelize(cube, {.resolution = 0}));
    PLC empty;
    CHECK(!voxelize::voxelize(empty, {.resolution = 8}));
    CHECK(voxelize::voxelize(empty, {.resolution = 8}).error().code ==
          MeshErrorCode::InvalidInput);
    // adjoint: one gradient value per cell
    std::vector<float> gs(s->cell_count(), 1.0f);
    auto grad = voxelize::sdf_backward(cube, {.resolution = 16, .pad = 2}, gs);
    CHECK(bool(grad) && grad->size() == cube.points.size());
    // test_foundation.cpp:14-31 — orientation and insphere signs
    Point3 pa{0, 0, 0}, pb{1, 0, 0}, pc{0, 1, 0}, pd{0, 0, 1};
    double o = robust::orient3d(pa, pb, pc, pd);
    CHECK(o != 0.0 && (o > 0) != (robust::orient3d(pa, pb, pd, pc) > 0));
    CHECK((robust::insphere(pa, pb, pc, pd, {0.25, 0.25, 0.25}) > 0) !=
          (robust::insphere(pa, pb, pc, pd, {100, 100, 100}) > 0));
    // robust.cpp:44-50 returns the determinant's value. For this tetrahedron (circumcentre
    // (1/2,1/2,1/2), r^2 = 3/4) insphere = orient3d(a,b,c,d) * (r^2 - |e - centre|^2),
    // and every operation of the naive formula is exact on these dyadic inputs.
    CHECK(o == -1.0);
    CHECK(robust::insphere(pa, pb, pc, pd, {0.25, 0.25, 0.25}) == -0.5625);
    CHECK(robust::insphere(pa, pb, pc, pd, {0.5, 0.5, 0.5}) == -0.75);
    CHECK(robust::insphere(pa, pb, pc, pd, {100, 100, 100}) == 29700.0);
    CHECK(robust::insphere(pa, pb, pc, pd, {1, 1, 0}) == 0.0);  // on the sphere
    // test_foundation.cpp:94-106 — batch equals per-item
    std::vector<Point3> pts{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, -1}};
    std::vector<Tetrahedron> cands{{0, 1, 2, 3}, {0, 1, 2, 4}};
    auto batch = robust::orient3d_batch(pts, cands);
    for (std::size_t i = 0; i < cands.size(); ++i)
        CHECK(batch[i] == robust::orient3d(pts[cands[i][0]], pts[cands[i][1]],
                                           pts[cands[i][2]], pts[cands[i][3]]));
    // test_backend.cpp:12-34 — signs equal the per-candidate orient3d sign
    unsigned sd = 5;
    auto nx = [&] { sd = sd * 1664525u + 1013904223u; return double((sd >> 8) / double(1u << 24)); };
    std::vector<Point3> cloud;
    for (int i = 0; i < 30; ++i) cloud.push_back({nx(), nx(), nx()});
    std::vector<Tetrahedron> q;
    for (int i = 0; i + 3 < 30; ++i) q.push_back({i, i + 1, i + 2, i + 3});
    auto signs = robust::orient3d_signs(cloud, q);
    for (std::size_t i = 0; i < q.size(); ++i) {
        double v = robust::orient3d(cloud[q[i][0]], cloud[q[i][1]], cloud[q[i][2]], cloud[q[i][3]]);
        CHECK(signs[i] == (v > 0 ? 1 : (v < 0 ? -1 : 0)));
    }
    CHECK(backend::last_backend() == backend::Backend::Hip);
    CHECK(backend::name(backend::Backend::Hip) == "hip");
    // tests/io/test_stl.cpp:35-109 — ascii round trip merges the 4 corners, one binary
    // triangle, truncated vertex record
    {
        const char* path = "/tmp/cmg_hip_cpp_api.stl";
        FILE* f = std::fopen(path, "w");
        const double P[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        const int F[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
        std::fprintf(f, "solid cmg\n");
        for (auto& t : F) {
            std::fprintf(f, "  facet normal 0 0 1\n    outer loop\n");
            for (int v : t) std::fprintf(f, "      vertex %g %g %g\n", P[v][0], P[v][1], P[v][2]);
            std::fprintf(f, "    endloop\n  endfacet\n");
        }
        std::fprintf(f, "endsolid cmg\n");
        std::fclose(f);
        auto r = io::read_stl(path);
        CHECK(bool(r) && r->facets.size() == 4 && r->points.size() == 4);
        auto r2 = io::read_plc(path);
        CHECK(bool(r2) && r2->points.size() == 4);
        const char* bpath = "/tmp/cmg_hip_cpp_api_bin.stl";
        f = std::fopen(bpath, "wb");
        unsigned char buf[134] = {0};
        const uint32_t count = 1;
        std::memcpy(buf + 80, &count, 4);
        const float coords[12] = {0, 0, 0, 0, 0, 0, 2, 0, 0, 0, 3, 0};
        std::memcpy(buf + 84, coords, sizeof coords);
        std::fwrite(buf, 1, sizeof buf, f);
        std::fclose(f);
        auto rb = io::read_stl(bpath);
        CHECK(bool(rb) && rb->facets.size() == 1 && rb->points.size() == 3);
        f = std::fopen(path, "w");
        std::fprintf(f, "solid bad\n  facet normal 0 0 1\n    outer loop\n      vertex 0 0\n");
        std::fclose(f);
        auto bad = io::read_stl(path);
        CHECK(!bad && bad.error().code == MeshErrorCode::InvalidInput);
    }
    // tests/io/test_obj.cpp:58-77 — v/vt/vn corners use the vertex index; missing file
    {
        const char* path = "/tmp/cmg_hip_cpp_api.obj";
        FILE* f = std::fopen(path, "w");
        std::fprintf(f, "# comment line\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nv 1 1 0\n"
                        "v 1 0 1\nv 0 1 1\nvt 0 0\nvn 0 0 1\nf 1/2/3 4/5/6 7/8/9\n");
        std::fclose(f);
        auto r = io::read_obj(path);
        CHECK(bool(r) && r->facets.size() == 1);
        if (r) {
            const auto& v = r->facets[0].polygons[0].vertices;
            CHECK(v.size() == 3 && v[0] == 0 && v[1] == 3 && v[2] == 6);
        }
        CHECK(!io::read_obj("/tmp/cmg_hip_cpp_api_missing.obj"));
    }
    // tests/test_metrics.cpp:79-101 — a single regular-ish tet
    {
        Mesh m;
        m.points = {{0, 0, 0}, {1, 0, 0}, {0.5, 0.866, 0}, {0.5, 0.289, 0.816}};
        m.tetrahedra = {Tetrahedron{0, 1, 2, 3}};
        auto q = quality::report(m, 10);
        CHECK(q.num_tets == 1);
        CHECK(q.min_dihedral >= 60.0 && q.min_dihedral <= 75.0);
        CHECK(q.min_radius_edge >= 0.5 && q.min_radius_edge <= 0.75);
        int hist = 0;
        for (int h : q.dihedral_histogram) hist += h;
        CHECK(hist == 6 && q.worst_tets.size() == 1 && q.worst_tets[0] == 0);
    }
    // dispatch.cpp:39-63: below an op's threshold the CPU serves it, and last_backend
    // records the decision
    CHECK(!backend::should_offload(backend::Op::OrientFilter, 10));
    CHECK(backend::last_backend() == backend::Backend::Cpu);
    CHECK(backend::should_offload(backend::Op::OrientFilter, std::size_t(1) << 16));
    CHECK(backend::last_backend() == backend::Backend::Hip);
    CHECK(!backend::should_offload(backend::Op::QualityScan, (std::size_t(1) << 17) - 1));
    CHECK(backend::last_backend() == backend::Backend::Cpu);
    CHECK(backend::should_offload(backend::Op::PointLocation, std::size_t(1) << 20));
    std::printf("C++ API: %s (%d failures)\n", failures ? "FAIL" : "OK", failures);
    return failures ? 1 : 0;
}
