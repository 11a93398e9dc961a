// meshio.cpp — the STL and OBJ surface readers: the reference's grammar and messages
// (src/io/stl.cpp, src/io/obj.cpp, include/cmg/io/detail.hpp), with the STL vertex merge
// and the binary STL decode on the GPU (weld.hip). Text parsing is host work: the
// numbers go through std::stod / std::stol exactly as the reference parses them.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "voxelize_internal.hpp"
#include "weld_internal.hpp"

namespace cmg::hip {
namespace {

// detail.hpp:24-38: the next line with something left after removing a '#' comment
bool data_line(std::istream& in, std::string& out) {
    std::string line;
    while (std::getline(in, line)) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line.erase(h);
        for (char c : line)
            if (!std::isspace(static_cast<unsigned char>(c))) {
                out = line;
                return true;
            }
    }
    return false;
}

// detail.hpp:41-54: tokens separated by whitespace or commas
void split(const std::string& line, std::vector<std::string>& tok) {
    tok.clear();
    std::string cur;
    for (char c : line) {
        if (std::isspace(static_cast<unsigned char>(c)) || c == ',') {
            if (!cur.empty()) tok.push_back(cur), cur.clear();
        } else {
            cur.push_back(c);
        }
    }
    if (!cur.empty()) tok.push_back(cur);
}

// detail.hpp:57-66: a number must use its whole token; range errors fail
bool real_of(const std::string& s, double& v) {
    try {
        size_t n = 0;
        v = std::stod(s, &n);
        return n == s.size();
    } catch (...) {
        return false;
    }
}
bool long_of(const std::string& s, long& v) {
    try {
        size_t n = 0;
        v = std::stol(s, &n);
        return n == s.size();
    } catch (...) {
        return false;
    }
}

[[noreturn]] void bad(const std::string& msg) { throw InvalidInput(msg); }

// Merges a host vertex soup on the GPU into f (stl.cpp:48-61: one triangle facet per
// three vertices, marker 0).
void weld_into(const std::vector<double>& soup, SurfaceFile& f) {
    const long long nv = (long long)soup.size() / 3;
    if (nv == 0 || nv % 3 != 0) bad(".stl: vertex count is not a positive multiple of 3");
    hipStream_t st = nullptr;
    DeviceArena A(st);
    double* d_soup = A.alloc<double>(soup.size());
    double* d_xyz = A.alloc<double>(soup.size());
    int* d_idx = A.alloc<int>((size_t)nv);
    CMG_HIP(hipMemcpyAsync(d_soup, soup.data(), sizeof(double) * soup.size(),
                           hipMemcpyHostToDevice, st));
    const long long np = weld_vertices(d_soup, nv, d_xyz, d_idx, st);
    std::vector<int> idx((size_t)nv);
    f.xyz.resize(3 * (size_t)np);
    CMG_HIP(hipMemcpyAsync(f.xyz.data(), d_xyz, sizeof(double) * f.xyz.size(),
                           hipMemcpyDeviceToHost, st));
    CMG_HIP(hipMemcpyAsync(idx.data(), d_idx, sizeof(int) * idx.size(), hipMemcpyDeviceToHost,
                           st));
    CMG_HIP(hipStreamSynchronize(st));
    f.polys.resize((size_t)nv / 3);
    for (size_t t = 0; t < f.polys.size(); ++t)
        f.polys[t] = {idx[3 * t], idx[3 * t + 1], idx[3 * t + 2]};
}

}  // namespace

SurfaceFile read_stl_file(const std::string& path) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) bad("cannot open file for reading: " + path);
    const std::streamoff sz = in.tellg();
    if (sz < 0) bad("cannot open file for reading: " + path);
    std::vector<char> buf((size_t)sz);
    in.seekg(0);
    if (sz > 0) in.read(buf.data(), sz);
    SurfaceFile f;
    if (buf.size() >= 84) {
        uint32_t n = 0;
        std::memcpy(&n, buf.data() + 80, 4);
        if (buf.size() == 84ull + 50ull * n) {
            // binary: decode and merge on the GPU
            if (n == 0) bad(".stl: vertex count is not a positive multiple of 3");
            hipStream_t st = nullptr;
            DeviceArena A(st);
            void* d_bytes = A.alloc<char>(buf.size());
            double* d_xyz = A.alloc<double>(9 * (size_t)n);
            int* d_tris = A.alloc<int>(3 * (size_t)n);
            CMG_HIP(hipMemcpyAsync(d_bytes, buf.data(), buf.size(), hipMemcpyHostToDevice, st));
            const long long np = stl_decode_weld(d_bytes, buf.size(), d_xyz, d_tris, st);
            std::vector<int> tris(3 * (size_t)n);
            f.xyz.resize(3 * (size_t)np);
            CMG_HIP(hipMemcpyAsync(f.xyz.data(), d_xyz, sizeof(double) * f.xyz.size(),
                                   hipMemcpyDeviceToHost, st));
            CMG_HIP(hipMemcpyAsync(tris.data(), d_tris, sizeof(int) * tris.size(),
                                   hipMemcpyDeviceToHost, st));
            CMG_HIP(hipStreamSynchronize(st));
            f.polys.resize(n);
            for (size_t t = 0; t < n; ++t)
                f.polys[t] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
            return f;
        }
    }
    // ASCII (stl.cpp:115-133)
    std::ifstream txt(path);
    if (!txt) bad("cannot open file for reading: " + path);
    std::vector<double> soup;
    std::string line;
    std::vector<std::string> tok;
    while (data_line(txt, line)) {
        split(line, tok);
        if (tok.empty()) continue;
        if (tok[0] == "vertex" || tok[0] == "VERTEX") {
            if (tok.size() < 4) bad(".stl: short vertex record");
            double x, y, z;
            if (!real_of(tok[1], x) || !real_of(tok[2], y) || !real_of(tok[3], z))
                bad(".stl: non-numeric vertex coordinate");
            soup.push_back(x);
            soup.push_back(y);
            soup.push_back(z);
        }
    }
    weld_into(soup, f);
    return f;
}

SurfaceFile read_obj_file(const std::string& path) {
    std::ifstream in(path);
    if (!in) bad("cannot open file for reading: " + path);
    SurfaceFile f;
    std::string line;
    std::vector<std::string> t;
    while (data_line(in, line)) {
        split(line, t);
        if (t.empty()) continue;
        if (t[0] == "v") {
            if (t.size() < 4) bad(".obj: short vertex record");
            double x, y, z;
            if (!real_of(t[1], x) || !real_of(t[2], y) || !real_of(t[3], z))
                bad(".obj: non-numeric vertex coordinate");
            f.xyz.push_back(x);
            f.xyz.push_back(y);
            f.xyz.push_back(z);
        } else if (t[0] == "f") {
            if (t.size() < 4) bad(".obj: face needs at least 3 corners");
            std::vector<int> poly;
            poly.reserve(t.size() - 1);
            for (size_t k = 1; k < t.size(); ++k) {
                long v = 0;
                if (!long_of(t[k].substr(0, t[k].find('/')), v))
                    bad(".obj: malformed face corner: " + t[k]);
                poly.push_back(static_cast<int>(v - 1));  // 1-based to 0-based
            }
            f.polys.push_back(std::move(poly));
        }
    }
    return f;
}

}  // namespace cmg::hip
