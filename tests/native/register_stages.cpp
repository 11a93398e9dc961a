// Host run of the register-only exact stages (the same source the GPU passes compile),
// for checking them against the oracle without a GPU.
// Usage: register_stages insphere|orient3d <in: n tuples of float64> <out: n x 8 int8>
// insphere tuples are pa pb pc pd pe (15 doubles); per tuple, for each stage in turn
// (fixed point with 62-bit differences, double-word bound, fixed point with 126-bit
// differences, fixed point with 31-bit differences): decided (0/1), sign. orient3d tuples
// are pa pb pc pd (12 doubles); the stages are fixed point with 31-bit differences, the
// double-word bound (the streaming kernel's ring drain, in that order) and fixed point
// with 62- and 126-bit differences (the exact pass). Unused slots read 0.
// Modes orient3d-exp and insphere-exp replay the exact pass instead: fixed point with 62-
// and 126-bit differences, then the expansion arithmetic with the first pass's small
// capacities and with the wide re-run's capacities (decided 0 = capacity exceeded).
#include <cstdio>
#include <cstring>
#include <vector>

#include "expansion.hpp"

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s insphere|orient3d in.bin out.bin\n", argv[0]);
        return 2;
    }
    const bool ins = std::strncmp(argv[1], "insphere", 8) == 0;
    const bool exp = std::strstr(argv[1], "-exp") != nullptr;
    const int per = ins ? 15 : 12;
    std::FILE* f = std::fopen(argv[2], "rb");
    if (!f) return 2;
    std::vector<double> v;
    double buf[15];
    while (std::fread(buf, sizeof(double), per, f) == (size_t)per) v.insert(v.end(), buf, buf + per);
    std::fclose(f);
    const size_t n = v.size() / per;
    std::vector<signed char> out(8 * n);
    for (size_t i = 0; i < n; ++i) {
        const double* t = v.data() + per * i;
        int s[4] = {0, 0, 0, 0};
        bool ok[4] = {false, false, false, false};
        if (exp) {
            double top = 0.0;
            if (ins) {
                ok[0] = cmg::hip::insphere_sign_fixed(t, t + 3, t + 6, t + 9, t + 12, s[0]);
                ok[1] = cmg::hip::insphere_sign_fixed_wide(t, t + 3, t + 6, t + 9, t + 12, s[1]);
                ok[2] = cmg::hip::insphere_exact<40, 96>(t, t + 3, t + 6, t + 9, t + 12, top);
                s[2] = top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
                ok[3] = cmg::hip::insphere_exact<200, 1024>(t, t + 3, t + 6, t + 9, t + 12, top);
                s[3] = top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
            } else {
                ok[0] = cmg::hip::orient3d_sign_fixed(t, t + 3, t + 6, t + 9, s[0]);
                ok[1] = cmg::hip::orient3d_sign_fixed_wide(t, t + 3, t + 6, t + 9, s[1]);
                ok[2] = cmg::hip::orient3d_exact<40>(t, t + 3, t + 6, t + 9, top);
                s[2] = top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
                ok[3] = cmg::hip::orient3d_exact<320>(t, t + 3, t + 6, t + 9, top);
                s[3] = top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
            }
        } else if (ins) {
            ok[0] = cmg::hip::insphere_sign_fixed(t, t + 3, t + 6, t + 9, t + 12, s[0]);
            ok[1] = cmg::hip::insphere_sign_dw(t, t + 3, t + 6, t + 9, t + 12, s[1]);
            ok[2] = cmg::hip::insphere_sign_fixed_wide(t, t + 3, t + 6, t + 9, t + 12, s[2]);
            ok[3] = cmg::hip::insphere_sign_fixed31(t, t + 3, t + 6, t + 9, t + 12, s[3]);
        } else {
            ok[0] = cmg::hip::orient3d_sign_fixed31(t, t + 3, t + 6, t + 9, s[0]);
            ok[1] = cmg::hip::orient3d_sign_dw(t, t + 3, t + 6, t + 9, s[1]);
            ok[2] = cmg::hip::orient3d_sign_fixed(t, t + 3, t + 6, t + 9, s[2]);
            ok[3] = cmg::hip::orient3d_sign_fixed_wide(t, t + 3, t + 6, t + 9, s[3]);
        }
        for (int k = 0; k < 4; ++k) {
            out[8 * i + 2 * k] = ok[k] ? 1 : 0;
            out[8 * i + 2 * k + 1] = (signed char)(ok[k] ? s[k] : 0);
        }
    }
    std::FILE* g = std::fopen(argv[3], "wb");
    if (!g) return 2;
    std::fwrite(out.data(), 1, out.size(), g);
    std::fclose(g);
    return 0;
}
