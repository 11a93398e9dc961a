// Host run of the register-only insphere stages (the same source the GPU exact passes
// compile), for checking them against the oracle without a GPU.
// Usage: fixed_insphere <in: n x 15 float64, tuples pa pb pc pd pe> <out: n x 6 int8>
// Per tuple, for each stage in turn (fixed point with 62-bit differences, double-word
// bound, fixed point with 126-bit differences): decided (0/1), sign.
#include <cstdio>
#include <vector>

#include "expansion.hpp"

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    std::FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<double> v;
    double buf[15];
    while (std::fread(buf, sizeof(double), 15, f) == 15) v.insert(v.end(), buf, buf + 15);
    std::fclose(f);
    const size_t n = v.size() / 15;
    std::vector<signed char> out(6 * n);
    for (size_t i = 0; i < n; ++i) {
        const double* t = v.data() + 15 * i;
        int s[3] = {0, 0, 0};
        const bool ok[3] = {
            cmg::hip::insphere_sign_fixed(t, t + 3, t + 6, t + 9, t + 12, s[0]),
            cmg::hip::insphere_sign_dw(t, t + 3, t + 6, t + 9, t + 12, s[1]),
            cmg::hip::insphere_sign_fixed_wide(t, t + 3, t + 6, t + 9, t + 12, s[2])};
        for (int k = 0; k < 3; ++k) {
            out[6 * i + 2 * k] = ok[k] ? 1 : 0;
            out[6 * i + 2 * k + 1] = (signed char)(ok[k] ? s[k] : 0);
        }
    }
    std::FILE* g = std::fopen(argv[2], "wb");
    if (!g) return 2;
    std::fwrite(out.data(), 1, out.size(), g);
    std::fclose(g);
    return 0;
}
