// Host check of fh::div_model_const (model.h) against IEEE single division,
// for the model constants the flight divides by. Every float pattern of
// |x| in [2^-104, 2^100] with both signs, taking one in `stride` (stride 1:
// all of them, about 3.4e9 per constant), plus the values that take the
// division instead (zeros, subnormals, the range's edges, huge values,
// infinities, NaN). Prints "<mismatches> <checked>" per constant; exits 1 on
// any mismatch (NaN results compare as NaN). With a second argument "plain"
// it checks the product x * rc alone instead, which must fail: the check can
// tell a missing correction.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>

#include "../murmuration-kernels_amd/csrc/model.h"

static float from_bits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
static uint32_t bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

static bool same(float a, float b) { return (std::isnan(a) && std::isnan(b)) || bits(a) == bits(b); }

int main(int argc, char **argv) {
    const uint32_t stride = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 997u;
    const bool plain = argc > 2 && std::strcmp(argv[2], "plain") == 0;
    auto div = [plain](float x, float c, float rc) { return plain ? x * rc : fh::div_model_const(x, c, rc); };
    const float consts[4] = {fh::MASS, fh::G, fh::TAU_ROLL, fh::V0};
    const char *names[4] = {"MASS", "G", "TAU_ROLL", "V0"};
    const uint32_t lo = bits(0x1p-104f), hi = bits(0x1p100f);
    const float edges[] = {0.0f, 1e-45f, 1e-40f, 1.17549435e-38f, 0x1p-105f, std::nextafter(0x1p-104f, 0.0f),
                           0x1p-104f, 0x1p100f, std::nextafter(0x1p100f, INFINITY), 1e35f, 3.4028235e38f,
                           INFINITY, NAN};
    int failed = 0;
    for (int k = 0; k < 4; ++k) {
        volatile float cv = consts[k];  // the division as the oracle does it
        const float c = consts[k], rc = 1.0f / c;
        unsigned long long bad = 0, n = 0;
        for (uint64_t u = lo; u <= hi; u += stride)
            for (uint32_t sign = 0; sign < 2; ++sign) {
                const float x = from_bits((uint32_t)u | (sign << 31));
                if (!same(div(x, c, rc), x / cv)) ++bad;
                ++n;
            }
        for (float e : edges)
            for (float x : {e, -e}) {
                if (!same(div(x, c, rc), x / cv)) ++bad;
                ++n;
            }
        std::printf("%s %llu %llu\n", names[k], bad, n);
        if (bad) failed = 1;
    }
    return failed;
}
