// Host-side check of the engine's fixed-point drift forms (common.h, compiled for the
// host): fx_drift_p2, the 32-bit form the gather uses on power-of-two grids, against
// fx_drift, the 64-bit form, on random and boundary inputs.  Prints one line per grid
// size: "ng cases mismatches".
#include "common.h"
#include <cstdio>
#include <cstdint>
#include <cmath>

static uint64_t state = 0x9E3779B97F4A7C15ull;
static uint64_t next() { // splitmix64
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static float uniform(float lo, float hi) { return lo + (hi - lo) * (float)((next() >> 40) * (1.0 / 16777216.0)); }

int main() {
    const int sizes[] = {16, 32, 64, 128, 256, 512, 1024};
    for (int ng : sizes) {
        const unsigned span_mask = ((unsigned)ng << HGPM_FRAC_BITS) - 1u;
        const uint64_t span = (uint64_t)ng << HGPM_FRAC_BITS;
        long cases = 0, bad = 0;
        auto check = [&](unsigned X, float v, float pf) {
            cases++;
            if (fx_drift(X, v, pf, ng) != fx_drift_p2(X, v, pf, ng, span_mask)) bad++;
        };
        const float f = (float)ng;
        const float special[] = {0.0f, 1e-30f, -1e-30f, 0.5f, -0.5f, 1.0f, -1.0f, f / 2, -f / 2, f, -f, 2 * f,
                                 -3 * f, std::nextafter(f, 0.0f), -std::nextafter(f, 0.0f), 511.99997f, 512.0f,
                                 -512.0f, -512.00006f, 1023.99994f, std::ldexp(1.0f, -23), std::ldexp(0.5f, -22)};
        const unsigned xs[] = {0u, 1u, (unsigned)(span - 1), (unsigned)(span / 2), (unsigned)(span / 2 - 1),
                               (unsigned)(1u << HGPM_FRAC_BITS), (unsigned)(span - (1u << HGPM_FRAC_BITS))};
        for (unsigned X : xs)
            for (float v : special)
                for (float pf : {1.0f, 0.75f, -0.4f, 3.0f}) check(X, v, pf);
        for (int i = 0; i < 2000000; i++) {
            const unsigned X = (unsigned)(next() % span);
            const int kind = i % 4;
            const float v = kind == 0 ? uniform(-3.0f, 3.0f) : kind == 1 ? uniform(-4.0f * f, 4.0f * f)
                          : kind == 2 ? uniform(-0.01f, 0.01f) : uniform(-f, f);
            const float pf = uniform(-2.0f, 2.0f);
            check(X, v, pf);
        }
        std::printf("%d %ld %ld\n", ng, cases, bad);
    }
    return 0;
}
