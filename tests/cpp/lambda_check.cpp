// Checks the closed forms of cgs_lambda.hpp against the sequential binary32
// loops they replace.  Prints "<mismatches> <comparisons> <runs of the
// 512-wide rows> <starts whose 64-step pass is exact> <starts of the
// one-crossing closed form> <rescales compared>".
#include <cstdio>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "cgs_lambda.hpp"

static bool same(float a, float b) { return std::memcmp(&a, &b, 4) == 0; }

int main() {
    std::mt19937 rng(12345);
    std::uniform_real_distribution<float> u01(0.0f, 1.0f);
    long bad = 0, n = 0, runs512 = 0;
    // lambda += 1.0f from starts in [0, 65536), including 0, -0, integers and
    // values just below powers of two
    std::vector<float> starts = {0.0f, -0.0f, 1.0f, 0.5f, 0.999999940f, 63.9999962f, 255.5f, 1023.99994f};
    for (int i = 0; i < 4000; ++i) {
        const int e = (int) (rng() % 17);
        starts.push_back(u01(rng) * (float) (1 << e));
    }
    for (float x0 : starts) {
        float seq = x0;
        for (int k = 0; k <= 600; ++k) {
            ++n;
            if (!same(cgs::advance_unit(x0, k), seq)) ++bad;
            seq = seq + 1.0f;
        }
    }
    // unit_pass_exact(x0): a pass of 64 steps is x0 + j exactly.  Starts:
    // integers (0 and the top of the range included), binade interiors and
    // edges, fractions below 1, -0; passes counts the starts it accepts
    long passes = 0;
    std::vector<float> pstarts = {0.0f, -0.0f, 1.0f, 63.0f, 64.0f, 0.5f, 1.5f, 1.25f, 16777152.0f, 16777153.0f,
                                  16777151.0f, 8388607.5f, 8388608.0f, 4194303.5f, 1.0000001f, 2.0000002f};
    for (int i = 0; i < 20000; ++i) {
        const int e = (int) (rng() % 25);
        const float v = u01(rng) * (float) (1 << e);
        pstarts.push_back(v);
        pstarts.push_back(__builtin_floorf(v));
    }
    for (float x0 : pstarts) {
        if (!cgs::unit_pass_exact(x0)) continue;
        ++passes;
        float seq = x0;
        for (int j = 0; j < 64; ++j) {
            ++n;
            if (!same(x0 + (float) j, seq)) ++bad;
            seq = seq + 1.0f;
        }
    }
    // unit_pass_cross(x0, j) for 32 <= x0 < 2^23: starts just below and above
    // powers of two, binade interiors, integers
    long crosses = 0;
    std::vector<float> cstarts = {32.0f, 32.5f, 63.9999962f, 64.0f, 127.3f, 127.99999f, 8388607.5f, 8388544.25f,
                                  4194303.75f, 65535.99f, 200.3f, 1000.7f};
    for (int e = 5; e < 23; ++e)
        for (int k = 1; k <= 70; ++k) {
            const float top = (float) (1 << (e + 1));
            cstarts.push_back(top - (float) k * 0.9f);
            cstarts.push_back(std::nextafter(top - (float) k, 0.0f));
            cstarts.push_back(top - (float) k + 0.5f);
        }
    for (int i = 0; i < 20000; ++i) {
        const int e = 5 + (int) (rng() % 18);
        cstarts.push_back((1.0f + u01(rng)) * (float) (1 << e));
    }
    for (float x0 : cstarts) {
        if (!(x0 >= 32.0f && x0 < 8388608.0f)) continue;
        ++crosses;
        float seq = x0;
        for (int j = 0; j < 64; ++j) {
            ++n;
            if (!same(cgs::unit_pass_cross(x0, j), seq)) ++bad;
            seq = seq + 1.0f;
        }
    }
    // rescale_pow2(x, 2^k) against the binary64 expression, x over every
    // magnitude and sign: zeros, subnormals, around 2^-31 and 2^52, huge
    // values and infinities (NaN left out: only its class matters)
    long rescales = 0;
    std::vector<float> xs = {0.0f, -0.0f, 0.5f, -0.5f, 1e-45f, -1e-45f, 1e-40f, 3.4e38f, -3.4e38f,
                             INFINITY, -INFINITY, 4503599627370496.0f, 9007199254740992.0f, 4.6566e-10f, -4.6566e-10f};
    for (int i = 0; i < 200000; ++i) {
        uint32_t b = (uint32_t) rng();
        if ((b & 0x7F800000u) == 0x7F800000u) continue;   // NaN / inf
        float x;
        std::memcpy(&x, &b, 4);
        xs.push_back(x);
        xs.push_back(u01(rng) - 0.5f);
        xs.push_back((u01(rng) - 0.5f) * 1e-8f);
    }
    for (int k = 0; k <= 14; ++k) {
        const float W = (float) (1 << k);
        for (float x : xs) {
            ++n;
            ++rescales;
            if (!same(cgs::rescale_pow2(x, W), (float) (((double) x + 0.5) * (double) W))) ++bad;
        }
    }
    // lambda += 1/(float)(W - 1) from starts in [0, 1), walked in runs (the
    // kernel's passes); runs counts the passes of the starts at 0
    std::vector<int> widths;
    for (int w = 2; w <= 300; ++w) widths.push_back(w);
    for (int i = 0; i < 300; ++i) widths.push_back(300 + (int) (rng() % 8000));
    for (int w : {256, 257, 384, 385, 511, 512, 513, 640, 1000, 1024, 1025, 4096, 4097, 6001}) widths.push_back(w);
    for (int w : widths) {
        const float s = 1 / ((float) (w - 1));
        std::vector<float> xs = {0.0f, -0.0f, s, 0.5f * s, 0.25f, 0.5f, 0.499999970f};
        for (int i = 0; i < 6; ++i) xs.push_back(u01(rng));
        for (int i = 0; i < 3; ++i) xs.push_back(u01(rng) * 1e-3f);
        for (float x0 : xs) {
            float seq = x0;
            std::vector<float> want;
            for (int k = 0; seq < 1.0f && k < w + 4; ++k) {
                want.push_back(seq);
                seq = seq + s;
            }
            float base = x0;
            size_t k = 0;
            while (k < want.size()) {
                if (x0 == 0.0f && w == 512) ++runs512;
                int32_t a, d;
                float U;
                const int run = cgs::step_run(base, s, &a, &d, &U);
                float last = base;
                for (int j = 0; j < run && k < want.size(); ++j, ++k) {
                    last = j == 0 ? base : (float) (a + j * d) * U;
                    ++n;
                    if (!same(last, want[k])) ++bad;
                }
                base = last + s;
            }
        }
    }
    std::printf("%ld %ld %ld %ld %ld %ld\n", bad, n, runs512, passes, crosses, rescales);
    return 0;
}
