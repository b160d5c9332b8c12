// cgs_lambda.hpp -- closed forms of the binary32 lambda sequences of the
// line-traversal costs, so that a lane can start at step j of a row without
// walking steps 0..j-1, and a binary32 form of calcMIAlt's binary64 rescale.
//
//   unit step   lambda = lambda + 1.0f     (src/gpu/cudaErrorFunctions.cuh:270, 346)
//   norm step   lambda = lambda + sc, sc = 1 / (float)(W_f - 1)
//                                          (src/gpu/cudaErrorFunction_mi.cuh:177)
//
// Both give exactly the values of the sequential loop; the host build (g++,
// -ffp-contract=off) is checked against that loop by
// tests/test_lambda_steps.py.  Within a binade [2^e, 2^(e+1)) with ulp U,
// x + s for x on the U grid rounds to x + delta with one delta per binade
// (delta = s rounded to the U grid), so runs of steps inside a binade are
// taken at once in integer units of U, and the step that leaves the binade
// is taken as the literal rounded add.
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define CGS_HD __host__ __device__
#else
#define CGS_HD
#endif

namespace cgs {

CGS_HD inline uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

CGS_HD inline float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// lambda after r steps of lambda = lambda + 1.0f from x (x >= 0 or -0,
// x < 2^23).  Every +1 inside a binade of x >= 1 is exact; below 1 every step
// leaves the binade.
CGS_HD inline float advance_unit(float x, int r) {
    while (r > 0) {
        if (!(x >= 1.0f)) {
            x = x + 1.0f;
            r -= 1;
            continue;
        }
        const float B = u2f((f2u(x) & 0x7F800000u) + 0x00800000u);   // top of x's binade
        const float t = __builtin_ceilf(B - x) - 1.0f;                // steps staying below B; B - x exact
        if ((float) r <= t) {
            x = x + (float) r;
            r = 0;
        } else {
            x = x + t;
            x = x + 1.0f;
            r -= (int) t + 1;
        }
    }
    return x;
}

// True when 64 steps of lambda = lambda + 1.0f from x0 are the exact x0 + j
// (j = 0..63): x0 + 63 stays inside x0's binade [1, 2^23), so every add is
// exact; or x0 is a non-negative integer (sign bit clear, so not -0) below
// 2^24 - 63, so every partial sum is an integer below 2^24.
CGS_HD inline bool unit_pass_exact(float x0) {
    return (x0 >= 1.0f && x0 + 63.0f < u2f((f2u(x0) & 0x7F800000u) + 0x00800000u)) ||
           (f2u(x0) < 0x4B7FFFC1u && x0 == __builtin_truncf(x0));
}

// Step j (0..63) of lambda = lambda + 1.0f from x0 when 32 <= x0 < 2^23.
// Then x0 + 63 < 2B for the top B of x0's binade, so the pass leaves the
// binade at most once: steps j <= t (t = ceil(B - x0) - 1; B - x0 is exact)
// are the exact x0 + j, step t + 1 is the one rounded add y = (x0 + t) + 1,
// and later steps are the exact y + (j - t - 1) inside [B, 2B), whose ulp is
// at most 1.  t and y are the same for every lane.
CGS_HD inline float unit_pass_cross(float x0, int j) {
    const float B = u2f((f2u(x0) & 0x7F800000u) + 0x00800000u);
    const float t = __builtin_ceilf(B - x0) - 1.0f;
    const float y = (x0 + t) + 1.0f;
    const float fj = (float) j;
    return fj <= t ? x0 + fj : y + (fj - (t + 1.0f));
}

// calcMIAlt scales a normalised coordinate back to pixels through binary64,
// (float) (((double) x + 0.5) * W).  For W = 2^k that is (x + 0.5f) * W in
// binary32: the scaling by 2^k is exact and commutes with both roundings
// (binary64 cannot overflow; binary32 overflows to the same infinity), and
// RN32(RN64(x + 0.5)) = RN32(x + 0.5), because x + 0.5 is exact in binary64
// unless |x| < 2^-31 (both give 0.5, far from a binary32 tie) or
// |x| >= 2^52 (both give x).
CGS_HD inline float rescale_pow2(float x, float W) { return (x + 0.5f) * W; }

// A run of the lambda = lambda + s sequence (x >= 0 or -0, s > 0, x and s
// normal) that a wavefront takes in one pass: returns n in [1, 64] such that
// step j < n of the run is (float) (a + j * d) * U (step 0 is x itself), all
// inside x's binade with one increment d * U; the step after the run is the
// literal add of s to step n - 1.  When s is an odd multiple of U / 2 every
// sum is a tie and ties-to-even picks the even one of the two neighbouring
// grid points: from an even a the increment is the even one of
// floor(s / U) and floor(s / U) + 1 at every step, so the run goes on; from an
// odd a the first step is taken alone.  A run of 1 also covers x == 0 and a
// step that leaves the binade.
CGS_HD inline int step_run(float x, float s, int32_t *a, int32_t *d, float *U) {
    *a = 0;
    *d = 0;
    *U = 1.0f;
    if (x == 0.0f) return 1;
    const uint32_t bits = f2u(x);
    const uint32_t expo = bits & 0x7F800000u;
    const float u = u2f(expo - (23u << 23));
    const float B = u2f(expo + 0x00800000u);
    const float y = x + s;
    const float delta = y - x;
    if (!(y < B) || delta == 0.0f) return 1;
    const float err = s - delta;                                   // exact, |err| <= U / 2
    const int32_t am = (int32_t) ((bits & 0x007FFFFFu) | 0x00800000u);  // x / U
    if ((err == 0.5f * u || err == -0.5f * u) && (am & 1)) return 1;
    *a = am;
    *d = (int32_t) (delta / u);
    *U = u;
    const int32_t room = (0x00FFFFFF - *a) / *d;
    return room + 1 < 64 ? room + 1 : 64;
}

}  // namespace cgs
