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
// tests/test_lambda_steps.py.  Unit steps: see unit_pass_direct.  Norm
// steps: within a binade [2^e, 2^(e+1)) with ulp U, x + s for x on the U
// grid rounds to x + delta with one delta per binade (delta = s rounded to
// the U grid), so runs of steps inside a binade are taken at once in integer
// units of U, and the step that leaves the binade is taken as the literal
// rounded add.
#pragma once

#include <math.h>
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

// Unit steps.  From x >= 1 a step that leaves a binade rounds the sum to a
// grid twice as coarse, dropping one low bit by ties-to-even.  Below 2^23
// every grid has spacing 1/2 or finer, so adding an integer (an even
// multiple of the grid) keeps a tie a tie and its even neighbour even: the
// rounding commutes with it.  Dropping one bit and then the next by
// ties-to-even gives what dropping both at once gives (the four values of
// the two bits), so below 2^23 a run of +1 steps from x that crosses at most
// two binade boundaries ends at the one rounded sum x + k.  Three can differ
// (...0|101 rounds down bit by bit and up at once), and so can a run into
// [2^23, 2^24), where +1 is an odd multiple of the grid:
// 8388607.5 -> 8388608 (a tie, to even) -> 8388609, but 8388607.5 + 2 ->
// 8388610.

// True when x0 + j, one rounded add, is step j (0..63) of lambda = lambda +
// 1.0f from x0: 16 <= x0 and x0 + 63 < 2^23 (the pass crosses at most two
// binade boundaries, all below 2^23), or x0 is a non-negative integer (sign
// bit clear, so not -0) below 2^24 - 63 (every sum is an exact integer).
// Both tests read the bit pattern, whose order is the value's order for
// non-negative floats, so that a wave-uniform x0 is tested with scalar
// instructions.
CGS_HD inline bool unit_int_start(float x0) {   // x0 a non-negative integer below 2^24 - 63
    const uint32_t b = f2u(x0);
    if (b == 0u) return true;                                  // +0
    if (b < 0x3F800000u || b >= 0x4B7FFFC1u) return false;     // (0, 1), >= 2^24 - 63, or sign bit set
    return (b & ((1u << (150u - (b >> 23))) - 1u)) == 0u;      // no fraction bits below the point
}
CGS_HD inline bool unit_pass_direct(float x0) {
    const uint32_t b = f2u(x0);
    return (b >= 0x41800000u && b < 0x4AFFFF82u) || unit_int_start(x0);   // 16 <= x0 < 2^23 - 63
}

// The same two tests in value form (vector compares).  A kernel whose scalar
// unit is busy runs faster with these; the host check holds both forms equal.
CGS_HD inline bool unit_int_start_v(float x0) { return f2u(x0) < 0x4B7FFFC1u && x0 == __builtin_truncf(x0); }
CGS_HD inline bool unit_pass_direct_v(float x0) { return (x0 >= 16.0f && x0 < 8388545.0f) || unit_int_start_v(x0); }

// Step j (0..63) of lambda = lambda + 1.0f from any x0 >= 0 or -0 below
// 2^23, the pass cut into the same segments for every lane.  A segment
// starts at lane jb with value base and holds the one rounded base + (j - jb)
// for t steps: up to t = ceil(4B - base) - 1, B the top of base's binade, so
// that at most two boundaries are crossed inside it, while 4B <= 2^23; from
// the binade below 2^23 up, to the binade's end (t = ceil(B - base) - 1; in
// [2^23, 2^24) every value is an integer and t runs past the pass); below 1
// a segment is base alone, because that first add drops many bits at once.
// The next segment starts at the literal add (base + t) + 1.  The segments
// depend on x0 only (at most 3: below 1, from [1, 2) to 8, from [8, 16) to
// 64; or up to 2^23 and past it), each lane keeps the value of the
// last segment that starts at or before it, and lane 0 keeps x0 itself
// (-0 + 0 is +0).
CGS_HD inline float unit_pass_chain(float x0, int j) {
    float base = x0, lam = x0;
    int jb = 0;
    while (true) {
        float t = 0.0f;
        if (base >= 1.0f) {
            const float B = u2f((f2u(base) & 0x7F800000u) + 0x00800000u);
            t = __builtin_ceilf(B - base) - 1.0f;   // B - base is exact
            if (B <= 2097152.0f) t += 3.0f * B;     // 4B <= 2^23
        }
        if (j >= jb) lam = base + (float) (j - jb);
        const float next = (float) jb + t + 1.0f;
        if (next > 63.0f) break;
        jb = (int) next;
        base = (base + t) + 1.0f;
    }
    return j == 0 ? x0 : lam;
}

// calcMIAlt scales a normalised coordinate back to pixels through binary64,
// (float) (((double) x + 0.5) * W).  For W = 2^k that is one binary32 fma,
// x * W + W / 2 rounded once: the scaling by 2^k is exact and commutes with
// the roundings (binary64 cannot overflow; binary32 overflows to the same
// infinity), so the binary64 form is RN32(RN64(x + 0.5)) * W and the fma
// RN32(x + 0.5) * W, and RN32(RN64(x + 0.5)) = RN32(x + 0.5) because x + 0.5
// is exact in binary64 unless |x| < 2^-31 (both give 0.5, far from a
// binary32 tie) or |x| >= 2^52 (both give x).
CGS_HD inline float rescale_pow2(float x, float W) { return fmaf(x, W, 0.5f * W); }

// A segment of the lambda = lambda + s sequence (s > 0 normal) from x: the
// steps k >= k0 with a + (k - k0) d <= lim (unsigned) are (float) (a + (k -
// k0) d) * U.  Inside x's binade, with ulp U, every add moves lambda by the
// same delta = s rounded to the U grid, so the steps up to the binade's top
// are a + j d in units of U (a = x / U, d = delta / U, lim = 2^24 - 1).  When
// s is an odd multiple of U / 2 every sum is a tie, which ties-to-even sends
// to the even neighbour: from an even a the increment is the even one of
// floor(s / U) and floor(s / U) + 1 at every step, so the segment goes on;
// from an odd a the first step is taken alone.  A step taken alone (x == 0, an
// add that leaves the binade or changes nothing, a tie from an odd a, a
// subnormal or non-finite x) is the one-step segment a = d = lim = 1, U = x,
// whose step is (float) 1 * x = x itself.  The step after a segment is the
// literal add of s to its last step.
struct LamSeg {
    int32_t k0;
    uint32_t a, d, lim;
    float U;
};

CGS_HD inline LamSeg lam_segment(float x, float s, int32_t k0) {
    LamSeg g{k0, 1u, 1u, 1u, x};
    if (x == 0.0f) return g;
    const uint32_t bits = f2u(x);
    const uint32_t expo = bits & 0x7F800000u;
    const float B = u2f(expo + 0x00800000u);
    const float y = x + s;
    const float delta = y - x;
    if (!(y < B) || delta == 0.0f) return g;
    const float u = u2f(expo - (23u << 23));
    const float err = s - delta;                                        // exact, |err| <= U / 2
    const uint32_t am = (bits & 0x007FFFFFu) | 0x00800000u;             // x / U
    if ((err == 0.5f * u || err == -0.5f * u) && (am & 1u)) return g;
    g.a = am;
    g.d = (uint32_t) (int32_t) __builtin_ldexpf(delta, 150 - (int32_t) (expo >> 23));   // delta / U, exact
    g.lim = 0x00FFFFFFu;
    g.U = u;
    return g;
}

// s + d added n times in binary32, the loop `area += d` of the
// line-traversal functors' moved area (src/gpu/cudaErrorFunctions.cuh:289,
// 365; src/gpu/cudaErrorFunctionsStreams.cuh:245;
// src/gpu/cudaErrorFunction_miStreams.cuh:110), for s >= 0 and d >= 0 (or
// NaN).  Inside a binade [2^e, 2^(e+1)) with grid U an add moves s by d
// rounded to the U grid, the same increment at every step (a tie goes to the
// even neighbour: from an even multiple of U that is the same increment
// again, so only a tie from an odd multiple is taken alone), so a run of adds
// that stays inside the binade is one integer multiply-add in units of U.
// The add that leaves the binade, the first add from 0 and adds to sums
// below 2^-103 are taken literally; an add that leaves s unchanged leaves it
// unchanged for good.  Returns exactly what the loop returns.
CGS_HD inline float add_repeat(float s, float d, int64_t n) {
    if (n <= 0) return s;
    if (!(s < __builtin_inff()) || !(d < __builtin_inff())) return s + d;   // inf and NaN stay
    while (n > 0) {
        const float y = s + d;
        if (y == s) return s;
        const uint32_t bits = f2u(s);
        const uint32_t expo = bits & 0x7F800000u;
        const float B = u2f(expo + 0x00800000u);
        if (expo <= (24u << 23) || !(y < B)) {   // from 0 or tiny sums, or leaving the binade
            s = y;
            --n;
            continue;
        }
        const float u = u2f(expo - (23u << 23));
        const float delta = y - s;                 // exact: both on the U grid
        const float err = d - delta;               // exact, |err| <= U / 2
        const int32_t a = (int32_t) ((bits & 0x007FFFFFu) | 0x00800000u);   // s / U
        if ((err == 0.5f * u || err == -0.5f * u) && (a & 1)) {
            s = y;
            --n;
            continue;
        }
        const int32_t dd = (int32_t) __builtin_ldexpf(delta, 150 - (int32_t) (expo >> 23));   // delta / U
        const int32_t room = (0x00FFFFFF - a) / dd;   // >= 1: a + dd < 2^24
        const int64_t k = n < (int64_t) room ? n : (int64_t) room;
        s = (float) (a + (int32_t) k * dd) * u;
        n -= k;
    }
    return s;
}

}  // namespace cgs
