// Checks the closed forms of cgs_lambda.hpp against the sequential binary32
// loops they replace.  Prints "<mismatches> <comparisons> <passes of the
// 512-wide rows from 0> <starts unit_pass_direct accepts> <starts through
// unit_pass_chain> <rescales compared> <areas compared> <calcMIAlt passes
// taken in closed form>".
#include <cstdio>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "cgs_lambda.hpp"

static bool same(float a, float b) { return std::memcmp(&a, &b, 4) == 0; }

// alt_norm_kernel's walk of one calcMIAlt row, lane by lane: passes of 64
// consecutive steps.  A pass whose current segment holds all 64 steps is
// taken in closed form, (lane d + b) U with b the segment's value at the
// pass's first step, tracked in 32 bits as the kernel does; any other pass is
// filled (the step before it taken as (b - d) U, checked against the walk) from
// segments (lam_segment) laid over the lanes in order, lane i of the pass
// taking the last segment that starts at or before step k0 + i and holds it
// (a + t d <= lim, with the 24-bit multiply of the kernel); the next segment
// starts at the literal add after the last covered step.  Unfilled lanes
// hold 2; the valid lanes (lambda < lend = min(le, 1), NaN kept) must form a
// prefix.  Returns the row's steps; *fast counts the closed-form passes.
static std::vector<float> packed_walk(float ls, float le, float s, int w, int *passes, long *segs, long *fast,
                                      long *bad) {
    std::vector<float> out;
    const float lend = le >= 1.0f ? 1.0f : le;
    cgs::LamSeg g = cgs::lam_segment(ls, s, 0);
    float carry = 0.0f;
    int k0 = 0;
    uint32_t b = g.a;   // (k0 - g.k0) d + a, kept in 32 bits as the kernel keeps it
    *passes = 0;
    for (int pass = 0; pass < w / 64 + 3; ++pass) {
        float lam[64];
        for (float &v : lam) v = 2.0f;
        if (g.k0 > k0 || (uint64_t) b + 63u * (uint64_t) g.d >= (1ull << 31)) ++*bad;   // the kernel's invariants
        const bool whole = b + 63u * g.d <= g.lim;
        if (whole) {
            for (int i = 0; i < 64; ++i) lam[i] = (float) ((uint32_t) i * g.d + b) * g.U;
            ++*fast;
        } else {
            // the kernel takes the step before the pass from g: (b - d) U
            if (pass > 0 && !same((float) (b - g.d) * g.U, carry)) ++*bad;
            uint64_t covered = 0;
            for (int fill = 0; fill <= 64; ++fill) {
                uint64_t m = 0;
                for (int i = 0; i < 64; ++i) {
                    const int kl = k0 + i;
                    const uint32_t t = (uint32_t) (kl - g.k0);
                    const uint32_t wv = (uint32_t) ((uint64_t) (t & 0xFFFFFFu) * (uint64_t) (g.d & 0xFFFFFFu)) + g.a;
                    if (kl >= g.k0 && wv <= g.lim) {
                        m |= 1ull << i;
                        lam[i] = (float) wv * g.U;
                    }
                }
                covered |= m;
                if (covered >> 63) break;
                const int last = covered ? 63 - __builtin_clzll(covered) : -1;
                const float x = (last < 0 ? carry : lam[last]) + s;
                if (!(x < lend)) break;
                g = cgs::lam_segment(x, s, k0 + last + 1);
                ++*segs;
            }
        }
        uint64_t m = 0;
        for (int i = 0; i < 64; ++i)
            if (lam[i] < lend) m |= 1ull << i;
        if (m == 0) break;
        ++*passes;
        if (m & (m + 1)) ++*bad;   // not a prefix
        for (int i = 0; i < 64 && (m >> i & 1); ++i) out.push_back(lam[i]);
        if (m != ~0ull) break;
        carry = lam[63];
        k0 += 64;
        b = whole ? b + 64u * g.d : (uint32_t) (k0 - g.k0) * g.d + g.a;
    }
    return out;
}

int main() {
    std::mt19937 rng(12345);
    std::uniform_real_distribution<float> u01(0.0f, 1.0f);
    long bad = 0, n = 0, runs512 = 0;
    // unit steps, 64 per pass: unit_pass_direct(x0) accepts a start only
    // when every x0 + j is the loop's value; unit_pass_chain must match the
    // loop from every start.  Starts: zeros, integers (the top of the range
    // included), fractions below 1, binade interiors and edges from 2^-6 to
    // 2^23, starts just below powers of two (a pass crossing one, two or
    // three boundaries), fractional starts just below 2^22 and 2^23 (where
    // the +1 steps become odd multiples of the grid)
    long direct = 0, chains = 0;
    std::vector<float> starts = {0.0f, -0.0f, 1e-30f, 1e-7f, 0.5f, 0.999999940f, 1.0f, 1.25f, 1.5f, 3.75f, 7.3f,
                                 15.9999990f, 16.0f, 16.5f, 31.7f, 31.9999981f, 32.0f, 32.5f, 63.0f, 64.0f, 100.25f,
                                 127.99999f, 255.5f, 1023.99994f, 8388544.5f, 8388607.5f, 8388608.0f, 4194303.75f,
                                 16777151.0f, 16777152.0f, 16777153.0f, 1.0000001f, 2.0000002f};
    for (int e = 0; e < 23; ++e)
        for (int k = 1; k <= 40; ++k) {
            const float top = (float) (1 << (e + 1));
            starts.push_back(top - (float) k * 0.9f);
            starts.push_back(std::nextafter(top - (float) k, 0.0f));
            starts.push_back(top - (float) k + 0.5f);
        }
    for (int i = 0; i < 2000; ++i) {   // fractional starts within 200 of 2^22 and 2^23
        starts.push_back(8388608.0f - 200.0f * u01(rng));
        starts.push_back(4194304.0f - 200.0f * u01(rng));
    }
    for (int i = 0; i < 40000; ++i) {
        const int e = (int) (rng() % 30) - 6;
        const float v = u01(rng) * std::ldexp(1.0f, e);
        starts.push_back(v);
        starts.push_back(std::floor(v));
    }
    for (float x0 : starts) {
        std::vector<float> seq(64);
        seq[0] = x0;
        for (int j = 1; j < 64; ++j) seq[j] = seq[j - 1] + 1.0f;
        // the bit-pattern predicates and the header's value forms against the
        // definition
        const bool int_start = !std::signbit(x0) && x0 == std::trunc(x0) && x0 < 16777153.0f;
        const bool direct_def = int_start || (x0 >= 16.0f && x0 < 8388545.0f);
        if (cgs::unit_int_start(x0) != int_start || cgs::unit_int_start_v(x0) != int_start) ++bad;
        if (cgs::unit_pass_direct(x0) != direct_def || cgs::unit_pass_direct_v(x0) != direct_def) ++bad;
        if (cgs::unit_pass_direct(x0)) {
            ++direct;
            for (int j = 0; j < 64; ++j, ++n)
                if (!same(x0 + (float) j, seq[j])) ++bad;
        }
        if (x0 < 8388608.0f) {
            ++chains;
            for (int j = 0; j < 64; ++j, ++n)
                if (!same(cgs::unit_pass_chain(x0, j), seq[j])) ++bad;
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
    // lambda += 1/(float)(W - 1) from starts in [0, 1), walked as the calcMIAlt
    // kernel walks a row (packed_walk): passes of 64 steps filled from
    // lam_segment's segments, against the literal loop with the row's end at
    // 1, inside (0, 1), at the start and below it, and NaN; passes512 counts
    // the passes of the 512-wide rows from 0 (+0 and -0) with the end at 1
    std::vector<int> widths;
    for (int w = 2; w <= 300; ++w) widths.push_back(w);
    for (int i = 0; i < 300; ++i) widths.push_back(300 + (int) (rng() % 8000));
    for (int w : {256, 257, 384, 385, 511, 512, 513, 640, 1000, 1024, 1025, 4096, 4097, 6001}) widths.push_back(w);
    long segs = 0, fastpass = 0;
    for (int w : widths) {
        const float s = 1 / ((float) (w - 1));
        std::vector<float> xs = {0.0f, -0.0f, s, 0.5f * s, 0.25f, 0.5f, 0.499999970f, 1e-40f, 0.999999940f};
        for (int i = 0; i < 6; ++i) xs.push_back(u01(rng));
        for (int i = 0; i < 3; ++i) xs.push_back(u01(rng) * 1e-3f);
        for (int i = 0; i < 3; ++i) xs.push_back(u01(rng) * 0.07f);
        for (float x0 : xs) {
            const std::vector<float> les = {1.0f, x0 + 0.5f * u01(rng), x0 + 0.03f * u01(rng), x0, x0 * 0.5f, NAN};
            for (float le : les) {
                std::vector<float> want;
                for (float seq = x0; seq < le && seq < 1.0f && (int) want.size() < w + 4; seq = seq + s)
                    want.push_back(seq);
                int passes = 0;
                const std::vector<float> got = packed_walk(x0, le, s, w, &passes, &segs, &fastpass, &bad);
                if (got.size() != want.size()) ++bad;
                for (size_t k = 0; k < want.size() && k < got.size(); ++k, ++n)
                    if (!same(got[k], want[k])) ++bad;
                if (w == 512 && x0 == 0.0f && le == 1.0f) runs512 += passes;
            }
        }
    }
    // add_repeat(s, d, n) against n literal adds, the moved area's
    // recurrence: increments |v| around 1 (the line lengths of near-identity
    // warps), exact halves and quarters of a grid (ties from odd and even
    // multiples), tiny and huge increments, starts at 0 and along the way,
    // and every intermediate count of each walk
    long areas = 0;
    std::vector<float> ds = {1.0f, 0.5f, 1.5f, 0.75f, 1.0000001f, 0.99999994f, 1e-3f, 1e-8f, 3.0f, 123.456f, 0.0f,
                             1.0f + 0x1p-10f, 1.0f + 0x1p-11f + 0x1p-12f, 0x1p-20f, 0x1p-130f, 1e30f};
    for (int i = 0; i < 300; ++i) ds.push_back(0.7f + 0.8f * u01(rng));
    for (int i = 0; i < 60; ++i) ds.push_back(std::ldexp(u01(rng) + 0.5f, (int) (rng() % 40) - 20));
    for (int i = 0; i < 60; ++i)   // (m + 1/2) * 2^-k: a tie on the grid of sums in [2^(24-k-1), 2^(24-k))
        ds.push_back(std::ldexp((float) (rng() % 64) + 0.5f, -(int) (rng() % 14)));
    for (float d : ds) {
        const std::vector<float> s0s = {0.0f, 1.0f, 3.0f, 1000.5f, 4096.0f, 65535.75f, 16777215.0f, 0x1p-110f,
                                        u01(rng) * 1000.0f};
        for (float s0 : s0s) {
            const int64_t len = d < 1e-6f ? 5000 : 70000;
            float seq = s0;
            int64_t next = 1;   // compare at n = 1, 2, 3, 5, 8, ... and at the end
            for (int64_t k = 1; k <= len; ++k) {
                seq = seq + d;
                if (k == next || k == len) {
                    ++n;
                    ++areas;
                    if (!same(cgs::add_repeat(s0, d, k), seq)) ++bad;
                    next = next + next / 2 + 1;
                }
            }
        }
    }
    // summing per-row partial sums as the stream functors do: the chained
    // closed form equals one long literal loop
    for (int t = 0; t < 200; ++t) {
        float lit = 0.0f, fast = 0.0f;
        for (int row = 0; row < 300; ++row) {
            const float d = 0.8f + 0.5f * u01(rng);
            const int k = (int) (rng() % 400);
            for (int j = 0; j < k; ++j) lit = lit + d;
            fast = cgs::add_repeat(fast, d, k);
        }
        ++n;
        ++areas;
        if (!same(lit, fast)) ++bad;
    }
    std::printf("%ld %ld %ld %ld %ld %ld %ld %ld\n", bad, n, runs512, direct, chains, rescales, areas, fastpass);
    return 0;
}
