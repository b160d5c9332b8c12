// cgs_alt.hip -- the line-traversal ("Alt") costs of the reference:
// calcSQDAlt, calcNCCAlt (src/gpu/cudaErrorFunctions.cuh:226-376),
// calcNCCstreamAlt (src/gpu/cudaErrorFunctionsStreams.cuh:165-281),
// calcMIstreamAlt (src/gpu/cudaErrorFunction_miStreams.cuh:9-210) and
// calcMIAlt (src/gpu/cudaErrorFunction_mi.cuh:117-255).
//
// These functors do not warp pixel by pixel.  Each fixed row maps to a line
// p0 + lambda v through the moving image (parametricAssignValues,
// src/gpu/cudaImageFunctions.cuh:315-441); the row's samples sit at lambda =
// start, start + step, ... below the line's end, with lambda accumulated in
// binary32 and the fixed image sampled at (row, lambda).
//
// Rows go in blocks of 64: lane r sets up row rb + r's line, then the
// wavefront walks the block's qualifying rows in order, taking 64
// consecutive lambda steps of a row per pass (lane j: step j), so the
// moving-image taps along the line and the fixed-image taps along the row
// are neighbours across lanes.  Lambda is the reference's binary32 sequence,
// evaluated in closed form per lane (cgs_lambda.hpp).
//   alt_step_kernel     calcSQDAlt / calcNCCAlt / calcNCCstreamAlt, one grid
//     point per wavefront.  Step k of a row accumulates into lane k mod 64
//     and the lanes combine in xor-butterfly order, as the oracle declares,
//     so the costs are the oracle's bit for bit.
//   alt_step_mi_kernel  calcMIstreamAlt, and
//   alt_norm_kernel     calcMIAlt (normalised H, rows from a host table,
//     lambda += 1/(W_f - 1) from segments inside one binade each, 64
//     consecutive steps per pass): one grid point per
//     workgroup, the four wavefronts taking row blocks in turn into one LDS
//     joint histogram.
// Every kernel asks for 8 wavefronts per SIMD (at most 64 VGPRs): the
// loops wait on dependent tap loads, and the extra wavefronts hide that
// latency; what the bound spills is the grid point's set-up, reloaded once
// per row block.
// The moved area is summed per row block in butterfly order and the block
// sums in block order (the oracle's rule).  Compiled with -ffp-contract=off
// (see cgs_device.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

#include "cgs.h"
#include "cgs_device.hpp"
#include "cgs_host.hpp"
#include "cgs_kernels.hpp"
#include "cgs_lambda.hpp"

namespace cgs {

struct AltParams {
    GridDev grid;
    const float2 *sincos;   // (cos, sin) per level of axis 0
    int cost;
    const void *mov;        // moving image + zero pad (several channels)
    const void *fix;        // fixed image + zero pad (several channels)
    const void *mq;         // moving image's exact tap table (one channel)
    const void *fq;         // fixed image's exact tap table (one channel)
    const void *fraw;       // fixed image itself (one channel): rows at integral lambda
    int mw, mh, fw, fh, C;
    int bins;
    const float *rows;      // calcMIAlt: the row values
    int nrows;
    const float *area;      // calcMIAlt: fixed area after n steps
    int narea;
    const float *lam0;      // calcMIAlt: the lambda sequence from 0 (below 1), then 2.0f up to n0pad entries
    int n0, n0pad;
    const uint8_t *fb0;     // calcMIAlt, one channel: fixed-image bin of row k at lam0[j], entry k * n0 + j
    float *result;
    unsigned long long *key;
    int tie_mode;
    uint32_t *hist_dump;    // histogram of grid point 0 (diagnostics)
};

struct AltLine {
    float ls, le, vx, vy, px, py;
    bool in;
};

// parametricAssignValues (src/gpu/cudaImageFunctions.cuh:315-376): the
// reference's expressions in their association order, repeated factors
// named once (dt, a0, aw, gy, gx).
__device__ __forceinline__ AltLine line_centred(const float *h, float MW, float MH, float FW, float FH, float r) {
    const float h11 = h[0], h12 = h[1], h13 = h[2], h21 = h[3], h22 = h[4], h23 = h[5], h31 = h[6], h32 = h[7];
    AltLine L;
    const float dt = h11*h22 - h12*h21 - h11*h23*h32 + h12*h23*h31 + h13*h21*h32 - h13*h22*h31;
    const float a0 = h11*h22 - h12*h21 - h11*h32*r + h12*h31*r;
    const float aw = h11*h22 - h12*h21 + FW*h21*h32 - FW*h22*h31 - h11*h32*r + h12*h31*r;
    const float gy = h21 - h31*r, gx = h22 - h32*r;
    L.vx = (gx*(dt))/(a0*aw);
    L.vy = -(gy*(dt))/(a0*aw);
    const float q = h12*h21 - h11*h22 + r*(h11*h32 - h12*h31);
    L.px = (h13*h22 - h12*h23 + r*(h12 - h13*h32))/q;
    L.py = -(h13*h21 - h11*h23 + r*(h11 - h13*h31))/q;
    const float lt = -((h11*h23 - h13*h21 - h11*r + h13*h31*r)*(aw))/(gy*(dt));
    const float xt = -(h23 - r)/gy;
    const float lb = -((aw)*(h11*h23 - h13*h21 - h11*r + h11*h22*MH - h12*h21*MH + h13*h31*r - h11*h32*MH*r
                             + h12*h31*MH*r))/(gy*(dt));
    const float xb = -(h23 - r + h22*MH - h32*MH*r)/gy;
    const float ll = -((h12*h23 - h13*h22 - h12*r + h13*h32*r)*(aw))/(gx*(dt));
    const float yl = -(h23 - r)/gx;
    const float lr = -((aw)*(h12*h23 - h13*h22 - h12*r - h11*h22*MW + h12*h21*MW + h13*h32*r + h11*h32*MW*r
                             - h12*h31*MW*r))/(gx*(dt));
    const float yr = -(h23 - r + h21*MW - h31*MW*r)/gx;
    L.ls = 0.0f;
    L.le = FW;
    L.in = false;
    if (xt >= 0 && xt < MW && lt >= 0 && lt < FW) { if (L.vy > 0) L.ls = lt; else L.le = lt; L.in = true; }
    if (xb >= 0 && xb < MW && lb >= 0 && lb < FW) { if (L.vy < 0) L.ls = lb; else L.le = lb; L.in = true; }
    if (yl >= 0 && yl < MH && ll >= 0 && ll < FH) { if (L.vx > 0) L.ls = ll; else L.le = ll; L.in = true; }
    if (yr >= 0 && yr < MH && lr >= 0 && lr < FH) { if (L.vx < 0) L.ls = lr; else L.le = lr; L.in = true; }
    return L;
}

// parametricAssignValuesNorm (src/gpu/cudaImageFunctions.cuh:378-441),
// column -0.5; ty / tx are the quotients shared with p0 = (tx, -ty).
__device__ __forceinline__ AltLine line_norm(const float *h, float row) {
    const float h11 = h[0], h12 = h[1], h13 = h[2], h21 = h[3], h22 = h[4], h23 = h[5], h31 = h[6], h32 = h[7];
    const float col = -0.5f;
    AltLine L;
    const float dt = h11*h22 - h12*h21 - h11*h23*h32 + h12*h23*h31 + h13*h21*h32 - h13*h22*h31;
    L.vy = -(h21 - h31*row)/dt;
    L.vx = (h22 - h32*row)/dt;
    const float q = h11*h22 - h12*h21 + col*(h21*h32 - h22*h31) - row*(h11*h32 - h12*h31);
    const float tx = (h12*h23 - h13*h22 + col*(h22 - h23*h32) - row*(h12 - h13*h32))/q;
    const float ty = (h11*h23 - h13*h21 + col*(h21 - h23*h31) - row*(h11 - h13*h31))/q;
    L.px = tx;
    L.py = -ty;
    const float gy = h21 - h31*row, gx = h22 - h32*row;
    const float lt = -((ty - 0.5f)*(dt))/gy;
    const float xt = (h22 - 2*h23 + 2*row - h32*row)/(2*gy);
    const float lb = -((ty + 0.5f)*(dt))/gy;
    const float xb = -(h22 + 2*h23 - 2*row - h32*row)/(2*gy);
    const float ll = -((tx + 0.5f)*(dt))/gx;
    const float yl = (h21 - 2*h23 + 2*row - h31*row)/(2*gx);
    const float lr = -((tx - 0.5f)*(dt))/gx;
    const float yr = -(h21 + 2*h23 - 2*row - h31*row)/(2*gx);
    L.ls = 0.0f;
    L.le = 1.0f;
    L.in = false;
    if (xt > -0.5f && xt < 0.5f && lt >= 0 && lt < 1) { if (L.vy > 0) L.ls = lt; else L.le = lt; L.in = true; }
    if (xb > -0.5f && xb < 0.5f && lb > 0 && lb < 1) { if (L.vy < 0) L.ls = lb; else L.le = lb; L.in = true; }
    if (yl > -0.5f && yl < 0.5f && ll > 0 && ll < 1) { if (L.vx > 0) L.ls = ll; else L.le = ll; L.in = true; }
    if (yr > -0.5f && yr < 0.5f && lr > 0 && lr < 1) { if (L.vx < 0) L.ls = lr; else L.le = lr; L.in = true; }
    return L;
}

// parametersToHomographyNorm (src/gpu/cudaImageFunctions.cuh:181-214): integer
// size ratios and centres, the moving centre's y taken from the width.
__device__ __forceinline__ void homography_norm(const float *p, float c, float s, int mw, int mh, int fw, int fh,
                                                float *h) {
    const float x_ratio = (float) (fw / mw), y_ratio = (float) (fh / mh);
    const float cfx = (float) (fw / 2), cfy = (float) (fh / 2), cmx = (float) (mw / 2), cmy = (float) (mw / 2);
    const float scaleX = p[1] / x_ratio, scaleY = p[2] / y_ratio, shearXY = p[3];
    h[0] = scaleX * c;
    h[1] = scaleY * shearXY * c - scaleY * s;
    h[2] = (p[4] - cfx + cmx) / (float) fw;
    h[3] = scaleX * s;
    h[4] = scaleY * shearXY * s + scaleY * c;
    h[5] = (p[5] - cfy + cmy) / (float) fh;
    h[6] = p[6];
    h[7] = p[7];
}

// v of lane j (j wavefront-uniform).
__device__ __forceinline__ float lane_value(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// CudaImage::valueAt (include/cudaTensor.cuh:552-575): the bilinear value,
// 0 outside the image.  img is a zero-padded copy.
template <typename T>
__device__ __forceinline__ float value_at(const T *img, float X, float Y, int c, int C, int W, int H) {
    if (!(X >= 0.0f && Y >= 0.0f && X < (float) W && Y < (float) H)) return 0.0f;
    return sample_bilinear(img, X, Y, c, C, W);
}

// valueAt(row, X) of the fixed image at an integral row inside the image and
// 0 <= X < W (row_off = row * W * C): ceil(row) == row, so the bottom lerp is
// the top one and the value is fmaf(0, top, 1 * top) -- top itself, or NaN
// where top is infinite, exactly as the full form.
template <typename T>
__device__ __forceinline__ float row_value(const T *img, float X, int row_off, int c, int C) {
    const float fx = floorf(X);
    const float ax = X - fx;
    const int b = row_off + (int) fx * C + c;
    const float tl = (float) img[b], tr = (float) img[b + (ax != 0.0f ? C : 0)];
    const float top = fmaf(ax, tr, (1.0f - ax) * tl);
    return fmaf(0.0f, top, top);
}

// mi_bin of a sample of image type T.  A uint8 image's bilinear value lies
// in [0, 255] up to a few ulps (each lerp is a rounded convex combination of
// taps in [0, 255]; 0 outside the image), so t = bins * (v / 256) is in
// [0, bins): mi_bin's clamps do nothing and the truncating conversion is its
// floor.  With kb = bins / 256 (exact, bins < 2^24), v * kb rounds the same
// real number bins * v / 256 as bins * (v / 256) does (v / 256 is exact for a
// normal v, and a v small enough for it to round gives 0 either way): one
// multiply.
template <typename T>
__device__ __forceinline__ int alt_bin(float v, int bins, float kb) {
    if constexpr (std::is_same<T, uint8_t>::value) return (int) (v * kb);
    else return mi_bin(v, bins);
}


// LDS address of the joint histogram's cell (fixed bin fb, moving bin mb):
// base + 4 mb + fb rowb, rows of bins + 1 counters (rowb = 4 (bins + 1)
// bytes), a shift-add and a 24-bit multiply-add (both factors below 2^24).
__device__ __forceinline__ uint32_t alt_lds_base(uint32_t *hist) { return lds_address(hist); }
__device__ __forceinline__ uint32_t alt_cell(int fb, int mb, uint32_t rowb, uint32_t base) {
    uint32_t r;
    asm("v_lshl_add_u32 %0, %1, 2, %2\n\tv_mad_u32_u24 %0, %3, %4, %0"
        : "=&v"(r)
        : "v"(mb), "s"(base), "v"(fb), "s"(rowb));
    return r;
}

// alt_cell with the moving bin trunc(t) (t >= 0 or NaN, which
// v_cvt_i32_f32 takes to 0): the conversion, the shift-add and the mad in
// one block.
__device__ __forceinline__ uint32_t alt_cell_trunc(int fb, float t, uint32_t rowb, uint32_t base) {
    uint32_t r;
    asm("v_cvt_i32_f32 %0, %1\n\tv_lshl_add_u32 %0, %0, 2, %2\n\tv_mad_u32_u24 %0, %3, %4, %0"
        : "=&v"(r)
        : "v"(t), "s"(base), "v"(fb), "s"(rowb));
    return r;
}

__device__ __forceinline__ void alt_count(uint32_t addr) { lds_count(addr); }

// MI of a wavefront's joint histogram h (row = fixed bin) over n samples,
// binary64: mi = sum pxy (log2 pxy - (log2 px + log2 py)), hx / hy the
// marginal entropies.  Every lane gets the totals.  Lane k handles column k.
// h: bins rows of rs >= bins counters (row = fixed-image bin).
__device__ __forceinline__ void wave_mi(const uint32_t *h, int bins, int rs, double n, double *mi_out, double *hx_out,
                                        double *hy_out) {
    const int lane = threadIdx.x & 63;
    double px = 0.0, py = 0.0;
    if (lane < bins) {
        uint32_t rsum = 0, csum = 0;
        for (int j = 0; j < bins; ++j) {
            rsum += h[lane * rs + j];
            csum += h[j * rs + lane];
        }
        px = rsum / n;
        py = csum / n;
    }
    const double lpx = px > 0 ? log2(px) : 0.0, lpy = py > 0 ? log2(py) : 0.0;
    double hx = px > 0 ? -(px * lpx) : 0.0, hy = py > 0 ? -(py * lpy) : 0.0;
    double mi = 0.0;
    for (int i = 0; i < bins; ++i) {
        const double lpx_i = __shfl(lpx, i), px_i = __shfl(px, i);
        if (lane < bins) {
            const double pxy = h[i * rs + lane] / n;
            if (pxy > 0 && px_i > 0 && py > 0) mi += pxy * (log2(pxy) - (lpx_i + lpy));
        }
    }
    *mi_out = wave_sum(mi);
    *hx_out = wave_sum(hx);
    *hy_out = wave_sum(hy);
}

// The overlap test of the unit-step functors (src/gpu/cudaErrorFunctions.cuh:296, 372;
// src/gpu/cudaErrorFunctionsStreams.cuh:272; src/gpu/cudaErrorFunction_miStreams.cuh:199),
// its two halves: moved area / (W_m H_m scaleY) > 0.25, fixed area / (W_f H_f) > 0.35.
__device__ __forceinline__ float alt_mov_denom(const AltParams &P, float scale_y) {
    return (float) (P.mw * P.mh) * scale_y;
}
__device__ __forceinline__ bool alt_mov_test(float mov_area, float denom) { return (double) (mov_area / denom) > 0.25; }
__device__ __forceinline__ bool alt_fix_test(float fix_area, const AltParams &P) {
    return (double) (fix_area / (float) (P.fw * P.fh)) > 0.35;
}

// The moved-area test on the reference's binary32 sum M, decided from S, the
// same steps * |v| terms summed in binary64 (exact to 2^-40 relative), where
// that is certain.  M is nadds rounded adds of non-negative terms, so every
// partial sum is at most Mmax = S (1 + nadds 2^-23) (for nadds <= 2^24) and
// |M - S| <= nadds * ulp(Mmax) / 2.  The test is monotone in M: equal at both
// ends of that interval, it is decided.  Returns 1 or 0, or -1 when the
// interval straddles the threshold (then the caller forms M the reference's
// way).
__device__ inline int alt_mov_test_band(double S, double nadds, float denom) {
    if (!(S >= 0.0) || !(S < 1e30) || nadds > 16777216.0) return -1;
    const double Mmax = S * (1.0 + nadds * 0x1p-23) + 0x1p-126;
    int e;
    (void) frexp(Mmax, &e);   // Mmax in [2^(e-1), 2^e): binary32 half-ulp 2^(e-25)
    const double E = nadds * ldexp(1.0, e - 25) + S * 0x1p-40;
    const double lo = S - E > 0.0 ? S - E : 0.0, hi = S + E;
    const bool tlo = alt_mov_test(__double2float_rd(lo), denom), thi = alt_mov_test(__double2float_ru(hi), denom);
    return tlo == thi ? (int) tlo : -1;
}

__device__ __forceinline__ void alt_grid_point(const AltParams &P, int32_t gi, float *p, float *c, float *s) {
    int32_t i0;
    grid_point(P.grid, gi, p, &i0);
    const float2 cs = P.sincos[i0];
    *c = cs.x;
    *s = cs.y;
}

// Writes the workgroup's costs (grid points base .. base + count - 1) and
// folds their minimum into the argmin key (thread 0).
__device__ __forceinline__ void alt_emit(const AltParams &P, int32_t base, const float *s_cost, int count) {
    unsigned long long best = ~0ull;
    for (int g = 0; g < count; ++g) {
        const int32_t gi = base + g;
        if (gi < P.grid.n) {
            if (P.result) P.result[gi] = s_cost[g];
            const unsigned long long k = argmin_key(s_cost[g], (uint32_t) gi, P.tie_mode);
            best = k < best ? k : best;
        }
    }
    if (best != ~0ull && best < __hip_atomic_load(P.key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMin(P.key, best);
}

// Image samplers of the kernels.  valueAt (include/cudaTensor.cuh:552-575)
// reads its taps at floor and ceil; at(X, Y) is that value, 0 outside the
// image; at_row(X, row) the value at an integral row inside the image with
// 0 <= X < W, where ceil(row) == row makes the bottom lerp the top one, so
// the value is fmaf(0, top, 1 * top).
//
// TabSampler: one-channel images through their exact tap table, one load per
// sample (entry L + W + 1 holds p[L], p[L+1], p[L+W], p[L+W+1], 0 outside);
// for a float image the ceil taps are picked by selects (see kFinite).
// Out-of-image lanes read past the table, where the buffer's range check
// returns zeros.
// The tap index fy * W + fx of an in-image sample is an integer below the
// image's element count (< 2^24, checked when the image is validated), so
// one binary32 fma forms it exactly.
template <typename T>
struct TabSampler {
    // A uint8 image's taps are finite, so the ceil taps need no selects: where
    // ceil equals floor the lerp weight is 0, and fmaf(0, t, (1 - 0) p) = p for
    // the neighbouring tap t the entry holds (0 * finite adds nothing), as for
    // the bottom row when the row weight is 0.  A float image may hold inf or
    // NaN there, which 0 * t would turn into NaN: it keeps the selects.
    static constexpr bool kFinite = std::is_same<T, uint8_t>::value;
    __amdgpu_buffer_rsrc_t rsrc;
    int w, h;
    uint32_t base;   // byte offset of entry 0 of the image: (w + 1) entries
    float W, H;
    __device__ TabSampler(const void *table, int width, int height) {
        rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void *) uniform(table), (short) 0,
            (int) (((size_t) width * height + 2 * (size_t) width + 2) * TapTable<T>::kBytes), 0x00020000);
        w = width;
        h = height;
        base = (uint32_t) ((width + 1) * TapTable<T>::kBytes);
        W = (float) width;
        H = (float) height;
    }
    __device__ __forceinline__ float at(float X, float Y, int) const {
        const bool in = X >= 0.0f && Y >= 0.0f && X < W && Y < H;
        const float fx = floorf(X), fy = floorf(Y);
        const float ax = X - fx, ay = Y - fy;
        const uint32_t idx = (uint32_t) (int) fmaf(fy, W, fx);
        const uint32_t off = in ? idx * TapTable<T>::kBytes + base : kOffsetOut;
        const float4 q = TapTable<T>::load(rsrc, off);
        float tr = q.y, bl = q.z, br = q.w;
        if constexpr (!kFinite) {
            tr = ax != 0.0f ? q.y : q.x;
            bl = ay != 0.0f ? q.z : q.x;
            br = ay != 0.0f ? (ax != 0.0f ? q.w : q.z) : tr;
        }
        const float top = fmaf(ax, tr, (1.0f - ax) * q.x);
        const float bottom = fmaf(ax, br, (1.0f - ax) * bl);
        const float v = fmaf(ay, bottom, (1.0f - ay) * top);
        return in ? v : 0.0f;
    }
    // at(X, Y) in two halves, so that a caller can issue several loads before
    // it waits on any: load_at forms the taps' offset by integer floors
    // (v_cvt_flr), the weights by v_fract and the offset as a shift-add and
    // a 24-bit mad -- for an in-image point (X, Y >= 0) floorf, X - floorf(X)
    // and the fma index exactly -- and loads the taps.
    // A uint8 image skips at's final select: off the image the taps read 0,
    // so v is 0 for finite X, Y and NaN otherwise, and both truncate to bin 0
    // (v_cvt_i32_f32 takes NaN to 0), the bin of the 0 that at returns;
    // cell_of gives the bin's histogram cell.  Its in-image test is two
    // unsigned compares of the integer floors: v_cvt_flr clamps infinities and
    // huge values to INT_MIN / INT_MAX (outside) and takes NaN to 0, whose
    // lane then reads taps inside the table but has the weight v_fract(NaN) =
    // NaN, so v is NaN and its bin 0 all the same.
    struct Taps {
        float4 q;
        float ax, ay;
        bool in;
    };
    __device__ __forceinline__ Taps load_at(float X, float Y) const {
        constexpr int EB = TapTable<T>::kBytes;
        Taps t;
        const int ix = floor_to_int(X), iy = floor_to_int(Y);
        if constexpr (kFinite) t.in = (uint32_t) ix < (uint32_t) w && (uint32_t) iy < (uint32_t) h;
        else t.in = X >= 0.0f && Y >= 0.0f && X < W && Y < H;
        t.ax = __builtin_amdgcn_fractf(X);
        t.ay = __builtin_amdgcn_fractf(Y);
        uint32_t off;
        asm("v_lshl_add_u32 %0, %1, %2, %3\n\tv_mad_i32_i24 %0, %4, %5, %0"
            : "=&v"(off)
            : "v"(ix), "i"(EB == 4 ? 2 : 4), "s"(base), "v"(iy), "s"(w * EB));
        t.q = TapTable<T>::load(rsrc, t.in ? off : kOffsetOut);
        return t;
    }
    // the histogram cell (fixed bin fb) of the value the taps give
    __device__ __forceinline__ uint32_t cell_of(const Taps &t, int fb, int bins, float kb, uint32_t rowb,
                                                uint32_t hbase) const {
        const float ax = t.ax, ay = t.ay;
        float tr = t.q.y, bl = t.q.z, br = t.q.w;
        if constexpr (!kFinite) {
            tr = ax != 0.0f ? t.q.y : t.q.x;
            bl = ay != 0.0f ? t.q.z : t.q.x;
            br = ay != 0.0f ? (ax != 0.0f ? t.q.w : t.q.z) : tr;
        }
        const float top = fmaf(ax, tr, (1.0f - ax) * t.q.x);
        const float bottom = fmaf(ax, br, (1.0f - ax) * bl);
        const float v = fmaf(ay, bottom, (1.0f - ay) * top);
        if constexpr (kFinite) return alt_cell_trunc(fb, v * kb, rowb, hbase);
        else return alt_cell(fb, alt_bin<T>(t.in ? v : 0.0f, bins, kb), rowb, hbase);
    }
    // at(X, Y) for a point the caller knows to be inside the image.
    __device__ __forceinline__ float at_inside(float X, float Y) const {
        const float fx = floorf(X), fy = floorf(Y);
        const float ax = X - fx, ay = Y - fy;
        const float4 q = TapTable<T>::load(rsrc, (uint32_t) (int) fmaf(fy, W, fx) * TapTable<T>::kBytes + base);
        float tr = q.y, bl = q.z, br = q.w;
        if constexpr (!kFinite) {
            tr = ax != 0.0f ? q.y : q.x;
            bl = ay != 0.0f ? q.z : q.x;
            br = ay != 0.0f ? (ax != 0.0f ? q.w : q.z) : tr;
        }
        const float top = fmaf(ax, tr, (1.0f - ax) * q.x);
        const float bottom = fmaf(ax, br, (1.0f - ax) * bl);
        return fmaf(ay, bottom, (1.0f - ay) * top);
    }
    // at_row for 0 <= X (every walked step): v_cvt_flr and v_fract are floorf
    // and X - floorf(X) there; a finite top (uint8 taps) is its own
    // fmaf(0, top, top).
    __device__ __forceinline__ float at_row(float X, int row, int) const {
        const int fx = floor_to_int(X);
        const float ax = __builtin_amdgcn_fractf(X);
        const float4 q = TapTable<T>::load(rsrc, (uint32_t) ((row * w + fx + w + 1) * TapTable<T>::kBytes));
        const float tr = kFinite || ax != 0.0f ? q.y : q.x;
        const float top = fmaf(ax, tr, (1.0f - ax) * q.x);
        if constexpr (kFinite) return top;
        else return fmaf(0.0f, top, top);
    }
};

// RawSampler: any channel count, on a zero-padded copy of the image.
template <typename T>
struct RawSampler {
    const T *img;
    int C, w, h;
    __device__ RawSampler(const void *p, int channels, int width, int height)
        : img((const T *) p), C(channels), w(width), h(height) {}
    __device__ __forceinline__ float at(float X, float Y, int c) const { return value_at(img, X, Y, c, C, w, h); }
    __device__ __forceinline__ float at_row(float X, int row, int c) const {
        return row_value(img, X, row * w * C, c, C);
    }
};

template <typename T, bool ONE>
using Sampler = typename std::conditional<ONE, TabSampler<T>, RawSampler<T>>::type;

// A one-channel fixed row at an integral lambda (0 <= lambda < W, as every
// walked step is): the lerp weight is 0, so at_row's value is
// fmaf(0, top, top) with top = fmaf(0, p, 1 * p) for the pixel p itself (p
// when finite, NaN otherwise), from one 4-byte load of a float image instead
// of a 16-byte tap entry.
template <typename T>
struct IntRow {
    __amdgpu_buffer_rsrc_t rsrc;
    int w;
    __device__ IntRow(const void *img, int width, int height) {
        rsrc = __builtin_amdgcn_make_buffer_rsrc((void *) uniform(img), (short) 0,
                                                 (int) ((size_t) width * height * sizeof(T)), 0x00020000);
        w = width;
    }
    __device__ __forceinline__ float at(float lam, int row) const {
        const float p = load_fix<T>(rsrc, (uint32_t) ((row * w + (int) lam) * (int) sizeof(T)));
        const float top = fmaf(0.0f, p, p);
        return fmaf(0.0f, top, top);
    }
};

template <typename T, bool ONE>
__device__ __forceinline__ Sampler<T, ONE> make_sampler(const void *tab, const void *raw, int C, int w, int h) {
    if constexpr (ONE) return TabSampler<T>(tab, w, h);
    else return RawSampler<T>(raw, C, w, h);
}

// Walks one row's unit-step line: 64 consecutive lambda steps per pass, lane
// j taking step j; calls at(lam, nx, ny, ipass) for the lane's step while it
// is below the line's end and the row's width, ipass telling (the same in
// every lane) that the pass's lambdas are integers.  Returns the row's step
// count (wavefront-uniform).  SCALAR_TESTS: the pass's wave-uniform tests
// in scalar instructions (the NCC/SQD kernel, whose scalar unit is busier,
// measured 4% faster with vector compares; the MI kernel 6% faster with
// scalar ones).
template <bool SCALAR_TESTS, typename F>
__device__ __forceinline__ int32_t walk_unit_row(float ls, float le, float FW, float vx, float vy, float px,
                                                 float py, F &&at) {
    const int lane = threadIdx.x & 63;
    const float flane = (float) lane;
    float x0 = ls;
    int32_t nrow = 0;
    while (true) {
        // x0 is the same in every lane: one branch per pass
        const bool direct = SCALAR_TESTS ? unit_pass_direct(x0) : unit_pass_direct_v(x0);
        const float lam = direct ? x0 + flane : unit_pass_chain(x0, lane);
        const bool ipass = SCALAR_TESTS ? unit_int_start(x0) : unit_int_start_v(x0);   // lam = x0 + j, integers
        // one ballot per compare: each is the compare's own lane mask, and
        // their AND is scalar (a ballot of the combined bool costs a
        // select and a compare per pass)
        const unsigned long long m =
            __builtin_amdgcn_ballot_w64(lam < le) & __builtin_amdgcn_ballot_w64(lam < FW);
        if (m == 0ull) break;
        if (lam < le && lam < FW) at(lam, lam * vx + px, lam * vy + py, ipass);   // the same compares
        nrow += __popcll(m);
        if (m != ~0ull) break;   // valid lanes form a prefix: the row ended in this pass
        x0 = lane_value(lam, 63) + 1.0f;
    }
    return nrow;
}

// The line of row rb + lane (centred H) and its |v|; in = false past the image.
__device__ __forceinline__ AltLine block_line(const float *h, const AltParams &P, int rb, float *mag) {
    const int lane = threadIdx.x & 63;
    AltLine L{};
    *mag = 0.0f;
    if (rb + lane < P.fh) {
        L = line_centred(h, (float) P.mw, (float) P.mh, (float) P.fw, (float) P.fh, (float) (rb + lane));
        *mag = sqrtf(L.vx * L.vx + L.vy * L.vy);
    }
    return L;
}

// The reference's binary32 moved area of one grid point (centred H h), the
// rows walked again for their step counts only: PER_ROW false adds |v| once
// per step into one sum over every row (calcSQDAlt, calcNCCAlt,
// src/gpu/cudaErrorFunctions.cuh:289, 365); true sums each row on its own,
// `adds` times per step (1 for calcNCCstreamAlt, the channel count for
// calcMIstreamAlt: src/gpu/cudaErrorFunctionsStreams.cuh:245,
// src/gpu/cudaErrorFunction_miStreams.cuh:110), and adds the row sums in row
// order.  Repeated adds go through add_repeat.  Wavefront-uniform; the rare
// path of the overlap test.
template <bool SCALAR_TESTS>
__device__ __forceinline__ float alt_exact_mov_area(const float *h, const AltParams &P, bool per_row, int adds) {
    float total = 0.0f;
    for (int rb = 0; rb < P.fh; rb += 64) {
        float magr;
        const AltLine Lr = block_line(h, P, rb, &magr);
        unsigned long long rows_in = __ballot(Lr.in);
        while (rows_in) {
            const int j = __builtin_amdgcn_readfirstlane(__builtin_ctzll(rows_in));
            rows_in &= rows_in - 1;
            const int32_t k = walk_unit_row<SCALAR_TESTS>(
                lane_value(Lr.ls, j), lane_value(Lr.le, j), (float) P.fw, lane_value(Lr.vx, j), lane_value(Lr.vy, j),
                lane_value(Lr.px, j), lane_value(Lr.py, j), [](float, float, float, bool) {});
            const float mag = lane_value(magr, j);
            if (per_row) total = total + add_repeat(0.0f, mag, (int64_t) k * adds);
            else total = add_repeat(total, mag, k);
        }
    }
    return total;
}

// calcSQDAlt (SQD true) / calcNCCAlt / calcNCCstreamAlt: one grid point per
// wavefront.  ONE: one-channel images through tap tables.
template <typename MT, typename FT, bool SQD, bool ONE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) void alt_step_kernel(const AltParams P) {
    __shared__ float s_cost[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int32_t base = (int32_t) blockIdx.x * 4;
    const int32_t gi = base + wave;
    const auto mov = make_sampler<MT, ONE>(P.mq, P.mov, P.C, P.mw, P.mh);
    const auto fix = make_sampler<FT, ONE>(P.fq, P.fix, P.C, P.fw, P.fh);
    // integral-lambda rows from the image itself for one-channel float images
    // (whose tap entries are 16 bytes; a uint8 entry is one word already)
    constexpr bool IROW = ONE && std::is_same<FT, float>::value;
    const IntRow<FT> irow(IROW ? P.fraw : P.fq, P.fw, P.fh);
    const int C = ONE ? 1 : P.C;
    float cost = 0.0f;
    if (gi < P.grid.n) {
        float p[kMaxDim], c, s, h[8];
        alt_grid_point(P, gi, p, &c, &s);
        homography_centred(p, c, s, (float) P.mw / 2, (float) P.mh / 2, h);
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, mov_area = 0.0;
        int32_t steps = 0;
        for (int rb = 0; rb < P.fh; rb += 64) {
            // lane r sets up row rb + r's line; the wavefront walks the block's
            // qualifying rows in order, reading each line from its lane
            float magr;
            const AltLine Lr = block_line(h, P, rb, &magr);
            double term = 0.0;   // lane r: row rb + r's steps * |v|
            unsigned long long rows_in = __ballot(Lr.in);
            while (rows_in) {
                const int j = __builtin_amdgcn_readfirstlane(__builtin_ctzll(rows_in));
                rows_in &= rows_in - 1;
                const int row = rb + j;
                const int32_t nrow = walk_unit_row<false>(
                    lane_value(Lr.ls, j), lane_value(Lr.le, j), (float) P.fw, lane_value(Lr.vx, j),
                    lane_value(Lr.vy, j), lane_value(Lr.px, j), lane_value(Lr.py, j),
                    [&](float lam, float nx, float ny, bool ipass) {
                        for (int ch = 0; ch < C; ++ch) {
                            const double md = (double) mov.at(nx, ny, ch);
                            float fv;
                            if constexpr (IROW) fv = ipass ? irow.at(lam, row) : fix.at_row(lam, row, ch);
                            else fv = fix.at_row(lam, row, ch);
                            const double fd = (double) fv;
                            if constexpr (SQD) s0 = fma(md - fd, md - fd, s0);
                            else s0 = fma(md, fd, s0);
                            s1 = fma(fd, fd, s1);
                            s2 = fma(md, md, s2);
                        }
                    });
                if (lane == j && nrow > 0) term = (double) nrow * (double) magr;
                steps += nrow;
            }
            mov_area += wave_sum(term);
        }
        // the fixed area counts steps in binary32, exactly (fewer than 2^24)
        const float denom = alt_mov_denom(P, p[2]);
        const bool stream = P.cost == CGS_COST_NCC_STREAM_ALT;
        int mt = alt_mov_test_band(mov_area, (double) steps + (double) P.fh, denom);
        if (mt < 0) mt = alt_mov_test(alt_exact_mov_area<false>(h, P, stream, 1), denom);
        const bool pass = mt && alt_fix_test((float) steps, P);
        const double S0 = wave_sum(s0), S1 = wave_sum(s1), S2 = wave_sum(s2);
        if (!pass) cost = P.cost == CGS_COST_NCC_STREAM_ALT ? 0.0f : 123123.0f;
        else if (SQD) cost = (float) (S0 / sqrt(S1 * S2));
        else cost = (float) (-S0 / sqrt(S1 * S2));
    }
    if (lane == 0) s_cost[wave] = cost;
    __syncthreads();
    if (threadIdx.x == 0) alt_emit(P, base, s_cost, 4);
}

// MI costs: one grid point per workgroup, the four wavefronts taking blocks
// of 64 rows in turn (wavefront w: blocks w, w + 4, ...) into one LDS joint
// histogram.  LDS, all of it dynamic so that the histogram starts at LDS
// address 0 and a cell's byte offset is its address: the bins^2 counters, one
// binary64 moved-area sum per row block (summed in block order), then the
// wavefronts' step counts and the cost (AltTail).
// LDS words of the joint histogram: bins rows padded to bins + 1 counters (a
// column spreads over the LDS banks), rounded up to keep what follows 8-byte
// aligned.
__host__ __device__ inline int alt_hist_words(int bins) { return (bins * (bins + 1) + 1) & ~1; }

struct AltTail {
    int32_t steps[4];
    float cost;
    int32_t pad;
};

__host__ __device__ inline size_t alt_lds_bytes(int bins, int fh) {
    return sizeof(uint32_t) * (size_t) alt_hist_words(bins) + sizeof(double) * (size_t) ((fh + 63) / 64) +
           sizeof(AltTail);
}

__device__ __forceinline__ uint32_t *alt_lds_hist() {
    extern __shared__ uint32_t s_dyn[];
    return s_dyn;
}

__device__ __forceinline__ AltTail *alt_tail(uint32_t *hist, int bins, int fh) {
    return (AltTail *) ((char *) hist + alt_lds_bytes(bins, fh) - sizeof(AltTail));
}

// calcMIstreamAlt.
template <typename MT, typename FT, bool ONE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) void alt_step_mi_kernel(const AltParams P) {
    uint32_t *hist = alt_lds_hist();
    AltTail *tail = alt_tail(hist, P.bins, P.fh);
    const int B2 = P.bins * P.bins;
    double *s_mov = (double *) (hist + alt_hist_words(P.bins));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int32_t gi = (int32_t) blockIdx.x;
    for (int i = threadIdx.x; i < alt_hist_words(P.bins); i += kThreads) hist[i] = 0u;
    __syncthreads();
    const auto mov = make_sampler<MT, ONE>(P.mq, P.mov, P.C, P.mw, P.mh);
    const auto fix = make_sampler<FT, ONE>(P.fq, P.fix, P.C, P.fw, P.fh);
    // integral-lambda rows from the image itself for one-channel float images
    // (whose tap entries are 16 bytes; a uint8 entry is one word already)
    constexpr bool IROW = ONE && std::is_same<FT, float>::value;
    const IntRow<FT> irow(IROW ? P.fraw : P.fq, P.fw, P.fh);
    const int C = ONE ? 1 : P.C;
    const int bins = P.bins;
    const float kb = (float) bins / 256.0f;
    const uint32_t rowb = 4u * (uint32_t) (bins + 1);
    const uint32_t hbase = __builtin_amdgcn_readfirstlane(alt_lds_base(hist));
    float p[kMaxDim], c, s, h[8];
    alt_grid_point(P, gi, p, &c, &s);
    homography_centred(p, c, s, (float) P.mw / 2, (float) P.mh / 2, h);
    int32_t steps = 0;
    for (int blk = wave; blk * 64 < P.fh; blk += 4) {
        const int rb = blk * 64;
        float magr;
        const AltLine Lr = block_line(h, P, rb, &magr);
        double term = 0.0;
        unsigned long long rows_in = __ballot(Lr.in);
        while (rows_in) {
            const int j = __builtin_amdgcn_readfirstlane(__builtin_ctzll(rows_in));
            rows_in &= rows_in - 1;
            const int row = rb + j;
            const int32_t nrow = walk_unit_row<true>(
                lane_value(Lr.ls, j), lane_value(Lr.le, j), (float) P.fw, lane_value(Lr.vx, j), lane_value(Lr.vy, j),
                lane_value(Lr.px, j), lane_value(Lr.py, j), [&](float lam, float nx, float ny, bool ipass) {
                    if constexpr (ONE) {   // the moving taps' load issued first (load_at / cell_of)
                        const auto tm = mov.load_at(nx, ny);
                        float fv;
                        if constexpr (IROW) fv = ipass ? irow.at(lam, row) : fix.at_row(lam, row, 0);
                        else fv = fix.at_row(lam, row, 0);
                        alt_count(mov.cell_of(tm, alt_bin<FT>(fv, bins, kb), bins, kb, rowb, hbase));
                    } else {
                        for (int ch = 0; ch < C; ++ch) {
                            const float mv = mov.at(nx, ny, ch);
                            const float fv = fix.at_row(lam, row, ch);
                            alt_count(alt_cell(alt_bin<FT>(fv, bins, kb), alt_bin<MT>(mv, bins, kb), rowb, hbase));
                        }
                    }
                });
            if (lane == j && nrow > 0) term = (double) nrow * (double) magr;
            steps += nrow;
        }
        const double bsum = wave_sum(term);
        if (lane == 0) s_mov[blk] = bsum;
    }
    if (lane == 0) tail->steps[wave] = steps;
    __syncthreads();
    if (wave == 0) {
        double mov_area = 0.0;
        for (int blk = 0; blk * 64 < P.fh; ++blk) mov_area += s_mov[blk];   // block order
        const int32_t total = tail->steps[0] + tail->steps[1] + tail->steps[2] + tail->steps[3];
        float cost = 123123.0f;
        // both areas grow once per step and channel; the fixed one in binary32
        // exactly (fewer than 2^24)
        const float denom = alt_mov_denom(P, p[2]);
        int mt = alt_mov_test_band((double) C * mov_area, (double) total * C + (double) P.fh, denom);
        if (mt < 0) mt = alt_mov_test(alt_exact_mov_area<true>(h, P, true, C), denom);
        if (mt && alt_fix_test((float) (total * C), P)) {
            double mi, hx, hy;
            wave_mi(hist, P.bins, P.bins + 1, (double) total * (double) C, &mi, &hx, &hy);
            cost = (float) (-2.0 * mi / (hx + hy));
        }
        if (P.hist_dump && gi == 0)
            for (int i = lane; i < B2; i += 64) P.hist_dump[i] = hist[(i / P.bins) * (P.bins + 1) + i % P.bins];
        if (lane == 0) tail->cost = cost;
    }
    __syncthreads();
    if (threadIdx.x == 0) alt_emit(P, gi, &tail->cost, 1);
}

// A segment computed from wavefront-uniform values, held in scalar registers.
__device__ __forceinline__ LamSeg uniform_seg(LamSeg g) {
    g.k0 = __builtin_amdgcn_readfirstlane(g.k0);
    g.a = __builtin_amdgcn_readfirstlane(g.a);
    g.d = __builtin_amdgcn_readfirstlane(g.d);
    g.lim = __builtin_amdgcn_readfirstlane(g.lim);
    g.U = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, g.U)));
    return g;
}

// A pass of 64 consecutive steps k0 .. k0 + 63 of the calcMIAlt sequence
// lambda += s (s = 1/(W_f - 1)) filled from its binade segments: each lane
// takes the last segment that starts at or before its step and still holds
// it; the next segment starts at the literal add after the last covered step
// (carry: the step before the pass, when none is covered), unless that add
// reaches lend.  Lanes past the row's last step hold 2.  g is the segment in
// force at the pass's start and, on return, at its end.
__device__ __forceinline__ float fill_pass(LamSeg &g, int32_t k0, float carry, float s, float lend) {
    const int32_t kl = k0 + (int) (threadIdx.x & 63);
    float lam = 2.0f;
    unsigned long long covered = 0ull;
    for (int fill = 0; fill <= 64; ++fill) {
        // t * d + a < 2^32 on every lane that can hold the segment
        // (t <= its length + 63, its length * d <= 2^24 + d)
        const uint32_t w = __umul24((uint32_t) (kl - g.k0), g.d) + g.a;
        const unsigned long long m =
            __builtin_amdgcn_ballot_w64(kl >= g.k0) & __builtin_amdgcn_ballot_w64(w <= g.lim);
        lam = __uint_as_float(mask_select(m, __float_as_uint(lam), __float_as_uint((float) w * g.U)));
        covered |= m;
        if (covered >> 63) break;
        const int last = covered ? 63 - __builtin_clzll(covered) : -1;
        const float x = (last < 0 ? carry : lane_value(lam, last)) + s;
        if (!(x < lend)) break;   // the row ends before it
        g = uniform_seg(lam_segment(x, s, k0 + last + 1));
    }
    return lam;
}

// calcMIAlt on one-channel images: the rows of wavefront `wave`'s blocks
// (blocks wave, wave + 4, ...) into the LDS joint histogram at byte address
// hbase; returns the steps taken.  Every walked step has lambda < lend =
// min(le, 1), the reference's lambda < le && lambda < 1 (NaN kept), and the
// steps of a row grow with the lane, so the walked lanes of a pass form a
// prefix and a pass that is not full ends the row.
//  * A row that starts at lambda = 0 (or -0, whose later steps are the same
//    and whose samples are those of +0) reads its steps from the lam0 table,
//    each pass loading the next pass's while it works, and its fixed-image
//    bins from the fb0 table.
//  * Any other row takes a pass in closed form, (lane d + b) U, when its
//    current segment holds all 64 of the pass's steps (a row from lambda >=
//    1/2 stays in one segment to its end), and otherwise fills the pass from
//    segments (fill_pass).  Its fixed sample lies on one fixed-image row:
//    ri = (row + 0.5) (H_f - 1) is the same at every step, so its floor row
//    and weight are taken once per row.
template <typename MT, typename FT, bool POW2>
__device__ __forceinline__ int32_t norm_rows_one(const AltParams &P, const float *h, uint32_t hbase) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const TabSampler<MT> mov(P.mq, P.mw, P.mh);
    constexpr int FEB = TapTable<FT>::kBytes;
    const int fw = uniform(P.fw), fh = uniform(P.fh);
    const __amdgpu_buffer_rsrc_t frsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void *) uniform(P.fq), (short) 0, (int) (((size_t) fw * fh + 2 * (size_t) fw + 2) * FEB), 0x00020000);
    const float sc = 1 / ((float) (fw - 1));
    const float wm1 = (float) (fw - 1);
    const float fmw = (float) P.mw, fmh = (float) P.mh;
    const double dmw = (double) P.mw, dmh = (double) P.mh;
    // lambda < 1 in steps of 1/(W_f - 1): at most W_f + 1 steps per row
    const int max_passes = fw / 64 + 3;
    const int n0 = uniform(P.n0);
    const __amdgpu_buffer_rsrc_t lam0_rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void *) uniform(P.lam0), (short) 0, uniform(P.n0pad) * (int) sizeof(float), 0x00020000);
    const __amdgpu_buffer_rsrc_t fb0_rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *) uniform(P.fb0), (short) 0, uniform(P.nrows) * n0, 0x00020000);
    const int bins = P.bins;
    const float kb = (float) bins / 256.0f;
    const uint32_t rowb = 4u * (uint32_t) (bins + 1);
    // the moving-image taps at step lambda of the line (px, py) + lambda (vx,
    // vy), scaled back to pixels as the reference does, through binary64
    // (rescale_pow2: one binary32 fma for power-of-two sizes)
    auto moving_taps = [&](float lam, float vx, float vy, float px, float py) {
        const float nx = lam * vx + px, ny = lam * vy + py;
        float X, Y;
        if constexpr (POW2) {
            X = rescale_pow2(nx, fmw);
            Y = rescale_pow2(ny, fmh);
        } else {
            X = (float) (((double) nx + 0.5) * dmw);
            Y = (float) (((double) ny + 0.5) * dmh);
        }
        return mov.load_at(X, Y);
    };
    int32_t steps = 0;
    for (int blk = wave; blk * 64 < P.nrows; blk += 4) {
        const int k = blk * 64 + lane;
        AltLine Lr{};
        float rir = 0.0f;
        if (k < P.nrows) {
            const float row = P.rows[k];
            Lr = line_norm(h, row);
            rir = (float) (((double) row + 0.5) * (double) (fh - 1));
        }
        unsigned long long rows_in = __ballot(Lr.in);
        // every row's first segment, one lane per row, for the rows that do
        // not start at lambda = 0 (those read the tables)
        const LamSeg g0 = lam_segment(Lr.ls, sc, 0);
        while (rows_in) {
            const int j = __builtin_amdgcn_readfirstlane(__builtin_ctzll(rows_in));
            rows_in &= rows_in - 1;
            const float le = lane_value(Lr.le, j);
            const float lend = le >= 1.0f ? 1.0f : le;
            const float vx = lane_value(Lr.vx, j), vy = lane_value(Lr.vy, j);
            const float px = lane_value(Lr.px, j), py = lane_value(Lr.py, j);
            const float ls = lane_value(Lr.ls, j);
            if (ls == 0.0f) {
                uint32_t koff = 4u * (uint32_t) lane;
                uint32_t fboff = (uint32_t) (blk * 64 + j) * (uint32_t) n0 + (uint32_t) lane;
                float lam = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(lam0_rsrc, koff, 0, 0));
                // every pass but the last is full; the table's 2.0f padding
                // ends the row at the latest in the pass that holds entry n0
                while (true) {
                    // the next pass's steps (a pass follows only a full one, so
                    // it lies inside the table)
                    const float nxt = __builtin_bit_cast(
                        float, __builtin_amdgcn_raw_buffer_load_b32(lam0_rsrc, koff + 256u, 0, 0));
                    const unsigned long long m = __builtin_amdgcn_ballot_w64(lam < lend);
                    if (lam < lend) {   // the same compare
                        const int fb = (int) __builtin_amdgcn_raw_buffer_load_b8(fb0_rsrc, fboff, 0, 0);
                        const auto tm = moving_taps(lam, vx, vy, px, py);
                        alt_count(mov.cell_of(tm, fb, bins, kb, rowb, hbase));
                    }
                    steps += __popcll(m);
                    if (m != ~0ull) break;   // the row ended in this pass
                    lam = nxt;
                    koff += 256u;
                    fboff += 64u;
                }
            } else {
                // (li, ri) lies inside the fixed image: 0 <= lambda < 1 gives
                // 0 <= li <= (1 - 2^-24) (W_f - 1) < W_f, and the table's rows,
                // -0.5 <= row < 0.5, give 0 <= ri < H_f ((row + 0.5) (H_f - 1)
                // < H_f - 1 rounds to at most H_f - 1); for such points
                // v_cvt_flr and v_fract are floorf and X - floorf(X) exactly
                const float ri = lane_value(rir, j);
                const float ay = __builtin_amdgcn_fractf(ri), cay = 1.0f - ay;
                const uint32_t frow =
                    (uint32_t) __builtin_amdgcn_readfirstlane((floor_to_int(ri) * fw + fw + 1) * FEB);
                // g: the segment in force at the pass's start (it starts at or
                // before step k0: fill_pass only starts segments inside the
                // pass it fills); b = (k0 - g.k0) d + a, its value at step k0.
                // While g holds a pass, b + 63 d <= lim < 2^24, so the next b is
                // below 2^25; after a filled pass the next b is at most 64 d +
                // a (g then starts inside the pass); d < 2^24 (an add that stays
                // in a binade moves less than the binade's 2^24 units): every
                // b + 63 d stays below 2^31
                // g covers step k0 - 1 at every pass after the first (the
                // segment in force at a pass's end holds its last lane), so the
                // step before a filled pass is (b - d) U.  Passes run while they
                // are full: lambda grows at every step (s > 2^-24, half an ulp
                // or more below 1), so a row ends; `filled` bounds the filled
                // passes as well
                LamSeg g;
                g.k0 = 0;
                g.a = (uint32_t) __builtin_amdgcn_readlane((int) g0.a, j);
                g.d = (uint32_t) __builtin_amdgcn_readlane((int) g0.d, j);
                g.lim = (uint32_t) __builtin_amdgcn_readlane((int) g0.lim, j);
                g.U = lane_value(g0.U, j);
                uint32_t laned = __umul24((uint32_t) lane, g.d);
                uint32_t b = g.a;
                int32_t k0 = 0, filled = 0;
                while (true) {
                    float lam;
                    if (b + 63u * g.d <= g.lim) {   // g holds all 64 steps
                        lam = (float) (laned + b) * g.U;
                        b += 64u * g.d;
                    } else {
                        if (++filled > max_passes) break;
                        lam = fill_pass(g, k0, (float) (b - g.d) * g.U, sc, lend);
                        laned = __umul24((uint32_t) lane, g.d);
                        b = (uint32_t) (k0 + 64 - g.k0) * g.d + g.a;
                    }
                    const unsigned long long m = __builtin_amdgcn_ballot_w64(lam < lend);
                    if (lam < lend) {
                        // both loads issued before either is waited on
                        const auto tm = moving_taps(lam, vx, vy, px, py);
                        const float li = lam * wm1;
                        const int ix = floor_to_int(li);
                        const float ax = __builtin_amdgcn_fractf(li);
                        const float4 q = TapTable<FT>::load(frsrc, (uint32_t) (ix * FEB) + frow);
                        float tr = q.y, bl = q.z, br = q.w;
                        if constexpr (!TabSampler<FT>::kFinite) {
                            tr = ax != 0.0f ? q.y : q.x;
                            bl = ay != 0.0f ? q.z : q.x;
                            br = ay != 0.0f ? (ax != 0.0f ? q.w : q.z) : tr;
                        }
                        const float top = fmaf(ax, tr, (1.0f - ax) * q.x);
                        const float bottom = fmaf(ax, br, (1.0f - ax) * bl);
                        const int fb = alt_bin<FT>(fmaf(ay, bottom, cay * top), bins, kb);
                        alt_count(mov.cell_of(tm, fb, bins, kb, rowb, hbase));
                    }
                    steps += __popcll(m);
                    if (m != ~0ull) break;
                    k0 += 64;
                }
            }
        }
    }
    return steps;
}

// calcMIAlt on images of several channels (zero-padded copies): the rows of
// wavefront `wave`'s blocks as norm_rows_one walks them from segments, each
// step sampling every channel.
template <typename MT, typename FT>
__device__ __forceinline__ int32_t norm_rows_raw(const AltParams &P, const float *h, uint32_t hbase) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const RawSampler<MT> mov(P.mov, P.C, P.mw, P.mh);
    const RawSampler<FT> fix(P.fix, P.C, P.fw, P.fh);
    const int C = P.C, bins = P.bins;
    const float sc = 1 / ((float) (P.fw - 1));
    const float wm1 = (float) (P.fw - 1);
    const double dmw = (double) P.mw, dmh = (double) P.mh;
    const float fmw = (float) P.mw, fmh = (float) P.mh;
    const bool pow2 = (P.mw & (P.mw - 1)) == 0 && (P.mh & (P.mh - 1)) == 0;
    const int max_passes = P.fw / 64 + 3;
    const float kb = (float) bins / 256.0f;
    const uint32_t rowb = 4u * (uint32_t) (bins + 1);
    int32_t steps = 0;
    for (int blk = wave; blk * 64 < P.nrows; blk += 4) {
        const int k = blk * 64 + lane;
        AltLine Lr{};
        float rir = 0.0f;
        if (k < P.nrows) {
            const float row = P.rows[k];
            Lr = line_norm(h, row);
            rir = (float) (((double) row + 0.5) * (double) (P.fh - 1));
        }
        unsigned long long rows_in = __ballot(Lr.in);
        while (rows_in) {
            const int j = __builtin_amdgcn_readfirstlane(__builtin_ctzll(rows_in));
            rows_in &= rows_in - 1;
            const float le = lane_value(Lr.le, j);
            const float lend = le >= 1.0f ? 1.0f : le;
            const float vx = lane_value(Lr.vx, j), vy = lane_value(Lr.vy, j);
            const float px = lane_value(Lr.px, j), py = lane_value(Lr.py, j);
            const float ri = lane_value(rir, j);
            LamSeg g = uniform_seg(lam_segment(lane_value(Lr.ls, j), sc, 0));
            float carry = 0.0f;
            int32_t k0 = 0;
            for (int pass = 0; pass < max_passes; ++pass) {
                const float lam = fill_pass(g, k0, carry, sc, lend);
                const unsigned long long m = __builtin_amdgcn_ballot_w64(lam < lend);
                if (m == 0ull) break;
                if (lam < lend) {
                    float nx = lam * vx + px, ny = lam * vy + py;
                    if (pow2) {
                        nx = rescale_pow2(nx, fmw);
                        ny = rescale_pow2(ny, fmh);
                    } else {
                        nx = (float) (((double) nx + 0.5) * dmw);
                        ny = (float) (((double) ny + 0.5) * dmh);
                    }
                    const float li = lam * wm1;   // (li, ri) inside the fixed image (see norm_rows_one)
                    for (int ch = 0; ch < C; ++ch) {
                        const int fb = alt_bin<FT>(fix.at(li, ri, ch), bins, kb);
                        alt_count(alt_cell(fb, alt_bin<MT>(mov.at(nx, ny, ch), bins, kb), rowb, hbase));
                    }
                }
                steps += __popcll(m);
                if (m != ~0ull) break;
                carry = lane_value(lam, 63);
                k0 += 64;
            }
        }
    }
    return steps;
}

// calcMIAlt: rows from the host's table, lambda += 1/(W_f - 1) taken 64
// consecutive steps per pass from the sequence's binade segments
// (lam_segment), so that the short segments near lambda = 0 share a pass; the
// fixed area after n steps comes from the host's table.  One grid point per
// workgroup, its four wavefronts taking blocks of 64 rows in turn.  POW2: the
// moving image's sides are powers of two (one-channel images only).
template <typename MT, typename FT, bool ONE, bool POW2>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8))) void alt_norm_kernel(const AltParams P) {
    uint32_t *hist = alt_lds_hist();
    AltTail *tail = alt_tail(hist, P.bins, P.fh);
    const int B2 = P.bins * P.bins;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int32_t gi = (int32_t) blockIdx.x;
    for (int i = threadIdx.x; i < alt_hist_words(P.bins); i += kThreads) hist[i] = 0u;
    __syncthreads();
    float p[kMaxDim], c, s, h[8];
    alt_grid_point(P, gi, p, &c, &s);
    homography_norm(p, c, s, P.mw, P.mh, P.fw, P.fh, h);
    const uint32_t hbase = __builtin_amdgcn_readfirstlane(alt_lds_base(hist));
    int32_t steps;
    if constexpr (ONE) steps = norm_rows_one<MT, FT, POW2>(P, h, hbase);
    else steps = norm_rows_raw<MT, FT>(P, h, hbase);
    if (lane == 0) tail->steps[wave] = steps;
    __syncthreads();
    if (wave == 0) {
        const int32_t total = tail->steps[0] + tail->steps[1] + tail->steps[2] + tail->steps[3];
        const float fa = P.area[total < P.narea ? total : P.narea - 1];
        double mi = 0.0, hx, hy;
        if (total > 0) wave_mi(hist, P.bins, P.bins + 1, (double) total * (double) P.C, &mi, &hx, &hy);
        if (P.hist_dump && gi == 0)
            for (int i = lane; i < B2; i += 64) P.hist_dump[i] = hist[(i / P.bins) * (P.bins + 1) + i % P.bins];
        if (lane == 0) tail->cost = (float) (-mi * (double) fa);
    }
    __syncthreads();
    if (threadIdx.x == 0) alt_emit(P, gi, &tail->cost, 1);
}

// calcMIAlt's fixed-image bin of row k (the row table's entry) at lambda =
// lam0[j], for the rows that start at lambda = 0: the arithmetic of
// alt_norm_kernel's fixed sample, entry by entry.
template <typename FT>
__global__ __launch_bounds__(256) void alt_fbin0_kernel(const AltParams P, uint8_t *fb0) {
    const TabSampler<FT> fix(P.fq, P.fw, P.fh);
    const float wm1 = (float) (P.fw - 1);
    const float kb = (float) P.bins / 256.0f;
    const size_t n = (size_t) P.nrows * (size_t) P.n0;
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x) {
        const int k = (int) (i / (size_t) P.n0), j = (int) (i % (size_t) P.n0);
        const float ri = (float) (((double) P.rows[k] + 0.5) * (double) (P.fh - 1));
        const float li = P.lam0[j] * wm1;
        fb0[i] = (uint8_t) alt_bin<FT>(fix.at_inside(li, ri), P.bins, kb);
    }
}

// ------------------------------------------------------------------ host

template <typename T>
__global__ void alt_pad_kernel(const T *src, T *dst, size_t n, size_t total) {
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t) gridDim.x * blockDim.x)
        dst[i] = i < n ? src[i] : (T) 0;
}

__global__ void alt_key_kernel(unsigned long long *key) {
    if (threadIdx.x == 0) *key = ~0ull;
}

static hipError_t pad_image(const cgs_image *im, void *dst, hipStream_t s) {
    const size_t n = (size_t) im->width * im->height * im->channels;
    const size_t total = n + mov_pad_elems(im->width, im->channels);
    if (im->pixtype == CGS_PIX_U8)
        hipLaunchKernelGGL(alt_pad_kernel<uint8_t>, dim3(256), dim3(256), 0, s, (const uint8_t *) im->data,
                           (uint8_t *) dst, n, total);
    else
        hipLaunchKernelGGL(alt_pad_kernel<float>, dim3(256), dim3(256), 0, s, (const float *) im->data,
                           (float *) dst, n, total);
    return hipGetLastError();
}

// Exact tap table of a one-channel image.
static hipError_t tap_table(const cgs_image *im, void *dst, hipStream_t s) {
    if (im->pixtype == CGS_PIX_U8)
        hipLaunchKernelGGL(tap_u8_kernel, dim3(1024), dim3(256), 0, s, (const uint8_t *) im->data, (uint32_t *) dst,
                           im->width, im->height);
    else
        hipLaunchKernelGGL(tap_f32_kernel, dim3(1024), dim3(256), 0, s, (const float *) im->data, (float4 *) dst,
                           im->width, im->height);
    return hipGetLastError();
}

template <typename MT, typename FT, bool ONE>
static hipError_t alt_launch(const AltParams &P, hipStream_t s) {
    const size_t lds = alt_lds_bytes(P.bins, P.fh);
    const bool pow2 = (P.mw & (P.mw - 1)) == 0 && (P.mh & (P.mh - 1)) == 0;
    if (P.cost == CGS_COST_MI_ALT) {
        // POW2 = ONE && pow2 (only the one-channel walk reads it)
        if (ONE && pow2)
            hipLaunchKernelGGL((alt_norm_kernel<MT, FT, ONE, ONE>), dim3(P.grid.n), dim3(kThreads), lds, s, P);
        else
            hipLaunchKernelGGL((alt_norm_kernel<MT, FT, ONE, false>), dim3(P.grid.n), dim3(kThreads), lds, s, P);
    } else if (P.cost == CGS_COST_MI_STREAM_ALT)
        hipLaunchKernelGGL((alt_step_mi_kernel<MT, FT, ONE>), dim3(P.grid.n), dim3(kThreads), lds, s, P);
    else if (P.cost == CGS_COST_SQD_ALT)
        hipLaunchKernelGGL((alt_step_kernel<MT, FT, true, ONE>), dim3((P.grid.n + 3) / 4), dim3(kThreads), 0, s, P);
    else
        hipLaunchKernelGGL((alt_step_kernel<MT, FT, false, ONE>), dim3((P.grid.n + 3) / 4), dim3(kThreads), 0, s, P);
    return hipGetLastError();
}

template <typename MT, typename FT>
static hipError_t alt_launch_c(const AltParams &P, hipStream_t s) {
    return P.C == 1 ? alt_launch<MT, FT, true>(P, s) : alt_launch<MT, FT, false>(P, s);
}

static hipError_t alt_dispatch(const AltParams &P, int mpix, int fpix, hipStream_t s) {
    if (mpix == CGS_PIX_U8 && fpix == CGS_PIX_U8) return alt_launch_c<uint8_t, uint8_t>(P, s);
    if (mpix == CGS_PIX_F32 && fpix == CGS_PIX_F32) return alt_launch_c<float, float>(P, s);
    if (mpix == CGS_PIX_U8) return alt_launch_c<uint8_t, float>(P, s);
    return alt_launch_c<float, uint8_t>(P, s);
}

// calcMIAlt's host tables, in the reference's own loops: the row values
// (row = -0.5f; row += 1/(float)(H_f - 1) while row < 0.5,
// src/gpu/cudaErrorFunction_mi.cuh:162) and the fixed area after n steps
// (fix_area += 1.0 / (W_f H_f) through binary64, rounded to binary32, :198)
// for n up to rows * (W_f + 2).
static void mi_alt_tables(int fw, int fh, std::vector<float> &rows, std::vector<float> &area,
                          std::vector<float> &lam0, int *n0) {
    rows.clear();
    const float step = 1 / ((float) (fh - 1));
    for (float row = -0.5f; row < 0.5; row += step) rows.push_back(row);
    // a row that starts at lambda = 0 walks this sequence (src/gpu/cudaErrorFunction_mi.cuh:177)
    lam0.clear();
    const float sc = 1 / ((float) (fw - 1));
    for (float lam = 0.0f; lam < 1.0f && lam0.size() <= (size_t) fw + 1; lam += sc) lam0.push_back(lam);
    // then 2.0f (no row reaches it) through the pass after the one that holds
    // entry n0, which the kernel loads while it walks that pass
    *n0 = (int) lam0.size();
    lam0.resize(((size_t) *n0 / 64 + 2) * 64, 2.0f);
    const size_t n = rows.size() * ((size_t) fw + 2) + 1;
    area.resize(n);
    const double d = 1.0 / (double) (fw * fh);
    float fa = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        area[i] = fa;
        fa = (float) ((double) fa + d);
    }
}

cgs_status run_alt_search(Workspace *ws, const cgs_grid *grid, const cgs_image *moved, const cgs_image *fixed,
                          const cgs_search_opts *o, float *result, cgs_argmin *out, hipStream_t stream,
                          uint32_t *hist_dump_host) {
    if (grid->dim != 8) return set_error(CGS_ERR_INVALID_ARGUMENT, "line-traversal costs take 8 parameters");
    const size_t esz_m = moved->pixtype == CGS_PIX_U8 ? 1 : 4, esz_f = fixed->pixtype == CGS_PIX_U8 ? 1 : 4;
    const size_t nm = (size_t) moved->width * moved->height * moved->channels + mov_pad_elems(moved->width, moved->channels);
    const size_t nf = (size_t) fixed->width * fixed->height * fixed->channels + mov_pad_elems(fixed->width, fixed->channels);
    const int levels0 = (int) grid->count[0];
    const bool mi = is_mi_cost(o->cost);
    const size_t B2 = (size_t) o->mi_bins * o->mi_bins;
    const bool one = moved->channels == 1;
    const size_t tm = one ? ((size_t) moved->width * moved->height + 2 * (size_t) moved->width + 2) *
                                (moved->pixtype == CGS_PIX_U8 ? 4 : 16) : 0;
    const size_t tf = one ? ((size_t) fixed->width * fixed->height + 2 * (size_t) fixed->width + 2) *
                                (fixed->pixtype == CGS_PIX_U8 ? 4 : 16) : 0;
    cgs_status st;
    if ((st = ws->key.ensure(64)) != CGS_OK || (st = ws->key_host.ensure(64, true)) != CGS_OK ||
        (st = ws->sincos.ensure(8 * (size_t) levels0)) != CGS_OK ||
        (st = ws->sincos_host.ensure(8 * (size_t) levels0, true)) != CGS_OK)
        return st;
    if (one ? ((st = ws->quads.ensure(tm)) != CGS_OK || (st = ws->fixq.ensure(tf)) != CGS_OK)
            : ((st = ws->mov.ensure(nm * esz_m)) != CGS_OK || (st = ws->fixp.ensure(nf * esz_f)) != CGS_OK))
        return st;
    if (hist_dump_host && mi && (st = ws->hist.ensure(B2 * sizeof(uint32_t))) != CGS_OK) return st;
    if (mi && alt_lds_bytes(o->mi_bins, fixed->height) > 64 * 1024)
        return set_error(CGS_ERR_TOO_LARGE, "line-traversal MI: the histogram and row-block sums exceed 64 KiB of LDS");
    unsigned long long *key = (unsigned long long *) ws->key.ptr;
    hipLaunchKernelGGL(alt_key_kernel, dim3(1), dim3(64), 0, stream, key);
    if (one) {
        HIP_TRY(tap_table(moved, ws->quads.ptr, stream));
        HIP_TRY(tap_table(fixed, ws->fixq.ptr, stream));
    } else {
        HIP_TRY(pad_image(moved, ws->mov.ptr, stream));
        HIP_TRY(pad_image(fixed, ws->fixp.ptr, stream));
    }

    AltParams P{};
    if (o->cost == CGS_COST_MI_ALT) {
        if (fixed->width < 2 || fixed->height < 2)
            return set_error(CGS_ERR_INVALID_ARGUMENT, "calcMIAlt needs a fixed image of at least 2 x 2");
        std::vector<float> rows, area;
        std::vector<float> lam0;
        int n0 = 0;
        mi_alt_tables(fixed->width, fixed->height, rows, area, lam0, &n0);
        if (area.size() > (size_t) INT32_MAX) return set_error(CGS_ERR_TOO_LARGE, "calcMIAlt: fixed image too large");
        const size_t bytes = sizeof(float) * (rows.size() + area.size() + lam0.size());
        if (ws->alt_tab_fw != fixed->width || ws->alt_tab_fh != fixed->height) {
            if ((st = ws->alt_tab.ensure(bytes)) != CGS_OK || (st = ws->alt_tab_host.ensure(bytes, true)) != CGS_OK)
                return st;
            HIP_TRY(hipStreamSynchronize(stream));   // the pinned staging buffer may still feed an earlier copy
            float *h = (float *) ws->alt_tab_host.ptr;
            std::copy(rows.begin(), rows.end(), h);
            std::copy(area.begin(), area.end(), h + rows.size());
            std::copy(lam0.begin(), lam0.end(), h + rows.size() + area.size());
            HIP_TRY(hipMemcpyAsync(ws->alt_tab.ptr, h, bytes, hipMemcpyHostToDevice, stream));
            ws->alt_tab_fw = fixed->width;
            ws->alt_tab_fh = fixed->height;
        }
        P.rows = (const float *) ws->alt_tab.ptr;
        P.nrows = (int) rows.size();
        P.area = P.rows + rows.size();
        P.narea = (int) area.size();
        P.lam0 = P.area + area.size();
        P.n0 = n0;
        P.n0pad = (int) lam0.size();
        if (one) {
            if ((st = ws->alt_fb0.ensure(rows.size() * (size_t) n0)) != CGS_OK) return st;
            P.fb0 = (const uint8_t *) ws->alt_fb0.ptr;
        }
    }
    {
        float *h = (float *) ws->sincos_host.ptr;
        for (int i = 0; i < levels0; ++i) {
            const float theta = fmaf((float) (grid->offset[0] + i), grid->resolution[0], grid->start[0]);
            h[2 * i] = (float) std::cos((double) theta);
            h[2 * i + 1] = (float) std::sin((double) theta);
        }
        HIP_TRY(hipMemcpyAsync(ws->sincos.ptr, h, sizeof(float) * 2 * levels0, hipMemcpyHostToDevice, stream));
    }

    fill_grid_dev(grid, &P.grid);
    P.sincos = (const float2 *) ws->sincos.ptr;
    P.cost = o->cost;
    P.mov = one ? nullptr : ws->mov.ptr;
    P.fix = one ? nullptr : ws->fixp.ptr;
    P.mq = one ? ws->quads.ptr : nullptr;
    P.fq = one ? ws->fixq.ptr : nullptr;
    P.fraw = one ? fixed->data : nullptr;
    P.mw = moved->width;
    P.mh = moved->height;
    P.fw = fixed->width;
    P.fh = fixed->height;
    P.C = moved->channels;
    P.bins = mi ? o->mi_bins : 0;
    P.result = result;
    P.key = key;
    P.tie_mode = o->tie_mode;
    P.hist_dump = (hist_dump_host && mi) ? (uint32_t *) ws->hist.ptr : nullptr;
    if (P.fb0) {   // calcMIAlt's fixed bins along lam0 for this image and bin count
        const size_t n = (size_t) P.nrows * (size_t) P.n0;
        const unsigned blocks = (unsigned) std::min<size_t>((n + 255) / 256, 4096);
        if (fixed->pixtype == CGS_PIX_U8)
            hipLaunchKernelGGL(alt_fbin0_kernel<uint8_t>, dim3(blocks), dim3(256), 0, stream, P, (uint8_t *) ws->alt_fb0.ptr);
        else
            hipLaunchKernelGGL(alt_fbin0_kernel<float>, dim3(blocks), dim3(256), 0, stream, P, (uint8_t *) ws->alt_fb0.ptr);
        HIP_TRY(hipGetLastError());
    }
    const bool timed = (o->flags & CGS_FLAG_TIME_KERNEL) != 0;
    if (timed) HIP_TRY(ws->record_start(stream));
    HIP_TRY(alt_dispatch(P, moved->pixtype, fixed->pixtype, stream));
    if (timed) HIP_TRY(ws->record_stop(stream));
    HIP_TRY(hipMemcpyAsync(ws->key_host.ptr, key, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    if (P.hist_dump)
        HIP_TRY(hipMemcpyAsync(hist_dump_host, ws->hist.ptr, B2 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (timed) HIP_TRY(ws->elapsed_ms());
    if (out) *out = decode_key(*(unsigned long long *) ws->key_host.ptr, o->tie_mode);
    return CGS_OK;
}

}  // namespace cgs
