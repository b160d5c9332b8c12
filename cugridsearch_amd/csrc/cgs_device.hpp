// cgs_device.hpp -- device building blocks of the gfx950 grid search.
//
// Every translation unit that includes this file is compiled with
// -ffp-contract=off: each + - * / below is one IEEE binary32 operation and
// fmaf() appears only where written.  Division and sqrt are correctly
// rounded (HIP's default).  These are the declared semantics the CPU oracle
// restates bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cgs {

constexpr int kMaxDim = 8;

// Grid decode state (CudaGrid, src/gpu/cudaGridSearch.cuh:74-187).
struct GridDev {
    float start[kMaxDim];
    float res[kMaxDim];
    int32_t incr[kMaxDim];
    int32_t off[kMaxDim];     // first sample index per axis (grid sub-boxes)
    int32_t count[kMaxDim];   // samples per axis
    int32_t wincr[kMaxDim];   // stride of each axis in work-index order
    int32_t permuted;         // 0: the work index is the grid index
    int32_t dim;
    int32_t n;
};

// indexToGridPoint (src/gpu/cudaGridSearch.cuh:160-187): mixed-radix decode,
// axis 0 fastest; the point is fmaf(off + i, res, start); *i0 is the local
// axis-0 index.
__device__ __forceinline__ void grid_point(const GridDev &g, int32_t index, float *p, int32_t *i0) {
    int32_t rem = index;
    int32_t idx[kMaxDim];
#pragma unroll
    for (int a = kMaxDim - 1; a >= 0; --a) {
        if (a < g.dim) {
            idx[a] = rem / g.incr[a];
            rem -= idx[a] * g.incr[a];
        } else {
            idx[a] = 0;
        }
    }
#pragma unroll
    for (int a = 0; a < kMaxDim; ++a)
        p[a] = (a < g.dim) ? fmaf((float) (idx[a] + g.off[a]), g.res[a], g.start[a]) : 0.0f;
    *i0 = idx[0];
}

// A workgroup's grid points are consecutive work indices.  The work index
// enumerates the grid with its own axis order (wincr), so that points sharing
// a workgroup can be chosen for locality of their taps; the grid index, which
// orders results and ties, is recovered from the axis indices.  Returns the
// grid index and fills the point as grid_point does.
__device__ __forceinline__ int32_t work_point(const GridDev &g, int32_t w, float *p, int32_t *i0) {
    if (!g.permuted) {
        grid_point(g, w, p, i0);
        return w;
    }
    int32_t idx[kMaxDim];
    int32_t gi = 0;
#pragma unroll
    for (int a = 0; a < kMaxDim; ++a) {
        idx[a] = (a < g.dim) ? (w / g.wincr[a]) % g.count[a] : 0;
        gi += (a < g.dim) ? idx[a] * g.incr[a] : 0;
    }
#pragma unroll
    for (int a = 0; a < kMaxDim; ++a)
        p[a] = (a < g.dim) ? fmaf((float) (idx[a] + g.off[a]), g.res[a], g.start[a]) : 0.0f;
    *i0 = idx[0];
    return gi;
}

enum ParamMode { kCentred = 0, kNonCentred = 1, kRawH = 2, kTranslation = 3 };

// parametersToHomography (src/gpu/cudaImageFunctions.cuh:216-252), the
// reference's expressions in their own association order; c, s are the
// binary32 roundings of cos/sin(theta) evaluated in binary64 on the host.
__device__ __forceinline__ void homography_centred(const float *p, float c, float s, float cx, float cy,
                                                   float *h) {
    const float scaleX = p[1], scaleY = p[2], shearXY = p[3];
    const float translateX = p[4], translateY = p[5], keystoneX = p[6], keystoneY = p[7];
    float h11 = scaleX*(c*(cx*keystoneX + 1) + cx*keystoneY*s);
    float h12 = scaleY*(shearXY*(c*(cx*keystoneX + 1) + cx*keystoneY*s) - s*(cx*keystoneX + 1) + cx*keystoneY*c);
    float h13 = cx + translateX*(cx*keystoneX + 1) - cx*scaleX*(c*(cx*keystoneX + 1) + cx*keystoneY*s) - cy*scaleY*(shearXY*(c*(cx*keystoneX + 1) + cx*keystoneY*s) - s*(cx*keystoneX + 1) + cx*keystoneY*c) + cx*keystoneY*translateY;
    float h21 = scaleX*(s*(cy*keystoneY + 1) + cy*keystoneX*c);
    float h22 = scaleY*(shearXY*(s*(cy*keystoneY + 1) + cy*keystoneX*c) + c*(cy*keystoneY + 1) - cy*keystoneX*s);
    float h23 = cy + translateY*(cy*keystoneY + 1) - cx*scaleX*(s*(cy*keystoneY + 1) + cy*keystoneX*c) - cy*scaleY*(shearXY*(s*(cy*keystoneY + 1) + cy*keystoneX*c) + c*(cy*keystoneY + 1) - cy*keystoneX*s) + cy*keystoneX*translateX;
    float h31 = scaleX*(keystoneX*c + keystoneY*s);
    float h32 = scaleY*(shearXY*(keystoneX*c + keystoneY*s) + keystoneY*c - keystoneX*s);
    float h33 = keystoneX*translateX + keystoneY*translateY - cy*scaleY*(shearXY*(keystoneX*c + keystoneY*s) + keystoneY*c - keystoneX*s) - cx*scaleX*(keystoneX*c + keystoneY*s) + 1;
    h[0] = h11 / h33; h[1] = h12 / h33; h[2] = h13 / h33;
    h[3] = h21 / h33; h[4] = h22 / h33; h[5] = h23 / h33;
    h[6] = h31 / h33; h[7] = h32 / h33;
}

// Coefficients {A, B, C} of the three projective forms, evaluated per pixel
// as F(x, y) = fmaf(B, y, fmaf(A, x, C)) for X (k=0), Y (k=1), W (k=2); the
// sample point is (X/W, Y/W).  Inverse warp: calcNewCoordH
// (src/gpu/cudaImageFunctions.cuh:277-285), whose (A*x + B*y + P) - Q is
// declared here with C = P - Q rounded once and the two products fused.
__device__ __forceinline__ void warp_coeffs(int mode, const float *p, float c, float s, float cx, float cy,
                                            float *coef) {
    if (mode == kTranslation) {
        coef[0] = 1.0f; coef[1] = 0.0f; coef[2] = p[0];
        coef[3] = 0.0f; coef[4] = 1.0f; coef[5] = p[1];
        coef[6] = 0.0f; coef[7] = 0.0f; coef[8] = 1.0f;
        return;
    }
    float h[8];
    if (mode == kCentred) {
        homography_centred(p, c, s, cx, cy, h);
    } else if (mode == kNonCentred) {
        // src/gpu/cudaErrorFunctions.cuh:50-57
        const float scaleX = p[1], scaleY = p[2], shearXY = p[3];
        h[0] = scaleX * c;
        h[1] = scaleY * shearXY * c - scaleY * s;
        h[2] = scaleX * p[4];
        h[3] = scaleX * s;
        h[4] = scaleY * shearXY * s + scaleY * c;
        h[5] = scaleY * p[5];
        h[6] = p[6];
        h[7] = p[7];
    } else {
        // forward raw H: src/gpu/cudaErrorFunctions.cuh:205-207
        coef[0] = p[0]; coef[1] = p[1]; coef[2] = p[2];
        coef[3] = p[3]; coef[4] = p[4]; coef[5] = p[5];
        coef[6] = p[6]; coef[7] = p[7]; coef[8] = 1.0f;
        return;
    }
    const float h11 = h[0], h12 = h[1], h13 = h[2], h21 = h[3], h22 = h[4], h23 = h[5], h31 = h[6], h32 = h[7];
    coef[0] = h22 - h23 * h32;       coef[1] = h13 * h32 - h12;       coef[2] = h12 * h23 - h13 * h22;
    coef[3] = h23 * h31 - h21;       coef[4] = h11 - h13 * h31;       coef[5] = h13 * h21 - h11 * h23;
    coef[6] = h21 * h32 - h22 * h31; coef[7] = h12 * h31 - h11 * h32; coef[8] = h11 * h22 - h12 * h21;
}

__device__ __forceinline__ float form(float a, float b, float c, float x, float y) {
    return fmaf(b, y, fmaf(a, x, c));
}

// 0 <= n/w < lim decided on the exact quotient: fmaf(-lim, w, n) is the
// exact n - lim*w rounded once, so its sign is exact.  CudaImage::inImage
// (include/cudaTensor.cuh:548-550) applied to the exact quotient.
__device__ __forceinline__ bool quotient_in(float n, float w, float lim) {
    if (w > 0.0f) return n >= 0.0f && fmaf(-lim, w, n) < 0.0f;
    if (w < 0.0f) return n <= 0.0f && fmaf(-lim, w, n) > 0.0f;
    return false;
}

// Order-preserving map of a binary32 cost to 32 bits; NaN and +inf map to
// the all-ones key, which never wins (reduceMin6's strict '<' against an
// initial +inf, include/cudaTensorKernels.cuh:148-160).
__device__ __forceinline__ uint32_t ordered_key(float v) {
    if (!(v < __builtin_huge_valf())) return 0xFFFFFFFFu;
    uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Rank of a grid index in find_extrema's tie order: block (idx / 1024), the
// bit-reversed 9-bit thread id, then the half of the block it was read in.
__device__ __forceinline__ uint32_t reference_rank(uint32_t idx) {
    const uint32_t t = idx & 511u, half = (idx >> 9) & 1u;
    return (idx & ~1023u) | ((__brev(t) >> 23) << 1) | half;
}

__device__ __forceinline__ unsigned long long argmin_key(float v, uint32_t idx, int tie_mode) {
    const uint32_t k = ordered_key(v);
    if (k == 0xFFFFFFFFu) return ~0ull;
    const uint32_t low = tie_mode ? reference_rank(idx) : idx;
    return ((unsigned long long) k << 32) | low;
}

template <typename T> __device__ __forceinline__ float to_f(T v) { return (float) v; }

}  // namespace cgs
