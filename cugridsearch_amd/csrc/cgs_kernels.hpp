// cgs_kernels.hpp -- the per-workgroup search bodies shared by the plain
// single-pair kernels (cgs_search.hip) and the persistent batch kernel
// (cgs_batch.hip).  Compiled with -ffp-contract=off (see cgs_device.hpp).
//
// Two sweep bodies:
//   fast_block -- bilinear SAD/NCC/SQD on one channel: the in-image decision
//     is exact (sign tests on the exact forms) while the sample point uses
//     v_rcp_f32 and the taps come from a precomputed float4 table; costs agree
//     with the oracle within FP32 rounding of the samples.
//   mi_block   -- bilinear MI on one channel, bit-exact sample values and
//     histogram counts.
// and the generic search_exact_kernel (every cost/sampler/channel count).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "cgs.h"
#include "cgs_device.hpp"

namespace cgs {

enum Family { kFamSAD = 0, kFamNCC = 1, kFamMI = 2 };

struct SearchParams {
    GridDev grid;
    const float2 *sincos;   // (cos, sin) per level of axis 0, or null
    int param_mode;
    int cost;
    const void *mov;        // exact kernel: moving image + zero pad
    const float4 *quads;    // fast kernel: tap table; entry j holds tap L = j - (W+1)
    int mw, mh, mc;
    const void *fix;
    int fw, fh;
    float cx, cy;
    float *result;
    unsigned long long *key;
    int tie_mode;
    int bins;
    uint32_t *hist_dump;    // histogram of grid point 0 of block 0 (diagnostics)
    int ceil_taps;          // MI kernel on a float tap table: read valueAt's ceil taps (a float
                            // source image, which may hold non-finite pixels); 0 for uint8 sources
};

constexpr int kThreads = 256;

// LDS words of the MI kernel's joint histograms: GPT histograms of bins rows,
// each row padded to bins + 1 counters (a column of the histogram then spreads
// over the LDS banks instead of hitting one).  Above 2 grid points per
// workgroup (at most 32 bins) every histogram takes the 32-bin stride, a
// compile-time constant: the sweep reaches point g's histogram through the LDS
// atomic's immediate offset instead of one address add per sample.
// Up to 4 points per workgroup and 32 bins the histograms are held twice,
// one copy for the even lanes and one for the odd (neighbouring pixels, whose
// bins often coincide, then never add to one counter in the same atomic), and
// the two are summed before the entropy step.  Larger histograms keep one
// copy: two would halve the workgroups a CU holds (64 bins: 79 ms with one,
// 115 with two, on 512^2 uint8).
constexpr int kMiHistFixed = 32 * 33;
__host__ __device__ inline int mi_hist_copies(int gpt, int bins) { return gpt <= 4 && bins <= 32 ? 2 : 1; }
__host__ __device__ inline size_t mi_hist_stride(int gpt, int bins) {
    return gpt > 2 ? (size_t) kMiHistFixed : (size_t) bins * (bins + 1);
}
__host__ __device__ inline size_t mi_hist_words(int gpt, int bins) {
    return (size_t) mi_hist_copies(gpt, bins) * gpt * mi_hist_stride(gpt, bins);
}

// Elements of zero padding after the moving image: covers every tap the
// samplers can address for a sample point in [0, W] x [0, H].
__host__ __device__ inline size_t mov_pad_elems(int W, int C) { return (size_t) C * (2 * (size_t) W + 4); }

// valueAt_nearest_neighbor (include/cudaTensor.cuh:557-559): the truncated
// float linear index of toIndex1d<float> (include/cudaTensor.cuh:171-180).
template <typename MT>
__device__ __forceinline__ float sample_nearest(const MT *img, float X, float Y, int c, int C, int W) {
    int32_t i = c;
    i = (int32_t) fmaf((float) C, X, (float) i);
    i = (int32_t) fmaf((float) (C * W), Y, (float) i);
    return (float) img[i];
}

// valueAt_bilinear (include/cudaTensor.cuh:561-575), each lerp declared as
// fmaf(a, q, (1-a)*p).  The right/bottom taps are read at ceil, as the
// reference does: at an integral coordinate that is the floor tap itself
// (weight 0, and a non-finite neighbour does not leak in).  ceil(x) == W
// lands on the next row through the linear index and ceil(y) == H on the
// zero pad, as in the reference's linear addressing.
template <typename MT>
__device__ __forceinline__ float sample_bilinear(const MT *img, float X, float Y, int c, int C, int W) {
    const float fx = floorf(X), fy = floorf(Y);
    const float ax = X - fx, ay = Y - fy;
    const int b = ((int) fy * W + (int) fx) * C + c;
    const int dx = ax != 0.0f ? C : 0, dy = ay != 0.0f ? C * W : 0;
    const float tl = (float) img[b], tr = (float) img[b + dx];
    const float bl = (float) img[b + dy], br = (float) img[b + dy + dx];
    const float top = fmaf(ax, tr, (1.0f - ax) * tl);
    const float bottom = fmaf(ax, br, (1.0f - ax) * bl);
    return fmaf(ay, bottom, (1.0f - ay) * top);
}

__device__ __forceinline__ int mi_bin(float v, int bins) {
    const float t = floorf((float) bins * (v / 256.0f));
    int b = (int) t;
    if (!(t >= 0.0f)) b = 0;
    if (b > bins - 1) b = bins - 1;
    return b;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Sums v over the workgroup; every thread gets the total.  scratch: >= 4 doubles.
__device__ __forceinline__ double block_sum(double v, double *scratch) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    return scratch[0] + scratch[1] + scratch[2] + scratch[3];
}

__device__ __forceinline__ float uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The sweeps read their parameters through uniform(): when the parameters
// come from a descriptor in memory (the batch kernel) rather than kernel
// arguments, this keeps pointers, sizes and the tap table's buffer resource
// in SGPRs instead of VGPRs with a waterfall loop around every buffer load.
template <typename T>
__device__ __forceinline__ T *uniform(T *p) {
    const uint64_t v = (uint64_t) p;
    const uint32_t lo = (uint32_t) __builtin_amdgcn_readfirstlane((int) (uint32_t) v);
    const uint32_t hi = (uint32_t) __builtin_amdgcn_readfirstlane((int) (uint32_t) (v >> 32));
    return (T *) (((uint64_t) hi << 32) | lo);
}

// Thread 0..GPT-1 of the workgroup: warp coefficients of work index base+tid
// (work index 0 for the padding slots past the end of the grid).  Slot 11
// holds the grid index as int bits, -1 for padding.
template <int GPT>
__device__ __forceinline__ void block_coefficients(const SearchParams &P, int32_t base, float (*s_coef)[12]) {
    const int tid = threadIdx.x;
    if (tid < GPT) {
        const bool live = base + tid < P.grid.n;
        float p[kMaxDim];
        int32_t i0;
        const int32_t gi = work_point(P.grid, live ? base + tid : 0, p, &i0);
        s_coef[tid][11] = __int_as_float(live ? gi : -1);
        float c = 1.0f, s = 0.0f;
        if (P.sincos) {
            const float2 cs = P.sincos[i0];
            c = cs.x;
            s = cs.y;
        }
        float co[9];
        warp_coeffs(P.param_mode, p, c, s, P.cx, P.cy, co);
#pragma unroll
        for (int k = 0; k < 9; ++k) s_coef[tid][k] = co[k];
        // W > 0 over the whole fixed image (W is affine: its minimum over the
        // rectangle is at a corner; the margin covers the rounding of the fmas)
        double wmin = 1e300;
        for (int k = 0; k < 4; ++k) {
            const double xx = (k & 1) ? P.fw - 1 : 0, yy = (k & 2) ? P.fh - 1 : 0;
            const double w = (double) co[7] * yy + ((double) co[6] * xx + (double) co[8]);
            wmin = w < wmin ? w : wmin;
        }
        double wmax = -1e300;
        for (int k = 0; k < 4; ++k) {
            const double xx = (k & 1) ? P.fw - 1 : 0, yy = (k & 2) ? P.fh - 1 : 0;
            const double w = (double) co[7] * yy + ((double) co[6] * xx + (double) co[8]);
            wmax = w > wmax ? w : wmax;
        }
        const double scale = fabs((double) co[6]) * P.fw + fabs((double) co[7]) * P.fh + fabs((double) co[8]);
        // bounded X and Y coefficients keep every form finite (no NaN), and
        // W >= 1e-20 keeps fmaf(-lim, W, n) clear of subnormal results
        bool bounded = true;
        for (int k = 0; k < 6; ++k) bounded = bounded && fabsf(co[k]) < 1e30f;
        const bool posw = wmin > 1e-5 * scale && wmin > 1e-20 && scale < 1e30 && bounded;
        s_coef[tid][9] = posw ? 1.0f : 0.0f;
        // W within [1/64, 64] over the image and X, Y numerators of 0 or of
        // magnitude in (2^-90, 2^90): the refined-reciprocal quotient (div_by)
        // then equals IEEE division on every pixel.  Every nonzero X/Y
        // coefficient at least 2^-60 in magnitude keeps the forms' nonzero
        // values (integer x, y) multiples of 2^-83, so at least 2^-83; the
        // corner bound keeps them under 2^89.
        bool divsafe = wmin >= 1.0 / 64 && wmax <= 64.0;
        for (int k = 0; k < 6; k += 3)
            divsafe = divsafe && fabs((double) co[k]) * P.fw + fabs((double) co[k + 1]) * P.fh +
                                         fabs((double) co[k + 2]) < 0x1p89;
        for (int k = 0; k < 6; ++k) divsafe = divsafe && (co[k] == 0.0f || fabsf(co[k]) >= 0x1p-60f);
        s_coef[tid][10] = (posw && divsafe) ? 1.0f : 0.0f;
    }
}

// Final cost of one grid point from its reduced sums (t[0..2]) and sample
// count t[3]; binary64 combine.
template <int FAM>
__device__ __forceinline__ float finish_cost(int cost, const double *t, bool exact_int) {
    if constexpr (FAM == kFamSAD) {
        if (cost == CGS_COST_SAD_SUM) return (float) t[0];
        if (cost == CGS_COST_SAD_STREAM && t[3] == 0.0) return 0.0f;
        if (exact_int) return (float) t[0] / (float) t[3];
        return (float) (t[0] / t[3]);
    } else {
        if (cost == CGS_COST_NCC) return t[3] > 0.0 ? (float) (-t[2] / sqrt(t[0] * t[1])) : 0.0f;
        if (cost == CGS_COST_NCC_STREAM) return (t[0] > 0.0 && t[1] > 0.0) ? (float) (-t[2] / sqrt(t[0] * t[1])) : 0.0f;
        if (cost == CGS_COST_SQD_STREAM) return (float) (t[2] / sqrt(t[0] * t[1]));
        return t[3] > 0.0 ? (float) (t[2] / sqrt(t[0] * t[1])) : 123123.0f;
    }
}

// Writes the costs of the workgroup's GPT points and folds their minimum
// into the global argmin key.  Called by thread 0.
template <int GPT>
__device__ __forceinline__ void emit(const SearchParams &P, const float (*s_coef)[12], const float *s_cost) {
    unsigned long long best = ~0ull;
#pragma unroll
    for (int g = 0; g < GPT; ++g) {
        const int32_t gi = __float_as_int(s_coef[g][11]);
        if (gi >= 0) {
            if (P.result) P.result[gi] = s_cost[g];
            const unsigned long long k = argmin_key(s_cost[g], (uint32_t) gi, P.tie_mode);
            best = k < best ? k : best;
        }
    }
    if (best != ~0ull && best < __hip_atomic_load(P.key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMin(P.key, best);
}

// ------------------------------------------------------------ exact kernel

template <int FAM, int SAMP, typename MT, typename FT, int GPT>
__global__ __launch_bounds__(kThreads) void search_exact_kernel(const SearchParams P) {
    __shared__ float s_coef[GPT][12];
    __shared__ double s_red[4][GPT][4];
    __shared__ float s_cost[GPT];
    __shared__ double s_scr[8];
    __shared__ uint32_t s_px[64], s_py[64];
    __shared__ double s_lpx[64], s_lpy[64];
    extern __shared__ uint32_t s_hist[];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int32_t base = (int32_t) blockIdx.x * GPT;
    const int B2 = P.bins * P.bins;

    block_coefficients<GPT>(P, base, s_coef);
    if constexpr (FAM == kFamMI) {
        for (int i = tid; i < GPT * B2; i += kThreads) s_hist[i] = 0u;
    }
    __syncthreads();

    float cf[GPT][9];
#pragma unroll
    for (int g = 0; g < GPT; ++g)
#pragma unroll
        for (int k = 0; k < 9; ++k) cf[g][k] = s_coef[g][k];

    const MT *mov = (const MT *) P.mov;
    const FT *fix = (const FT *) P.fix;
    const int C = P.mc;
    const float Wm = (float) P.mw, Hm = (float) P.mh;
    const int npix = P.fw * P.fh;
    constexpr bool kExactInt = (FAM == kFamSAD) && (SAMP == CGS_SAMPLER_NEAREST) &&
                               std::is_same<MT, uint8_t>::value && std::is_same<FT, uint8_t>::value;

    double acc0[GPT], acc1[GPT], acc2[GPT];
    uint32_t isum[GPT], cnt[GPT];
#pragma unroll
    for (int g = 0; g < GPT; ++g) { acc0[g] = acc1[g] = acc2[g] = 0.0; isum[g] = cnt[g] = 0u; }

    int x = tid % P.fw, y = tid / P.fw;
    const int step_x = kThreads % P.fw, step_y = kThreads / P.fw;
    for (int p = tid; p < npix; p += kThreads) {
        const float fxp = (float) x, fyp = (float) y;
#pragma unroll
        for (int g = 0; g < GPT; ++g) {
            const float Wn = form(cf[g][6], cf[g][7], cf[g][8], fxp, fyp);
            const float Xn = form(cf[g][0], cf[g][1], cf[g][2], fxp, fyp);
            const float Yn = form(cf[g][3], cf[g][4], cf[g][5], fxp, fyp);
            const float X = Xn / Wn, Y = Yn / Wn;
            // inImage: MI on the rounded quotients (the reference's own test),
            // SAD/NCC/SQD on the exact quotient (the declared rule)
            const bool in = FAM == kFamMI ? (X >= 0.0f && Y >= 0.0f && X < Wm && Y < Hm)
                                          : (quotient_in(Xn, Wn, Wm) && quotient_in(Yn, Wn, Hm));
            if (in) {
                for (int c = 0; c < C; ++c) {
                    const float m = SAMP == CGS_SAMPLER_NEAREST ? sample_nearest(mov, X, Y, c, C, P.mw)
                                                                : sample_bilinear(mov, X, Y, c, C, P.mw);
                    const float f = (float) fix[p * C + c];
                    cnt[g] += 1u;
                    if constexpr (FAM == kFamSAD) {
                        if constexpr (kExactInt) isum[g] += (uint32_t) fabsf(m - f);
                        else acc0[g] += (double) fabsf(m - f);
                    } else if constexpr (FAM == kFamNCC) {
                        const double md = (double) m, fd = (double) f;
                        acc0[g] = fma(fd, fd, acc0[g]);
                        acc1[g] = fma(md, md, acc1[g]);
                        if (P.cost == CGS_COST_NCC || P.cost == CGS_COST_NCC_STREAM) acc2[g] = fma(fd, md, acc2[g]);
                        else acc2[g] = fma(md - fd, md - fd, acc2[g]);
                    } else {
                        atomicAdd(&s_hist[g * B2 + mi_bin(f, P.bins) * P.bins + mi_bin(m, P.bins)], 1u);
                    }
                }
            }
        }
        x += step_x;
        y += step_y;
        if (x >= P.fw) { x -= P.fw; ++y; }
    }

#pragma unroll
    for (int g = 0; g < GPT; ++g) {
        const double v0 = wave_sum(kExactInt ? (double) isum[g] : acc0[g]);
        const double v1 = wave_sum(acc1[g]);
        const double v2 = wave_sum(acc2[g]);
        const double v3 = wave_sum((double) cnt[g]);
        if (lane == 0) {
            s_red[wave][g][0] = v0;
            s_red[wave][g][1] = v1;
            s_red[wave][g][2] = v2;
            s_red[wave][g][3] = v3;
        }
    }
    __syncthreads();

    if constexpr (FAM != kFamMI) {
        if (tid < GPT) {
            double t[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) t[k] = s_red[0][tid][k] + s_red[1][tid][k] + s_red[2][tid][k] + s_red[3][tid][k];
            s_cost[tid] = finish_cost<FAM>(P.cost, t, kExactInt);
        }
    } else {
        // calcMIstream's entropy step (src/gpu/cudaErrorFunction_miStreams.cuh:306-365), binary64.
        const int bins = P.bins;
        for (int g = 0; g < GPT; ++g) {
            const uint32_t *h = s_hist + g * B2;
            const double n = s_red[0][g][3] + s_red[1][g][3] + s_red[2][g][3] + s_red[3][g][3];
            if (tid < bins) {
                uint32_t r = 0;
                for (int j = 0; j < bins; ++j) r += h[tid * bins + j];
                s_px[tid] = r;
            } else if (tid < 2 * bins) {
                uint32_t r = 0;
                const int col = tid - bins;
                for (int i = 0; i < bins; ++i) r += h[i * bins + col];
                s_py[col] = r;
            }
            __syncthreads();
            double hxy = 0.0;   // -p log2 p of the marginals
            if (tid < bins) {
                const double pxi = s_px[tid] / n;
                s_lpx[tid] = pxi > 0 ? log2(pxi) : 0.0;
                if (pxi > 0) hxy -= pxi * s_lpx[tid];
            } else if (tid < 2 * bins) {
                const double pyi = s_py[tid - bins] / n;
                s_lpy[tid - bins] = pyi > 0 ? log2(pyi) : 0.0;
                if (pyi > 0) hxy -= pyi * s_lpy[tid - bins];
            }
            __syncthreads();
            double mi = 0.0;
            for (int j = tid; j < B2; j += kThreads) {
                const int i = j / bins, k = j - i * bins;
                const double pxy = h[j] / n, pxi = s_px[i] / n, pyj = s_py[k] / n;
                if (pxy > 0 && pxi > 0 && pyj > 0) mi += pxy * (log2(pxy) - (s_lpx[i] + s_lpy[k]));
            }
            const double hx = block_sum(tid < bins ? hxy : 0.0, s_scr);
            const double hy = block_sum(tid >= bins && tid < 2 * bins ? hxy : 0.0, s_scr + 4);
            const double mis = block_sum(mi, s_scr);
            if (tid == 0) {
                double r = P.cost == CGS_COST_MI ? -2.0 * mis / (hx + hy) : -mis;
                if (n == 0.0) r = P.cost == CGS_COST_MI ? __builtin_nan("") : -0.0;
                s_cost[g] = (float) r;
            }
            __syncthreads();
        }
        if (P.hist_dump && blockIdx.x == 0)
            for (int j = tid; j < B2; j += kThreads) P.hist_dump[j] = s_hist[j];
    }
    __syncthreads();
    if (tid == 0) emit<GPT>(P, s_coef, s_cost);
}

// ------------------------------------------------------------- fast kernel
//
// Work split: a thread owns one fixed-image column x for a whole sweep, so
// fmaf(A, x, C) of every form is computed once per grid point and each pixel
// costs one fmaf per form.  For W <= 256 the 256 lanes cover 256/W rows per
// step; wider images are swept in 256-column groups.  The B coefficients are
// wave-uniform and held in SGPRs.  Per-thread FP32 partial sums are folded
// into FP64 every kFlush rows.
//
// Taps are read with raw buffer loads from the tap table: the offset of tap
// L is (L + W + 1) * 16 bytes, and a lane outside the image gets an offset
// past the table's end, which the hardware range check turns into zeros --
// so the positive-W sweep has no per-lane branch and issues all GPT loads
// before it consumes any.

constexpr int kFlush = 32;
constexpr uint32_t kOffsetOut = 0x80000000u;   // beyond any tap table

__device__ __forceinline__ int floor_to_int(float v) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// Byte offset of tap (ix, iy) in the tap table: 16 * ((iy + 1) * W + ix + 1)
// = iy * 16W + (16 ix + 16 (W + 1)), one shift-add and one 24-bit mad.
__device__ __forceinline__ uint32_t tap_offset(int ix, int iy, int w16, int base16) {
    uint32_t r;
    const uint32_t t = (uint32_t) ((ix << 4) + base16);
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(iy), "s"(w16), "v"(t));
    return r;
}

// Lane select on a wave mask held in SGPRs: mask bit set -> b, else a.
__device__ __forceinline__ uint32_t mask_select(unsigned long long m, uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_cndmask_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(m));
    return r;
}


__device__ __forceinline__ float4 load_tap(__amdgpu_buffer_rsrc_t rsrc, uint32_t offset) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, offset, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}

// One fixed-image pixel through a buffer resource: offsets past the image
// read 0.
template <typename FT> __device__ __forceinline__ float load_fix(__amdgpu_buffer_rsrc_t rsrc, uint32_t offset);
template <> __device__ __forceinline__ float load_fix<float>(__amdgpu_buffer_rsrc_t rsrc, uint32_t offset) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, offset, 0, 0));
}
template <> __device__ __forceinline__ float load_fix<uint8_t>(__amdgpu_buffer_rsrc_t rsrc, uint32_t offset) {
    return (float) __builtin_amdgcn_raw_buffer_load_b8(rsrc, offset, 0, 0);
}

// acc += v on the lanes of wave mask m (a subset of the active lanes); the
// other lanes keep acc.  One VALU add under a narrowed EXEC instead of a
// select and an add.
__device__ __forceinline__ void masked_add(float &acc, float v, unsigned long long m) {
    unsigned long long saved;
    asm volatile("s_and_saveexec_b64 %1, %3\n\tv_add_f32 %0, %0, %2\n\ts_mov_b64 exec, %1"
                 : "+v"(acc), "=&s"(saved)
                 : "v"(v), "s"(m)
                 : "scc");
}

// The NCC sums of one sample on the lanes of wave mask m, under one narrowed
// EXEC: s0 += f2, s1 = fmaf(v, v, s1), s2 = fmaf(f, v, s2).
__device__ __forceinline__ void masked_ncc(float &s0, float &s1, float &s2, float f2, float v, float f,
                                           unsigned long long m) {
    unsigned long long saved;
    asm volatile("s_and_saveexec_b64 %3, %7\n\tv_add_f32 %0, %0, %4\n\tv_fmac_f32 %1, %5, %5\n\t"
                 "v_fmac_f32 %2, %6, %5\n\ts_mov_b64 exec, %3"
                 : "+v"(s0), "+v"(s1), "+v"(s2), "=&s"(saved)
                 : "v"(f2), "v"(v), "v"(f), "s"(m)
                 : "scc");
}

// The SQD sums of one sample on the lanes of wave mask m: s0 += f2,
// s1 = fmaf(v, v, s1), s2 += dd.
__device__ __forceinline__ void masked_sqd(float &s0, float &s1, float &s2, float f2, float v, float dd,
                                           unsigned long long m) {
    unsigned long long saved;
    asm volatile("s_and_saveexec_b64 %3, %7\n\tv_add_f32 %0, %0, %4\n\tv_fmac_f32 %1, %5, %5\n\t"
                 "v_add_f32 %2, %2, %6\n\ts_mov_b64 exec, %3"
                 : "+v"(s0), "+v"(s1), "+v"(s2), "=&s"(saved)
                 : "v"(f2), "v"(v), "v"(dd), "s"(m)
                 : "scc");
}

// Positive-W in-image test from sign bits.  With Wn > 0, Xn never -0 and no
// NaN: Xn >= 0 and fmaf(-Wm, Wn, Xn) < 0 (the exact-quotient test) holds
// exactly when their sign bits differ, since Xn < 0 forces the second
// negative too (and it is never -0).  Bit 31 of the result is set for lanes
// inside in both axes: ((Yn ^ b) & (Xn ^ a)), the v_bitop3 table 0x28 of
// (src0 ^ src1) & src2.
__device__ __forceinline__ bool inside_posw(float Xn, float Yn, float Wn, float Wm, float Hm) {
    const uint32_t sx = __float_as_uint(Xn) ^ __float_as_uint(fmaf(-Wm, Wn, Xn));
    const uint32_t w =
        __builtin_amdgcn_bitop3_b32(__float_as_uint(Yn), __float_as_uint(fmaf(-Hm, Wn, Yn)), sx, 0x28);
    return (int32_t) w < 0;
}

// Column chunks of the fixed image for the sweeps.  A lane owns one column of
// a chunk and walks its rows; a chunk of cw columns runs kThreads / cw row
// phases at once.  The image is cut into full chunks of kThreads columns (one
// phase), then the remaining columns into k equal chunks of ceil(rem / k),
// k = 1..8 chosen for the fewest row steps (a 300-wide image: 256 columns, then
// 44 columns in 5 phases, instead of a second pass with 44 of 256 lanes busy;
// an 86-wide one: 2 chunks of 43 columns in 5 phases instead of 86 in 2).
// Every value is workgroup-uniform.
struct ColChunks {
    int nfull, ksplit, cwsplit;
    __device__ ColChunks(int fw, int fh) {
        nfull = fw / kThreads;
        const int rem = fw - nfull * kThreads;
        ksplit = rem ? 1 : 0;
        cwsplit = rem;
        int best = rem ? (fh + kThreads / rem - 1) / (kThreads / rem) : 0;
        for (int k = 2; k <= 8 && k <= rem; ++k) {
            const int cw = (rem + k - 1) / k, rpi = kThreads / cw;
            const int chunks = (rem + cw - 1) / cw;   // no empty chunk at the end
            const int steps = chunks * ((fh + rpi - 1) / rpi);
            if (steps < best) {
                best = steps;
                ksplit = chunks;
                cwsplit = cw;
            }
        }
    }
    __device__ int count() const { return nfull + ksplit; }
    __device__ int x0(int c) const { return c < nfull ? c * kThreads : nfull * kThreads + (c - nfull) * cwsplit; }
    __device__ int width(int c) const { return c < nfull ? kThreads : cwsplit; }
};

template <int COSTK, typename FT, int GPT, bool POSW>
__device__ __forceinline__ void fast_sweep(const SearchParams &P, const float (*s_coef)[12], double (*acc)[3],
                                           uint32_t *cnt) {
    const int tid = threadIdx.x;
    const int fw = uniform(P.fw), fh = uniform(P.fh);
    const ColChunks chunks(fw, fh);
    const int mw = uniform(P.mw), mh = uniform(P.mh);
    const float Wm = (float) mw, Hm = (float) mh;
    const FT *fix = uniform((const FT *) P.fix);
    const int w16 = mw * 16;
    const int base16 = (mw + 1) * 16;
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void *) uniform(P.quads), (short) 0, (int) (((size_t) mw * mh + 2 * (size_t) mw + 2) * 16), 0x00020000);
    const __amdgpu_buffer_rsrc_t frsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *) fix, (short) 0, (int) ((size_t) fw * fh * sizeof(FT)), 0x00020000);

    float B[GPT][3];
#pragma unroll
    for (int g = 0; g < GPT; ++g)
#pragma unroll
        for (int k = 0; k < 3; ++k) B[g][k] = uniform(s_coef[g][3 * k + 1]);

    for (int c = 0; c < chunks.count(); ++c) {
        const int x0 = chunks.x0(c), cols = chunks.width(c);
        const int rpi = kThreads / cols;
        const int r0 = tid / cols, xc = tid - r0 * cols;
        const int full = fh / rpi;   // steps in which every lane's row is inside the image
        const int x = x0 + xc;
        const bool valid = (r0 < rpi) && (x < fw);
        const float xf = (float) x;
        float col[GPT][3];
        // + 0.0f turns a -0 into +0 (and changes nothing else), so that the
        // forms fmaf(B, y, col) are never -0
#pragma unroll
        for (int g = 0; g < GPT; ++g)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                col[g][k] = fmaf(uniform(s_coef[g][3 * k]), xf, uniform(s_coef[g][3 * k + 2])) + 0.0f;
        float part[GPT][3];
#pragma unroll
        for (int g = 0; g < GPT; ++g) part[g][0] = part[g][1] = part[g][2] = 0.0f;
        int since = 0;
        // the same trip count on every lane keeps the ballot counts wave-uniform
        const int steps = (fh + rpi - 1) / rpi;
        const unsigned long long valid_mask = __builtin_amdgcn_ballot_w64(valid);
        // byte offset of the lane's pixel; lanes off the image start past its end
        uint32_t foff = valid ? (uint32_t) ((r0 * fw + x) * (int) sizeof(FT)) : kOffsetOut;
        const uint32_t fstep = (uint32_t) (rpi * fw * (int) sizeof(FT));
        // only the step after the full ones (if any) has lanes below the image
        const unsigned long long tail_mask = __builtin_amdgcn_ballot_w64(valid && r0 + full * rpi < fh);
        float yf = (float) r0;   // the row, counted in binary32 (exact below 2^24)
        for (int it = 0; it < steps; ++it, foff += fstep, yf += (float) rpi) {
            const unsigned long long vmask = it < full ? valid_mask : tail_mask;
            const float f = load_fix<FT>(frsrc, foff);   // 0 on lanes off the image
            const float f2 = f * f;
            unsigned long long in[GPT];   // wave mask of lanes whose sample is in the image
            float ax[GPT], ay[GPT];
            float4 q[GPT];
#pragma unroll
            for (int g = 0; g < GPT; ++g) {
                const float Xn = fmaf(B[g][0], yf, col[g][0]);
                const float Yn = fmaf(B[g][1], yf, col[g][1]);
                const float Wn = fmaf(B[g][2], yf, col[g][2]);
                unsigned long long m;
                if constexpr (POSW) {
                    m = __builtin_amdgcn_ballot_w64(inside_posw(Xn, Yn, Wn, Wm, Hm));
                } else {
                    const bool inside = quotient_in(Xn, Wn, Wm) && quotient_in(Yn, Wn, Hm);
                    m = __builtin_amdgcn_ballot_w64(inside);
                }
                m &= vmask;
                in[g] = m;
                cnt[g] += (uint32_t) __builtin_popcountll(m);
                const float r = __builtin_amdgcn_rcpf(Wn);
                const float X = Xn * r, Y = Yn * r;
                int ix = floor_to_int(X), iy = floor_to_int(Y);
                if constexpr (!POSW) {
                    // W may be tiny or negative here: keep the taps inside the table
                    ix = min(max(ix, -1), mw);
                    iy = min(max(iy, -1), mh);
                }
                ax[g] = __builtin_amdgcn_fractf(X);
                ay[g] = __builtin_amdgcn_fractf(Y);
                // lanes outside the image load from any offset (past the table
                // the range check returns zeros); every sum below is masked, so
                // their values, finite or not, reach no sum
                q[g] = load_tap(rsrc, tap_offset(ix, iy, w16, base16));
            }
#pragma unroll
            for (int g = 0; g < GPT; ++g) {
                const float v = fmaf(ay[g], fmaf(ax[g], q[g].w, q[g].z), fmaf(ax[g], q[g].y, q[g].x));
                if constexpr (COSTK == 0) {
                    masked_add(part[g][0], fabsf(v - f), in[g]);
                } else if constexpr (COSTK == 1) {
                    masked_ncc(part[g][0], part[g][1], part[g][2], f2, v, f, in[g]);
                } else {
                    masked_sqd(part[g][0], part[g][1], part[g][2], f2, v, (v - f) * (v - f), in[g]);
                }
            }
            if (++since == kFlush) {
                since = 0;
#pragma unroll
                for (int g = 0; g < GPT; ++g)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        acc[g][k] += (double) part[g][k];
                        part[g][k] = 0.0f;
                    }
            }
        }
#pragma unroll
        for (int g = 0; g < GPT; ++g)
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[g][k] += (double) part[g][k];
    }
}

// Shared memory of one workgroup's GPT grid points (fast and MI paths).
template <int GPT>
struct BlockShared {
    float coef[GPT][12];
    double red[4][GPT][4];
    float cost[GPT];
    double scr[8];
    uint32_t px[64], py[64];
    double lpx[64], lpy[64];
};

// One workgroup's GPT grid points [base, base+GPT) with the fast sweep.
template <int COSTK, typename FT, int GPT>
__device__ __forceinline__ void fast_block(const SearchParams &P, int32_t base, BlockShared<GPT> &S) {
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    block_coefficients<GPT>(P, base, S.coef);
    __syncthreads();
    bool posw = true;
#pragma unroll
    for (int g = 0; g < GPT; ++g) posw = posw && S.coef[g][9] != 0.0f;

    double acc[GPT][3];
    uint32_t cnt[GPT];
#pragma unroll
    for (int g = 0; g < GPT; ++g) { acc[g][0] = acc[g][1] = acc[g][2] = 0.0; cnt[g] = 0u; }
    if (posw) fast_sweep<COSTK, FT, GPT, true>(P, S.coef, acc, cnt);
    else fast_sweep<COSTK, FT, GPT, false>(P, S.coef, acc, cnt);

#pragma unroll
    for (int g = 0; g < GPT; ++g) {
        const double v0 = wave_sum(acc[g][0]);
        const double v1 = wave_sum(acc[g][1]);
        const double v2 = wave_sum(acc[g][2]);
        if (lane == 0) {
            S.red[wave][g][0] = v0;
            S.red[wave][g][1] = v1;
            S.red[wave][g][2] = v2;
            S.red[wave][g][3] = (double) cnt[g];   // wave-uniform ballot count
        }
    }
    __syncthreads();
    if (tid < GPT) {
        double t[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) t[k] = S.red[0][tid][k] + S.red[1][tid][k] + S.red[2][tid][k] + S.red[3][tid][k];
        if constexpr (COSTK == 0) S.cost[tid] = finish_cost<kFamSAD>(P.cost, t, false);
        else S.cost[tid] = finish_cost<kFamNCC>(P.cost, t, false);
    }
    __syncthreads();
    if (tid == 0) emit<GPT>(P, S.coef, S.cost);
    __syncthreads();
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8).  This bijection hands each XCD a contiguous run of tiles instead, so
// the cost-tensor lines that neighbouring tiles fill meet in one XCD's L2
// (and write back whole) rather than being written back in pieces by eight.
__device__ __forceinline__ int32_t xcd_tile(int32_t b, int32_t n) {
    constexpr int32_t kXcds = 8;
    const int32_t xcd = b % kXcds, local = b / kXcds;
    const int32_t per = n / kXcds, rem = n % kXcds;   // XCDs below rem hold per + 1 tiles
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

template <int COSTK, typename FT, int GPT>
__global__ __launch_bounds__(kThreads) void search_fast_kernel(const SearchParams P) {
    __shared__ BlockShared<GPT> S;
    fast_block<COSTK, FT, GPT>(P, xcd_tile((int32_t) blockIdx.x, (int32_t) gridDim.x) * GPT, S);
}

// --------------------------------------------------------------- MI kernel
//
// calcMIstream on one channel with the column-per-thread sweep of the fast
// kernel, but bit-exact: sample points are correctly rounded quotients and the
// bilinear value follows the declared lerps, so every joint-histogram count
// equals the oracle's.  The taps of a one-byte image are packed four to a
// 32-bit word (tl, tr, bl, br), those of a float image four to a float4; both
// tables are indexed like the fast kernel's.  Lanes outside the image load
// whatever taps their offsets reach (zeros past the table) and their counter
// adds are masked off: each grid point's atomic runs under EXEC = its in-image
// lane mask.

struct Recip {
    float d, r;
};

// Reciprocal of d refined by one Newton step (v_rcp_f32, then
// r += r (1 - d r)), and the quotient n / d from it by one correction step.
// The refined r is RN(1/d), the correctly rounded reciprocal, for every
// binary32 mantissa (cgs_check_rcp checks all 2^23 of [1, 2) on the device;
// v_rcp_f32 alone misses about one in nine), and the steps scale exactly
// with d's exponent while 1/d stays normal.  With r = RN(1/d), q = RN(n r) is
// within an ulp of n/d, rem = n - d q is exact (one fma), and RN(q + rem r)
// = RN(n/d) (Markstein's theorem), so div_by is IEEE division for d in
// [2^-6, 2^6] and n == 0 or 2^-90 < |n| < 2^90 (no intermediate underflows
// or overflows); the kernel checks the W range per grid point and uses '/'
// otherwise.
__device__ __forceinline__ Recip refine_rcp(float d) {
    float r = __builtin_amdgcn_rcpf(d);
    const float e = fmaf(-d, r, 1.0f);
    r = fmaf(e, r, r);
    return {d, r};
}

__device__ __forceinline__ float div_by(float n, Recip R) {
    const float q = n * R.r;
    const float rem = fmaf(-R.d, q, n);
    return fmaf(rem, R.r, q);
}

// Exact-bilinear tap tables of a one-channel image (the MI and line-traversal
// kernels): entry j (tap L = j - (W+1))
// holds p[L], p[L+1], p[L+W], p[L+W+1] (0 outside the image) -- four bytes
// packed in a word for uint8 images, a float4 for float images.
static __global__ void tap_u8_kernel(const uint8_t *src, uint32_t *dst, int W, int H) {
    const long n = (long) W * H;
    const long total = n + 2L * W + 2;
    for (long j = (long) blockIdx.x * blockDim.x + threadIdx.x; j < total; j += (long) gridDim.x * blockDim.x) {
        const long L = j - (W + 1);
        auto P = [&](long i) { return (i >= 0 && i < n) ? (uint32_t) src[i] : 0u; };
        dst[j] = P(L) | (P(L + 1) << 8) | (P(L + W) << 16) | (P(L + W + 1) << 24);
    }
}

static __global__ void tap_f32_kernel(const float *src, float4 *dst, int W, int H) {
    const long n = (long) W * H;
    const long total = n + 2L * W + 2;
    for (long j = (long) blockIdx.x * blockDim.x + threadIdx.x; j < total; j += (long) gridDim.x * blockDim.x) {
        const long L = j - (W + 1);
        auto P = [&](long i) { return (i >= 0 && i < n) ? src[i] : 0.0f; };
        dst[j] = make_float4(P(L), P(L + 1), P(L + W), P(L + W + 1));
    }
}

template <typename MT> struct TapTable;
template <> struct TapTable<uint8_t> {
    static constexpr int kBytes = 4;
    __device__ static float4 load(__amdgpu_buffer_rsrc_t rsrc, uint32_t off) {
        const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(rsrc, off, 0, 0);
        return make_float4((float) (w & 255u), (float) ((w >> 8) & 255u), (float) ((w >> 16) & 255u),
                           (float) (w >> 24));
    }
};
template <> struct TapTable<float> {
    static constexpr int kBytes = 16;
    __device__ static float4 load(__amdgpu_buffer_rsrc_t rsrc, uint32_t off) { return load_tap(rsrc, off); }
};

__device__ __forceinline__ int clamp_int(int v, int lo, int hi) {
    int r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(lo), "v"(hi));
    return r;
}

// floor(bins * (v / 256)) clamped to [0, bins-1], as v * scale with scale =
// bins / 256 (exact for bins < 2^24): v / 256 is exact for a normal v, so
// both forms round the same real number bins * v / 256 once; where v / 256
// would round (|v| < 2^-118) both give a t in (-1, 1) whose clamped floor is
// 0, and a NaN gives NaN either way.  U8: v is a bilinear value of a uint8
// image, in [0, 255] up to a few ulps (rounded convex combinations of taps in
// [0, 255]), so t lies in [0, bins) and the clamp does nothing (lanes outside
// the image, whose v may be anything, have their counter replaced by the
// caller).
template <bool U8 = false>
__device__ __forceinline__ int mi_bin_fast(float v, int bins, float scale) {
    const float t = v * scale;
    if constexpr (U8) return floor_to_int(t);
    else return clamp_int(floor_to_int(t), 0, bins - 1);
}

// LDS byte address of a shared-memory pointer, and a 32-bit counter add at
// an LDS byte address.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
__device__ __forceinline__ uint32_t lds_address(uint32_t *p) { return (uint32_t) (uintptr_t) (lds_u32 *) p; }
__device__ __forceinline__ void lds_count(uint32_t addr) { atomicAdd((uint32_t *) (lds_u32 *) (uintptr_t) addr, 1u); }

template <typename MT, typename FT, int GPT, bool FASTDIV>
__device__ __forceinline__ void mi_sweep(const SearchParams &P, const float (*s_coef)[12], uint32_t *hist,
                                         uint32_t *cnt) {
    const int tid = threadIdx.x;
    const int fw = uniform(P.fw), fh = uniform(P.fh);
    const ColChunks chunks(fw, fh);
    const int mw = uniform(P.mw), mh = uniform(P.mh);
    const float Wm = (float) mw, Hm = (float) mh;
    const FT *fix = uniform((const FT *) P.fix);
    constexpr int EB = TapTable<MT>::kBytes;
    const int wE = mw * EB, baseE = (mw + 1) * EB;
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void *) uniform(P.quads), (short) 0, (int) (((size_t) mw * mh + 2 * (size_t) mw + 2) * EB), 0x00020000);
    const __amdgpu_buffer_rsrc_t frsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *) fix, (short) 0, (int) ((size_t) fw * fh * sizeof(FT)), 0x00020000);
    const int bins = uniform(P.bins), RS = bins + 1, HS = (int) mi_hist_stride(GPT, bins);   // rows padded
    const bool ceil_taps = uniform(P.ceil_taps) != 0;
    const float bscale = (float) bins / 256.0f;
    // LDS byte addresses: the lane's copy of the histograms, one row, one histogram
    const uint32_t hbase = __builtin_amdgcn_readfirstlane(lds_address(hist)) +
                           (mi_hist_copies(GPT, bins) == 2 && (threadIdx.x & 1) ? 4u * (uint32_t) (GPT * HS) : 0u);
    const uint32_t row4 = 4u * (uint32_t) RS, hist4 = 4u * (uint32_t) HS;

    float B[GPT][3];
#pragma unroll
    for (int g = 0; g < GPT; ++g)
#pragma unroll
        for (int k = 0; k < 3; ++k) B[g][k] = uniform(s_coef[g][3 * k + 1]);

    for (int c = 0; c < chunks.count(); ++c) {
        const int x0 = chunks.x0(c), cols = chunks.width(c);
        const int rpi = kThreads / cols;
        const int r0 = tid / cols, xc = tid - r0 * cols;
        const int full = fh / rpi;   // steps in which every lane's row is inside the image
        const int x = x0 + xc;
        const bool valid = (r0 < rpi) && (x < fw);
        const float xf = (float) x;
        float col[GPT][3];
        // + 0.0f: the forms are never -0 (a -0 sample coordinate would only
        // change the sign of a zero weight, which no bin sees)
#pragma unroll
        for (int g = 0; g < GPT; ++g)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                col[g][k] = fmaf(uniform(s_coef[g][3 * k]), xf, uniform(s_coef[g][3 * k + 2])) + 0.0f;
        const int steps = (fh + rpi - 1) / rpi;
        const unsigned long long valid_mask = __builtin_amdgcn_ballot_w64(valid);
        const unsigned long long tail_mask = __builtin_amdgcn_ballot_w64(valid && r0 + full * rpi < fh);
        uint32_t foff = valid ? (uint32_t) ((r0 * fw + x) * (int) sizeof(FT)) : kOffsetOut;
        const uint32_t fstep = (uint32_t) (rpi * fw * (int) sizeof(FT));
        float yf = (float) r0;
        for (int it = 0; it < steps; ++it, foff += fstep, yf += (float) rpi) {
            const unsigned long long vmask = it < full ? valid_mask : tail_mask;
            const float f = load_fix<FT>(frsrc, foff);   // 0 on lanes off the image
            unsigned long long in[GPT];
            float ax[GPT], ay[GPT];
            float4 q[GPT];
#pragma unroll
            for (int g = 0; g < GPT; ++g) {
                const float Xn = fmaf(B[g][0], yf, col[g][0]);
                const float Yn = fmaf(B[g][1], yf, col[g][1]);
                const float Wn = fmaf(B[g][2], yf, col[g][2]);
                // the sample point: the correctly rounded quotients; in the
                // image (CudaImage::inImage on them, the MI costs' rule) when
                // 0 <= X < W and 0 <= Y < H
                float X, Y;
                if constexpr (FASTDIV) {
                    const Recip R = refine_rcp(Wn);
                    X = div_by(Xn, R);
                    Y = div_by(Yn, R);
                } else {
                    X = Xn / Wn;
                    Y = Yn / Wn;
                }
                int ix = floor_to_int(X), iy = floor_to_int(Y);
                unsigned long long m;
                if constexpr (FASTDIV) {
                    // X and Y are the IEEE quotients here, below 2^96 in
                    // magnitude: 0 <= X < W exactly when 0 <= floor(X) < W, and
                    // v_cvt_flr saturates what lies beyond int32.  One ballot
                    // per compare: each is a v_cmp into an SGPR pair, where a
                    // ballot of the && costs a v_cndmask and a v_cmp more
                    m = __builtin_amdgcn_ballot_w64((uint32_t) ix < (uint32_t) mw) &
                        __builtin_amdgcn_ballot_w64((uint32_t) iy < (uint32_t) mh);
                } else {
                    // W may be tiny, zero or negative: the quotients may be huge,
                    // infinite or NaN (no compare holds); the floors stay near the
                    // table for the lanes whose taps are discarded
                    m = __builtin_amdgcn_ballot_w64(X >= 0.0f) & __builtin_amdgcn_ballot_w64(Y >= 0.0f) &
                        __builtin_amdgcn_ballot_w64(X < Wm) & __builtin_amdgcn_ballot_w64(Y < Hm);
                    ix = clamp_int(ix, -1, mw);
                    iy = clamp_int(iy, -1, mh);
                }
                m &= vmask;
                in[g] = m;
                cnt[g] += (uint32_t) __builtin_popcountll(m);
                ax[g] = __builtin_amdgcn_fractf(X);
                ay[g] = __builtin_amdgcn_fractf(Y);
                uint32_t t, off;
                asm("v_lshl_add_u32 %0, %1, %2, %3" : "=v"(t) : "v"(ix), "i"(EB == 4 ? 2 : 4), "s"(baseE));
                asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(off) : "v"(iy), "s"(wE), "v"(t));
                // lanes outside the image load from any offset (past the table
                // the range check returns zeros); their counts are masked
                q[g] = TapTable<MT>::load(rsrc, off);
            }
            uint32_t raddr;   // the fixed bin's row of the first histogram
            asm("v_mad_u32_u24 %0, %1, %2, %3"
                : "=v"(raddr)
                : "v"(mi_bin_fast(f, bins, bscale)), "s"(row4), "v"(hbase));   // one scalar operand per VOP3
#pragma unroll
            for (int g = 0; g < GPT; ++g) {
                float tr = q[g].y, bl = q[g].z, br = q[g].w;
                if (std::is_same<MT, float>::value && ceil_taps) {
                    // valueAt's ceil taps: at an integral coordinate the floor tap
                    // itself, so a non-finite neighbour stays out (taps from a
                    // uint8 image are finite and non-negative: there the weight-0
                    // tap adds exactly +0 and the selects change nothing)
                    tr = ax[g] != 0.0f ? q[g].y : q[g].x;
                    bl = ay[g] != 0.0f ? q[g].z : q[g].x;
                    br = ay[g] != 0.0f ? (ax[g] != 0.0f ? q[g].w : q[g].z) : tr;
                }
                const float top = fmaf(ax[g], tr, (1.0f - ax[g]) * q[g].x);
                const float bottom = fmaf(ax[g], br, (1.0f - ax[g]) * bl);
                const float v = fmaf(ay[g], bottom, (1.0f - ay[g]) * top);
                uint32_t a;
                asm("v_lshl_add_u32 %0, %1, 2, %2"
                    : "=v"(a)
                    : "v"(mi_bin_fast<std::is_same<MT, uint8_t>::value>(v, bins, bscale)), "v"(raddr));
                if constexpr (GPT > 2) {
                    // EXEC = the in-image lanes; point g's histogram at a constant offset
                    if (__builtin_amdgcn_inverse_ballot_w64(in[g])) lds_count(a + 4u * kMiHistFixed * g);
                } else {
                    if (__builtin_amdgcn_inverse_ballot_w64(in[g])) lds_count(a);
                    raddr += hist4;   // the next grid point's histogram
                }
            }
        }
    }
}

// calcMIstream's entropy step (src/gpu/cudaErrorFunction_miStreams.cuh:306-365)
// in binary64 for grid point g of the workgroup; n = in-image sample count.
// Every thread calls it; thread 0 returns the cost.
// h: bins rows of rs >= bins counters (row = fixed-image bin).
__device__ inline double mi_entropy_cost(const uint32_t *h, int bins, int rs, double n, int cost, uint32_t *s_px,
                                  uint32_t *s_py, double *s_lpx, double *s_lpy, double *s_scr) {
    const int tid = threadIdx.x;
    const int B2 = bins * bins;
    if (tid < bins) {
        uint32_t r = 0;
        for (int j = 0; j < bins; ++j) r += h[tid * rs + j];
        s_px[tid] = r;
    } else if (tid < 2 * bins) {
        uint32_t r = 0;
        const int col = tid - bins;
        for (int i = 0; i < bins; ++i) r += h[i * rs + col];
        s_py[col] = r;
    }
    __syncthreads();
    double hxy = 0.0;   // -p log2 p of the marginals
    if (tid < bins) {
        const double pxi = s_px[tid] / n;
        s_lpx[tid] = pxi > 0 ? log2(pxi) : 0.0;
        if (pxi > 0) hxy -= pxi * s_lpx[tid];
    } else if (tid < 2 * bins) {
        const double pyi = s_py[tid - bins] / n;
        s_lpy[tid - bins] = pyi > 0 ? log2(pyi) : 0.0;
        if (pyi > 0) hxy -= pyi * s_lpy[tid - bins];
    }
    __syncthreads();
    double mi = 0.0;
    for (int j = tid; j < B2; j += kThreads) {
        const int i = j / bins, k = j - i * bins;
        const double pxy = h[i * rs + k] / n, pxi = s_px[i] / n, pyj = s_py[k] / n;
        if (pxy > 0 && pxi > 0 && pyj > 0) mi += pxy * (log2(pxy) - (s_lpx[i] + s_lpy[k]));
    }
    const double hx = block_sum(tid < bins ? hxy : 0.0, s_scr);
    const double hy = block_sum(tid >= bins && tid < 2 * bins ? hxy : 0.0, s_scr + 4);
    const double mis = block_sum(mi, s_scr);
    double r = cost == CGS_COST_MI ? -2.0 * mis / (hx + hy) : -mis;
    if (n == 0.0) r = cost == CGS_COST_MI ? __builtin_nan("") : -0.0;
    __syncthreads();
    return r;
}

// One workgroup's GPT grid points with the MI sweep; hist: mi_hist_words(GPT,
// bins) words of LDS.
template <typename MT, typename FT, int GPT>
__device__ __forceinline__ void mi_block(const SearchParams &P, int32_t base, BlockShared<GPT> &S, uint32_t *hist) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int B2 = P.bins * P.bins, RS = P.bins + 1, HS = (int) mi_hist_stride(GPT, P.bins);
    block_coefficients<GPT>(P, base, S.coef);
    for (int i = tid; i < (int) mi_hist_words(GPT, P.bins); i += kThreads) hist[i] = 0u;
    __syncthreads();
    bool fastdiv = true;
#pragma unroll
    for (int g = 0; g < GPT; ++g) fastdiv = fastdiv && S.coef[g][10] != 0.0f;
    uint32_t cnt[GPT];
#pragma unroll
    for (int g = 0; g < GPT; ++g) cnt[g] = 0u;
    if (fastdiv) mi_sweep<MT, FT, GPT, true>(P, S.coef, hist, cnt);
    else mi_sweep<MT, FT, GPT, false>(P, S.coef, hist, cnt);
    if (lane == 0)
#pragma unroll
        for (int g = 0; g < GPT; ++g) S.red[wave][g][3] = (double) cnt[g];
    __syncthreads();
    if (mi_hist_copies(GPT, P.bins) == 2) {   // the odd lanes' copy into the even lanes'
        for (int i = tid; i < GPT * HS; i += kThreads) hist[i] += hist[GPT * HS + i];
        __syncthreads();
    }
    for (int g = 0; g < GPT; ++g) {
        const double n = S.red[0][g][3] + S.red[1][g][3] + S.red[2][g][3] + S.red[3][g][3];
        const double r = mi_entropy_cost(hist + g * HS, P.bins, RS, n, P.cost, S.px, S.py, S.lpx, S.lpy, S.scr);
        if (tid == 0) S.cost[g] = (float) r;
    }
    if (P.hist_dump && base == 0)
        for (int j = tid; j < B2; j += kThreads) P.hist_dump[j] = hist[(j / P.bins) * RS + j % P.bins];
    __syncthreads();
    if (tid == 0) emit<GPT>(P, S.coef, S.cost);
    __syncthreads();
}

template <typename MT, typename FT, int GPT>
__global__ __launch_bounds__(kThreads) void search_mi_kernel(const SearchParams P) {
    __shared__ BlockShared<GPT> S;
    extern __shared__ uint32_t s_hist[];   // GPT histograms
    mi_block<MT, FT, GPT>(P, xcd_tile((int32_t) blockIdx.x, (int32_t) gridDim.x) * GPT, S, s_hist);
}

}  // namespace cgs
