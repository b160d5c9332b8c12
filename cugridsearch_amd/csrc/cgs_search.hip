// cgs_search.hip -- plain single-pair grid search for gfx950 and its C-ABI.
//
// One kernel launch evaluates every grid point: each 256-thread workgroup
// owns GPT consecutive grid indices, computes their warp coefficients once,
// sweeps the fixed image (every lane a pixel, every pixel against all GPT
// points so the fixed-pixel load and its coordinates are shared), reduces the
// cost terms (wavefront shuffles -> LDS -> one lane), writes the cost tensor
// and folds its minimum into a 64-bit (ordered cost, index) key with one
// global atomicMin.  That replaces evaluationKernel_by_value + the separate
// find_extrema pass of the reference (src/gpu/cudaGridSearch.cuh:287-314,
// include/cudaTensor.cuh:215-265).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "cgs.h"
#include "cgs_device.hpp"
#include "cgs_host.hpp"

namespace cgs {

enum Family { kFamSAD = 0, kFamNCC = 1, kFamMI = 2 };

struct SearchParams {
    GridDev grid;
    const float2 *sincos;   // (cos, sin) per level of axis 0, or null
    int param_mode;
    int cost;
    const void *mov;        // moving image, zero padded by C*(W+1) elements
    int mw, mh, mc;
    const void *fix;
    int fw, fh;
    float cx, cy;
    float *result;
    unsigned long long *key;
    int tie_mode;
    int bins;
    uint32_t *hist_dump;    // histogram of grid point 0 of block 0 (diagnostics)
};

constexpr int kThreads = 256;

// valueAt_nearest_neighbor (include/cudaTensor.cuh:557-559): the truncated
// float linear index of toIndex1d<float> (include/cudaTensor.cuh:171-180).
template <typename MT>
__device__ __forceinline__ float sample_nearest(const MT *img, float X, float Y, int c, int C, int W) {
    int32_t i = c;
    i = (int32_t) fmaf((float) C, X, (float) i);
    i = (int32_t) fmaf((float) (C * W), Y, (float) i);
    return (float) img[i];
}

// valueAt_bilinear (include/cudaTensor.cuh:561-575).  The right/bottom taps
// are read at floor+1; where the coordinate is integral the reference reads
// ceil == floor instead, and the tap's weight is exactly 0, so the value is
// the same.  ceil(x) == W lands on the next row through the linear index
// and ceil(y) == H on the zero pad, as in the reference's linear addressing.
template <typename MT>
__device__ __forceinline__ float sample_bilinear(const MT *img, float X, float Y, int c, int C, int W) {
    const float fx = floorf(X), fy = floorf(Y);
    const int b = ((int) fy * W + (int) fx) * C + c;
    const float tl = (float) img[b], tr = (float) img[b + C];
    const float bl = (float) img[b + C * W], br = (float) img[b + C * W + C];
    const float ax = X - fx, ay = Y - fy;
    const float top = (1.0f - ax) * tl + ax * tr;
    const float bottom = (1.0f - ax) * bl + ax * br;
    return (1.0f - ay) * top + ay * bottom;
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

template <int FAM, int SAMP, typename MT, typename FT, int GPT>
__global__ __launch_bounds__(kThreads) void search_kernel(const SearchParams P) {
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

    if (tid < GPT) {
        const int32_t gi = base + tid < P.grid.n ? base + tid : 0;
        float p[kMaxDim];
        int32_t i0;
        grid_point(P.grid, gi, p, &i0);
        float c = 1.0f, s = 0.0f;
        if (P.sincos) {
            const float2 cs = P.sincos[i0];
            c = cs.x;
            s = cs.y;
        }
        float co[12];
        warp_coeffs(P.param_mode, p, c, s, P.cx, P.cy, co);
#pragma unroll
        for (int k = 0; k < 12; ++k) s_coef[tid][k] = co[k];
    }
    if constexpr (FAM == kFamMI) {
        for (int i = tid; i < GPT * B2; i += kThreads) s_hist[i] = 0u;
    }
    __syncthreads();

    float cf[GPT][12];
#pragma unroll
    for (int g = 0; g < GPT; ++g)
#pragma unroll
        for (int k = 0; k < 12; ++k) cf[g][k] = s_coef[g][k];

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
            const float W = form(cf[g][8], cf[g][9], cf[g][10], cf[g][11], fxp, fyp);
            const float X = form(cf[g][0], cf[g][1], cf[g][2], cf[g][3], fxp, fyp) / W;
            const float Y = form(cf[g][4], cf[g][5], cf[g][6], cf[g][7], fxp, fyp) / W;
            if (X >= 0.0f && Y >= 0.0f && X < Wm && Y < Hm) {
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
                        if (P.cost == CGS_COST_NCC) acc2[g] = fma(fd, md, acc2[g]);
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

    // ---- reduce per grid point: count, and the family's sums
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
            float cost;
            if constexpr (FAM == kFamSAD) {
                if (P.cost == CGS_COST_SAD_SUM) cost = (float) t[0];
                else if (kExactInt) cost = (float) t[0] / (float) t[3];
                else cost = (float) (t[0] / t[3]);
            } else {
                if (P.cost == CGS_COST_NCC) cost = t[3] > 0.0 ? (float) (-t[2] / sqrt(t[0] * t[1])) : 0.0f;
                else cost = t[3] > 0.0 ? (float) (t[2] / sqrt(t[0] * t[1])) : 123123.0f;
            }
            s_cost[tid] = cost;
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

    if (tid == 0) {
        unsigned long long best = ~0ull;
#pragma unroll
        for (int g = 0; g < GPT; ++g) {
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
}

__global__ void fill_u64_kernel(unsigned long long *p, int n, unsigned long long v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_u8_kernel(uint8_t *p, size_t n, uint8_t v) {
    const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// find_extrema over an arbitrary device array: grid-stride keys, workgroup
// minimum, one atomicMin per workgroup.
__global__ __launch_bounds__(256) void extrema_kernel(const float *v, int64_t n, int tie_mode,
                                                      unsigned long long *key) {
    __shared__ unsigned long long s[4];
    unsigned long long best = ~0ull;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        const unsigned long long k = argmin_key(v[i], (uint32_t) i, tie_mode);
        best = k < best ? k : best;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(best, o);
        best = t < best ? t : best;
    }
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) best = s[w] < best ? s[w] : best;
        if (best != ~0ull) atomicMin(key, best);
    }
}

// ------------------------------------------------------------------ host

template <int FAM, int SAMP, typename MT, typename FT, int GPT>
static hipError_t launch_family(const SearchParams &P, hipStream_t stream) {
    const int blocks = (P.grid.n + GPT - 1) / GPT;
    const size_t lds = FAM == kFamMI ? (size_t) GPT * P.bins * P.bins * sizeof(uint32_t) : 0;
    hipLaunchKernelGGL((search_kernel<FAM, SAMP, MT, FT, GPT>), dim3(blocks), dim3(kThreads), lds, stream, P);
    return hipGetLastError();
}

template <int FAM, int SAMP, int G>
static hipError_t dispatch_pix_g(const SearchParams &P, int mpix, int fpix, hipStream_t stream) {
    if (mpix == CGS_PIX_U8 && fpix == CGS_PIX_U8) return launch_family<FAM, SAMP, uint8_t, uint8_t, G>(P, stream);
    if (mpix == CGS_PIX_F32 && fpix == CGS_PIX_F32) return launch_family<FAM, SAMP, float, float, G>(P, stream);
    if (mpix == CGS_PIX_U8 && fpix == CGS_PIX_F32) return launch_family<FAM, SAMP, uint8_t, float, G>(P, stream);
    return launch_family<FAM, SAMP, float, uint8_t, G>(P, stream);
}

// Grid points per workgroup: 8 for the register-accumulating costs; for MI the
// LDS histograms (GPT * bins^2 * 4 B) stay at <= 32 KiB: 8 up to 32 bins, 2 above.
template <int FAM, int SAMP>
static hipError_t dispatch_pix(const SearchParams &P, int mpix, int fpix, hipStream_t stream) {
    if (FAM == kFamMI && P.bins > 32) return dispatch_pix_g<FAM, SAMP, 2>(P, mpix, fpix, stream);
    return dispatch_pix_g<FAM, SAMP, 8>(P, mpix, fpix, stream);
}

static hipError_t dispatch(const SearchParams &P, int sampler, int mpix, int fpix, hipStream_t stream) {
    int fam = kFamSAD;
    if (P.cost == CGS_COST_NCC || P.cost == CGS_COST_SQD) fam = kFamNCC;
    if (P.cost == CGS_COST_MI || P.cost == CGS_COST_MI_PLAIN) fam = kFamMI;
    if (sampler == CGS_SAMPLER_NEAREST) {
        if (fam == kFamSAD) return dispatch_pix<kFamSAD, CGS_SAMPLER_NEAREST>(P, mpix, fpix, stream);
        if (fam == kFamNCC) return dispatch_pix<kFamNCC, CGS_SAMPLER_NEAREST>(P, mpix, fpix, stream);
        return dispatch_pix<kFamMI, CGS_SAMPLER_NEAREST>(P, mpix, fpix, stream);
    }
    if (fam == kFamSAD) return dispatch_pix<kFamSAD, CGS_SAMPLER_BILINEAR>(P, mpix, fpix, stream);
    if (fam == kFamNCC) return dispatch_pix<kFamNCC, CGS_SAMPLER_BILINEAR>(P, mpix, fpix, stream);
    return dispatch_pix<kFamMI, CGS_SAMPLER_BILINEAR>(P, mpix, fpix, stream);
}

// Runs one plain search: pads the moving image into the workspace, uploads
// the (cos, sin) table, clears the key, launches, and reads the key back.
static cgs_status run_search(const cgs_grid *grid, const cgs_image *moved, const cgs_image *fixed,
                             const cgs_search_opts *o, float *result, cgs_argmin *out, hipStream_t stream,
                             uint32_t *hist_dump_host) {
    cgs_status st = validate_search(grid, moved, fixed, o);
    if (st != CGS_OK) return st;
    Workspace *ws = nullptr;
    st = workspace_for_current_device(&ws);
    if (st != CGS_OK) return st;
    std::lock_guard<std::mutex> lock(ws->mu);

    const size_t esz = moved->pixtype == CGS_PIX_U8 ? 1 : 4;
    const size_t nmov = (size_t) moved->width * moved->height * moved->channels;
    const size_t npad = (size_t) moved->channels * (moved->width + 1);
    const int levels0 = (o->param_mode == CGS_PARAM_CENTRED || o->param_mode == CGS_PARAM_NONCENTRED)
                            ? (int) grid->count[0] : 0;
    const size_t B2 = (size_t) o->mi_bins * o->mi_bins;
    if ((st = ws->reserve((nmov + npad) * esz, levels0, hist_dump_host ? B2 : 0)) != CGS_OK) return st;

    HIP_TRY(hipMemcpyAsync(ws->mov, moved->data, nmov * esz, hipMemcpyDeviceToDevice, stream));
    {
        const size_t nb = npad * esz;
        hipLaunchKernelGGL(fill_u8_kernel, dim3((unsigned) ((nb + 255) / 256)), dim3(256), 0, stream,
                           (uint8_t *) ws->mov + nmov * esz, nb, (uint8_t) 0);
    }
    if (levels0 > 0) {
        for (int i = 0; i < levels0; ++i) {
            const float theta = fmaf((float) i, grid->resolution[0], grid->start[0]);
            ws->sincos_host[2 * i] = (float) std::cos((double) theta);
            ws->sincos_host[2 * i + 1] = (float) std::sin((double) theta);
        }
        HIP_TRY(hipMemcpyAsync(ws->sincos, ws->sincos_host, sizeof(float) * 2 * levels0, hipMemcpyHostToDevice,
                               stream));
    }
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, stream, ws->key, 1, ~0ull);

    SearchParams P{};
    fill_grid_dev(grid, &P.grid);
    P.sincos = levels0 > 0 ? (const float2 *) ws->sincos : nullptr;
    P.param_mode = o->param_mode;
    P.cost = o->cost;
    P.mov = ws->mov;
    P.mw = moved->width;
    P.mh = moved->height;
    P.mc = moved->channels;
    P.fix = fixed->data;
    P.fw = fixed->width;
    P.fh = fixed->height;
    P.cx = (float) moved->width / 2;
    P.cy = (float) moved->height / 2;
    P.result = result;
    P.key = ws->key;
    P.tie_mode = o->tie_mode;
    P.bins = o->mi_bins;
    P.hist_dump = hist_dump_host ? ws->hist : nullptr;
    const bool timed = (o->flags & CGS_FLAG_TIME_KERNEL) != 0;
    if (timed) HIP_TRY(ws->record_start(stream));
    HIP_TRY(dispatch(P, o->sampler, moved->pixtype, fixed->pixtype, stream));
    if (timed) HIP_TRY(ws->record_stop(stream));
    HIP_TRY(hipMemcpyAsync(ws->key_host, ws->key, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    if (hist_dump_host)
        HIP_TRY(hipMemcpyAsync(hist_dump_host, ws->hist, B2 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (timed) HIP_TRY(ws->elapsed_ms());
    if (out) *out = decode_key(*ws->key_host, o->tie_mode);
    return CGS_OK;
}

}  // namespace cgs

using namespace cgs;

extern "C" {

const char *cgs_version(void) { return "cgs 0.1.0 gfx950"; }

void cgs_default_opts(cgs_search_opts *o) {
    if (!o) return;
    o->cost = CGS_COST_SAD;
    o->sampler = CGS_SAMPLER_BILINEAR;
    o->param_mode = CGS_PARAM_CENTRED;
    o->tie_mode = CGS_TIE_LOWEST;
    o->mi_bins = 32;
    o->flags = 0;
}

cgs_status cgs_search(const cgs_grid *grid, const cgs_image *moved, const cgs_image *fixed,
                      const cgs_search_opts *opts, float *result, cgs_argmin *argmin_out, void *stream) {
    clear_error();
    return run_search(grid, moved, fixed, opts, result, argmin_out, (hipStream_t) stream, nullptr);
}

cgs_status cgs_eval_point(const cgs_image *moved, const cgs_image *fixed, const cgs_search_opts *opts,
                          int32_t dim, const float *params, float *cost_out, uint32_t *hist_out, void *stream) {
    clear_error();
    if (!params || !cost_out) return set_error(CGS_ERR_INVALID_ARGUMENT, "null params/cost_out");
    cgs_grid g;
    float cnt[CGS_MAX_DIM];
    for (int i = 0; i < CGS_MAX_DIM; ++i) cnt[i] = 1.0f;
    cgs_status st = cgs_grid_init(&g, dim, params, params, cnt);
    if (st != CGS_OK) return st;
    cgs_argmin am;
    float *dres = nullptr;
    if (hipMalloc(&dres, sizeof(float)) != hipSuccess) return set_error(CGS_ERR_HIP, "hipMalloc");
    st = run_search(&g, moved, fixed, opts, dres, &am, (hipStream_t) stream,
                    (opts && (opts->cost == CGS_COST_MI || opts->cost == CGS_COST_MI_PLAIN)) ? hist_out : nullptr);
    if (st == CGS_OK && hipMemcpy(cost_out, dres, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        st = set_error(CGS_ERR_HIP, "hipMemcpy");
    hipFree(dres);
    return st;
}

cgs_status cgs_find_extrema(const float *values, int64_t n, int32_t tie_mode, cgs_argmin *out, void *stream) {
    clear_error();
    if (!values || !out || n <= 0) return set_error(CGS_ERR_INVALID_ARGUMENT, "find_extrema: bad arguments");
    if (n > 2147483647LL) return set_error(CGS_ERR_GRID_TOO_LARGE, "find_extrema: n beyond int32");
    if (tie_mode != CGS_TIE_LOWEST && tie_mode != CGS_TIE_REFERENCE)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "bad tie mode");
    Workspace *ws = nullptr;
    cgs_status st = workspace_for_current_device(&ws);
    if (st != CGS_OK) return st;
    std::lock_guard<std::mutex> lock(ws->mu);
    hipStream_t s = (hipStream_t) stream;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, ws->key, 1, ~0ull);
    int blocks = (int) std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(extrema_kernel, dim3(blocks), dim3(256), 0, s, values, n, tie_mode, ws->key);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ws->key_host, ws->key, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *out = decode_key(*ws->key_host, tie_mode);
    return CGS_OK;
}

}  // extern "C"
