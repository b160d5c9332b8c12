// cgs_search.hip -- plain single-pair grid search for gfx950 and its C-ABI.
//
// One launch evaluates every grid point.  Each 256-thread workgroup owns GPT
// consecutive grid indices, computes their warp coefficients once, sweeps the
// fixed image (a lane per pixel, every pixel against all GPT points so the
// fixed-pixel load and the column part of the warp are shared), reduces the
// cost terms (wavefront shuffles -> LDS -> one lane), writes the cost tensor
// and folds its minimum into a 64-bit (ordered cost, index) key with one
// global atomicMin.  That replaces evaluationKernel_by_value and the separate
// find_extrema pass of the reference (src/gpu/cudaGridSearch.cuh:287-314,
// include/cudaTensor.cuh:215-265).
//
// The per-workgroup bodies live in cgs_kernels.hpp; this file holds the
// staging kernels, the dispatch and the C-ABI entry points.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "cgs.h"
#include "cgs_device.hpp"
#include "cgs_host.hpp"
#include "cgs_kernels.hpp"

namespace cgs {

// ----------------------------------------------------------- prep kernels

__global__ void fill_u64_kernel(unsigned long long *p, int n, unsigned long long v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Copies the moving image into the workspace followed by its zero pad.
template <typename T>
__global__ void pad_copy_kernel(const T *src, T *dst, size_t n, size_t total) {
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t) gridDim.x * blockDim.x)
        dst[i] = i < n ? src[i] : (T) 0;
}

// Bilinear tap table of a one-channel moving image: entry L (for L in
// [-(W+1), W*H + W]) holds (p[L], p[L+1]-p[L], p[L+W]-p[L],
// p[L+W+1]-p[L+W]-p[L+1]+p[L]) with p = 0 outside [0, W*H), so a sample at
// linear tap L with weights (ax, ay) is q.x + ax q.y + ay (q.z + ax q.w).
template <typename T>
__global__ void quad_kernel(const T *src, float4 *dst, int W, int H) {
    const long n = (long) W * H;
    const long total = n + 2L * W + 2;
    for (long j = (long) blockIdx.x * blockDim.x + threadIdx.x; j < total; j += (long) gridDim.x * blockDim.x) {
        const long L = j - (W + 1);
        auto P = [&](long i) { return (i >= 0 && i < n) ? (float) src[i] : 0.0f; };
        const float tl = P(L), tr = P(L + 1), bl = P(L + W), br = P(L + W + 1);
        dst[j] = make_float4(tl, tr - tl, bl - tl, (br - bl) - (tr - tl));
    }
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
static hipError_t launch_exact(const SearchParams &P, hipStream_t stream) {
    const int blocks = (P.grid.n + GPT - 1) / GPT;
    const size_t lds = FAM == kFamMI ? (size_t) GPT * P.bins * P.bins * sizeof(uint32_t) : 0;
    hipLaunchKernelGGL((search_exact_kernel<FAM, SAMP, MT, FT, GPT>), dim3(blocks), dim3(kThreads), lds, stream, P);
    return hipGetLastError();
}

template <int FAM, int SAMP, int G>
static hipError_t exact_pix(const SearchParams &P, int mpix, int fpix, hipStream_t stream) {
    if (mpix == CGS_PIX_U8 && fpix == CGS_PIX_U8) return launch_exact<FAM, SAMP, uint8_t, uint8_t, G>(P, stream);
    if (mpix == CGS_PIX_F32 && fpix == CGS_PIX_F32) return launch_exact<FAM, SAMP, float, float, G>(P, stream);
    if (mpix == CGS_PIX_U8 && fpix == CGS_PIX_F32) return launch_exact<FAM, SAMP, uint8_t, float, G>(P, stream);
    return launch_exact<FAM, SAMP, float, uint8_t, G>(P, stream);
}

// Grid points per workgroup: 8 for the register-accumulating costs; for MI the
// LDS histograms (GPT * bins^2 * 4 B) stay at <= 32 KiB: 8 up to 32 bins, 2 above.
template <int FAM, int SAMP>
static hipError_t exact_gpt(const SearchParams &P, int mpix, int fpix, hipStream_t stream) {
    if (FAM == kFamMI && P.bins > 32) return exact_pix<FAM, SAMP, 2>(P, mpix, fpix, stream);
    return exact_pix<FAM, SAMP, 8>(P, mpix, fpix, stream);
}

static int family_of(int cost) {
    if (cost == CGS_COST_NCC || cost == CGS_COST_SQD || cost == CGS_COST_NCC_STREAM || cost == CGS_COST_SQD_STREAM)
        return kFamNCC;
    if (cost == CGS_COST_MI || cost == CGS_COST_MI_PLAIN) return kFamMI;
    return kFamSAD;
}

static hipError_t dispatch_exact(const SearchParams &P, int sampler, int mpix, int fpix, hipStream_t stream) {
    const int fam = family_of(P.cost);
    if (sampler == CGS_SAMPLER_NEAREST) {
        if (fam == kFamSAD) return exact_gpt<kFamSAD, CGS_SAMPLER_NEAREST>(P, mpix, fpix, stream);
        if (fam == kFamNCC) return exact_gpt<kFamNCC, CGS_SAMPLER_NEAREST>(P, mpix, fpix, stream);
        return exact_gpt<kFamMI, CGS_SAMPLER_NEAREST>(P, mpix, fpix, stream);
    }
    if (fam == kFamSAD) return exact_gpt<kFamSAD, CGS_SAMPLER_BILINEAR>(P, mpix, fpix, stream);
    if (fam == kFamNCC) return exact_gpt<kFamNCC, CGS_SAMPLER_BILINEAR>(P, mpix, fpix, stream);
    return exact_gpt<kFamMI, CGS_SAMPLER_BILINEAR>(P, mpix, fpix, stream);
}


template <int COSTK, int G>
static hipError_t fast_launch(const SearchParams &P, int fpix, hipStream_t stream) {
    const int blocks = (P.grid.n + G - 1) / G;
    if (fpix == CGS_PIX_U8)
        hipLaunchKernelGGL((search_fast_kernel<COSTK, uint8_t, G>), dim3(blocks), dim3(kThreads), 0, stream, P);
    else
        hipLaunchKernelGGL((search_fast_kernel<COSTK, float, G>), dim3(blocks), dim3(kThreads), 0, stream, P);
    return hipGetLastError();
}

template <int COSTK>
static hipError_t fast_pix(const SearchParams &P, int fpix, int gpt, hipStream_t stream) {
    if (gpt == 8) return fast_launch<COSTK, 8>(P, fpix, stream);
    if (gpt == 6) return fast_launch<COSTK, 6>(P, fpix, stream);
    return fast_launch<COSTK, 4>(P, fpix, stream);
}

static hipError_t dispatch_fast(const SearchParams &P, int fpix, int gpt, hipStream_t stream) {
    if (P.cost == CGS_COST_NCC || P.cost == CGS_COST_NCC_STREAM) return fast_pix<1>(P, fpix, gpt, stream);
    if (P.cost == CGS_COST_SQD || P.cost == CGS_COST_SQD_STREAM) return fast_pix<2>(P, fpix, gpt, stream);
    return fast_pix<0>(P, fpix, gpt, stream);
}

template <typename MT, typename FT, int G>
static hipError_t mi_launch(const SearchParams &P, hipStream_t stream) {
    const int blocks = (P.grid.n + G - 1) / G;
    const size_t lds = mi_hist_words(G, P.bins) * sizeof(uint32_t);
    hipLaunchKernelGGL((search_mi_kernel<MT, FT, G>), dim3(blocks), dim3(kThreads), lds, stream, P);
    return hipGetLastError();
}

// 4 grid points per workgroup up to 32 bins (their histograms in two copies,
// 34 KiB; on 512^2 uint8 at 32 bins 67.5 ms against 71.0 at 6 points and
// 72.0 at 8, which hold one copy), 2 above, and 2 at 8 bins or fewer (where
// most lanes of an atomic share a counter: 69.9 ms against 73.6 at 4); tile_gpt
// 2, 4, 6 or 8 overrides up to 32 bins.
template <typename MT, typename FT>
static hipError_t mi_gpt(const SearchParams &P, int gpt, hipStream_t stream) {
    if (P.bins > 32 || gpt == 2 || (gpt == 0 && P.bins <= 8)) return mi_launch<MT, FT, 2>(P, stream);
    if (gpt == 8) return mi_launch<MT, FT, 8>(P, stream);
    if (gpt == 6) return mi_launch<MT, FT, 6>(P, stream);
    return mi_launch<MT, FT, 4>(P, stream);
}

static hipError_t dispatch_mi(const SearchParams &P, int mpix, int fpix, int gpt, hipStream_t stream) {
    if (mpix == CGS_PIX_U8 && fpix == CGS_PIX_U8) return mi_gpt<uint8_t, uint8_t>(P, gpt, stream);
    if (mpix == CGS_PIX_F32 && fpix == CGS_PIX_F32) return mi_gpt<float, float>(P, gpt, stream);
    if (mpix == CGS_PIX_U8) return mi_gpt<uint8_t, float>(P, gpt, stream);
    return mi_gpt<float, uint8_t>(P, gpt, stream);
}

// Which kernel runs a search: the fast kernel for bilinear SAD/NCC/SQD on one
// channel, the MI kernel for bilinear MI on one channel, the exact kernel for
// everything else (nearest neighbour, several channels).
static KernelKind kernel_kind(const cgs_image *moved, const cgs_search_opts *o) {
    if (o->sampler != CGS_SAMPLER_BILINEAR || moved->channels != 1 || moved->width > 65536 ||
        moved->height > 65536)
        return kKernelExact;
    return family_of(o->cost) == kFamMI ? kKernelMI : kKernelFast;
}

// Device and host buffers one plain search stages its scratch in: the
// argmin key, the moving image's tap table (fast and MI kernels) or padded
// copy (exact kernel), the (cos, sin) table and, for cgs_eval_point, the
// histogram dump.  Host buffers hold the (cos, sin) upload and the key
// read-back until the stream is synchronised.
struct SearchBuffers {
    unsigned long long *key = nullptr;
    void *taps = nullptr;
    void *mov = nullptr;
    float2 *sincos = nullptr;
    uint32_t *hist = nullptr;
    float *sincos_host = nullptr;
    unsigned long long *key_host = nullptr;
};

SearchSizes search_sizes(const cgs_grid *grid, const cgs_image *moved, const cgs_search_opts *o) {
    SearchSizes z;
    z.kind = kernel_kind(moved, o);
    const size_t tap_bytes = z.kind == kKernelFast ? sizeof(float4) : (moved->pixtype == CGS_PIX_U8 ? 4 : 16);
    const size_t esz = moved->pixtype == CGS_PIX_U8 ? 1 : 4;
    const size_t nmov = (size_t) moved->width * moved->height * moved->channels;
    z.nmov = nmov;
    z.ntot = nmov + mov_pad_elems(moved->width, moved->channels);
    const size_t nquad = (size_t) moved->width * moved->height + 2 * (size_t) moved->width + 2;
    z.tap_bytes = z.kind != kKernelExact ? nquad * tap_bytes : 0;
    z.mov_bytes = z.kind == kKernelExact ? z.ntot * esz : 0;
    z.levels0 = (o->param_mode == CGS_PARAM_CENTRED || o->param_mode == CGS_PARAM_NONCENTRED)
                    ? (int) grid->count[0] : 0;
    z.hist_bytes = (size_t) o->mi_bins * o->mi_bins * sizeof(uint32_t);
    return z;
}

// Enqueues one plain search on the stream with its scratch in B: stages the
// moving image, uploads the (cos, sin) table, clears the key, launches and
// copies the key (and the histogram dump) back.  ws, when not null, times the
// search kernel with its events.  The caller synchronises the stream.
static cgs_status enqueue_search(const SearchBuffers &B, const SearchSizes &z, const cgs_grid *grid,
                                 const cgs_image *moved, const cgs_image *fixed, const cgs_search_opts *o,
                                 float *result, hipStream_t stream, uint32_t *hist_dump_host, Workspace *ws) {
    const unsigned prep_blocks = 1024;
    if (z.kind == kKernelMI) {
        if (moved->pixtype == CGS_PIX_U8)
            hipLaunchKernelGGL(tap_u8_kernel, dim3(prep_blocks), dim3(256), 0, stream, (const uint8_t *) moved->data,
                               (uint32_t *) B.taps, moved->width, moved->height);
        else
            hipLaunchKernelGGL(tap_f32_kernel, dim3(prep_blocks), dim3(256), 0, stream, (const float *) moved->data,
                               (float4 *) B.taps, moved->width, moved->height);
    } else if (z.kind == kKernelFast) {
        if (moved->pixtype == CGS_PIX_U8)
            hipLaunchKernelGGL(quad_kernel<uint8_t>, dim3(prep_blocks), dim3(256), 0, stream,
                               (const uint8_t *) moved->data, (float4 *) B.taps, moved->width, moved->height);
        else
            hipLaunchKernelGGL(quad_kernel<float>, dim3(prep_blocks), dim3(256), 0, stream,
                               (const float *) moved->data, (float4 *) B.taps, moved->width, moved->height);
    } else if (moved->pixtype == CGS_PIX_U8) {
        hipLaunchKernelGGL(pad_copy_kernel<uint8_t>, dim3(prep_blocks), dim3(256), 0, stream,
                           (const uint8_t *) moved->data, (uint8_t *) B.mov, z.nmov, z.ntot);
    } else {
        hipLaunchKernelGGL(pad_copy_kernel<float>, dim3(prep_blocks), dim3(256), 0, stream,
                           (const float *) moved->data, (float *) B.mov, z.nmov, z.ntot);
    }
    HIP_TRY(hipGetLastError());
    if (z.levels0 > 0) {
        float *h = B.sincos_host;
        for (int i = 0; i < z.levels0; ++i) {
            const float theta = fmaf((float) (grid->offset[0] + i), grid->resolution[0], grid->start[0]);
            h[2 * i] = (float) std::cos((double) theta);
            h[2 * i + 1] = (float) std::sin((double) theta);
        }
        HIP_TRY(hipMemcpyAsync(B.sincos, h, sizeof(float) * 2 * z.levels0, hipMemcpyHostToDevice, stream));
    }
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, stream, B.key, 1, ~0ull);

    SearchParams P{};
    fill_grid_dev(grid, &P.grid);
    if (z.kind != kKernelExact) {
        int order[kMaxDim];
        if (work_order(grid->dim, o->param_mode, order)) set_work_order(&P.grid, order);
    }
    const int gpt = o->tile_gpt ? o->tile_gpt : default_fast_gpt(P.grid);
    P.sincos = z.levels0 > 0 ? B.sincos : nullptr;
    P.param_mode = o->param_mode;
    P.cost = o->cost;
    P.mov = B.mov;
    P.quads = z.kind != kKernelExact ? (const float4 *) B.taps : nullptr;   // entry 0: tap L = -(W+1)
    P.mw = moved->width;
    P.mh = moved->height;
    P.mc = moved->channels;
    P.fix = fixed->data;
    P.fw = fixed->width;
    P.fh = fixed->height;
    P.cx = (float) moved->width / 2;
    P.cy = (float) moved->height / 2;
    P.result = result;
    P.key = B.key;
    P.tie_mode = o->tie_mode;
    P.bins = o->mi_bins;
    P.hist_dump = hist_dump_host ? B.hist : nullptr;
    P.ceil_taps = moved->pixtype == CGS_PIX_F32;
    const bool timed = ws && (o->flags & CGS_FLAG_TIME_KERNEL) != 0;
    if (timed) HIP_TRY(ws->record_start(stream));
    if (z.kind == kKernelFast) HIP_TRY(dispatch_fast(P, fixed->pixtype, gpt, stream));
    else if (z.kind == kKernelMI) HIP_TRY(dispatch_mi(P, moved->pixtype, fixed->pixtype, o->tile_gpt, stream));
    else HIP_TRY(dispatch_exact(P, o->sampler, moved->pixtype, fixed->pixtype, stream));
    if (timed) HIP_TRY(ws->record_stop(stream));
    HIP_TRY(hipMemcpyAsync(B.key_host, B.key, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    if (hist_dump_host)
        HIP_TRY(hipMemcpyAsync(hist_dump_host, B.hist, z.hist_bytes, hipMemcpyDeviceToHost, stream));
    return CGS_OK;
}

// Runs one plain search with its scratch in the per-device workspace (grown
// on demand), synchronises the stream and decodes the key.
static cgs_status run_search(const cgs_grid *grid, const cgs_image *moved, const cgs_image *fixed,
                             const cgs_search_opts *o, float *result, cgs_argmin *out, hipStream_t stream,
                             uint32_t *hist_dump_host) {
    cgs_status st = validate_search(grid, moved, fixed, o);
    if (st != CGS_OK) return st;
    Workspace *ws = nullptr;
    st = workspace_for_current_device(&ws);
    if (st != CGS_OK) return st;
    std::lock_guard<std::mutex> lock(ws->mu);
    if (is_alt_cost(o->cost)) return run_alt_search(ws, grid, moved, fixed, o, result, out, stream, hist_dump_host);

    const SearchSizes z = search_sizes(grid, moved, o);
    if ((st = ws->key.ensure(64)) != CGS_OK || (st = ws->key_host.ensure(64, true)) != CGS_OK) return st;
    if (z.tap_bytes && (st = ws->quads.ensure(z.tap_bytes)) != CGS_OK) return st;
    if (z.mov_bytes && (st = ws->mov.ensure(z.mov_bytes)) != CGS_OK) return st;
    if (z.levels0 > 0 && ((st = ws->sincos.ensure(8 * (size_t) z.levels0)) != CGS_OK ||
                          (st = ws->sincos_host.ensure(8 * (size_t) z.levels0, true)) != CGS_OK))
        return st;
    if (hist_dump_host && (st = ws->hist.ensure(z.hist_bytes)) != CGS_OK) return st;
    SearchBuffers B;
    B.key = (unsigned long long *) ws->key.ptr;
    B.taps = ws->quads.ptr;
    B.mov = ws->mov.ptr;
    B.sincos = (float2 *) ws->sincos.ptr;
    B.hist = (uint32_t *) ws->hist.ptr;
    B.sincos_host = (float *) ws->sincos_host.ptr;
    B.key_host = (unsigned long long *) ws->key_host.ptr;
    if ((st = enqueue_search(B, z, grid, moved, fixed, o, result, stream, hist_dump_host, ws)) != CGS_OK) return st;
    HIP_TRY(hipStreamSynchronize(stream));
    if (o->flags & CGS_FLAG_TIME_KERNEL) HIP_TRY(ws->elapsed_ms());
    if (out) *out = decode_key(*B.key_host, o->tie_mode);
    return CGS_OK;
}

// The search of cgs_search_noalloc: the same kernels with their scratch
// carved from `scratch` (device, scratch_bytes long, 256-byte aligned), host
// staging on this call's stack, nothing allocated.
cgs_status run_search_in(void *scratch, size_t scratch_bytes, const cgs_grid *grid, const cgs_image *moved,
                         const cgs_image *fixed, const cgs_search_opts *o, float *result, cgs_argmin *out,
                         hipStream_t stream) {
    cgs_status st = validate_search(grid, moved, fixed, o);
    if (st != CGS_OK) return st;
    if (is_alt_cost(o->cost))
        return set_error(CGS_ERR_UNSUPPORTED, "search_noalloc: the line-traversal costs are not supported");
    if (o->flags & CGS_FLAG_TIME_KERNEL)
        return set_error(CGS_ERR_UNSUPPORTED, "search_noalloc: CGS_FLAG_TIME_KERNEL is not supported");
    const SearchSizes z = search_sizes(grid, moved, o);
    constexpr size_t kKeyBytes = 256, kSinCosBytes = 8 * (size_t) kNoAllocMaxLevels0;
    const size_t stage = z.tap_bytes + z.mov_bytes;
    if (z.levels0 > kNoAllocMaxLevels0)
        return set_error(CGS_ERR_TOO_LARGE, "search_noalloc: more than " + std::to_string(kNoAllocMaxLevels0) +
                                                " levels on axis 0");
    if (kKeyBytes + kSinCosBytes + stage > scratch_bytes)
        return set_error(CGS_ERR_TOO_LARGE, "search_noalloc: the moving image needs " + std::to_string(stage) +
                                                " bytes of staging, beyond the " +
                                                std::to_string(scratch_bytes - kKeyBytes - kSinCosBytes) +
                                                " the static scratch holds");
    char *base = (char *) scratch;
    float sincos_host[2 * kNoAllocMaxLevels0];
    unsigned long long key_host = ~0ull;
    SearchBuffers B;
    B.key = (unsigned long long *) base;
    B.sincos = (float2 *) (base + kKeyBytes);
    B.taps = z.tap_bytes ? base + kKeyBytes + kSinCosBytes : nullptr;
    B.mov = z.mov_bytes ? base + kKeyBytes + kSinCosBytes : nullptr;
    B.sincos_host = sincos_host;
    B.key_host = &key_host;
    // the stack buffers stay alive until the stream is synchronised below
    st = enqueue_search(B, z, grid, moved, fixed, o, result, stream, nullptr, nullptr);
    const hipError_t e = hipStreamSynchronize(stream);
    if (st != CGS_OK) return st;
    if (e != hipSuccess) return set_error(CGS_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    if (out) *out = decode_key(key_host, o->tie_mode);
    return CGS_OK;
}

}  // namespace cgs

using namespace cgs;

extern "C" {

// CGS_SOURCE_HASH: the build's hash of the library's sources
// (cugridsearch_amd/build.py source_hash), so that a caller can tell which
// sources a libcgs.so was built from.
#ifndef CGS_SOURCE_HASH
#define CGS_SOURCE_HASH "unknown"
#endif

const char *cgs_version(void) { return "cgs 0.4.0 gfx950 src " CGS_SOURCE_HASH; }

int32_t cgs_abi_version(void) { return CGS_ABI_VERSION; }

void cgs_default_opts(cgs_search_opts *o) {
    if (!o) return;
    o->cost = CGS_COST_SAD;
    o->sampler = CGS_SAMPLER_BILINEAR;
    o->param_mode = CGS_PARAM_CENTRED;
    o->tie_mode = CGS_TIE_LOWEST;
    o->mi_bins = 32;
    o->flags = 0;
    o->tile_gpt = 0;
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
    st = run_search(&g, moved, fixed, opts, nullptr, &am, (hipStream_t) stream,
                    (opts && is_mi_cost(opts->cost)) ? hist_out : nullptr);
    if (st == CGS_OK) *cost_out = am.index < 0 ? NAN : am.value;
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
    if ((st = ws->key.ensure(64)) != CGS_OK || (st = ws->key_host.ensure(64, true)) != CGS_OK) return st;
    hipStream_t s = (hipStream_t) stream;
    unsigned long long *key = (unsigned long long *) ws->key.ptr;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, key, 1, ~0ull);
    const int blocks = (int) std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(extrema_kernel, dim3(blocks), dim3(256), 0, s, values, n, tie_mode, key);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ws->key_host.ptr, key, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *out = decode_key(*(unsigned long long *) ws->key_host.ptr, tie_mode);
    return CGS_OK;
}

}  // extern "C"
