// cgs_image.hip -- the image preprocessing of the reference's multispectral
// matcher (src/cpu/multispectral_ImageMatcher.cu:568-632):
//   CudaImage::filter  (include/cudaTensor.cuh:577-593) with every channel
//     filtered, kernel filterProcess (include/cudaTensorKernels.cuh:353-392);
//   transformPixel     (src/cpu/multispectral_ImageMatcher.cu:411-430), the
//     intensity quantisation.
// One thread per pixel; both are a single pass over the image, bound by HBM
// (one read of the filter footprint from L2/L1, one write per element).
// Compiled with -ffp-contract=off: the filter's "pixVal += f * p" is the
// declared fmaf(f, p, pixVal) (nvcc contracts that statement by default).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cgs.h"
#include "cgs_host.hpp"

namespace cgs {

constexpr int kMaxFilterTaps = 256;

struct FilterTaps {
    float f[kMaxFilterTaps];
};

// The filter's output conversion.  The reference converts with (pixType)
// pixVal; for uint8 an out-of-range float conversion is undefined in C++, and
// is declared here as truncation toward zero saturated to [0, 255].
template <typename T> __device__ __forceinline__ T to_pixel(float v);
template <> __device__ __forceinline__ uint8_t to_pixel<uint8_t>(float v) {
    const float t = truncf(v);
    return (uint8_t) (!(t >= 0.0f) ? 0.0f : (t > 255.0f ? 255.0f : t));   // negatives and NaN -> 0
}
template <> __device__ __forceinline__ float to_pixel<float>(float v) { return v; }

// filterProcess: the output pixel (x, y) correlates the filter with the
// window whose top-left tap is (x - (fw-1)/2, y - (fh-1)/2); a tap outside
// the image reads the centre pixel instead.  The running sum is not reset
// between channels (the reference declares it once per pixel).  Channel c is
// filtered when bit c of filter_mask is set (FILTER) and copied otherwise
// (COPY, which leaves the running sum alone).
template <typename T>
__global__ void filter_kernel(const T *in, T *out, int W, int H, int C, const FilterTaps taps, int fw, int fh,
                              uint32_t filter_mask) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const int x0 = x - (fw - 1) / 2, y0 = y - (fh - 1) / 2;
    const size_t centre = ((size_t) y * W + x) * C;
    float pix = 0.0f;
    for (int c = 0; c < C; ++c) {
        if (!((filter_mask >> c) & 1u)) {
            out[centre + c] = in[centre + c];
            continue;
        }
        for (int fr = 0; fr < fh; ++fr) {
            const int yy = y0 + fr;
            for (int fc = 0; fc < fw; ++fc) {
                const int xx = x0 + fc;
                const bool in_img = xx >= 0 && yy >= 0 && xx < W && yy < H;
                const float p = (float) in[in_img ? ((size_t) yy * W + xx) * C + c : centre + c];
                pix = fmaf(taps.f[fr * fw + fc], p, pix);
            }
        }
        out[centre + c] = to_pixel<T>(pix);
    }
}

// transformPixel: each byte is read through a signed char, divided by mod
// and multiplied back in binary32, truncating to int after each step, and
// stored modulo 256 (mod >= 2^-20 keeps both quotients inside the int range).
__global__ void quantize_kernel(uint8_t *img, size_t n, float mod) {
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x) {
        int t = (int) (int8_t) img[i];
        t = (int) ((float) t / mod);
        t = (int) ((float) t * mod);
        img[i] = (uint8_t) t;
    }
}

// stbir_resize_uint8 (stb_image_resize v0.97: linear colour space, no alpha,
// clamped edges; Mitchell on a shrinking axis, Catmull-Rom on a growing one),
// as a separable gather.  The host lays out, per output column and per output
// row, the inputs (clamped indices) and weights in the order the library
// accumulates them; the two passes below add them from 0 in that order, with
// the library's binary32 decode table, so every byte equals the library's.
struct ResizeAxis {
    std::vector<int> first;   // output o gathers entries first[o] .. first[o+1]
    std::vector<int> src;     // clamped input index
    std::vector<float> wt;
};

static float resize_kernel(bool grow, float x) {
    x = std::fabs(x);
    if (grow) {   // Catmull-Rom
        if (x < 1.0f) return 1.0f - x * x * (2.5f - 1.5f * x);
        if (x < 2.0f) return 2.0f - x * (4.0f + x * (0.5f * x - 2.5f));
        return 0.0f;
    }
    if (x < 1.0f) return (16.0f + x * x * (21.0f * x - 36.0f)) / 18.0f;   // Mitchell
    if (x < 2.0f) return (32.0f + x * (-60.0f + x * (36.0f - 7.0f * x))) / 18.0f;
    return 0.0f;
}

// The library's filter set-up for one axis (stbir__calculate_filters and the
// coefficient / normalisation steps it calls), turned into per-output lists.
static ResizeAxis resize_axis(int in_size, int out_size) {
    ResizeAxis A;
    A.first.assign((size_t) out_size + 1, 0);
    const float scale = (float) out_size / (float) in_size;
    auto clamp_in = [&](int k) { return k < 0 ? 0 : (k >= in_size ? in_size - 1 : k); };
    constexpr int kWidth = 4;   // ceil(2 * support): outputs per input (shrink) or inputs per output (grow)
    if (scale > 1.0f) {
        // each output pixel's inputs: the support around its centre, leading
        // exact zeros dropped, normalised by 1/sum, trailing zeros dropped
        const float radius = 2.0f * scale;
        for (int o = 0; o < out_size; ++o) {
            const float centre = (float) o + 0.5f;
            const float in_centre = centre / scale;
            int lo = (int) std::floor((double) ((centre - radius) / scale) + 0.5);
            const int hi = (int) std::floor((double) ((centre + radius) / scale) - 0.5);
            std::vector<float> c;
            float total = 0.0f;
            for (int k = lo; k <= hi; ++k) {
                const float w = resize_kernel(true, in_centre - ((float) k + 0.5f));
                if (c.empty() && w == 0.0f) {
                    lo = k + 1;
                    continue;
                }
                c.push_back(w);
                total += w;
            }
            const float s = 1.0f / total;
            for (float &w : c) w *= s;
            while (!c.empty() && c.back() == 0.0f) c.pop_back();
            A.first[o] = (int) A.src.size();
            for (size_t i = 0; i < c.size(); ++i) {
                A.src.push_back(clamp_in(lo + (int) i));
                A.wt.push_back(c[i]);
            }
        }
        A.first[out_size] = (int) A.src.size();
        return A;
    }
    // shrinking: every input pixel (with a margin of edge pixels) spreads over
    // the outputs its support covers; each output's weights are normalised by
    // 1/sum over the inputs that reach it
    const int margin = (int) std::ceil((double) (2.0f * 2.0f / scale)) / 2;
    const int n = in_size + 2 * margin;
    const float radius = 2.0f / scale;
    std::vector<int> n0(n), n1(n);
    // contiguous groups as the library lays them out: a fifth weight of pixel
    // j (its output range can span 5) lands in group j+1's first slot and is
    // read back from there; a spare group takes the last pixel's
    std::vector<float> cf((size_t) (n + 1) * kWidth, 0.0f);
    for (int j = 0; j < n; ++j) {
        const float centre = (float) (j - margin) + 0.5f;
        const float out_centre = centre * scale;
        n0[j] = (int) std::floor((double) ((centre - radius) * scale) + 0.5);
        n1[j] = (int) std::floor((double) ((centre + radius) * scale) - 0.5);
        float *c = &cf[(size_t) j * kWidth];
        const int cnt = n1[j] - n0[j] + 1;
        for (int i = 0; i < cnt; ++i) c[i] = resize_kernel(false, ((float) (n0[j] + i) + 0.5f) - out_centre) * scale;
        for (int i = cnt - 1; i >= 0 && c[i] == 0.0f; --i) n1[j] = n0[j] + i - 1;
    }
    for (int o = 0; o < out_size; ++o) {
        float total = 0.0f;
        for (int j = 0; j < n && o >= n0[j]; ++j)
            if (o <= n1[j]) total += cf[(size_t) j * kWidth + (o - n0[j])];
        const float s = 1.0f / total;
        for (int j = 0; j < n && o >= n0[j]; ++j)
            if (o <= n1[j]) cf[(size_t) j * kWidth + (o - n0[j])] *= s;
    }
    for (int j = 0; j < n; ++j) {   // drop leading zeros and outputs below 0
        float *c = &cf[(size_t) j * kWidth];
        int skip = 0;
        while (skip < kWidth && c[skip] == 0.0f) ++skip;
        if (skip == kWidth) {
            n1[j] = n0[j] - 1;
            continue;
        }
        n0[j] += skip;
        while (n0[j] < 0) {
            ++n0[j];
            ++skip;
        }
        const int range = n1[j] - n0[j] + 1;
        for (int i = 0; i < std::min(range, kWidth) && i + skip < kWidth; ++i) c[i] = c[i + skip];
        n1[j] = std::min(n1[j], out_size - 1);
    }
    for (int o = 0; o < out_size; ++o) {
        A.first[o] = (int) A.src.size();
        for (int j = 0; j < n; ++j)
            if (o >= n0[j] && o <= n1[j]) {
                A.src.push_back(clamp_in(j - margin));
                A.wt.push_back(cf[(size_t) j * kWidth + (o - n0[j])]);
            }
    }
    A.first[out_size] = (int) A.src.size();
    return A;
}

// Horizontal pass over every input row: hb[r][x][c] for x < ow.
__global__ void resize_h_kernel(const uint8_t *in, float *hb, int iw, int ih, int ow, int C, const int *first,
                                const int *src, const float *wt, const float *decode) {
    const size_t total = (size_t) ih * ow * C;
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t) gridDim.x * blockDim.x) {
        const int c = (int) (i % C);
        const size_t px = i / C;
        const int x = (int) (px % ow);
        const size_t r = px / ow;
        const uint8_t *row = in + r * iw * C;
        float acc = 0.0f;
        for (int t = first[x]; t < first[x + 1]; ++t) acc += decode[row[(size_t) src[t] * C + c]] * wt[t];
        hb[i] = acc;
    }
}

// Vertical pass and the library's encode: (uint8) (int) ((double) (sat(v) * 255) + 0.5).
__global__ void resize_v_kernel(const float *hb, uint8_t *out, int ow, int oh, int C, const int *first, const int *src,
                                const float *wt) {
    const size_t total = (size_t) oh * ow * C;
    const size_t rowlen = (size_t) ow * C;
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t) gridDim.x * blockDim.x) {
        const size_t col = i % rowlen;
        const int y = (int) (i / rowlen);
        float acc = 0.0f;
        for (int t = first[y]; t < first[y + 1]; ++t) acc += hb[(size_t) src[t] * rowlen + col] * wt[t];
        const float v = acc < 0.0f ? 0.0f : (acc > 1.0f ? 1.0f : acc);
        out[i] = (uint8_t) (int) ((double) (v * 255.0f) + 0.5);
    }
}

// ------------------------------------------------------------ image writers
// The output images of the reference's drivers as one-pass kernels:
//   transformImage           src/gpu/cudaImageFunctions.cuh:40-62 (the _local
//                            copy, src/cpu/multispectral_ImageMatcher.cu:181-203,
//                            has the same body)
//   fuseAlignedImages        src/gpu/cudaImageFunctions.cuh:64-114
//   fuseAlignedImages_local  src/cpu/multispectral_ImageMatcher.cu:115-179
// The reference runs each as one thread looping over the image; here a thread
// owns an output element and reproduces what the loop leaves in it.  The
// inverse map is the reference's adjugate expression of H with every product
// and sum rounded on its own (this file compiles with -ffp-contract=off) and an
// IEEE quotient; the in-image test is inImage on that quotient; samples are
// valueAt_bilinear with the declared lerps (taps past the end read 0).

struct WarpH {
    float h[8];
};

__device__ __forceinline__ void inverse_point(const WarpH &H, float x, float y, float &nx, float &ny) {
    const float h11 = H.h[0], h12 = H.h[1], h13 = H.h[2], h21 = H.h[3], h22 = H.h[4], h23 = H.h[5], h31 = H.h[6],
                h32 = H.h[7];
    const float w = (h21 * h32 - h22 * h31) * x + (h12 * h31 - h11 * h32) * y + h11 * h22 - h12 * h21;
    nx = ((h22 - h23 * h32) * x + (h13 * h32 - h12) * y + h12 * h23 - h13 * h22) / w;
    ny = ((h23 * h31 - h21) * x + (h11 - h13 * h31) * y + h13 * h21 - h11 * h23) / w;
}

struct SrcImage {
    const void *data;
    int w, h, c, u8;
};

__device__ __forceinline__ bool src_in(const SrcImage &S, float x, float y) {
    return x >= 0.0f && y >= 0.0f && x < (float) S.w && y < (float) S.h;
}

__device__ __forceinline__ float src_tap(const SrcImage &S, long i) {
    if (i >= (long) S.w * S.h * S.c) return 0.0f;
    return S.u8 ? (float) ((const uint8_t *) S.data)[i] : ((const float *) S.data)[i];
}

// valueAt_bilinear at an in-image point: taps at floor and ceil (the ceil tap
// is the floor tap at an integral coordinate), linear addressing.
__device__ __forceinline__ float src_bilinear(const SrcImage &S, float x, float y, int c) {
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const long b = ((long) (int) fy * S.w + (int) fx) * S.c + c;
    const long dx = ax != 0.0f ? S.c : 0, dy = ay != 0.0f ? (long) S.c * S.w : 0;
    const float top = fmaf(ax, src_tap(S, b + dx), (1.0f - ax) * src_tap(S, b));
    const float bottom = fmaf(ax, src_tap(S, b + dy + dx), (1.0f - ax) * src_tap(S, b + dy));
    return fmaf(ay, bottom, (1.0f - ay) * top);
}

template <typename OT>
__global__ void warp_kernel(SrcImage S, OT *out, int ow, int oh, WarpH H) {
    const size_t n = (size_t) ow * oh;
    for (size_t p = (size_t) blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t) gridDim.x * blockDim.x) {
        const int x = (int) (p % ow), y = (int) (p / ow);
        float nx, ny;
        inverse_point(H, (float) x, (float) y, nx, ny);
        if (!src_in(S, nx, ny)) continue;   // the pixel keeps what out holds
        for (int c = 0; c < S.c; ++c) out[p * S.c + c] = to_pixel<OT>(src_bilinear(S, nx, ny, c));
    }
}

// fuseAlignedImages: channel 0 the moving image under est, 1 the reference
// under the identity, 2 the moving image under truth, each sampled at `ch`.
template <typename OT>
__global__ void fuse_kernel(SrcImage ref, SrcImage mov, OT *out, int ow, int oh, WarpH est, WarpH truth, int ch) {
    const size_t n = (size_t) ow * oh * 3;
    WarpH ident{};
    ident.h[0] = ident.h[4] = 1.0f;
    for (size_t e = (size_t) blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t) gridDim.x * blockDim.x) {
        const int c = (int) (e % 3);
        const size_t p = e / 3;
        const int x = (int) (p % ow), y = (int) (p / ow);
        const SrcImage &S = c == 1 ? ref : mov;
        float nx, ny;
        inverse_point(c == 0 ? est : (c == 1 ? ident : truth), (float) x, (float) y, nx, ny);
        if (src_in(S, nx, ny)) out[e] = to_pixel<OT>(src_bilinear(S, nx, ny, ch));
    }
}

// fuseAlignedImages_local steps normalised coordinates t = -0.5, -0.5 + 1/n, ...
// in binary32 and writes the sample of (x_i, y_j) to pixel (round(rows (y_j +
// 0.5)), round(cols (x_i + 0.5))) by linear index, so a target column of cols
// lands on the next row's first pixel and one past the end is not written.
// Loop order is channel, i, j: an element holds the sample of the last (i, j)
// in that order that targets it and whose point is in the source image.  The
// host tabulates x_i, y_j and, per target t, the (contiguous, as the targets
// rise with i) range of i and of j that hit it.
struct LocalTables {
    const float *xs, *ys;
    const int *ilo, *jlo;   // target t's steps: [lo[t], lo[t + 1])
    int cols, rows;
};

__device__ __forceinline__ bool local_sample(const SrcImage &S, const SrcImage &mov, const WarpH &H, float x, float y,
                                             int ch, float &v) {
    float nx, ny;
    inverse_point(H, x, y, nx, ny);
    nx = (float) (((double) nx + 0.5) * (double) mov.w);   // scaled by the moving image's size for every channel
    ny = (float) (((double) ny + 0.5) * (double) mov.h);
    if (!src_in(S, nx, ny)) return false;
    v = src_bilinear(S, nx, ny, ch);
    return true;
}

template <typename OT>
__global__ void fuse_local_kernel(SrcImage ref, SrcImage mov, OT *out, LocalTables T, WarpH est, WarpH truth, int ch) {
    const size_t n = (size_t) T.cols * T.rows * 3;
    WarpH ident{};
    ident.h[0] = ident.h[4] = 1.0f;
    for (size_t e = (size_t) blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t) gridDim.x * blockDim.x) {
        const int c = (int) (e % 3);
        const size_t p = e / 3;
        const int tx = (int) (p % T.cols), ty = (int) (p / T.cols);
        const SrcImage &S = c == 1 ? ref : mov;
        const WarpH &H = c == 0 ? est : (c == 1 ? ident : truth);
        float v = 0.0f;
        bool hit = false;
        // the wrapped writes (target column cols of the row above) come after
        // every write that targets column 0 directly
        for (int pass = 0; pass < 2 && !hit; ++pass) {
            int gx = tx, gy = ty;
            if (pass == 0) {
                if (tx != 0 || ty == 0) continue;
                gx = T.cols;
                gy = ty - 1;
            }
            for (int i = T.ilo[gx + 1] - 1; i >= T.ilo[gx] && !hit; --i)
                for (int j = T.jlo[gy + 1] - 1; j >= T.jlo[gy] && !hit; --j) hit = local_sample(S, mov, H, T.xs[i], T.ys[j], ch, v);
        }
        if (hit) out[e] = to_pixel<OT>(v);
    }
}

// The binary32 steps of one normalised axis and, per target 0..size, the
// first step that hits it (size + 2 entries: lo[t] .. lo[t + 1]).
static void local_axis(int size, std::vector<float> &ts, std::vector<int> &lo) {
    ts.clear();
    const float step = 1.0f / (float) size;
    for (float t = -0.5f; t < 0.5f; t += step) ts.push_back(t);
    std::vector<int> tgt(ts.size());
    for (size_t i = 0; i < ts.size(); ++i) tgt[i] = (int) std::round((double) size * ((double) ts[i] + 0.5));
    lo.assign((size_t) size + 2, 0);
    size_t i = 0;
    for (int t = 0; t <= size + 1; ++t) {
        while (i < tgt.size() && tgt[i] < t) ++i;
        lo[(size_t) t] = (int) i;
    }
}

}  // namespace cgs

using namespace cgs;

extern "C" {

cgs_status cgs_image_filter(const cgs_image *in, const cgs_image *out, const float *filter, int32_t fw, int32_t fh,
                            void *stream) {
    return cgs_image_filter_channels(in, out, filter, fw, fh, 0xFu, stream);
}

cgs_status cgs_image_filter_channels(const cgs_image *in, const cgs_image *out, const float *filter, int32_t fw,
                                     int32_t fh, uint32_t filter_mask, void *stream) {
    clear_error();
    cgs_status st;
    if ((st = validate_image(in, "in")) != CGS_OK || (st = validate_image(out, "out")) != CGS_OK) return st;
    if (in->width != out->width || in->height != out->height || in->channels != out->channels ||
        in->pixtype != out->pixtype)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "filter: in and out differ in shape or pixel type");
    if (in->data == out->data) return set_error(CGS_ERR_INVALID_ARGUMENT, "filter: in and out must not alias");
    if (!filter || fw < 1 || fh < 1 || fw * fh > kMaxFilterTaps)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "filter: 1..256 taps");
    FilterTaps taps{};
    for (int i = 0; i < fw * fh; ++i) taps.f[i] = filter[i];
    const dim3 block(16, 16), grid((in->width + 15) / 16, (in->height + 15) / 16);
    hipStream_t s = (hipStream_t) stream;
    if (in->pixtype == CGS_PIX_U8)
        hipLaunchKernelGGL(filter_kernel<uint8_t>, grid, block, 0, s, (const uint8_t *) in->data,
                           (uint8_t *) out->data, in->width, in->height, in->channels, taps, fw, fh, filter_mask);
    else
        hipLaunchKernelGGL(filter_kernel<float>, grid, block, 0, s, (const float *) in->data, (float *) out->data,
                           in->width, in->height, in->channels, taps, fw, fh, filter_mask);
    HIP_TRY(hipGetLastError());
    return CGS_OK;
}

cgs_status cgs_image_resize(const cgs_image *in, const cgs_image *out, void *stream) {
    clear_error();
    cgs_status st;
    if ((st = validate_image(in, "in")) != CGS_OK || (st = validate_image(out, "out")) != CGS_OK) return st;
    if (in->pixtype != CGS_PIX_U8 || out->pixtype != CGS_PIX_U8)
        return set_error(CGS_ERR_UNSUPPORTED, "resize: uint8 images only (stbir_resize_uint8)");
    if (in->channels != out->channels) return set_error(CGS_ERR_INVALID_ARGUMENT, "resize: channel counts differ");
    if (in->data == out->data) return set_error(CGS_ERR_INVALID_ARGUMENT, "resize: in and out must not alias");
    const ResizeAxis X = resize_axis(in->width, out->width), Y = resize_axis(in->height, out->height);
    // one staging block: decode table, both axes' lists, then the row buffer
    std::vector<char> host;
    // copies `bytes` from p and leaves `pad` zero bytes after them (a list
    // may be empty; the kernels never read its padding)
    auto put = [&](const void *p, size_t bytes, size_t pad) {
        const size_t at = (host.size() + 255) / 256 * 256;
        host.resize(at + bytes + pad, 0);
        if (bytes) std::memcpy(host.data() + at, p, bytes);
        return at;
    };
    float decode[256];
    for (int p = 0; p < 256; ++p) decode[p] = (float) p / 255.0f;
    const size_t o_dec = put(decode, sizeof(decode), 0);
    const size_t o_xf = put(X.first.data(), X.first.size() * sizeof(int), 0);
    const size_t o_xs = put(X.src.data(), X.src.size() * sizeof(int), 4);
    const size_t o_xw = put(X.wt.data(), X.wt.size() * sizeof(float), 4);
    const size_t o_yf = put(Y.first.data(), Y.first.size() * sizeof(int), 0);
    const size_t o_ys = put(Y.src.data(), Y.src.size() * sizeof(int), 4);
    const size_t o_yw = put(Y.wt.data(), Y.wt.size() * sizeof(float), 4);
    const size_t o_hb = (host.size() + 255) / 256 * 256;
    const size_t hb_bytes = (size_t) in->height * out->width * in->channels * sizeof(float);
    char *dev = nullptr;
    HIP_TRY(hipMalloc(&dev, o_hb + hb_bytes));
    hipStream_t s = (hipStream_t) stream;
    hipError_t e = hipMemcpyAsync(dev, host.data(), host.size(), hipMemcpyHostToDevice, s);
    const unsigned blocks = 2048;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(resize_h_kernel, dim3(blocks), dim3(256), 0, s, (const uint8_t *) in->data,
                           (float *) (dev + o_hb), in->width, in->height, out->width, in->channels,
                           (const int *) (dev + o_xf), (const int *) (dev + o_xs), (const float *) (dev + o_xw),
                           (const float *) (dev + o_dec));
        hipLaunchKernelGGL(resize_v_kernel, dim3(blocks), dim3(256), 0, s, (const float *) (dev + o_hb),
                           (uint8_t *) out->data, out->width, out->height, in->channels, (const int *) (dev + o_yf),
                           (const int *) (dev + o_ys), (const float *) (dev + o_yw));
        e = hipGetLastError();
    }
    const hipError_t e2 = hipStreamSynchronize(s);   // the staging block is freed below
    (void) hipFree(dev);
    HIP_TRY(e);
    HIP_TRY(e2);
    return CGS_OK;
}

static SrcImage src_of(const cgs_image *im) {
    return SrcImage{im->data, im->width, im->height, im->channels, im->pixtype == CGS_PIX_U8 ? 1 : 0};
}

static WarpH warp_of(const float *h) {
    WarpH H;
    for (int k = 0; k < 8; ++k) H.h[k] = h[k];
    return H;
}

cgs_status cgs_image_warp(const cgs_image *in, const cgs_image *out, const float *h, void *stream) {
    clear_error();
    cgs_status st;
    if ((st = validate_image(in, "in")) != CGS_OK || (st = validate_image(out, "out")) != CGS_OK) return st;
    if (!h) return set_error(CGS_ERR_INVALID_ARGUMENT, "warp: null homography");
    if (in->channels != out->channels) return set_error(CGS_ERR_INVALID_ARGUMENT, "warp: channel counts differ");
    if (in->data == out->data) return set_error(CGS_ERR_INVALID_ARGUMENT, "warp: in and out must not alias");
    const SrcImage S = src_of(in);
    const WarpH H = warp_of(h);
    hipStream_t s = (hipStream_t) stream;
    const unsigned blocks = (unsigned) std::min<size_t>(((size_t) out->width * out->height + 255) / 256, 4096);
    if (out->pixtype == CGS_PIX_U8)
        hipLaunchKernelGGL(warp_kernel<uint8_t>, dim3(blocks), dim3(256), 0, s, S, (uint8_t *) out->data, out->width,
                           out->height, H);
    else
        hipLaunchKernelGGL(warp_kernel<float>, dim3(blocks), dim3(256), 0, s, S, (float *) out->data, out->width,
                           out->height, H);
    HIP_TRY(hipGetLastError());
    return CGS_OK;
}

cgs_status cgs_image_fuse(const cgs_image *ref, const cgs_image *mov, const float *est_h, const float *true_h,
                          int32_t channel, int32_t local, const cgs_image *fused, void *stream) {
    clear_error();
    cgs_status st;
    if ((st = validate_image(ref, "ref")) != CGS_OK || (st = validate_image(mov, "mov")) != CGS_OK ||
        (st = validate_image(fused, "fused")) != CGS_OK)
        return st;
    if (!est_h || !true_h) return set_error(CGS_ERR_INVALID_ARGUMENT, "fuse: null homography");
    if (fused->channels != 3) return set_error(CGS_ERR_INVALID_ARGUMENT, "fuse: the fused image has 3 channels");
    if (channel < 0 || channel >= ref->channels || channel >= mov->channels)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "fuse: channel outside the source images");
    if (fused->data == ref->data || fused->data == mov->data)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "fuse: the fused image must not alias a source");
    const SrcImage R = src_of(ref), M = src_of(mov);
    const WarpH E = warp_of(est_h), T = warp_of(true_h);
    hipStream_t s = (hipStream_t) stream;
    const size_t n = (size_t) fused->width * fused->height * 3;
    const unsigned blocks = (unsigned) std::min<size_t>((n + 255) / 256, 4096);
    const bool u8 = fused->pixtype == CGS_PIX_U8;
    if (!local) {
        if (u8)
            hipLaunchKernelGGL(fuse_kernel<uint8_t>, dim3(blocks), dim3(256), 0, s, R, M, (uint8_t *) fused->data,
                               fused->width, fused->height, E, T, (int) channel);
        else
            hipLaunchKernelGGL(fuse_kernel<float>, dim3(blocks), dim3(256), 0, s, R, M, (float *) fused->data,
                               fused->width, fused->height, E, T, (int) channel);
        HIP_TRY(hipGetLastError());
        return CGS_OK;
    }
    std::vector<float> xs, ys;
    std::vector<int> ilo, jlo;
    local_axis(fused->width, xs, ilo);
    local_axis(fused->height, ys, jlo);
    std::vector<char> host;
    auto put = [&](const void *p, size_t bytes) {
        const size_t at = (host.size() + 255) / 256 * 256;
        host.resize(at + bytes + 4, 0);
        if (bytes) std::memcpy(host.data() + at, p, bytes);
        return at;
    };
    const size_t o_xs = put(xs.data(), xs.size() * sizeof(float)), o_ys = put(ys.data(), ys.size() * sizeof(float));
    const size_t o_il = put(ilo.data(), ilo.size() * sizeof(int)), o_jl = put(jlo.data(), jlo.size() * sizeof(int));
    char *dev = nullptr;
    HIP_TRY(hipMalloc(&dev, host.size()));
    hipError_t e = hipMemcpyAsync(dev, host.data(), host.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        const LocalTables L{(const float *) (dev + o_xs), (const float *) (dev + o_ys), (const int *) (dev + o_il),
                            (const int *) (dev + o_jl), fused->width, fused->height};
        if (u8)
            hipLaunchKernelGGL(fuse_local_kernel<uint8_t>, dim3(blocks), dim3(256), 0, s, R, M,
                               (uint8_t *) fused->data, L, E, T, (int) channel);
        else
            hipLaunchKernelGGL(fuse_local_kernel<float>, dim3(blocks), dim3(256), 0, s, R, M, (float *) fused->data, L,
                               E, T, (int) channel);
        e = hipGetLastError();
    }
    const hipError_t e2 = hipStreamSynchronize(s);   // the tables are freed below
    (void) hipFree(dev);
    HIP_TRY(e);
    HIP_TRY(e2);
    return CGS_OK;
}

// A PNG of 8-bit samples (1: grey, 2: grey + alpha, 3: RGB, 4: RGBA) with the
// image data in stored (uncompressed) deflate blocks: what the reference's
// drivers write through stb_image_write, pixel for pixel.
static uint32_t png_crc(const unsigned char *p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    for (size_t i = 0; i < n; ++i) {
        c ^= p[i];
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1u)));
    }
    return c;
}

cgs_status cgs_write_png(const char *path, const void *pixels, int32_t width, int32_t height, int32_t channels) {
    clear_error();
    if (!path || !pixels || width < 1 || height < 1 || channels < 1 || channels > 4)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "write_png: bad arguments");
    const size_t row = (size_t) width * channels;
    std::vector<unsigned char> raw;
    raw.reserve((row + 1) * height);
    for (int y = 0; y < height; ++y) {
        raw.push_back(0);   // filter: none
        const unsigned char *r = (const unsigned char *) pixels + row * y;
        raw.insert(raw.end(), r, r + row);
    }
    std::vector<unsigned char> z = {0x78, 0x01};
    for (size_t at = 0; at < raw.size() || at == 0;) {
        const size_t len = std::min<size_t>(65535, raw.size() - at);
        const bool last = at + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((unsigned char) (len & 255));
        z.push_back((unsigned char) (len >> 8));
        z.push_back((unsigned char) (~len & 255));
        z.push_back((unsigned char) ((~len >> 8) & 255));
        z.insert(z.end(), raw.begin() + at, raw.begin() + at + len);
        at += len;
        if (last) break;
    }
    uint32_t a = 1, b = 0;
    for (unsigned char v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    const uint32_t adler = (b << 16) | a;
    for (int k = 3; k >= 0; --k) z.push_back((unsigned char) (adler >> (8 * k)));
    std::vector<unsigned char> f = {137, 80, 78, 71, 13, 10, 26, 10};
    auto be32 = [&](uint32_t v) {
        for (int k = 3; k >= 0; --k) f.push_back((unsigned char) (v >> (8 * k)));
    };
    auto chunk = [&](const char *type, const std::vector<unsigned char> &data) {
        be32((uint32_t) data.size());
        const size_t start = f.size();
        f.insert(f.end(), type, type + 4);
        f.insert(f.end(), data.begin(), data.end());
        be32(png_crc(f.data() + start, f.size() - start) ^ 0xFFFFFFFFu);
    };
    static const unsigned char kColour[5] = {0, 0, 4, 2, 6};
    std::vector<unsigned char> ihdr;
    for (uint32_t v : {(uint32_t) width, (uint32_t) height})
        for (int k = 3; k >= 0; --k) ihdr.push_back((unsigned char) (v >> (8 * k)));
    for (unsigned char v : {(unsigned char) 8, kColour[channels], (unsigned char) 0, (unsigned char) 0, (unsigned char) 0})
        ihdr.push_back(v);
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", {});
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return set_error(CGS_ERR_INVALID_ARGUMENT, std::string("write_png: cannot open ") + path);
    const size_t wrote = std::fwrite(f.data(), 1, f.size(), fp);
    const int closed = std::fclose(fp);
    if (wrote != f.size() || closed != 0)
        return set_error(CGS_ERR_INVALID_ARGUMENT, std::string("write_png: cannot write ") + path);
    return CGS_OK;
}

cgs_status cgs_image_quantize(const cgs_image *img, float mod, void *stream) {
    clear_error();
    cgs_status st;
    if ((st = validate_image(img, "img")) != CGS_OK) return st;
    if (img->pixtype != CGS_PIX_U8) return set_error(CGS_ERR_UNSUPPORTED, "quantize: uint8 images only");
    // the reference calls it for mod > 0; below 2^-20 the quotient could leave the int range
    if (!(mod >= 1.0f / 1048576.0f) || !std::isfinite(mod))
        return set_error(CGS_ERR_INVALID_ARGUMENT, "quantize: mod must be in [2^-20, inf)");
    const size_t n = (size_t) img->width * img->height * img->channels;
    const unsigned blocks = (unsigned) std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(quantize_kernel, dim3(blocks), dim3(256), 0, (hipStream_t) stream, (uint8_t *) img->data, n,
                       mod);
    HIP_TRY(hipGetLastError());
    return CGS_OK;
}

}  // extern "C"
