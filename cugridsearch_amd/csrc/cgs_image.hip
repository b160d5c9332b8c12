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
#include <cstring>
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
// between channels (the reference declares it once per pixel).
template <typename T>
__global__ void filter_kernel(const T *in, T *out, int W, int H, int C, const FilterTaps taps, int fw, int fh) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= W || y >= H) return;
    const int x0 = x - (fw - 1) / 2, y0 = y - (fh - 1) / 2;
    const size_t centre = ((size_t) y * W + x) * C;
    float pix = 0.0f;
    for (int c = 0; c < C; ++c) {
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

}  // namespace cgs

using namespace cgs;

extern "C" {

cgs_status cgs_image_filter(const cgs_image *in, const cgs_image *out, const float *filter, int32_t fw, int32_t fh,
                            void *stream) {
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
                           (uint8_t *) out->data, in->width, in->height, in->channels, taps, fw, fh);
    else
        hipLaunchKernelGGL(filter_kernel<float>, grid, block, 0, s, (const float *) in->data, (float *) out->data,
                           in->width, in->height, in->channels, taps, fw, fh);
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
    auto put = [&](const void *p, size_t bytes) {
        const size_t at = (host.size() + 255) / 256 * 256;
        host.resize(at + bytes);
        std::memcpy(host.data() + at, p, bytes);
        return at;
    };
    float decode[256];
    for (int p = 0; p < 256; ++p) decode[p] = (float) p / 255.0f;
    const size_t o_dec = put(decode, sizeof(decode));
    const size_t o_xf = put(X.first.data(), X.first.size() * sizeof(int));
    const size_t o_xs = put(X.src.data(), X.src.size() * sizeof(int) + 4);
    const size_t o_xw = put(X.wt.data(), X.wt.size() * sizeof(float) + 4);
    const size_t o_yf = put(Y.first.data(), Y.first.size() * sizeof(int));
    const size_t o_ys = put(Y.src.data(), Y.src.size() * sizeof(int) + 4);
    const size_t o_yw = put(Y.wt.data(), Y.wt.size() * sizeof(float) + 4);
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
