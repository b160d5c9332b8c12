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
