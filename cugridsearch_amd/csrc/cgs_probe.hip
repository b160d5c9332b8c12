// cgs_probe.hip -- the device's measured non-packed FP32 FMA throughput, the
// VALU roofline the search kernels are held against beside the published
// peak (which counts packed FP32, an instruction this build does not emit).
#include <hip/hip_runtime.h>

#include <string>

#include "cgs.h"
#include "cgs_host.hpp"

namespace cgs {

constexpr int kProbeChains = 16;

// Independent fmaf chains per thread, one store per thread (out holds
// gridDim.x * blockDim.x floats).
__global__ __launch_bounds__(256) void fma_chain_kernel(float *out, int iters, float a, float b) {
    float acc[kProbeChains];
#pragma unroll
    for (int c = 0; c < kProbeChains; ++c) acc[c] = (float) (threadIdx.x + c);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < kProbeChains; ++c) acc[c] = fmaf(acc[c], a, b);
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kProbeChains; ++c) s += acc[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace cgs

using namespace cgs;

extern "C" cgs_status cgs_fma_peak(float *tflops_out, void *stream) {
    clear_error();
    if (!tflops_out) return set_error(CGS_ERR_INVALID_ARGUMENT, "fma_peak: null output");
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int block = 256, blocks = cus * 8, iters = 8192;
    float *out = nullptr;
    HIP_TRY(hipMalloc(&out, sizeof(float) * (size_t) blocks * block));
    hipStream_t s = (hipStream_t) stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float best = 0.0f;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
        e = hipEventRecord(e0, s);
        hipLaunchKernelGGL(fma_chain_kernel, dim3(blocks), dim3(block), 0, s, out, iters, 1.0000001f, 1e-7f);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        const double flop = 2.0 * (double) blocks * block * iters * kProbeChains;
        if (e == hipSuccess && ms > 0.0f && flop / (ms * 1e9) > best) best = (float) (flop / (ms * 1e9));
    }
    if (e0) (void) hipEventDestroy(e0);
    if (e1) (void) hipEventDestroy(e1);
    (void) hipFree(out);
    if (e != hipSuccess) return set_error(CGS_ERR_HIP, std::string("fma_peak: ") + hipGetErrorString(e));
    *tflops_out = best;
    return CGS_OK;
}
