// cgs_tensor.hip -- the reference's CudaTensor / CudaMatrix utilities that its
// drivers and the MEX gateway use (include/cudaTensor.cuh:197-206, 322-360,
// 474-496; kernels include/cudaTensorKernels.cuh:230-237, 341-351, 396-440):
// fill, element-wise add / subtract / multiply, matrix product.  Not on the
// search path; grid-stride kernels bound by HBM (fill, element-wise) and a
// one-thread-per-element product kept in the reference's summation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "cgs.h"
#include "cgs_host.hpp"

namespace cgs {

template <typename T>
__global__ void fill_kernel(T *d, int64_t n, T v) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x)
        d[i] = v;
}

template <typename T, int OP>
__global__ void binary_kernel(T *a, const T *b, int64_t n) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        if constexpr (OP == CGS_OP_ADD) a[i] = (T) (a[i] + b[i]);
        else if constexpr (OP == CGS_OP_SUB) a[i] = (T) (a[i] - b[i]);
        else a[i] = (T) (a[i] * b[i]);
    }
}

// C[r][c] = sum over j ascending of A[r][j] * B[j][c], each step one fma from
// 0: the value the reference's 4x4-tiled kernel leaves (its zero-padded tile
// tails add exact zeros).  A wave's lanes take consecutive columns, so B's rows
// and C are read and written coalesced; A's row element is wave-uniform.
template <typename T>
__global__ void multiply_kernel(const T *A, const T *B, T *C, int m, int k, int n) {
    const int64_t total = (int64_t) m * n;
    for (int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t) gridDim.x * blockDim.x) {
        const int r = (int) (e / n), c = (int) (e % n);
        T acc = 0;
        for (int j = 0; j < k; ++j) acc = fma(A[(int64_t) r * k + j], B[(int64_t) j * n + c], acc);
        C[e] = acc;
    }
}

static unsigned blocks_for(int64_t n) { return (unsigned) std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

static int dtype_bytes(int32_t dt) {
    switch (dt) {
    case CGS_DT_F32: return 4;
    case CGS_DT_F64: return 8;
    case CGS_DT_I32: return 4;
    case CGS_DT_U8: return 1;
    default: return 0;
    }
}

template <int OP>
static void launch_binary(int32_t dt, void *a, const void *b, int64_t n, hipStream_t s) {
    const unsigned g = blocks_for(n);
    switch (dt) {
    case CGS_DT_F32: hipLaunchKernelGGL((binary_kernel<float, OP>), dim3(g), dim3(256), 0, s, (float *) a, (const float *) b, n); break;
    case CGS_DT_F64: hipLaunchKernelGGL((binary_kernel<double, OP>), dim3(g), dim3(256), 0, s, (double *) a, (const double *) b, n); break;
    case CGS_DT_I32: hipLaunchKernelGGL((binary_kernel<int32_t, OP>), dim3(g), dim3(256), 0, s, (int32_t *) a, (const int32_t *) b, n); break;
    default: hipLaunchKernelGGL((binary_kernel<uint8_t, OP>), dim3(g), dim3(256), 0, s, (uint8_t *) a, (const uint8_t *) b, n); break;
    }
}

}  // namespace cgs

using namespace cgs;

extern "C" {

cgs_status cgs_tensor_fill(void *data, int64_t n, int32_t dtype, double value, void *stream) {
    clear_error();
    if (!dtype_bytes(dtype)) return set_error(CGS_ERR_INVALID_ARGUMENT, "fill: bad element type");
    if (n < 0 || (n > 0 && !data)) return set_error(CGS_ERR_INVALID_ARGUMENT, "fill: bad array");
    if (n == 0) return CGS_OK;
    hipStream_t s = (hipStream_t) stream;
    const unsigned g = blocks_for(n);
    switch (dtype) {
    case CGS_DT_F32: hipLaunchKernelGGL(fill_kernel<float>, dim3(g), dim3(256), 0, s, (float *) data, n, (float) value); break;
    case CGS_DT_F64: hipLaunchKernelGGL(fill_kernel<double>, dim3(g), dim3(256), 0, s, (double *) data, n, value); break;
    case CGS_DT_I32: hipLaunchKernelGGL(fill_kernel<int32_t>, dim3(g), dim3(256), 0, s, (int32_t *) data, n, (int32_t) value); break;
    default: hipLaunchKernelGGL(fill_kernel<uint8_t>, dim3(g), dim3(256), 0, s, (uint8_t *) data, n, (uint8_t) value); break;
    }
    HIP_TRY(hipGetLastError());
    return CGS_OK;
}

cgs_status cgs_tensor_binary(int32_t op, void *a, const void *b, int64_t n, int32_t dtype, void *stream) {
    clear_error();
    if (!dtype_bytes(dtype)) return set_error(CGS_ERR_INVALID_ARGUMENT, "tensor op: bad element type");
    if (op < CGS_OP_ADD || op > CGS_OP_MUL) return set_error(CGS_ERR_INVALID_ARGUMENT, "tensor op: bad operation");
    if (n < 0 || (n > 0 && (!a || !b))) return set_error(CGS_ERR_INVALID_ARGUMENT, "tensor op: bad arrays");
    if (n == 0) return CGS_OK;
    hipStream_t s = (hipStream_t) stream;
    if (op == CGS_OP_ADD) launch_binary<CGS_OP_ADD>(dtype, a, b, n, s);
    else if (op == CGS_OP_SUB) launch_binary<CGS_OP_SUB>(dtype, a, b, n, s);
    else launch_binary<CGS_OP_MUL>(dtype, a, b, n, s);
    HIP_TRY(hipGetLastError());
    return CGS_OK;
}

cgs_status cgs_matrix_multiply(const void *A, const void *B, void *C, int32_t m, int32_t k, int32_t n, int32_t dtype,
                               void *stream) {
    clear_error();
    if (dtype != CGS_DT_F32 && dtype != CGS_DT_F64)
        return set_error(CGS_ERR_UNSUPPORTED, "multiply: float or double matrices");
    if (m < 1 || k < 1 || n < 1 || !A || !B || !C) return set_error(CGS_ERR_INVALID_ARGUMENT, "multiply: bad matrices");
    if (C == A || C == B) return set_error(CGS_ERR_INVALID_ARGUMENT, "multiply: C must not alias A or B");
    hipStream_t s = (hipStream_t) stream;
    const unsigned g = blocks_for((int64_t) m * n);
    if (dtype == CGS_DT_F32)
        hipLaunchKernelGGL(multiply_kernel<float>, dim3(g), dim3(256), 0, s, (const float *) A, (const float *) B,
                           (float *) C, m, k, n);
    else
        hipLaunchKernelGGL(multiply_kernel<double>, dim3(g), dim3(256), 0, s, (const double *) A, (const double *) B,
                           (double *) C, m, k, n);
    HIP_TRY(hipGetLastError());
    return CGS_OK;
}

}  // extern "C"
