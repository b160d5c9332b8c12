// cgs_covariance.hip -- covariance of the cost surface around its argmin
// (CudaGridSearcher::covariance_by_value, src/gpu/cudaGridSearch.cuh:768-794,
// kernel calculateCovariance_by_value :317-419).
//
// The reference runs one GPU thread over every grid point.  Here every
// thread takes a strided share of the points and keeps the D(D+1)/2 distinct
// sums in binary64; workgroups reduce through wave shuffles and LDS, and a
// second kernel adds the workgroups' partial sums in workgroup order, so the
// result is deterministic.  Per point the arithmetic is the reference's, in
// binary32 with every operation rounded on its own:
//   f   = result[i], with f == 0 replaced by -1e-9
//   dx  = p(i) - p(min)                        per axis
//   len = sqrt(sum_d dx_d * dx_d)              summed in axis order
//   s   = (log|min| - log|f|) * (1 / len)      skipped when s < 0
//   g_d = dx_d * s / len
//   C  += g g^T
// log is declared as the binary64 log rounded to binary32, on the host and
// the device alike (the reference calls CUDA's logf).  The minimum follows
// the reference's scan: strict < starting from the value 0 at index 0, i.e.
// the lowest index of the smallest negative cost, else index 0 with value 0.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "cgs.h"
#include "cgs_device.hpp"
#include "cgs_host.hpp"

namespace cgs {

constexpr int kCovPairs = kMaxDim * (kMaxDim + 1) / 2;   // 36 distinct entries of a symmetric 8x8
constexpr int kCovBlocks = 1024;

__device__ __forceinline__ double cov_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void covariance_kernel(GridDev g, const float *result, int32_t n, int32_t min_idx,
                                                         float logmin, double *part) {
    double acc[kCovPairs];
#pragma unroll
    for (int k = 0; k < kCovPairs; ++k) acc[k] = 0.0;
    float xmin[kMaxDim];
    int32_t i0;
    grid_point(g, min_idx, xmin, &i0);
    const int dim = g.dim;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i == min_idx) continue;
        float p[kMaxDim];
        grid_point(g, i, p, &i0);
        float f = result[i];
        if (f == 0.0f) f = -1e-9f;
        float dx[kMaxDim], len2 = 0.0f;
#pragma unroll
        for (int d = 0; d < kMaxDim; ++d) {
            dx[d] = d < dim ? p[d] - xmin[d] : 0.0f;
            if (d < dim) len2 = len2 + dx[d] * dx[d];
        }
        const float len = sqrtf(len2);
        const float slope = (logmin - (float) log((double) fabsf(f))) * (1.0f / len);
        if (slope < 0.0f) continue;
        float gr[kMaxDim];
#pragma unroll
        for (int d = 0; d < kMaxDim; ++d) gr[d] = dx[d] * slope / len;
        int k = 0;
#pragma unroll
        for (int r = 0; r < kMaxDim; ++r)
#pragma unroll
            for (int c = r; c < kMaxDim; ++c, ++k) acc[k] += (double) (gr[r] * gr[c]);
    }
    __shared__ double red[4][kCovPairs];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kCovPairs; ++k) {
        const double v = cov_wave_sum(acc[k]);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kCovPairs) {
        const int k = threadIdx.x;
        part[(size_t) blockIdx.x * kCovPairs + k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
    }
}

__global__ void covariance_final_kernel(const double *part, int blocks, double *out) {
    const int k = threadIdx.x;
    if (k >= kCovPairs) return;
    double s = 0.0;
    for (int b = 0; b < blocks; ++b) s += part[(size_t) b * kCovPairs + k];
    out[k] = s;
}

}  // namespace cgs

using namespace cgs;

extern "C" {

cgs_status cgs_covariance(const cgs_grid *grid, const float *result, float *covar, cgs_argmin *min_out, void *stream) {
    clear_error();
    if (!grid || !result || !covar) return set_error(CGS_ERR_INVALID_ARGUMENT, "covariance: null argument");
    if (grid->dim < 1 || grid->dim > CGS_MAX_DIM || grid->num_elements <= 0)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "covariance: bad grid");
    // the reference's scan: smallest negative cost, lowest index; else index 0, value 0
    cgs_argmin m;
    cgs_status st = cgs_find_extrema(result, grid->num_elements, CGS_TIE_LOWEST, &m, stream);
    if (st != CGS_OK) return st;
    if (!(m.value < 0.0f)) {
        m.value = 0.0f;
        m.index = 0;
    }
    if (min_out) *min_out = m;
    Workspace *ws = nullptr;
    if ((st = workspace_for_current_device(&ws)) != CGS_OK) return st;
    std::lock_guard<std::mutex> lock(ws->mu);
    const size_t part_bytes = sizeof(double) * kCovPairs * (kCovBlocks + 1);
    if ((st = ws->cov.ensure(part_bytes)) != CGS_OK || (st = ws->cov_host.ensure(sizeof(double) * kCovPairs, true)) != CGS_OK)
        return st;
    hipStream_t s = (hipStream_t) stream;
    GridDev g;
    fill_grid_dev(grid, &g);
    const float logmin = (float) std::log((double) std::fabs(m.value));
    const int64_t n = grid->num_elements;
    const int blocks = (int) std::min<int64_t>((n + 255) / 256, kCovBlocks);
    double *part = (double *) ws->cov.ptr, *sum = part + (size_t) kCovBlocks * kCovPairs;
    hipLaunchKernelGGL(covariance_kernel, dim3(blocks), dim3(256), 0, s, g, result, (int32_t) n, m.index, logmin, part);
    hipLaunchKernelGGL(covariance_final_kernel, dim3(1), dim3(64), 0, s, (const double *) part, blocks, sum);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ws->cov_host.ptr, sum, sizeof(double) * kCovPairs, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const double *h = (const double *) ws->cov_host.ptr;
    const int D = grid->dim;
    int k = 0;
    for (int r = 0; r < kMaxDim; ++r)
        for (int c = r; c < kMaxDim; ++c, ++k)
            if (r < D && c < D) covar[r * D + c] = covar[c * D + r] = (float) h[k];
    return CGS_OK;
}

}  // extern "C"
