// cuGridSearch.hpp -- C++ host-side drop-in for the reference's search API,
// implemented over the C-ABI of cgs.h (libcgs.so, gfx950).
//
// Same class names, member names and argument meaning as the reference's
// header templates, so a driver ports by swapping the include and the functor
// binding:
//   CudaGrid<P, D>          src/gpu/cudaGridSearch.cuh:74-231
//   CudaTensor<P, D>        include/cudaTensor.cuh:108-361
//   CudaImage<P, C>         include/cudaTensor.cuh:507-596
//   CudaGridSearcher<F,G,D> src/gpu/cudaGridSearch.cuh:540-855
//   write*ImageToDisk(_local) src/gpu/cudaImageFunctions.cuh:129-179,
//                           src/cpu/multispectral_ImageMatcher.cu:205-255
//   umbrella include        include/cuGridSearch.cuh:8-11
// Containers stay shallow views over caller-allocated device memory
// (hipMalloc(&x.data(), x.bytesSize())), as in the reference.
//
// Functors come two ways:
//   * tags naming a cost of the C-ABI (calcNCC, calcMIstream,
//     averageAbsoluteDifference, ...): the library's kernels, specialised per
//     cost, with the fused argmin;
//   * a user's own __device__ functor, bound as in the reference (a
//     __device__ func_byvalue_t / func_byreference_t variable read back with
//     hipMemcpyFromSymbol, src/gpu/cudaGridSearch_main.cu:69-72, 270-271) and
//     evaluated once per grid point by the kernels defined below, reading its
//     images through CudaImage's device accessors;
// Differences, by design:
//   * errors throw cgs_error (the reference prints and exits the process);
//   * the grid's start / end / resolution / count rows are kept on the host
//     and mirrored into the caller's 4 x D device matrix when one is set.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <typeinfo>
#include <vector>

#include "cgs.h"

struct cgs_error : std::runtime_error {
    cgs_status status;
    cgs_error(cgs_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void cgs_check(cgs_status s, const char *what) {
    if (s != CGS_OK) throw cgs_error(s, std::string(what) + ": " + cgs_last_error());
}

inline void cgs_check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) throw cgs_error(CGS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ functors

// A cost functor: the C-ABI cost, how a grid point becomes the warp, which
// sampler reads the moving image, and the MI bin count.
struct CgsFunctor {
    cgs_cost cost;
    cgs_param_mode mode;
    cgs_sampler sampler;
    int bins;
    constexpr CgsFunctor with_sampler(cgs_sampler s) const { return {cost, mode, s, bins}; }
    constexpr CgsFunctor with_bins(int b) const { return {cost, mode, sampler, b}; }
};

// The reference functors (src/gpu/cudaErrorFunctions.cuh, cudaErrorFunction_mi*.cuh).
constexpr CgsFunctor averageAbsoluteDifference{CGS_COST_SAD, CGS_PARAM_NONCENTRED, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor averageAbsoluteDifferenceH{CGS_COST_SAD, CGS_PARAM_RAW_H, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor sumOfAbsoluteDifferences{CGS_COST_SAD_SUM, CGS_PARAM_TRANSLATION, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor calcSQD{CGS_COST_SQD, CGS_PARAM_CENTRED, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor calcNCC{CGS_COST_NCC, CGS_PARAM_CENTRED, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor calcMIstream{CGS_COST_MI, CGS_PARAM_CENTRED, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor calcMI{CGS_COST_MI_PLAIN, CGS_PARAM_CENTRED, CGS_SAMPLER_BILINEAR, 64};
// the thread-per-row "stream" functors (src/gpu/cudaErrorFunctionsStreams.cuh)
constexpr CgsFunctor averageAbsoluteDifference_stream{CGS_COST_SAD_STREAM, CGS_PARAM_TRANSLATION, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor calcNCCstream{CGS_COST_NCC_STREAM, CGS_PARAM_CENTRED, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor calcSQDstream{CGS_COST_SQD_STREAM, CGS_PARAM_CENTRED, CGS_SAMPLER_BILINEAR, 64};
// the line-traversal functors (src/gpu/cudaErrorFunctions.cuh:226-376,
// cudaErrorFunctionsStreams.cuh:165-281, cudaErrorFunction_miStreams.cuh:9-210,
// cudaErrorFunction_mi.cuh:117-255)
constexpr CgsFunctor calcSQDAlt{CGS_COST_SQD_ALT, CGS_PARAM_CENTRED, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor calcNCCAlt{CGS_COST_NCC_ALT, CGS_PARAM_CENTRED, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor calcNCCstreamAlt{CGS_COST_NCC_STREAM_ALT, CGS_PARAM_CENTRED, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor calcMIstreamAlt{CGS_COST_MI_STREAM_ALT, CGS_PARAM_CENTRED, CGS_SAMPLER_BILINEAR, 64};
constexpr CgsFunctor calcMIAlt{CGS_COST_MI_ALT, CGS_PARAM_NORM, CGS_SAMPLER_BILINEAR, 64};
// centred-homography SAD (the search workloads' SAD)
constexpr CgsFunctor centredAbsoluteDifference{CGS_COST_SAD, CGS_PARAM_CENTRED, CGS_SAMPLER_BILINEAR, 64};

// ------------------------------------------------------------------ tensors

template <typename T> struct CgsPixType;
template <> struct CgsPixType<uint8_t> { static constexpr cgs_pixtype value = CGS_PIX_U8; };
template <> struct CgsPixType<float> { static constexpr cgs_pixtype value = CGS_PIX_F32; };

// Element type of a tensor for the C-ABI's utilities (-1: host-side fallback).
template <typename T> struct CgsDType { static constexpr int value = -1; };
template <> struct CgsDType<float> { static constexpr int value = CGS_DT_F32; };
template <> struct CgsDType<double> { static constexpr int value = CGS_DT_F64; };
template <> struct CgsDType<int32_t> { static constexpr int value = CGS_DT_I32; };
template <> struct CgsDType<uint8_t> { static constexpr int value = CGS_DT_U8; };

// CudaTensor<P, D> (include/cudaTensor.cuh:108-361): a shallow view over
// caller-allocated device memory; axis 0 varies fastest.
template <typename precision, unsigned int D>
struct CudaTensor {
    precision *_data = nullptr;
    int32_t _dims[D] = {};
    int32_t _dimensional_increments[D] = {};
    int32_t _total_size = 0;

    CudaTensor() = default;

    template <typename T>
    __host__ __device__ explicit CudaTensor(const T (&dims)[D]) {
        _total_size = 1;
        for (unsigned d = 0; d < D; ++d) {
            _dims[d] = (int32_t) dims[d];
            _dimensional_increments[d] = d == 0 ? 1 : _dims[d - 1] * _dimensional_increments[d - 1];
            _total_size *= _dims[d];
        }
    }

    __host__ __device__ precision *&data() { return _data; }
    __host__ __device__ uint32_t size(int dim = -1) const {
        return dim == -1 ? (uint32_t) _total_size : (dim < (int) D ? (uint32_t) _dims[dim] : 0u);
    }
    __host__ __device__ uint32_t bytesSize() const { return size() * sizeof(precision); }

    // toIndex1d (include/cudaTensor.cuh:171-180): axis 0 fastest.  With
    // floating-point indices the reference adds each float term to the running
    // int index and truncates; here each step is one fmaf (the library's
    // declared "truncated float linear index"), so device code indexes exactly
    // as the library's nearest-neighbour sampler.
    template <typename T>
    __host__ __device__ int toIndex1d(const T (&axis_index)[D]) const {
        int idx = (int) axis_index[0];
        for (unsigned a = 1; a < D; ++a) {
            if constexpr (std::is_floating_point<T>::value)
                idx = (int) fmaf((float) _dimensional_increments[a], (float) axis_index[a], (float) idx);
            else
                idx += _dimensional_increments[a] * (int) axis_index[a];
        }
        return idx;
    }

    // toIndexNd (include/cudaTensor.cuh:182-188)
    __host__ __device__ void toIndexNd(uint32_t index1d, uint32_t (&indexNd)[D]) const {
        for (int a = (int) D - 1; a >= 0; --a) {
            indexNd[a] = index1d / (uint32_t) _dimensional_increments[a];
            index1d -= indexNd[a] * (uint32_t) _dimensional_increments[a];
        }
    }

    // element access on the device (include/cudaTensor.cuh:147-150, 162-164)
    __device__ precision &operator[](int index) const { return _data[index]; }
    __device__ precision &at(const uint32_t (&pt)[D]) const { return _data[toIndex1d(pt)]; }

    // display (include/cudaTensor.cuh:273-300): "Tensor <name> d0 x d1 ... elements
    // of <type>", then every value in one brace group, a line break after each
    // row_stride values.
    void display(const std::string &name = "", uint32_t row_stride = 0xFFFFFFFFu) const {
        const std::vector<precision> h = values();
        std::cout << "Tensor " << name << " ";
        for (unsigned d = 0; d < D; ++d) std::cout << _dims[d] << (d + 1 < D ? " x " : "");
        std::cout << " elements of " << typeid(precision).name() << "\n\n{ ";
        for (int32_t i = 0; i < _total_size; ++i) {
            std::cout << h[i] << (i + 1 < _total_size ? ", " : " ");
            if ((uint32_t) (i + 1) % row_stride == 0) std::cout << std::endl;
        }
        std::cout << "} " << std::endl;
    }

    // transform (include/cudaTensor.cuh:322-332, kernel transformProcess
    // include/cudaTensorKernels.cuh:341-351): *this[i] = fn(*this[i], A[i]) with a
    // __device__ callable fn, one thread per element.
    template <typename T>
    CudaTensor transform(const CudaTensor &A, T fn);

    void setValuesFromVector(const std::vector<precision> &vals) const {
        cgs_check_hip(hipMemcpy(_data, vals.data(), vals.size() * sizeof(precision), hipMemcpyHostToDevice),
                      "setValuesFromVector");
    }

    std::vector<precision> values() const {
        std::vector<precision> h(size());
        cgs_check_hip(hipMemcpy(h.data(), _data, bytesSize(), hipMemcpyDeviceToHost), "values");
        return h;
    }

    // include/cudaTensor.cuh:197-206 (fillProcess)
    void fill(precision value) {
        if constexpr (CgsDType<precision>::value >= 0)
            cgs_check(cgs_tensor_fill(_data, _total_size, CgsDType<precision>::value, (double) value, nullptr), "fill");
        else
            setValuesFromVector(std::vector<precision>(size(), value));
    }

    // operator+= / -= / *= (include/cudaTensor.cuh:322-360: transform with the
    // element-wise lambdas); a transform with an arbitrary device lambda has
    // no C-ABI form and is not offered.
    CudaTensor &operator+=(const CudaTensor &m) { return elementwise(CGS_OP_ADD, m); }
    CudaTensor &operator-=(const CudaTensor &m) { return elementwise(CGS_OP_SUB, m); }
    CudaTensor &operator*=(const CudaTensor &m) { return elementwise(CGS_OP_MUL, m); }

    // operator= (include/cudaTensor.cuh:334-348): takes m's shape and copies its
    // values device to device into this tensor's memory.
    CudaTensor &copyFrom(const CudaTensor &m) {
        for (unsigned d = 0; d < D; ++d) _dims[d] = m._dims[d];
        _total_size = m._total_size;
        cgs_check_hip(hipMemcpy(_data, m._data, m.bytesSize(), hipMemcpyDeviceToDevice), "operator=");
        return *this;
    }

  private:
    CudaTensor &elementwise(int op, const CudaTensor &m) {
        static_assert(CgsDType<precision>::value >= 0, "element-wise operations take float, double, int32 or uint8");
        if (m._total_size != _total_size) throw cgs_error(CGS_ERR_INVALID_ARGUMENT, "tensor sizes differ");
        cgs_check(cgs_tensor_binary(op, _data, m._data, _total_size, CgsDType<precision>::value, nullptr), "transform");
        return *this;
    }

  public:

    // include/cudaTensor.cuh:215-265: global minimum and its int32 index (-1
    // when no value is below +inf); prints the reference's line.
    void find_extrema(precision &extrema_value, int32_t &extrema_index1d, cgs_tie_mode tie = CGS_TIE_LOWEST) {
        cgs_argmin a;
        cgs_check(cgs_find_extrema((const float *) _data, _total_size, tie, &a, nullptr), "find_extrema");
        extrema_value = a.value;
        extrema_index1d = a.index;
        std::printf("\n Grid MIN value has idx: %d  value: %f\n", extrema_index1d, (double) extrema_value);
    }
};

// CudaMatrix<P>(rows, columns) (include/cudaTensor.cuh:363-505): dims
// {columns, rows}, row-major.
template <typename precision>
struct CudaMatrix : CudaTensor<precision, 2> {
    CudaMatrix(uint32_t rows, uint32_t columns) : CudaTensor<precision, 2>({columns, rows}) {}
    int32_t width() const { return this->_dims[0]; }
    int32_t height() const { return this->_dims[1]; }

    // include/cudaTensor.cuh:386-409: "Matrix name R x C elements of <type>" then one
    // "{ a, b, ... }" line per row
    void display(const std::string &name = "") const {
        const std::vector<precision> h = this->values();
        std::cout << "Matrix " << name << " " << height() << " x " << width() << " elements of "
                  << typeid(precision).name() << "\n\n";
        for (int r = 0; r < height(); ++r) {
            std::cout << "{ ";
            for (int c = 0; c < width() - 1; ++c) std::cout << h[(size_t) r * width() + c] << ", ";
            std::cout << h[(size_t) r * width() + width() - 1] << " }\n";
        }
        std::cout << std::endl;
    }

    // include/cudaTensor.cuh:419-423: the first height() diagonal elements
    void setDiagonal(const precision value) {
        const std::vector<precision> v((size_t) height(), value);
        cgs_check_hip(hipMemcpy2D(this->_data, (width() + 1) * sizeof(precision), v.data(), sizeof(precision),
                                  sizeof(precision), (size_t) std::min(height(), width()), hipMemcpyHostToDevice),
                      "setDiagonal");
    }

    // include/cudaTensor.cuh:433-441
    void setColumnFromVector(int column_index, const std::vector<precision> &values) {
        if (column_index < 0 || column_index >= width() || (int) values.size() != height())
            throw cgs_error(CGS_ERR_INVALID_ARGUMENT, "setColumnFromVector: bad column or length");
        cgs_check_hip(hipMemcpy2D(this->_data + column_index, width() * sizeof(precision), values.data(),
                                  sizeof(precision), sizeof(precision), (size_t) height(), hipMemcpyHostToDevice),
                      "setColumnFromVector");
    }

    // include/cudaTensor.cuh:451-456
    void setRowFromVector(int row_index, const std::vector<precision> &values) {
        if (row_index < 0 || row_index >= height() || (int) values.size() != width())
            throw cgs_error(CGS_ERR_INVALID_ARGUMENT, "setRowFromVector: bad row or length");
        cgs_check_hip(hipMemcpy(this->_data + (size_t) row_index * width(), values.data(),
                                width() * sizeof(precision), hipMemcpyHostToDevice),
                      "setRowFromVector");
    }

    // include/cudaTensor.cuh:464-470: zeros with ones at every (width + 1)-th element
    void setIdentity() {
        std::vector<precision> v(this->size(), (precision) 0);
        for (size_t d = 0; d < v.size(); d += (size_t) width() + 1) v[d] = (precision) 1;
        this->setValuesFromVector(v);
    }

    // include/cudaTensor.cuh:474-496: C = A * B
    static void multiply(const CudaMatrix &A, const CudaMatrix &B, CudaMatrix &C) {
        if (A.width() != B.height() || C.height() != A.height() || C.width() != B.width())
            throw cgs_error(CGS_ERR_INVALID_ARGUMENT, "multiply: shapes do not chain");
        cgs_check(cgs_matrix_multiply(A._data, B._data, C._data, A.height(), A.width(), B.width(),
                                      CgsDType<precision>::value, nullptr),
                  "multiply");
    }
};

// Per-channel action of CudaImage::filter (include/cudaTensor.cuh:64-67).
enum CHANNEL_ACTION { COPY = 0, FILTER = 1 };

// CudaImage<P, C>(H, W) (include/cudaTensor.cuh:507-596): dims {C, W, H}.
template <typename precision, uint32_t CHANNELS = 1>
struct CudaImage : CudaTensor<precision, 3> {
    __host__ __device__ CudaImage(uint32_t height, uint32_t width) : CudaTensor<precision, 3>({CHANNELS, width, height}) {}
    __host__ __device__ int32_t channels() const { return this->_dims[0]; }
    __host__ __device__ int32_t width() const { return this->_dims[1]; }
    __host__ __device__ int32_t height() const { return this->_dims[2]; }
    cgs_image view() const {
        return cgs_image{this->_data, width(), height(), channels(), CgsPixType<precision>::value};
    }

    // ---- device accessors (include/cudaTensor.cuh:531-575), for user functors.
    // They follow the library's declared arithmetic, so a functor that samples
    // through them sees what the library's kernels see: the truncated float
    // linear index with one fmaf per axis, bilinear taps at floor and ceil with
    // each lerp as fmaf(a, q, (1 - a) * p), and 0 for a tap past the image (the
    // reference reads past its buffer there).

    // toIndex (include/cudaTensor.cuh:531-534): element (y, x, c)
    template <typename T>
    __host__ __device__ int toIndex(T y, T x, T c = 0) const {
        const T ix[3] = {c, x, y};
        return this->toIndex1d(ix);
    }

    // at (include/cudaTensor.cuh:536-539)
    template <typename T>
    __device__ precision &at(T row_index, T column_index, T channel_index = 0) const {
        return this->_data[toIndex(row_index, column_index, channel_index)];
    }

    // inImage (include/cudaTensor.cuh:548-550)
    __host__ __device__ bool inImage(float y, float x) const {
        return x >= 0 && y >= 0 && x < (float) width() && y < (float) height();
    }

    // valueAt (include/cudaTensor.cuh:552-555): bilinear
    __device__ float valueAt(float y, float x, float c = 0) const { return valueAt_bilinear(y, x, c); }

    // valueAt_nearest_neighbor (include/cudaTensor.cuh:557-559)
    __device__ precision valueAt_nearest_neighbor(float y, float x, float c = 0) const {
        return inImage(y, x) ? (precision) tap(toIndex(y, x, c)) : (precision) 0;
    }

    // valueAt_bilinear (include/cudaTensor.cuh:561-575)
    __device__ float valueAt_bilinear(float y, float x, float c = 0) const {
#pragma clang fp contract(off)
        if (!inImage(y, x)) return 0.0f;
        const float fx = floorf(x), fy = floorf(y);
        const float ax = x - fx, ay = y - fy;
        const int C = channels(), W = width();
        const int b = ((int) fy * W + (int) fx) * C + (int) c;
        const int dx = ax != 0.0f ? C : 0, dy = ay != 0.0f ? C * W : 0;   // ceil == floor at an integer
        const float tl = tap(b), tr = tap(b + dx), bl = tap(b + dy), br = tap(b + dy + dx);
        const float top = fmaf(ax, tr, (1.0f - ax) * tl);
        const float bottom = fmaf(ax, br, (1.0f - ax) * bl);
        return fmaf(ay, bottom, (1.0f - ay) * top);
    }

    // filter (include/cudaTensor.cuh:577-596): filterProcess with a per-channel
    // COPY / FILTER action (CHANNEL_ACTION), through cgs_image_filter_channels;
    // the filter image's values are read back to the host first.
    CudaImage filter(const CudaImage<float, 1> &filt, CudaImage &out, const CHANNEL_ACTION actions[CHANNELS]) {
        const std::vector<float> taps = filt.values();
        uint32_t mask = 0;
        for (uint32_t c = 0; c < CHANNELS; ++c) mask |= (actions[c] == COPY ? 0u : 1u) << c;
        const cgs_image in = view(), ov = out.view();
        cgs_check(cgs_image_filter_channels(&in, &ov, taps.data(), filt.width(), filt.height(), mask, nullptr),
                  "CudaImage::filter");
        cgs_check_hip(hipDeviceSynchronize(), "CudaImage::filter");
        return out;
    }

  private:
    __device__ float tap(int i) const {
        return (i >= 0 && i < this->_total_size) ? (float) this->_data[i] : 0.0f;
    }
};

// nv_ext::Vec<T, D> (include/nvVectorNd.h): the fixed-size vector the
// reference's drivers pass homographies and grid points in.
namespace nv_ext {
template <typename T, unsigned int D>
struct Vec {
    T v[D] = {};
    Vec() = default;
    __host__ __device__ explicit Vec(const T (&a)[D]) {
        for (unsigned i = 0; i < D; ++i) v[i] = a[i];
    }
    __host__ __device__ explicit Vec(T fill) {
        for (unsigned i = 0; i < D; ++i) v[i] = fill;
    }
    __host__ __device__ T &operator[](unsigned i) { return v[i]; }
    __host__ __device__ const T &operator[](unsigned i) const { return v[i]; }
    __host__ __device__ constexpr unsigned size() const { return D; }
};
}  // namespace nv_ext

// --------------------------------------------------------------- image files

inline bool endsWithCaseInsensitive(const std::string &s, const std::string &suffix) {
    if (s.size() < suffix.size()) return false;
    for (size_t i = 0; i < suffix.size(); ++i) {
        const char a = s[s.size() - suffix.size() + i], b = suffix[i];
        if (std::tolower((unsigned char) a) != std::tolower((unsigned char) b)) return false;
    }
    return true;
}

// The drivers' suffix rule (src/gpu/cudaImageFunctions.cuh:140-149): .png is
// written (cgs_write_png); .jpg is not (this build has no JPEG encoder) and is
// reported; anything else prints the reference's message.
inline void cgs_write_image_file(const std::string &name, const std::vector<uint8_t> &pixels, int w, int h, int c) {
    if (endsWithCaseInsensitive(name, ".png"))
        cgs_check(cgs_write_png(name.c_str(), pixels.data(), w, h, c), "stbi_write_png");
    else if (endsWithCaseInsensitive(name, ".jpg"))
        std::cout << "JPEG output is not part of this build; write a .png file instead." << std::endl;
    else
        std::cout << "Filename suffix has image format not recognized." << std::endl;
}

template <typename HV>
inline void cgs_h8(const HV &H, float (&h)[8]) {
    for (int k = 0; k < 8; ++k) h[k] = (float) H[k];
}

// writeTransformedImageToDisk (src/gpu/cudaImageFunctions.cuh:129-152): image
// under H on a rowsf x colsf grid, zero where its point falls outside, as an
// 8-bit file.  The _local twin (src/cpu/multispectral_ImageMatcher.cu:205-229)
// runs the same kernel.
template <typename pixType, uint32_t CHANNELS, typename HV>
void writeTransformedImageToDisk(CudaImage<pixType, CHANNELS> image, int rowsf, int colsf, const HV &H,
                                 std::string img_out_filename) {
    CudaImage<uint8_t, CHANNELS> out((uint32_t) rowsf, (uint32_t) colsf);
    cgs_check_hip(hipMalloc(&out.data(), out.bytesSize()), "hipMalloc");
    float h[8];
    cgs_h8(H, h);
    const cgs_image iv = image.view(), ov = out.view();
    cgs_status st = CGS_OK;
    std::vector<uint8_t> host;
    if (hipMemset(out.data(), 0, out.bytesSize()) == hipSuccess && (st = cgs_image_warp(&iv, &ov, h, nullptr)) == CGS_OK)
        host = out.values();
    (void) hipFree(out.data());
    cgs_check(st, "writeTransformedImageToDisk");
    cgs_write_image_file(img_out_filename, host, colsf, rowsf, (int) CHANNELS);
}

template <typename pixType, uint32_t CHANNELS, typename HV>
void writeTransformedImageToDisk_local(CudaImage<pixType, CHANNELS> image, int rowsf, int colsf, const HV &H,
                                       std::string img_out_filename) {
    writeTransformedImageToDisk<pixType, CHANNELS>(image, rowsf, colsf, H, img_out_filename);
}

// writeAlignedAndFusedImageToDisk (src/gpu/cudaImageFunctions.cuh:154-179) and
// its _local twin (src/cpu/multispectral_ImageMatcher.cu:231-255): the
// 3-channel overlay of image_fix's size (fuseAlignedImages / _local).
template <typename pixType, uint32_t CHANNELS, typename HV>
void cgs_write_fused(CudaImage<pixType, CHANNELS> image_fix, CudaImage<pixType, CHANNELS> image_mov, const HV &est,
                     const HV &truth, const std::string &name, int local) {
    CudaImage<uint8_t, 3> fused((uint32_t) image_fix.height(), (uint32_t) image_fix.width());
    cgs_check_hip(hipMalloc(&fused.data(), fused.bytesSize()), "hipMalloc");
    float e[8], t[8];
    cgs_h8(est, e);
    cgs_h8(truth, t);
    const cgs_image rv = image_fix.view(), mv = image_mov.view(), fv = fused.view();
    cgs_status st = CGS_OK;
    std::vector<uint8_t> host;
    if (hipMemset(fused.data(), 0, fused.bytesSize()) == hipSuccess &&
        (st = cgs_image_fuse(&rv, &mv, e, t, 0, local, &fv, nullptr)) == CGS_OK)
        host = fused.values();
    (void) hipFree(fused.data());
    cgs_check(st, "writeAlignedAndFusedImageToDisk");
    cgs_write_image_file(name, host, image_fix.width(), image_fix.height(), 3);
}

template <typename pixType, uint32_t CHANNELS, typename HV>
void writeAlignedAndFusedImageToDisk(CudaImage<pixType, CHANNELS> image_fix, CudaImage<pixType, CHANNELS> image_mov,
                                     const HV &estimatedH, const HV &trueH, std::string img_fused_filename) {
    cgs_write_fused<pixType, CHANNELS>(image_fix, image_mov, estimatedH, trueH, img_fused_filename, 0);
}

template <typename pixType, uint32_t CHANNELS, typename HV>
void writeAlignedAndFusedImageToDisk_local(CudaImage<pixType, CHANNELS> image_fix, CudaImage<pixType, CHANNELS> image_mov,
                                           const HV &estimatedH, const HV &trueH, std::string img_fused_filename) {
    cgs_write_fused<pixType, CHANNELS>(image_fix, image_mov, estimatedH, trueH, img_fused_filename, 1);
}

// ------------------------------------------------- device callables, functors

// func_byvalue_t / func_byreference_t (src/gpu/cudaGridSearch.cuh:284-285,
// 484-485): the signatures of a user's cost functor.  Bind one as the
// reference's drivers do:
//   __device__ func_byvalue_t<float, float, 8, CudaImage<uint8_t>, CudaImage<uint8_t>> dev_ptr = myCost<...>;
//   hipMemcpyFromSymbol(&host_ptr, HIP_SYMBOL(dev_ptr), sizeof(host_ptr));
//   searcher.search_by_value(host_ptr, image_mov, image_fix);
template <typename func_precision, typename grid_precision, unsigned int D, typename... Types>
using func_byvalue_t = func_precision (*)(nv_ext::Vec<grid_precision, D> &gridpt, Types... arg_vals);
template <typename func_precision, typename grid_precision, unsigned int D, typename... Types>
using func_byreference_t = func_precision (*)(nv_ext::Vec<grid_precision, D> &gridpt, Types *... arg_ptrs);

namespace cgs_facade {

// indexToGridPoint (src/gpu/cudaGridSearch.cuh:160-187) as the library's
// kernels compute it: mixed-radix decode, axis 0 fastest, each coordinate
// fmaf(offset + i, resolution, start).
template <typename G, unsigned D>
__device__ inline void grid_point(const cgs_grid &g, int32_t index, nv_ext::Vec<G, D> &pt) {
    int32_t rem = index;
    for (int a = (int) D - 1; a >= 0; --a) {
        const int32_t i = rem / g.increments[a];
        rem -= i * g.increments[a];
        pt[a] = (G) fmaf((float) (i + g.offset[a]), g.resolution[a], g.start[a]);
    }
}

// evaluationKernel_by_value / _by_reference (src/gpu/cudaGridSearch.cuh:287-314,
// 487-511): one thread per grid point.
template <typename F, typename G, unsigned D, typename... Types>
__global__ void by_value_kernel(cgs_grid g, F *result, int32_t n, func_byvalue_t<F, G, D, Types...> op,
                                Types... args) {
    const int32_t i = (int32_t) (blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    nv_ext::Vec<G, D> pt;
    grid_point(g, i, pt);
    result[i] = (*op)(pt, args...);
}

template <typename F, typename G, unsigned D, typename... Types>
__global__ void by_reference_kernel(cgs_grid g, F *result, int32_t n, func_byreference_t<F, G, D, Types...> op,
                                    Types *... args) {
    const int32_t i = (int32_t) (blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    nv_ext::Vec<G, D> pt;
    grid_point(g, i, pt);
    result[i] = (*op)(pt, args...);
}

// evaluationKernel_by_value_block / _stream and _by_reference_stream
// (src/gpu/cudaGridSearch.cuh:421-482, 513-538): one workgroup per grid point,
// every thread calls the functor (the cooperative "stream" functors take
// threadIdx.x as their row and reduce through shared memory), and thread 0's
// value is stored.  The reference lets every thread store its own return
// value, racing with thread 0's; thread 0's is the one its functors compute.
template <typename F, typename G, unsigned D, typename... Types>
__global__ void by_value_block_kernel(cgs_grid g, F *result, int32_t n, func_byvalue_t<F, G, D, Types...> op,
                                      Types... args) {
    const int32_t i = (int32_t) blockIdx.x;
    if (i >= n) return;
    nv_ext::Vec<G, D> pt;
    grid_point(g, i, pt);
    const F v = (*op)(pt, args...);
    if (threadIdx.x == 0) result[i] = v;
}

template <typename F, typename G, unsigned D, typename... Types>
__global__ void by_reference_block_kernel(cgs_grid g, F *result, int32_t n,
                                          func_byreference_t<F, G, D, Types...> op, Types *... args) {
    const int32_t i = (int32_t) blockIdx.x;
    if (i >= n) return;
    nv_ext::Vec<G, D> pt;
    grid_point(g, i, pt);
    const F v = (*op)(pt, args...);
    if (threadIdx.x == 0) result[i] = v;
}

// transformProcess (include/cudaTensorKernels.cuh:341-351)
template <typename P, unsigned D, typename T>
__global__ void transform_kernel(CudaTensor<P, D> A, CudaTensor<P, D> B, T fn) {
    const int32_t i = (int32_t) (blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= A._total_size) return;
    A._data[i] = fn(A._data[i], B._data[i]);
}

}  // namespace cgs_facade

template <typename precision, unsigned int D>
template <typename T>
CudaTensor<precision, D> CudaTensor<precision, D>::transform(const CudaTensor &A, T fn) {
    if (A._total_size != _total_size) throw cgs_error(CGS_ERR_INVALID_ARGUMENT, "transform: tensor sizes differ");
    if (_total_size > 0) {
        hipLaunchKernelGGL((cgs_facade::transform_kernel<precision, D, T>), dim3((_total_size + 255) / 256), dim3(256),
                           0, nullptr, *this, A, fn);
        cgs_check_hip(hipGetLastError(), "transform");
        cgs_check_hip(hipDeviceSynchronize(), "transform");
    }
    return *this;
}

// --------------------------------------------------------------------- grid

// CudaGrid<P, D> (src/gpu/cudaGridSearch.cuh:74-231).
template <typename precision, unsigned int D>
struct CudaGrid {
    precision *_data = nullptr;   // optional 4 x D device mirror (start, end, resolution, count rows)
    cgs_grid _g{};
    float _start[D] = {}, _end[D] = {}, _count[D];

    CudaGrid() {
        if (cgs_abi_version() != CGS_ABI_VERSION)
            throw cgs_error(CGS_ERR_UNSUPPORTED, "libcgs was built with another cgs.h (ABI version mismatch)");
        for (unsigned d = 0; d < D; ++d) _count[d] = 1.0f;
        update();
    }

    precision *&data() { return _data; }
    uint32_t bytesSize() const { return 4 * D * sizeof(precision); }
    int getDimension() const { return D; }

    // update() -> updateProcess (src/gpu/cudaGridSearch.cuh:100-105, 256-270)
    void update(int = -1) {
        cgs_check(cgs_grid_init(&_g, D, _start, _end, _count), "CudaGrid::update");
        if (_data) {
            precision rows[4 * D];
            for (unsigned d = 0; d < D; ++d) {
                rows[d] = _g.start[d];
                rows[D + d] = _g.end[d];
                rows[2 * D + d] = _g.resolution[d];
                rows[3 * D + d] = _g.count[d];
            }
            cgs_check_hip(hipMemcpy(_data, rows, sizeof(rows), hipMemcpyHostToDevice), "CudaGrid::update");
        }
    }

    template <typename T> void setStartPoint(std::vector<T> v) { set_row(_start, v); }
    template <typename T> void setEndPoint(std::vector<T> v) { set_row(_end, v); }
    template <typename T> void setNumSamples(std::vector<T> v) { set_row(_count, v); }

    uint32_t numElements() const { return (uint32_t) _g.num_elements; }

    void getAxisSampleCounts(precision *axis_sample_counts) const {
        for (unsigned d = 0; d < D; ++d) axis_sample_counts[d] = _g.count[d];
    }

    void getGridPoint(precision (&grid_point)[D], int index) const {
        float p[CGS_MAX_DIM];
        cgs_check(cgs_grid_point(&_g, index, p), "getGridPoint");
        for (unsigned d = 0; d < D; ++d) grid_point[d] = p[d];
    }

    // CudaMatrix::display format (include/cudaTensor.cuh:395-415); the first
    // four {..} groups are what the MATLAB harness parses.
    void display(const std::string &name = "") const {
        std::cout << "Matrix " << name << " 4 x " << D << " elements of f\n\n";
        const float *rows[4] = {_g.start, _g.end, _g.resolution, _g.count};
        for (auto r : rows) {
            std::cout << "{ ";
            for (unsigned d = 0; d + 1 < D; ++d) std::cout << r[d] << ", ";
            std::cout << r[D - 1] << " }\n";
        }
        std::cout << std::endl;
    }

  private:
    template <typename T> void set_row(float *row, const std::vector<T> &v) {
        if (v.size() != D) throw cgs_error(CGS_ERR_INVALID_ARGUMENT, "CudaGrid: vector size differs from D");
        for (unsigned d = 0; d < D; ++d) row[d] = (float) v[d];
        update();
    }
};

// --------------------------------------------------------------- searcher

// CudaGridSearcher<F, G, D> (src/gpu/cudaGridSearch.cuh:540-855).  Every
// search variant fills the result tensor with the functor's cost per grid
// point (grid-index order) and synchronises; the fused argmin of the last
// search is kept in last_argmin.
template <typename func_precision, typename grid_precision, unsigned int D>
struct CudaGridSearcher {
    CudaGrid<grid_precision, D> *_grid;
    CudaTensor<func_precision, D> *_result;
    cgs_argmin last_argmin{INFINITY, -1};
    cgs_tie_mode tie_mode = CGS_TIE_LOWEST;
    hipStream_t stream = nullptr;

    CudaGridSearcher(CudaGrid<grid_precision, D> &grid, CudaTensor<func_precision, D> &result)
        : _grid(&grid), _result(&result) {}

    template <typename Img> void search(const CgsFunctor &f, Img &moved, Img &fixed) {
        search_by_value(f, moved, fixed);
    }

    template <typename Img> void search_by_value(const CgsFunctor &f, Img &moved, Img &fixed) {
        if (_result->size() != _grid->numElements())
            throw cgs_error(CGS_ERR_INVALID_ARGUMENT, "result tensor size differs from the grid's point count");
        std::cout << "CudaGrid has " << _grid->numElements() << " search samples." << std::endl;
        cgs_search_opts o;
        cgs_default_opts(&o);
        o.cost = f.cost;
        o.param_mode = f.mode;
        o.sampler = f.sampler;
        o.mi_bins = f.bins;
        o.tie_mode = tie_mode;
        const cgs_image m = moved.view(), x = fixed.view();
        cgs_check(cgs_search(&_grid->_g, &m, &x, &o, (float *) _result->data(), &last_argmin, stream),
                  "search_by_value");
    }

    // The launch-policy variants compute the same costs with the one fused
    // kernel; their tuning arguments are accepted and unused.
    // CudaGridSearcher::covariance_by_value (src/gpu/cudaGridSearch.cuh:768-794):
    // adds the D x D covariance around the result's minimum into covar_matrix
    // (host, row-major; the reference copies (D+1)^2 values into a D^2 array,
    // here exactly D^2 are touched).
    void covariance_by_value(float *covar_matrix) {
        std::cout << "CudaGrid has " << _grid->numElements() << " search samples." << std::endl;
        float c[CGS_MAX_DIM * CGS_MAX_DIM];
        cgs_check(cgs_covariance(&_grid->_g, (const float *) _result->data(), c, nullptr, stream),
                  "covariance_by_value");
        for (unsigned i = 0; i < D * D; ++i) covar_matrix[i] += c[i];
    }

    // ---- a user's own __device__ functor (src/gpu/cudaGridSearch.cuh:540-765):
    // the costs land in the result tensor in grid order and last_argmin holds
    // find_extrema's minimum of them (tie_mode).  Synchronises the stream.
    template <typename... Types>
    void search(func_byvalue_t<func_precision, grid_precision, D, Types...> fn, Types &... args) {
        search_by_value(fn, args...);
    }

    // evaluationKernel_by_value: one thread per grid point
    template <typename... Types>
    void search_by_value(func_byvalue_t<func_precision, grid_precision, D, Types...> fn, Types &... args) {
        const int32_t n = begin_user_search();
        hipLaunchKernelGGL((cgs_facade::by_value_kernel<func_precision, grid_precision, D, Types...>),
                           dim3((n + 255) / 256), dim3(256), 0, stream, _grid->_g, _result->data(), n, fn, args...);
        finish_user_search("search_by_value");
    }

    // evaluationKernel_by_value_stream: one workgroup of numCores threads per
    // grid point (at most 1024); STREAM_BLOCK_DIM, the reference's stream
    // count, is accepted and unused (one launch covers every point)
    template <typename... Types>
    void search_by_value_stream(func_byvalue_t<func_precision, grid_precision, D, Types...> fn, int STREAM_BLOCK_DIM,
                                int numCores, Types &... args) {
        (void) STREAM_BLOCK_DIM;
        search_by_value_block(fn, numCores, args...);
    }

    // evaluationKernel_by_value_block: one workgroup of numThreads threads per point
    template <typename... Types>
    void search_by_value_block(func_byvalue_t<func_precision, grid_precision, D, Types...> fn, int numThreads,
                               Types &... args) {
        const int32_t n = begin_user_search();
        hipLaunchKernelGGL((cgs_facade::by_value_block_kernel<func_precision, grid_precision, D, Types...>), dim3(n),
                           dim3(block_threads(numThreads)), 0, stream, _grid->_g, _result->data(), n, fn, args...);
        finish_user_search("search_by_value_block");
    }

    // evaluationKernel_by_reference: the arguments are device pointers
    template <typename... Types>
    void search_by_reference(func_byreference_t<func_precision, grid_precision, D, Types...> fn, Types *... args) {
        const int32_t n = begin_user_search();
        hipLaunchKernelGGL((cgs_facade::by_reference_kernel<func_precision, grid_precision, D, Types...>),
                           dim3((n + 255) / 256), dim3(256), 0, stream, _grid->_g, _result->data(), n, fn, args...);
        finish_user_search("search_by_reference");
    }

    template <typename... Types>
    void search_by_reference_stream(func_byreference_t<func_precision, grid_precision, D, Types...> fn,
                                    int STREAM_BLOCK_DIM, int numCores, Types *... args) {
        (void) STREAM_BLOCK_DIM;
        const int32_t n = begin_user_search();
        hipLaunchKernelGGL((cgs_facade::by_reference_block_kernel<func_precision, grid_precision, D, Types...>),
                           dim3(n), dim3(block_threads(numCores)), 0, stream, _grid->_g, _result->data(), n, fn,
                           args...);
        finish_user_search("search_by_reference_stream");
    }

  private:
    static int block_threads(int t) { return t < 1 ? 1 : (t > 1024 ? 1024 : t); }

    int32_t begin_user_search() {
        if (_result->size() != _grid->numElements())
            throw cgs_error(CGS_ERR_INVALID_ARGUMENT, "result tensor size differs from the grid's point count");
        if (!_result->data()) throw cgs_error(CGS_ERR_INVALID_ARGUMENT, "result tensor has no device memory");
        std::cout << "CudaGrid has " << _grid->numElements() << " search samples." << std::endl;
        return (int32_t) _grid->numElements();
    }

    void finish_user_search(const char *what) {
        cgs_check_hip(hipGetLastError(), what);
        cgs_check_hip(hipStreamSynchronize(stream), what);
        if constexpr (std::is_same<func_precision, float>::value) {
            cgs_check(cgs_find_extrema((const float *) _result->data(), _result->size(), tie_mode, &last_argmin,
                                       stream),
                      what);
        } else {   // the C-ABI's argmin reads floats: the lowest-index minimum on the host
            const std::vector<func_precision> v = _result->values();
            last_argmin = cgs_argmin{INFINITY, -1};
            for (size_t i = 0; i < v.size(); ++i)
                if ((float) v[i] < last_argmin.value) last_argmin = cgs_argmin{(float) v[i], (int32_t) i};
        }
    }

  public:
    template <typename Img> void search_by_value_stream(const CgsFunctor &f, int, int, Img &m, Img &x) {
        search_by_value(f, m, x);
    }
    template <typename Img> void search_by_value_block(const CgsFunctor &f, int, Img &m, Img &x) {
        search_by_value(f, m, x);
    }
    template <typename Img> void search_by_reference(const CgsFunctor &f, Img *m, Img *x) {
        search_by_value(f, *m, *x);
    }
    template <typename Img> void search_by_reference_stream(const CgsFunctor &f, int, int, Img *m, Img *x) {
        search_by_value(f, *m, *x);
    }
};
