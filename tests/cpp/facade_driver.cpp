// facade_driver.cpp -- a search driver written against include/cuGridSearch.hpp
// the way the reference's drivers are written against its headers
// (src/gpu/cudaGridSearch_main_mi.cu:300-345, src/gpu/cudaGridSearch_example1.cu:97-143).
//
// usage: facade_driver <spec.txt> <moving.raw> <fixed.raw> <costs_out.bin> [png_prefix]
//        facade_driver --matrix   (the CudaMatrix / CudaTensor utilities)
// With png_prefix it also writes the output images <prefix>_out.png,
// <prefix>_fused.png and <prefix>_fused_local.png.
// spec.txt holds whitespace-separated fields:
//   functor sampler pixtype width height channels tie bins start[8] end[8] count[8]
// functor: ncc | mi | sad | sad_sum | sqd ; sampler: nearest | bilinear ; pixtype: u8 | f32
// Prints one JSON line with the searcher's fused argmin and find_extrema's.
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "cuGridSearch.hpp"

static CgsFunctor functor_named(const std::string &n) {
    if (n == "ncc") return calcNCC;
    if (n == "mi") return calcMIstream;
    if (n == "sqd") return calcSQD;
    if (n == "sad_sum") return sumOfAbsoluteDifferences;
    return centredAbsoluteDifference;
}

template <typename pixel_precision>
static int run(std::ifstream &spec, const std::string &fn, const std::string &samp, int W, int H, int tie, int bins,
               const char *mov_path, const char *fix_path, const char *out_path, const char *png_prefix) {
    constexpr unsigned grid_dimension = 8;
    typedef float grid_precision;
    typedef float func_precision;
    std::vector<grid_precision> start_point(grid_dimension), end_point(grid_dimension), num_samples(grid_dimension);
    for (auto &v : start_point) spec >> v;
    for (auto &v : end_point) spec >> v;
    for (auto &v : num_samples) spec >> v;

    std::vector<pixel_precision> mov((size_t) W * H), fix((size_t) W * H);
    std::ifstream(mov_path, std::ios::binary).read((char *) mov.data(), mov.size() * sizeof(pixel_precision));
    std::ifstream(fix_path, std::ios::binary).read((char *) fix.data(), fix.size() * sizeof(pixel_precision));

    CudaImage<pixel_precision> image_mov(H, W), image_fix(H, W);
    cgs_check_hip(hipMalloc(&image_mov.data(), image_mov.bytesSize()), "hipMalloc");
    cgs_check_hip(hipMalloc(&image_fix.data(), image_fix.bytesSize()), "hipMalloc");
    image_mov.setValuesFromVector(mov);
    image_fix.setValuesFromVector(fix);

    CudaGrid<grid_precision, grid_dimension> grid;
    cgs_check_hip(hipMalloc(&grid.data(), grid.bytesSize()), "hipMalloc");
    grid.setStartPoint(start_point);
    grid.setEndPoint(end_point);
    grid.setNumSamples(num_samples);
    grid.display("grid");

    grid_precision axis_sample_counts[grid_dimension];
    grid.getAxisSampleCounts(axis_sample_counts);
    CudaTensor<func_precision, grid_dimension> func_values(axis_sample_counts);
    cgs_check_hip(hipMalloc(&func_values.data(), func_values.bytesSize()), "hipMalloc");

    CudaGridSearcher<func_precision, grid_precision, grid_dimension> searcher(grid, func_values);
    searcher.tie_mode = (cgs_tie_mode) tie;
    CgsFunctor f = functor_named(fn).with_bins(bins);
    f = f.with_sampler(samp == "nearest" ? CGS_SAMPLER_NEAREST : CGS_SAMPLER_BILINEAR);
    searcher.search_by_value(f, image_mov, image_fix);

    func_precision min_value;
    int32_t min_value_index1d;
    func_values.find_extrema(min_value, min_value_index1d, (cgs_tie_mode) tie);
    grid_precision min_grid_point[grid_dimension];
    grid.getGridPoint(min_grid_point, min_value_index1d);

    // covariance_by_value adds into the caller's matrix (src/cpu/multispectral_ImageMatcher.cu:818-925
    // zeroes it first)
    float covar[grid_dimension * grid_dimension] = {};
    searcher.covariance_by_value(covar);
    double covar_trace = 0.0;
    for (unsigned d = 0; d < grid_dimension; ++d) covar_trace += covar[d * grid_dimension + d];

    if (png_prefix) {
        // the drivers' output images (src/gpu/cudaGridSearch_main.cu:312-316), here
        // under a fixed shift by (2, 1) and the identity
        const float shift_vals[8] = {1, 0, 2, 0, 1, 1, 0, 0}, ident_vals[8] = {1, 0, 0, 0, 1, 0, 0, 0};
        nv_ext::Vec<float, 8> shift(shift_vals), ident(ident_vals);
        const std::string pre(png_prefix);
        writeTransformedImageToDisk<pixel_precision, 1>(image_mov, H, W, shift, pre + "_out.png");
        writeAlignedAndFusedImageToDisk<pixel_precision, 1>(image_fix, image_mov, shift, ident, pre + "_fused.png");
        writeAlignedAndFusedImageToDisk_local<pixel_precision, 1>(image_fix, image_mov, ident, ident,
                                                                  pre + "_fused_local.png");
    }

    const std::vector<func_precision> costs = func_values.values();
    std::ofstream(out_path, std::ios::binary).write((const char *) costs.data(), costs.size() * sizeof(float));
    std::printf("{\"fused_index\": %d, \"fused_value\": %.9g, \"extrema_index\": %d, \"extrema_value\": %.9g, "
                "\"num_elements\": %u, \"point0\": %.9g, \"covar_trace\": %.9g}\n",
                searcher.last_argmin.index, (double) searcher.last_argmin.value, min_value_index1d,
                (double) min_value, grid.numElements(), (double) min_grid_point[0], covar_trace);
    (void) hipFree(image_mov.data());
    (void) hipFree(image_fix.data());
    (void) hipFree(grid.data());
    (void) hipFree(func_values.data());
    return 0;
}

// --matrix: the CudaMatrix / CudaTensor utilities on fixed small operands;
// prints one JSON line with the results.
static int matrix_selftest() {
    CudaMatrix<float> A(3, 4), B(4, 2), C(3, 2);
    cgs_check_hip(hipMalloc(&A.data(), A.bytesSize()), "hipMalloc");
    cgs_check_hip(hipMalloc(&B.data(), B.bytesSize()), "hipMalloc");
    cgs_check_hip(hipMalloc(&C.data(), C.bytesSize()), "hipMalloc");
    A.setIdentity();
    A.setDiagonal(2.0f);
    A.setRowFromVector(2, {1.0f, -2.0f, 3.0f, 0.5f});
    A.setColumnFromVector(3, {4.0f, 5.0f, 6.0f});
    B.setValuesFromVector({1, 2, 3, 4, 5, 6, 7, 8});
    CudaMatrix<float>::multiply(A, B, C);
    C.display("C");
    CudaTensor<float, 2> T({2u, 3u});
    cgs_check_hip(hipMalloc(&T.data(), T.bytesSize()), "hipMalloc");
    T.copyFrom(C);
    T += C;   // 2C
    T *= C;   // 2C^2
    T -= C;   // 2C^2 - C
    CudaTensor<uint8_t, 1> u({5u});
    cgs_check_hip(hipMalloc(&u.data(), u.bytesSize()), "hipMalloc");
    u.fill(200);
    u += u;   // 400 mod 256
    const std::vector<float> a = A.values(), c = C.values(), t = T.values();
    const std::vector<uint8_t> uv = u.values();
    auto list = [](const auto &v) {
        std::string r = "[";
        for (size_t i = 0; i < v.size(); ++i) r += (i ? ", " : "") + std::to_string((double) v[i]);
        return r + "]";
    };
    std::printf("{\"A\": %s, \"C\": %s, \"T\": %s, \"u\": %s}\n", list(a).c_str(), list(c).c_str(),
                list(t).c_str(), list(uv).c_str());
    for (void *p : {(void *) A.data(), (void *) B.data(), (void *) C.data(), (void *) T.data(), (void *) u.data()})
        (void) hipFree(p);
    return 0;
}

int main(int argc, char **argv) {
    if (argc == 2 && std::string(argv[1]) == "--matrix") return matrix_selftest();
    if (argc != 5 && argc != 6) {
        std::cerr << "usage: facade_driver <spec.txt> <moving.raw> <fixed.raw> <costs_out.bin>\n";
        return 2;
    }
    std::ifstream spec(argv[1]);
    std::string fn, samp, pix;
    int W, H, C, tie, bins;
    spec >> fn >> samp >> pix >> W >> H >> C >> tie >> bins;
    try {
        const char *png = argc == 6 ? argv[5] : nullptr;
        if (pix == "u8") return run<uint8_t>(spec, fn, samp, W, H, tie, bins, argv[2], argv[3], argv[4], png);
        return run<float>(spec, fn, samp, W, H, tie, bins, argv[2], argv[3], argv[4], png);
    } catch (const cgs_error &e) {
        std::cerr << "cgs_error: " << e.what() << "\n";
        return 1;
    }
}
