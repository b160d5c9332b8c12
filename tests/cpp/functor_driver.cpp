// functor_driver.cpp -- user-written __device__ cost functors searched through
// the facade's generic path, bound the way the reference's drivers bind theirs
// (a __device__ func_byvalue_t / func_byreference_t variable read back with
// hipMemcpyFromSymbol; src/gpu/cudaGridSearch_main.cu:69-72, 270-271,
// src/gpu/cudaGridSearch_example2.cu:35-40, 115-126).
//
// usage: functor_driver <spec.txt> <moving.raw> <fixed.raw> <outdir>
// spec.txt: W H, then the translation grid (start end count for x and y), the
// probe grid (the same for x and y), then start[8] end[8] count[8] of an 8-D
// grid.  uint8 images.  Writes <outdir>/<name>.bin (float costs, grid order)
// and prints one JSON line of {name: [argmin index, value]}.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "cuGridSearch.hpp"

typedef CudaImage<uint8_t, 1> Img;
typedef float func_precision;
typedef float grid_precision;

// ---------------------------------------------------------------- functors

// Sum of absolute differences under a 2-D translation, sampled with valueAt
// as sumOfAbsoluteDifferences does (x outer, y inner), summed in binary64.
__device__ float user_sad(nv_ext::Vec<grid_precision, 2> &t, Img moved, Img fixed) {
    double sum = 0.0;
    for (int x = 0; x < fixed.width(); x++)
        for (int y = 0; y < fixed.height(); y++)
            if (moved.inImage(y + t[1], x + t[0]))
                sum += fabsf(moved.valueAt(y + t[1], x + t[0]) - fixed.valueAt((float) y, (float) x));
    return (float) sum;
}

// The same cost through pointers to device copies of the image structures.
__device__ float user_sad_ref(nv_ext::Vec<grid_precision, 2> &t, Img *moved, Img *fixed) {
    return user_sad(t, *moved, *fixed);
}

// A cooperative functor: thread r of the workgroup sums rows r, r + blockDim.x,
// ...; thread 0 adds the row sums in row order and returns the total (the
// other threads' values are not stored).
__device__ float user_sad_rows(nv_ext::Vec<grid_precision, 2> &t, Img moved, Img fixed) {
    __shared__ double rows[1024];
    for (int y = threadIdx.x; y < fixed.height(); y += blockDim.x) {
        double s = 0.0;
        for (int x = 0; x < fixed.width(); x++)
            if (moved.inImage(y + t[1], x + t[0]))
                s += fabsf(moved.valueAt(y + t[1], x + t[0]) - fixed.valueAt((float) y, (float) x));
        rows[y] = s;
    }
    __syncthreads();
    double total = 0.0;
    if (threadIdx.x == 0)
        for (int y = 0; y < fixed.height(); y++) total += rows[y];
    __syncthreads();
    return (float) total;
}

// Sampler probes: the moving image at (y, x) = (p[1], p[0]).
__device__ float probe_bilinear(nv_ext::Vec<grid_precision, 2> &p, Img moved, Img) {
    return moved.valueAt_bilinear(p[1], p[0]);
}
__device__ float probe_nearest(nv_ext::Vec<grid_precision, 2> &p, Img moved, Img) {
    return (float) moved.valueAt_nearest_neighbor(p[1], p[0]);
}

// The grid point's coordinate A, for checking the 8-D decode.
template <int A>
__device__ float axis_value(nv_ext::Vec<grid_precision, 8> &p, Img, Img) {
    return p[A];
}

typedef func_byvalue_t<func_precision, grid_precision, 2, Img, Img> fn2_t;
typedef func_byreference_t<func_precision, grid_precision, 2, Img, Img> fn2_ref_t;
typedef func_byvalue_t<func_precision, grid_precision, 8, Img, Img> fn8_t;

__device__ fn2_t dev_user_sad = user_sad;
__device__ fn2_ref_t dev_user_sad_ref = user_sad_ref;
__device__ fn2_t dev_user_sad_rows = user_sad_rows;
__device__ fn2_t dev_probe_bilinear = probe_bilinear;
__device__ fn2_t dev_probe_nearest = probe_nearest;
__device__ fn8_t dev_axis[8] = {axis_value<0>, axis_value<1>, axis_value<2>, axis_value<3>,
                                axis_value<4>, axis_value<5>, axis_value<6>, axis_value<7>};

// ------------------------------------------------------------------ driver

static std::string g_json = "{";

template <unsigned D>
static void emit(const std::string &dir, const std::string &name, CudaTensor<float, D> &res, const cgs_argmin &am) {
    const std::vector<float> v = res.values();
    std::ofstream(dir + "/" + name + ".bin", std::ios::binary).write((const char *) v.data(), v.size() * sizeof(float));
    g_json += (g_json.size() > 1 ? ", \"" : "\"") + name + "\": [" + std::to_string(am.index) + ", " +
              std::to_string((double) am.value) + "]";
}

template <unsigned D>
static void make_grid(CudaGrid<grid_precision, D> &g, const std::vector<float> &s, const std::vector<float> &e,
                      const std::vector<float> &c) {
    g.setStartPoint(s);
    g.setEndPoint(e);
    g.setNumSamples(c);
}

int main(int argc, char **argv) {
    if (argc != 5) {
        std::cerr << "usage: functor_driver <spec.txt> <moving.raw> <fixed.raw> <outdir>\n";
        return 2;
    }
    try {
        std::ifstream spec(argv[1]);
        int W, H;
        spec >> W >> H;
        std::vector<float> ts(2), te(2), tc(2), ps(2), pe(2), pc(2), s8(8), e8(8), c8(8);
        for (int a = 0; a < 2; ++a) spec >> ts[a] >> te[a] >> tc[a];
        for (int a = 0; a < 2; ++a) spec >> ps[a] >> pe[a] >> pc[a];
        for (auto &v : s8) spec >> v;
        for (auto &v : e8) spec >> v;
        for (auto &v : c8) spec >> v;
        const std::string dir(argv[4]);

        std::vector<uint8_t> mov((size_t) W * H), fix((size_t) W * H);
        std::ifstream(argv[2], std::ios::binary).read((char *) mov.data(), mov.size());
        std::ifstream(argv[3], std::ios::binary).read((char *) fix.data(), fix.size());
        Img image_mov(H, W), image_fix(H, W);
        cgs_check_hip(hipMalloc(&image_mov.data(), image_mov.bytesSize()), "hipMalloc");
        cgs_check_hip(hipMalloc(&image_fix.data(), image_fix.bytesSize()), "hipMalloc");
        image_mov.setValuesFromVector(mov);
        image_fix.setValuesFromVector(fix);

        // 2-D translation grid: the user's SAD by value, by reference, cooperatively,
        // and the built-in sumOfAbsoluteDifferences tag
        {
            CudaGrid<grid_precision, 2> grid;
            make_grid(grid, ts, te, tc);
            grid.display("translation_grid");
            grid_precision counts[2];
            grid.getAxisSampleCounts(counts);
            CudaTensor<func_precision, 2> res(counts);
            cgs_check_hip(hipMalloc(&res.data(), res.bytesSize()), "hipMalloc");
            CudaGridSearcher<func_precision, grid_precision, 2> searcher(grid, res);

            fn2_t host_fn;
            cgs_check_hip(hipMemcpyFromSymbol(&host_fn, HIP_SYMBOL(dev_user_sad), sizeof(fn2_t)), "hipMemcpyFromSymbol");
            searcher.search_by_value(host_fn, image_mov, image_fix);
            emit(dir, "user_sad", res, searcher.last_argmin);

            res.fill(-1.0f);
            searcher.search(host_fn, image_mov, image_fix);
            emit(dir, "user_sad_search", res, searcher.last_argmin);

            Img *d_mov, *d_fix;
            cgs_check_hip(hipMalloc((void **) &d_mov, sizeof(Img)), "hipMalloc");
            cgs_check_hip(hipMalloc((void **) &d_fix, sizeof(Img)), "hipMalloc");
            cgs_check_hip(hipMemcpy(d_mov, &image_mov, sizeof(Img), hipMemcpyHostToDevice), "hipMemcpy");
            cgs_check_hip(hipMemcpy(d_fix, &image_fix, sizeof(Img), hipMemcpyHostToDevice), "hipMemcpy");
            fn2_ref_t host_ref;
            cgs_check_hip(hipMemcpyFromSymbol(&host_ref, HIP_SYMBOL(dev_user_sad_ref), sizeof(fn2_ref_t)),
                          "hipMemcpyFromSymbol");
            res.fill(-1.0f);
            searcher.search_by_reference(host_ref, d_mov, d_fix);
            emit(dir, "user_sad_ref", res, searcher.last_argmin);
            res.fill(-1.0f);
            searcher.search_by_reference_stream(host_ref, 10000, 1, d_mov, d_fix);
            emit(dir, "user_sad_ref_stream", res, searcher.last_argmin);

            fn2_t host_rows;
            cgs_check_hip(hipMemcpyFromSymbol(&host_rows, HIP_SYMBOL(dev_user_sad_rows), sizeof(fn2_t)),
                          "hipMemcpyFromSymbol");
            res.fill(-1.0f);
            searcher.search_by_value_stream(host_rows, 10000, image_fix.height(), image_mov, image_fix);
            emit(dir, "user_sad_rows_stream", res, searcher.last_argmin);
            res.fill(-1.0f);
            searcher.search_by_value_block(host_rows, 64, image_mov, image_fix);
            emit(dir, "user_sad_rows_block", res, searcher.last_argmin);

            res.fill(-1.0f);
            searcher.search_by_value(sumOfAbsoluteDifferences, image_mov, image_fix);
            emit(dir, "tag_sad_sum", res, searcher.last_argmin);
            (void) hipFree(d_mov);
            (void) hipFree(d_fix);
            (void) hipFree(res.data());
        }

        // 2-D probe grid: the device accessors at fractional, integral and outside points
        {
            CudaGrid<grid_precision, 2> grid;
            make_grid(grid, ps, pe, pc);
            grid_precision counts[2];
            grid.getAxisSampleCounts(counts);
            CudaTensor<func_precision, 2> res(counts);
            cgs_check_hip(hipMalloc(&res.data(), res.bytesSize()), "hipMalloc");
            CudaGridSearcher<func_precision, grid_precision, 2> searcher(grid, res);
            fn2_t host_fn;
            cgs_check_hip(hipMemcpyFromSymbol(&host_fn, HIP_SYMBOL(dev_probe_bilinear), sizeof(fn2_t)),
                          "hipMemcpyFromSymbol");
            searcher.search_by_value(host_fn, image_mov, image_fix);
            emit(dir, "probe_bilinear", res, searcher.last_argmin);
            cgs_check_hip(hipMemcpyFromSymbol(&host_fn, HIP_SYMBOL(dev_probe_nearest), sizeof(fn2_t)),
                          "hipMemcpyFromSymbol");
            searcher.search_by_value(host_fn, image_mov, image_fix);
            emit(dir, "probe_nearest", res, searcher.last_argmin);
            (void) hipFree(res.data());
        }

        // 8-D grid: each coordinate of every grid point
        {
            CudaGrid<grid_precision, 8> grid;
            make_grid(grid, s8, e8, c8);
            grid_precision counts[8];
            grid.getAxisSampleCounts(counts);
            CudaTensor<func_precision, 8> res(counts);
            cgs_check_hip(hipMalloc(&res.data(), res.bytesSize()), "hipMalloc");
            CudaGridSearcher<func_precision, grid_precision, 8> searcher(grid, res);
            fn8_t host_axis[8];
            cgs_check_hip(hipMemcpyFromSymbol(host_axis, HIP_SYMBOL(dev_axis), sizeof(host_axis)),
                          "hipMemcpyFromSymbol");
            for (int a = 0; a < 8; ++a) {
                searcher.search_by_value(host_axis[a], image_mov, image_fix);
                emit(dir, "axis" + std::to_string(a), res, searcher.last_argmin);
            }
            (void) hipFree(res.data());
        }

        // CudaTensor::transform with a device lambda, display, toIndex1d / toIndexNd
        {
            CudaTensor<float, 2> A({3u, 2u}), B({3u, 2u});
            cgs_check_hip(hipMalloc(&A.data(), A.bytesSize()), "hipMalloc");
            cgs_check_hip(hipMalloc(&B.data(), B.bytesSize()), "hipMalloc");
            A.setValuesFromVector({1, 2, 3, 4, 5, 6});
            B.setValuesFromVector({10, 20, 30, 40, 50, 60});
            A.transform(B, [] __device__(float x, float y) { return 2.0f * x + y; });
            A.display("A", 3);
            const std::vector<float> av = A.values();
            uint32_t nd[2];
            A.toIndexNd(5, nd);
            const uint32_t ix[2] = {2, 1};
            g_json += ", \"transform\": [" + std::to_string(av[0]) + ", " + std::to_string(av[5]) + "]" +
                      ", \"toIndexNd5\": [" + std::to_string(nd[0]) + ", " + std::to_string(nd[1]) + "]" +
                      ", \"toIndex1d21\": " + std::to_string(A.toIndex1d(ix));
            (void) hipFree(A.data());
            (void) hipFree(B.data());
        }
        (void) hipFree(image_mov.data());
        (void) hipFree(image_fix.data());
        std::printf("%s}\n", g_json.c_str());
        return 0;
    } catch (const cgs_error &e) {
        std::cerr << "cgs_error: " << e.what() << "\n";
        return 1;
    }
}
