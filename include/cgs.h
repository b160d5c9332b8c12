/*
 * cgs.h -- C-ABI of the MI355X (gfx950) brute-force homography grid search.
 *
 * Plain C: pointers, sizes, status codes.  Device buffers are caller-owned
 * (as in the reference, where the caller cudaMalloc's every CudaTensor /
 * CudaImage / CudaGrid); the library never frees a caller pointer and never
 * calls exit().  Streams are passed as void* (a hipStream_t, or NULL for the
 * null stream).
 *
 * Each entry point names the reference interface it replaces (file:line,
 * relative to the uncc-visionlab/cuGridSearch source root).
 */
#ifndef CGS_H
#define CGS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGS_MAX_DIM 8

/* Layout version of the structures below; cgs_abi_version() returns the
 * value the library was built with, so a caller compiled against another
 * header can refuse to run instead of passing mis-sized structures. */
#define CGS_ABI_VERSION 6

typedef enum {
    CGS_OK = 0,
    CGS_ERR_INVALID_ARGUMENT = 1,   /* null pointer, bad dimension, bad enum */
    CGS_ERR_GRID_TOO_LARGE = 2,     /* point count leaves the int32 range */
    CGS_ERR_UNSUPPORTED = 3,        /* combination not implemented */
    CGS_ERR_HIP = 4,                /* a HIP runtime call failed */
    CGS_ERR_NO_DEVICE = 5,
    CGS_ERR_TOO_LARGE = 6           /* a size beyond a stated maximum */
} cgs_status;

/* Cost functors of the reference (by-value device functions bound through a
 * __device__ function pointer, src/gpu/cudaGridSearch.cuh:284-285). */
typedef enum {
    CGS_COST_SAD = 0,       /* averageAbsoluteDifference  src/gpu/cudaErrorFunctions.cuh:31-87  */
    CGS_COST_NCC = 1,       /* calcNCC (-uncentred NCC)   src/gpu/cudaErrorFunctions.cuh:448-507 */
    CGS_COST_MI = 2,        /* calcMIstream (-2MI/(Hx+Hy)) src/gpu/cudaErrorFunction_miStreams.cuh:212-368 */
    CGS_COST_SQD = 3,       /* calcSQD                    src/gpu/cudaErrorFunctions.cuh:378-446 */
    CGS_COST_SAD_SUM = 4,   /* sumOfAbsoluteDifferences   src/gpu/cudaErrorFunctions.cuh:89-114  */
    CGS_COST_MI_PLAIN = 5,  /* calcMI (-MI)               src/gpu/cudaErrorFunction_mi.cuh:8-115  */
    /* the thread-per-row "stream" functors: same sums, their own final combine */
    CGS_COST_NCC_STREAM = 6, /* calcNCCstream: 0 unless both sums > 0   src/gpu/cudaErrorFunctionsStreams.cuh:85-163 */
    CGS_COST_SQD_STREAM = 7, /* calcSQDstream: no overlap sentinel      src/gpu/cudaErrorFunctionsStreams.cuh:283-351 */
    CGS_COST_SAD_STREAM = 8, /* averageAbsoluteDifference_stream: no overlap -> 0
                                                              src/gpu/cudaErrorFunctionsStreams.cuh:31-83 */
    /* line-traversal ("Alt") costs: each fixed row walks its line through the
       moving image in unit lambda steps; bilinear only; their own overlap test */
    CGS_COST_SQD_ALT = 9,         /* calcSQDAlt (123123)        src/gpu/cudaErrorFunctions.cuh:226-300 */
    CGS_COST_NCC_ALT = 10,        /* calcNCCAlt (123123)        src/gpu/cudaErrorFunctions.cuh:302-376 */
    CGS_COST_NCC_STREAM_ALT = 11, /* calcNCCstreamAlt (0)       src/gpu/cudaErrorFunctionsStreams.cuh:165-281 */
    CGS_COST_MI_STREAM_ALT = 12,  /* calcMIstreamAlt (123123)   src/gpu/cudaErrorFunction_miStreams.cuh:9-210 */
    CGS_COST_MI_ALT = 13          /* calcMIAlt, -MI * fixed area, CGS_PARAM_NORM only
                                                              src/gpu/cudaErrorFunction_mi.cuh:117-255 */
} cgs_cost;

/* CudaImage samplers (include/cudaTensor.cuh:557-575). */
typedef enum {
    CGS_SAMPLER_NEAREST = 0,   /* valueAt_nearest_neighbor, truncated float linear index */
    CGS_SAMPLER_BILINEAR = 1   /* valueAt_bilinear (= valueAt) */
} cgs_sampler;

/* How the grid point becomes the warp. */
typedef enum {
    CGS_PARAM_CENTRED = 0,      /* parametersToHomography, cx=W_m/2, cy=H_m/2
                                   src/gpu/cudaImageFunctions.cuh:216-252, inverse warp :277-285 */
    CGS_PARAM_NONCENTRED = 1,   /* averageAbsoluteDifference's H  src/gpu/cudaErrorFunctions.cuh:50-64 */
    CGS_PARAM_RAW_H = 2,        /* averageAbsoluteDifferenceH, forward  src/gpu/cudaErrorFunctions.cuh:194-224 */
    CGS_PARAM_TRANSLATION = 3,  /* 2-D (tx, ty)  src/gpu/cudaErrorFunctions.cuh:89-114 */
    CGS_PARAM_NORM = 4          /* parametersToHomographyNorm (calcMIAlt only)
                                   src/gpu/cudaImageFunctions.cuh:181-214 */
} cgs_param_mode;

/* Argmin tie order. */
typedef enum {
    CGS_TIE_LOWEST = 0,      /* lowest grid index among equal minima */
    CGS_TIE_REFERENCE = 1    /* find_extrema's order: include/cudaTensor.cuh:215-260 */
} cgs_tie_mode;

typedef enum { CGS_PIX_U8 = 0, CGS_PIX_F32 = 1 } cgs_pixtype;

/* CudaGrid<float, D> (src/gpu/cudaGridSearch.cuh:74-231): start / end /
 * resolution / sample-count rows, held on the host.  Filled by
 * cgs_grid_init; read-only afterwards. */
typedef struct {
    int32_t dim;
    float start[CGS_MAX_DIM];
    float end[CGS_MAX_DIM];
    float count[CGS_MAX_DIM];       /* after the 1e-7 rule of updateProcess */
    float resolution[CGS_MAX_DIM];
    int32_t increments[CGS_MAX_DIM];
    int64_t num_elements;           /* numElements(), int32 semantics */
    int32_t offset[CGS_MAX_DIM];    /* first sample index per axis: 0 for a grid from
                                       cgs_grid_init, lo for a cgs_grid_subbox */
} cgs_grid;

/* CudaImage<pix, C>(H, W) view (include/cudaTensor.cuh:507-550): row-major,
 * channels interleaved, element (y, x, c) at c + C*x + C*W*y.  data is a
 * device pointer. */
typedef struct {
    const void *data;
    int32_t width;
    int32_t height;
    int32_t channels;
    int32_t pixtype;   /* cgs_pixtype */
} cgs_image;

typedef struct {
    int32_t cost;        /* cgs_cost */
    int32_t sampler;     /* cgs_sampler */
    int32_t param_mode;  /* cgs_param_mode */
    int32_t tie_mode;    /* cgs_tie_mode */
    int32_t mi_bins;     /* joint-histogram bins per axis for the MI costs: 8..64 (reference: 64) */
    int32_t flags;       /* CGS_FLAG_* bits */
    int32_t tile_gpt;    /* grid points per workgroup: 4, 6 or 8 for SAD/NCC/SQD, 2, 4, 6 or 8 for MI
                            (2 above 32 bins); 0 = library default (6 or 8 for SAD/NCC/SQD, 4 for
                            MI from 9 to 32 bins, 2 otherwise) */
} cgs_search_opts;

/* opts.flags: time the search kernel with HIP events on the caller's
 * stream; read the duration with cgs_last_kernel_ms(). */
#define CGS_FLAG_TIME_KERNEL 1

/* find_extrema's outputs (include/cudaTensor.cuh:215): min value, int32
 * index (-1 when no value is below +inf). */
typedef struct {
    float value;
    int32_t index;
} cgs_argmin;

/* Library version string: "cgs <version> gfx950 src <hash>", the hash being
 * the build's digest of the library's source files. */
const char *cgs_version(void);

/* ---------------------------------------------------------- preprocessing
 * The multispectral matcher's image preparation
 * (src/cpu/multispectral_ImageMatcher.cu:568-632).                        */

/* CudaImage::filter with every channel filtered (include/cudaTensor.cuh:577-593,
 * kernel filterProcess include/cudaTensorKernels.cuh:353-392): out(x, y) =
 * sum of filter[r * fw + c] * in(x - (fw-1)/2 + c, y - (fh-1)/2 + r), a tap
 * outside the image reading the centre pixel; uint8 output truncated and
 * saturated to [0, 255].  filter: host array of fw * fh <= 256 floats, row
 * major.  in and out: same shape and pixel type, device memory, no aliasing.
 * Asynchronous on stream. */
cgs_status cgs_image_filter(const cgs_image *in, const cgs_image *out, const float *filter, int32_t fw, int32_t fh,
                            void *stream);

/* CudaImage::filter with per-channel actions (include/cudaTensor.cuh:577-596,
 * CHANNEL_ACTION :64-67): channel c is filtered as above when bit c of
 * filter_mask is set (FILTER) and copied from `in` otherwise (COPY, which
 * leaves the running sum untouched).  cgs_image_filter is filter_mask 0xF.
 * The reference writes the output through at<pixType>(row, col), which for
 * uint8 images truncates row and column to 8 bits (images over 256 pixels
 * wrap); here every pixel is written at its own position.  Asynchronous. */
cgs_status cgs_image_filter_channels(const cgs_image *in, const cgs_image *out, const float *filter, int32_t fw,
                                     int32_t fh, uint32_t filter_mask, void *stream);

/* transformPixel (src/cpu/multispectral_ImageMatcher.cu:411-430) in place on
 * a uint8 image: each byte read as a signed char, (int)(v / mod), then
 * (int)(v * mod), stored modulo 256.  mod >= 2^-20.  Asynchronous. */
cgs_status cgs_image_quantize(const cgs_image *img, float mod, void *stream);

/* stbir_resize_uint8 of stb_image_resize v0.97 (vendored by the reference at
 * include/third_party/stb_image_resize.h:228, 2462-2470; called by the
 * matcher at src/cpu/multispectral_ImageMatcher.cu:531 and :550 to bring the
 * larger image to the smaller one's size): resamples `in` to `out`'s width and
 * height, byte for byte as the library does (linear colour space, clamped
 * edges, Mitchell filter on a shrinking axis and Catmull-Rom on a growing one,
 * channels alike).  uint8 images with equal channel counts, device memory, no
 * aliasing.  Synchronises the stream (the filter tables are staged from the
 * host). */
cgs_status cgs_image_resize(const cgs_image *in, const cgs_image *out, void *stream);

/* transformImage (src/gpu/cudaImageFunctions.cuh:40-62; transformImage_local,
 * src/cpu/multispectral_ImageMatcher.cu:181-203, has the same body), which
 * writeTransformedImageToDisk (src/gpu/cudaImageFunctions.cuh:129-152) runs:
 * every pixel (x, y) of `out` whose inverse-mapped point (the adjugate of H,
 * h = h11 h12 h13 h21 h22 h23 h31 h32 with h33 = 1) lies in `in` takes
 * valueAt_bilinear there, per channel (a uint8 out truncates); the other
 * pixels keep what `out` holds (the reference zero-fills it first).  Equal
 * channel counts, uint8 or float each, no aliasing.  Asynchronous. */
cgs_status cgs_image_warp(const cgs_image *in, const cgs_image *out, const float *h, void *stream);

/* fuseAlignedImages (local = 0, src/gpu/cudaImageFunctions.cuh:64-114) and
 * fuseAlignedImages_local (local = 1, src/cpu/multispectral_ImageMatcher.cu:
 * 115-179), which writeAlignedAndFusedImageToDisk(_local) run: a 3-channel
 * overlay, channel 0 `mov` under est_h, 1 `ref` under the identity, 2 `mov`
 * under true_h, each sampled at `channel`; elements whose point falls outside
 * their source keep what `fused` holds.  The local form steps normalised
 * coordinates in binary32 and writes by rounded target index, as the
 * reference's loop does, including its overwrites.  Asynchronous for
 * local = 0; local = 1 synchronises the stream (its step tables are staged
 * from the host). */
cgs_status cgs_image_fuse(const cgs_image *ref, const cgs_image *mov, const float *est_h, const float *true_h,
                          int32_t channel, int32_t local, const cgs_image *fused, void *stream);

/* The drivers' stbi_write_png (src/gpu/cudaImageFunctions.cuh:141-143, 167-169):
 * writes host pixels (row-major, channels interleaved, 8-bit; 1 grey, 2 grey +
 * alpha, 3 RGB, 4 RGBA) as a PNG file, the image data stored uncompressed. */
cgs_status cgs_write_png(const char *path, const void *pixels, int32_t width, int32_t height, int32_t channels);

/* stbi_load for PNG files (stb_image v2.27, include/third_party/stb_image.h:
 * 5023-5229), which the matcher loads its inputs with
 * (src/cpu/multispectral_ImageMatcher.cu:503-514): 8-bit samples, row-major,
 * channels interleaved, req_comp channels per pixel (1 grey, 2 grey + alpha,
 * 3 RGB, 4 RGBA; 0 = the file's own), converted as stb converts (grey from
 * colour as (77 R + 150 G + 29 B) >> 8, 16-bit samples to their high byte,
 * sub-8-bit grey scaled to 0..255, palettes expanded, tRNS keys to alpha).
 * width, height and channels_in_file (stb's *n) are written when not NULL.
 * pixels NULL: only the header facts are read.  Otherwise pixels is host
 * memory of capacity bytes, at least width * height * channels.  Host-only. */
cgs_status cgs_read_png(const char *path, int32_t req_comp, void *pixels, int64_t capacity, int32_t *width,
                        int32_t *height, int32_t *channels_in_file);

/* CGS_ABI_VERSION of the library's build. */
int32_t cgs_abi_version(void);

/* Fills the options with the defaults (SAD, bilinear, centred, lowest-index
 * ties, 32 MI bins). */
void cgs_default_opts(cgs_search_opts *opts);

/* CudaGrid::setStartPoint / setEndPoint / setNumSamples + update()
 * (src/gpu/cudaGridSearch.cuh:198-215, 100-105, 235-272) and numElements
 * (:89-98).  Host-only. */
cgs_status cgs_grid_init(cgs_grid *grid, int32_t dim, const float *start, const float *end,
                         const float *count);

/* CudaGrid::indexToGridPoint / getGridPoint (src/gpu/cudaGridSearch.cuh:160-196).
 * Host-only; out has grid->dim floats.  Point a = start + (offset + i) * res
 * as fmaf((float) (offset + i), res, start). */
cgs_status cgs_grid_point(const cgs_grid *grid, int64_t index, float *out);

/* A box of a grid's samples: axis a keeps the grid's start and resolution
 * and takes count[a] samples from sample index lo[a] on.  Its points are
 * exactly the grid's points with those indices (the same binary32
 * operations), in the same axis-0-fastest order, so a search of the box
 * returns bit-identical costs.  No reference counterpart (the reference
 * narrows a search by building a new grid, src/cpu/multispectral_ImageMatcher.cu:751-807,
 * whose points are not the old grid's).  Host-only. */
cgs_status cgs_grid_subbox(const cgs_grid *grid, const int32_t *lo, const int32_t *count, cgs_grid *out);

/* CudaGridSearcher::search / search_by_value (src/gpu/cudaGridSearch.cuh:553-556,
 * 736-765) with the functor chosen by opts->cost, fused with
 * CudaTensor::find_extrema (include/cudaTensor.cuh:215-265).
 * result: device float[num_elements] in grid-index order (axis 0 fastest), or
 * NULL to skip writing the cost tensor.  argmin_out: host pointer, or NULL.
 * Synchronises the stream before returning (the reference synchronises after
 * every search). */
cgs_status cgs_search(const cgs_grid *grid, const cgs_image *moved, const cgs_image *fixed,
                      const cgs_search_opts *opts, float *result, cgs_argmin *argmin_out,
                      void *stream);

/* Bytes of the static device scratch cgs_search_noalloc stages in (8 MiB per
 * GPU the library runs on, held until the process ends). */
#define CGS_NOALLOC_SCRATCH_BYTES (8u << 20)

/* cgs_search without any allocation.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and stream synchronisation as cgs_search, but the
 * argmin key, the moving image's tap table or padded copy and the (cos, sin)
 * table live in one static device array of CGS_NOALLOC_SCRATCH_BYTES per GPU,
 * and no hipMalloc, hipFree or host allocation is made.  Limits: plain costs
 * only (the line-traversal costs return CGS_ERR_UNSUPPORTED), no
 * CGS_FLAG_TIME_KERNEL, at most 1024 levels on axis 0, and a moving image
 * whose staging fits in the scratch (a one-channel bilinear SAD/NCC/SQD
 * search stages 16 bytes per pixel: up to about 720 x 720 pixels); larger
 * inputs return CGS_ERR_TOO_LARGE.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other threads, other streams) race
 * and must be serialised by the caller. */
cgs_status cgs_search_noalloc(const cgs_grid *grid, const cgs_image *moved, const cgs_image *fixed,
                              const cgs_search_opts *opts, float *result, cgs_argmin *argmin_out,
                              void *stream);

/* CudaTensor::find_extrema (include/cudaTensor.cuh:215-265) over a device
 * float array of n values. */
cgs_status cgs_find_extrema(const float *values, int64_t n, int32_t tie_mode, cgs_argmin *out,
                            void *stream);

/* Element types of the tensor utilities below. */
typedef enum { CGS_DT_F32 = 0, CGS_DT_F64 = 1, CGS_DT_I32 = 2, CGS_DT_U8 = 3 } cgs_dtype;

/* Element-wise operations of CudaTensor::operator+= / -= / *= (transform with
 * the reference's lambdas, include/cudaTensor.cuh:322-360, kernel
 * transformProcess include/cudaTensorKernels.cuh:341-351). */
typedef enum { CGS_OP_ADD = 0, CGS_OP_SUB = 1, CGS_OP_MUL = 2 } cgs_tensor_op;

/* CudaTensor::fill (include/cudaTensor.cuh:197-206, kernel fillProcess
 * include/cudaTensorKernels.cuh:230-237): n elements of `data` (device) set to
 * value converted to the element type.  Asynchronous. */
cgs_status cgs_tensor_fill(void *data, int64_t n, int32_t dtype, double value, void *stream);

/* a[i] = a[i] op b[i] for i < n in the element type's arithmetic (a uint8
 * result wraps modulo 256, as the reference's store does).  Device arrays,
 * asynchronous. */
cgs_status cgs_tensor_binary(int32_t op, void *a, const void *b, int64_t n, int32_t dtype, void *stream);

/* CudaMatrix::multiply (include/cudaTensor.cuh:474-496, kernel multiplyProcess
 * include/cudaTensorKernels.cuh:396-440): C (m x n) = A (m x k) * B (k x n),
 * row-major device matrices of CGS_DT_F32 or CGS_DT_F64.  Each element is the
 * reference's dot product in its order, k ascending from 0, with each step a
 * fused multiply-add (nvcc contracts the reference's Csub += a * b).
 * Asynchronous. */
cgs_status cgs_matrix_multiply(const void *A, const void *B, void *C, int32_t m, int32_t k, int32_t n, int32_t dtype,
                               void *stream);

/* Evaluates a single grid point and returns its cost and, for the MI costs,
 * the bins x bins joint histogram (row = fixed-image bin, column = moving
 * bin) in hist_out (host, may be NULL).  Counts are exact integers.
 * Diagnostic counterpart of calling the functor once, as
 * src/gpu/cudaGridSearch_main_mi.cu:388 does on the host. */
cgs_status cgs_eval_point(const cgs_image *moved, const cgs_image *fixed, const cgs_search_opts *opts,
                          int32_t dim, const float *params, float *cost_out, uint32_t *hist_out,
                          void *stream);

/* CudaGridSearcher::covariance_by_value (src/gpu/cudaGridSearch.cuh:768-794,
 * kernel calculateCovariance_by_value :317-419) over a cost tensor `result`
 * (device, grid->num_elements floats in grid order).  The minimum is the
 * reference's: the lowest index of the smallest negative cost, else index 0
 * with value 0 (returned in min_out, host, may be NULL).  For every other
 * point i with f = result[i] (0 read as -1e-9), dx = p(i) - p(min),
 * len = |dx| and s = (log|min| - log|f|) / len >= 0, the sum of g g^T with
 * g = dx s / len goes to covar (host, dim x dim, row-major, overwritten;
 * the reference adds into the caller's matrix, which the facade does).
 * Per-point arithmetic in binary32 as the reference writes it, log as the
 * binary64 log rounded to binary32, sums in binary64.  Synchronous. */
cgs_status cgs_covariance(const cgs_grid *grid, const float *result, float *covar, cgs_argmin *min_out,
                          void *stream);

/* ------------------------------------------------------------------------
 * Batched ragged search.  No reference counterpart: the reference searches
 * one pair per CudaGridSearcher call (src/gpu/cudaGridSearch.cuh:736-765).
 * One persistent kernel takes (pair, grid-tile) work items of uneven size
 * from a device work counter; every pair's costs and argmin are identical to
 * cgs_search on that pair.  Pairs: one channel, bilinear sampling, any cost
 * (SAD, SAD_SUM, NCC, SQD, MI, MI_PLAIN), uint8 or float images of any size.
 * ---------------------------------------------------------------------- */
typedef struct {
    cgs_image moved;
    cgs_image fixed;
    cgs_grid grid;           /* from cgs_grid_init */
    cgs_search_opts opts;    /* sampler must be CGS_SAMPLER_BILINEAR */
    float *result;           /* device cost tensor (grid order) or NULL */
} cgs_batch_pair;

/* Tile shapes of the persistent kernel; 0 fields take the defaults. */
typedef struct {
    int32_t fast_gpt;        /* grid points per workgroup for SAD/NCC/SQD: 4 or 6 (default 6) */
    int32_t mi_gpt;          /* grid points per workgroup for MI: 2, 4 or 8 (default 4; 2 above 32 bins) */
    int32_t wg_per_cu;       /* resident workgroups per compute unit: 1..8 */
    int32_t items_per_claim; /* work items taken per atomic claim: 1..64 */
} cgs_tile_config;

typedef struct cgs_batch_s cgs_batch;

cgs_status cgs_batch_create(cgs_batch **out);

/* Validates the pairs, records their views and allocates the batch's device
 * workspace (tap tables, descriptors, keys).  The images may change between
 * runs; their shapes, grids and options may not (plan again). */
cgs_status cgs_batch_plan(cgs_batch *batch, const cgs_batch_pair *pairs, int32_t n, const cgs_tile_config *tiles);

/* Enqueues one batched search on the stream: restages every pair's tap
 * table, clears the keys and the work counter, launches the persistent
 * kernel.  use_graph != 0: the first run on a stream captures this sequence
 * into a hipGraph and later runs replay it.  Asynchronous. */
cgs_status cgs_batch_run(cgs_batch *batch, int32_t use_graph, void *stream);

/* Waits for the stream and returns every pair's argmin (n entries). */
cgs_status cgs_batch_results(cgs_batch *batch, cgs_argmin *out, void *stream);

/* Total work items and grid points of the planned batch (diagnostics). */
cgs_status cgs_batch_info(const cgs_batch *batch, int64_t *items, int64_t *points, int64_t *pixel_evals);

/* Local refinement with full-grid fallback.  No reference counterpart (the
 * reference's coarse-to-fine loop, src/cpu/multispectral_ImageMatcher.cu:751-807,
 * builds new grids instead).  Each pair searches the box of its grid within
 * `radius` samples of its centre on every axis (3^8 points for radius 1 on
 * an 8-D grid; the box is cut at the grid's faces).  A box minimum on a box
 * face that is not a grid face re-centres the box there, up to max_steps
 * boxes; the pair then falls back (outcome BOUNDARY).  A minimum inside its
 * box is a local minimum of the grid; it is kept when its cost is at most
 * `threshold`, else the pair falls back (outcome THRESHOLD).  With
 * fallback != 0 each pair that falls back is searched over its whole grid
 * in the same call, so its answer equals cgs_search's.  Box points are the
 * grid's own points (cgs_grid_subbox), so every cost is bit-identical to
 * the full search's cost of that point.  With starts > 1 every pair runs
 * that many descents from its own centres and keeps the lowest local
 * minimum (lower index on ties); a descent that reaches a centre an earlier
 * start of the pair holds stops there.  It falls back (BOUNDARY) only when
 * no descent found a local minimum.  All descents still moving share one
 * batched launch per step. */
typedef struct {
    int32_t radius;        /* 1..16 */
    int32_t max_steps;     /* boxes per descent before it counts as stuck on a face: 1..1024 */
    float threshold;       /* largest accepted local-minimum cost (+inf: any); not NaN */
    int32_t fallback;      /* != 0: run the full-grid searches of the pairs that fall back */
    int32_t starts;        /* descents per pair, each from its own centre: 1..64 */
} cgs_refine_opts;

typedef enum {
    CGS_REFINE_LOCAL = 0,               /* answer = the local minimum */
    CGS_REFINE_FALLBACK_BOUNDARY = 1,   /* still on a box face after max_steps boxes */
    CGS_REFINE_FALLBACK_THRESHOLD = 2   /* local minimum costs more than threshold */
} cgs_refine_outcome;

typedef struct {
    int32_t index;          /* grid index of the answer (-1: none, when a fallback was not run) */
    float value;            /* its cost */
    int32_t steps;          /* boxes searched, over all starts */
    int32_t outcome;        /* cgs_refine_outcome */
    int32_t local_index;    /* grid index of the last box minimum */
    float local_value;
    int64_t points;         /* grid points evaluated for this pair, fallback included */
} cgs_refine_result;

/* pairs: as for cgs_batch_plan (result pointers are ignored); centres:
 * n x starts x grid.dim sample indices (clamped to the grid); out: n results.
 * Replaces any plan the batch held.  Synchronous. */
cgs_status cgs_batch_refine(cgs_batch *batch, const cgs_batch_pair *pairs, int32_t n, const int32_t *centres,
                            const cgs_refine_opts *opts, cgs_refine_result *out, void *stream);

cgs_status cgs_batch_destroy(cgs_batch *batch);

/* ------------------------------------------------------------------------
 * Proposal network.  No reference counterpart: the reference always searches
 * the full grid (src/gpu/cudaGridSearch.cuh:736-765).  A small CNN regresses
 * the homography's 4 corner offsets from a downsampled pair; a local grid
 * around its proposal then replaces most of the full search.
 *
 * Fixed architecture: input 2 x 128 x 128 bf16, NHWC (channel 0 = moving,
 * 1 = fixed); six 3x3 convolutions, zero padding 1, channels
 * 32-32-64-64-128-128, stride 2 on the 2nd, 4th and 6th, ReLU after each;
 * global average pooling; a fully connected layer to 8 outputs.  BF16
 * weights and activations, FP32 accumulation (matrix cores), FP32 biases,
 * FC, loss and optimiser.
 *
 * Parameters are one FP32 vector of cgs_cnn_param_count() values: for each
 * conv layer its weight [Cout][ky][kx][Cin] then its bias [Cout]; then the
 * FC weight [8][128] and bias [8].
 * ---------------------------------------------------------------------- */
typedef struct cgs_cnn_s cgs_cnn;

int64_t cgs_cnn_param_count(void);

/* Allocates a network (all parameters zero) for batches of up to max_batch
 * (1..4096) pairs on the current device. */
cgs_status cgs_cnn_create(int32_t max_batch, cgs_cnn **out);

cgs_status cgs_cnn_destroy(cgs_cnn *net);

/* Host FP32 parameter vectors in and out.  set_params also resets the
 * optimiser state. */
cgs_status cgs_cnn_set_params(cgs_cnn *net, const float *params);
cgs_status cgs_cnn_get_params(const cgs_cnn *net, float *params);

/* The network input of n pairs: each one-channel 128x128 or 256x256 image
 * (uint8 or float) is reduced to 128x128 (2x2 box mean), standardised to
 * zero mean and unit variance and rounded to bf16.  input: device
 * uint16_t[n][128][128][2] (bf16 bit patterns).  Synchronous. */
cgs_status cgs_cnn_prepare(const cgs_image *moved, const cgs_image *fixed, int32_t n, uint16_t *input,
                           void *stream);

/* Forward pass of n pairs (n <= max_batch).  offsets: device float[n][8]
 * or NULL.  Keeps the activations for cgs_cnn_backward.  Asynchronous. */
cgs_status cgs_cnn_forward(cgs_cnn *net, const uint16_t *input, int32_t n, float *offsets, void *stream);

/* Gradient of the mean squared error sum((out - target)^2) / (8n) of the
 * last forward pass with respect to every parameter; target: device
 * float[n][8].  loss_out (host, may be NULL: then asynchronous) receives
 * the loss.  Read the gradient with cgs_cnn_get_grads. */
cgs_status cgs_cnn_backward(cgs_cnn *net, const float *target, float *loss_out, void *stream);
cgs_status cgs_cnn_get_grads(const cgs_cnn *net, float *grads);

/* One Adam step (bias-corrected) on the master weights with the last
 * gradient, then the bf16 working copies are refreshed.  Asynchronous. */
cgs_status cgs_cnn_adam_step(cgs_cnn *net, float lr, float beta1, float beta2, float eps, void *stream);

/* The current device's measured FP32 FMA throughput in TFLOP/s (2 FLOP per
 * fmaf): a kernel of independent v_fmac_f32 chains on VGPR operands (no
 * packed FP32), timed with HIP events, best of three.  The VALU roofline the
 * search kernels are measured against beside the published peak.  No
 * reference counterpart.  Allocates a small output buffer; synchronous. */
cgs_status cgs_fma_peak(float *tflops_out, void *stream);

/* VALU issue rates of the current device, in wave64 instructions per CU and
 * per cycle of the nominal (maximum) clock, one float per form:
 *   0 v_fmac_f32 on VGPR operands, 1 v_fma_f32 with an SGPR operand,
 *   2 v_cvt_f32_ubyte1, 3 v_mad_u32_u24 with an SGPR operand, 4 v_rcp_f32,
 *   5 v_fmac_f32 alternating with v_cvt_f32_ubyte1.
 * Each from a kernel of independent instructions of that form, 8 waves per
 * SIMD, best of three.  Forms 0 and 5 run well above 1 per CU-cycle (two
 * VALU instructions can issue per cycle when one of them is a plain fp32 or
 * int32 operation on VGPRs), forms 1-3 near 1, form 4 near 1/2: the ceiling
 * a kernel's VALU stream meets depends on that mix.  No reference
 * counterpart.  rates_out holds n >= CGS_VALU_RATE_COUNT floats. */
#define CGS_VALU_RATE_COUNT 6
cgs_status cgs_valu_rates(float *rates_out, int n, void *stream);

/* Counts, on the current device, the binary32 mantissas d in [1, 2) whose
 * reciprocal refined by one Newton step (v_rcp_f32, then r += r (1 - d r))
 * is not the correctly rounded 1/d.  The MI kernels' exact quotients take
 * one correction step from that reciprocal, which is IEEE division when the
 * count is 0.  No reference counterpart.  Synchronous. */
cgs_status cgs_check_rcp(uint32_t *mismatches_out, void *stream);

/* Duration in ms of the calling thread's last timed search kernel (flags &
 * CGS_FLAG_TIME_KERNEL), measured with HIP events around that launch; -1
 * when none was timed.  No reference counterpart (the reference prints wall
 * time, src/gpu/cudaGridSearch.cuh:755-764). */
float cgs_last_kernel_ms(void);

/* Last error message of the calling thread ("" when none). */
const char *cgs_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CGS_H */
