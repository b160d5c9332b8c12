/*
 * cgs_oracle.c -- CPU oracle for the brute-force homography grid search.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  The product path (cugridsearch_amd/csrc) never links it.
 *
 * It restates, in plain C, the arithmetic of the reference search path
 * (uncc-visionlab/cuGridSearch) with every floating-point step declared:
 *   - compile with -ffp-contract=off: every + - * / is one IEEE-754 binary32
 *     (or binary64 where stated) operation; FMA appears only where fmaf() is
 *     written explicitly;
 *   - the three projective forms of the warp are evaluated per pixel as
 *     F(x, y) = fmaf(B, y, fmaf(A, x, C)) with C = P - Q rounded once (the
 *     reference writes (A*x + B*y + P) - Q and leaves contraction to nvcc);
 *   - inImage is decided on the exact rational quotient of those rounded
 *     forms (0 <= X/W < width, 0 <= Y/W < height, by sign and fmaf tests)
 *     for SAD/NCC/SQD, and on the correctly rounded quotients themselves
 *     (the reference's own test) for the MI costs; the sample point is the
 *     correctly rounded quotient;
 *   - sin/cos of the rotation angle: (float)sin((double)theta), likewise cos;
 *   - division and sqrt: IEEE correctly rounded;
 *   - NCC / SQD / bilinear-SAD sums: binary64 accumulation of binary32 terms;
 *   - MI entropies: binary64 with log2 from libm.
 *
 * Reference citations (file:line, relative to the reference root):
 *   grid            src/gpu/cudaGridSearch.cuh:89-98, 160-187, 235-272
 *   homography      src/gpu/cudaImageFunctions.cuh:216-252 (centred)
 *                   src/gpu/cudaErrorFunctions.cuh:50-57 (non-centred)
 *                   src/gpu/cudaErrorFunctions.cuh:205-207 (raw H, forward)
 *                   src/gpu/cudaErrorFunctions.cuh:99-100 (2-D translation)
 *   inverse warp    src/gpu/cudaImageFunctions.cuh:277-285
 *   image layout    include/cudaTensor.cuh:507-550, 171-180
 *   samplers        include/cudaTensor.cuh:557-575
 *   SAD             src/gpu/cudaErrorFunctions.cuh:31-87, 89-114
 *   SQD             src/gpu/cudaErrorFunctions.cuh:378-446
 *   NCC             src/gpu/cudaErrorFunctions.cuh:448-507
 *   MI              src/gpu/cudaErrorFunction_miStreams.cuh:212-368 (normalised)
 *                   src/gpu/cudaErrorFunction_mi.cuh:8-115 (plain -MI)
 *   argmin          include/cudaTensor.cuh:215-265,
 *                   include/cudaTensorKernels.cuh:133-220
 *   line traversal  src/gpu/cudaErrorFunctions.cuh:226-376 (calcSQDAlt, calcNCCAlt)
 *                   src/gpu/cudaErrorFunctionsStreams.cuh:165-281 (calcNCCstreamAlt)
 *                   src/gpu/cudaErrorFunction_miStreams.cuh:9-210 (calcMIstreamAlt)
 *                   src/gpu/cudaErrorFunction_mi.cuh:117-255 (calcMIAlt)
 *                   src/gpu/cudaImageFunctions.cuh:181-214, 315-441 (norm H, lines)
 *
 * Declared decisions where the reference is undefined (each is pinned by a
 * test in tests/test_oracle_*.py):
 *   - an axis with count 1 gets resolution 0 (the reference reads
 *     end[-1]-start[-1], out of bounds);
 *   - grid point = fmaf(i, res, start);
 *   - the nearest-neighbour index keeps the reference's truncated float
 *     linear index, with fmaf for each "index += incr * coord" step;
 *   - reads past the last pixel (bilinear ceil on the last row, the
 *     nearest-neighbour index overflow) return 0;
 *   - bilinear: each of the three lerps (1-a)*p + a*q is evaluated as
 *     fmaf(a, q, (1-a)*p) (the reference writes the unfused form and leaves
 *     contraction to nvcc);
 *   - MI bins are clamped to [0, bins-1] (the reference writes out of bounds
 *     for values outside [0, 256)).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXD 8

enum { ORC_COST_SAD = 0, ORC_COST_NCC = 1, ORC_COST_MI = 2, ORC_COST_SQD = 3,
       ORC_COST_SAD_SUM = 4, ORC_COST_MI_PLAIN = 5, ORC_COST_NCC_STREAM = 6, ORC_COST_SQD_STREAM = 7,
       ORC_COST_SAD_STREAM = 8, ORC_COST_SQD_ALT = 9, ORC_COST_NCC_ALT = 10, ORC_COST_NCC_STREAM_ALT = 11,
       ORC_COST_MI_STREAM_ALT = 12, ORC_COST_MI_ALT = 13 };
enum { ORC_NEAREST = 0, ORC_BILINEAR = 1 };
enum { ORC_PARAM_CENTRED = 0, ORC_PARAM_NONCENTRED = 1, ORC_PARAM_RAW_H = 2,
       ORC_PARAM_TRANSLATION = 3, ORC_PARAM_NORM = 4 };
enum { ORC_TIE_LOWEST = 0, ORC_TIE_REFERENCE = 1 };
enum { ORC_PIX_U8 = 0, ORC_PIX_F32 = 1 };

typedef struct {
    const void *data;
    int width, height, channels, pixtype;
} orc_image;

/* ------------------------------------------------------------------ grid */

/* CudaGrid::update() -> updateProcess (cudaGridSearch.cuh:256-270), the
 * changed_axis = -1 branch, which every setter uses.  count is updated in
 * place (the 1e-7 rule forces it to 1).  The comparison is in double, as
 * "abs(float) < 1.0e-7 * float" promotes to double. */
void orc_grid_update(int D, const float *start, const float *end,
                     float *count, float *res) {
    for (int a = 0; a < D; a++) {
        float diff = end[a] - start[a];
        if ((double) fabsf(diff) < 1.0e-7 * (double) count[a]) count[a] = 1.0f;
        if (count[a] == 1.0f) res[a] = 0.0f;
        else res[a] = diff / (count[a] - 1.0f);
    }
}

/* CudaGrid::numElements (cudaGridSearch.cuh:89-98): int32 total, each step
 * (float)total * count truncated to int.  Returns -1 when a step leaves the
 * int32 range (the reference overflows silently). */
int64_t orc_grid_num_elements(int D, const float *count) {
    int32_t total = 1;
    for (int a = 0; a < D; a++) {
        float p = (float) total * count[a];
        if (!(p < 2147483648.0f) || p < 0.0f) return -1;
        total = (int32_t) p;
    }
    return total;
}

/* dimensional increments (cudaGridSearch.cuh:168-172): int32 from a float
 * product of the previous count and increment. */
void orc_grid_increments(int D, const float *count, int32_t *incr) {
    incr[0] = 1;
    for (int a = 1; a < D; a++) incr[a] = (int32_t) (count[a - 1] * (float) incr[a - 1]);
}

/* indexToGridPoint (cudaGridSearch.cuh:160-187); axis 0 varies fastest. */
void orc_grid_indices(int D, const int32_t *incr, int64_t index, int32_t *idx) {
    int32_t rem = (int32_t) index;
    for (int a = D - 1; a >= 0; a--) {
        idx[a] = rem / incr[a];
        rem -= idx[a] * incr[a];
    }
}

void orc_grid_point(int D, const float *start, const float *res, const float *count,
                    int64_t index, float *out) {
    int32_t incr[ORC_MAXD], idx[ORC_MAXD];
    orc_grid_increments(D, count, incr);
    orc_grid_indices(D, incr, index, idx);
    for (int a = 0; a < D; a++) out[a] = fmaf((float) idx[a], res[a], start[a]);
}

/* ------------------------------------------------------------ homography */

/* parametersToHomography (cudaImageFunctions.cuh:216-252), every expression
 * in the reference's own association order. */
static void homography_centred(const float *p, float cx, float cy, float *h) {
    const float theta = p[0], scaleX = p[1], scaleY = p[2], shearXY = p[3];
    const float translateX = p[4], translateY = p[5], keystoneX = p[6], keystoneY = p[7];
    const float c = (float) cos((double) theta);
    const float s = (float) sin((double) theta);
    float h11, h12, h13, h21, h22, h23, h31, h32, h33;
    h11 = scaleX*(c*(cx*keystoneX + 1) + cx*keystoneY*s);
    h12 = scaleY*(shearXY*(c*(cx*keystoneX + 1) + cx*keystoneY*s) - s*(cx*keystoneX + 1) + cx*keystoneY*c);
    h13 = cx + translateX*(cx*keystoneX + 1) - cx*scaleX*(c*(cx*keystoneX + 1) + cx*keystoneY*s) - cy*scaleY*(shearXY*(c*(cx*keystoneX + 1) + cx*keystoneY*s) - s*(cx*keystoneX + 1) + cx*keystoneY*c) + cx*keystoneY*translateY;
    h21 = scaleX*(s*(cy*keystoneY + 1) + cy*keystoneX*c);
    h22 = scaleY*(shearXY*(s*(cy*keystoneY + 1) + cy*keystoneX*c) + c*(cy*keystoneY + 1) - cy*keystoneX*s);
    h23 = cy + translateY*(cy*keystoneY + 1) - cx*scaleX*(s*(cy*keystoneY + 1) + cy*keystoneX*c) - cy*scaleY*(shearXY*(s*(cy*keystoneY + 1) + cy*keystoneX*c) + c*(cy*keystoneY + 1) - cy*keystoneX*s) + cy*keystoneX*translateX;
    h31 = scaleX*(keystoneX*c + keystoneY*s);
    h32 = scaleY*(shearXY*(keystoneX*c + keystoneY*s) + keystoneY*c - keystoneX*s);
    h33 = keystoneX*translateX + keystoneY*translateY - cy*scaleY*(shearXY*(keystoneX*c + keystoneY*s) + keystoneY*c - keystoneX*s) - cx*scaleX*(keystoneX*c + keystoneY*s) + 1;
    h[0] = h11 / h33; h[1] = h12 / h33; h[2] = h13 / h33;
    h[3] = h21 / h33; h[4] = h22 / h33; h[5] = h23 / h33;
    h[6] = h31 / h33; h[7] = h32 / h33;
}

/* averageAbsoluteDifference's inline non-centred H (cudaErrorFunctions.cuh:50-57). */
static void homography_noncentred(const float *p, float *h) {
    const float theta = p[0], scaleX = p[1], scaleY = p[2], shearXY = p[3];
    const float c = (float) cos((double) theta);
    const float s = (float) sin((double) theta);
    h[0] = scaleX * c;
    h[1] = scaleY * shearXY * c - scaleY * s;
    h[2] = scaleX * p[4];
    h[3] = scaleX * s;
    h[4] = scaleY * shearXY * s + scaleY * c;
    h[5] = scaleY * p[5];
    h[6] = p[6];
    h[7] = p[7];
}

void orc_homography(int mode, const float *params, float cx, float cy, float *h) {
    if (mode == ORC_PARAM_CENTRED) homography_centred(params, cx, cy, h);
    else if (mode == ORC_PARAM_NONCENTRED) homography_noncentred(params, h);
    else for (int i = 0; i < 8; i++) h[i] = params[i];
}

/* Per-grid-point coefficients of the three projective forms, evaluated per
 * pixel as F(x, y) = fmaf(B, y, fmaf(A, x, C)); coef[3*k + {0,1,2}] =
 * {A, B, C} for k = 0 (X), 1 (Y), 2 (W).  The moving-image sample point is
 * (X/W, Y/W).
 *   inverse warp, calcNewCoordH (cudaImageFunctions.cuh:282-284):
 *     W = (h21h32 - h22h31)x + (h12h31 - h11h32)y + h11h22 - h12h21
 *     X = (h22 - h23h32)x + (h13h32 - h12)y + h12h23 - h13h22
 *     Y = (h23h31 - h21)x + (h11 - h13h31)y + h13h21 - h11h23
 *   forward raw H (cudaErrorFunctions.cuh:205-207):
 *     W = h7x + h8y + 1,  X = h1x + h2y + h3,  Y = h4x + h5y + h6
 *   translation (cudaErrorFunctions.cuh:99-100): X = x + t0, Y = y + t1, W = 1
 */
void orc_warp_coeffs(int mode, const float *params, float cx, float cy, float *coef) {
    if (mode == ORC_PARAM_TRANSLATION) {
        const float c[9] = {1.0f, 0.0f, params[0],
                            0.0f, 1.0f, params[1],
                            0.0f, 0.0f, 1.0f};
        memcpy(coef, c, sizeof(c));
        return;
    }
    float h[8];
    orc_homography(mode, params, cx, cy, h);
    const float h11 = h[0], h12 = h[1], h13 = h[2], h21 = h[3], h22 = h[4], h23 = h[5];
    const float h31 = h[6], h32 = h[7];
    if (mode == ORC_PARAM_RAW_H) {
        const float c[9] = {h11, h12, h13,
                            h21, h22, h23,
                            h31, h32, 1.0f};
        memcpy(coef, c, sizeof(c));
        return;
    }
    coef[0] = h22 - h23 * h32;       coef[1] = h13 * h32 - h12;       coef[2] = h12 * h23 - h13 * h22;
    coef[3] = h23 * h31 - h21;       coef[4] = h11 - h13 * h31;       coef[5] = h13 * h21 - h11 * h23;
    coef[6] = h21 * h32 - h22 * h31; coef[7] = h12 * h31 - h11 * h32; coef[8] = h11 * h22 - h12 * h21;
}

static inline float form(const float *c, float x, float y) {
    return fmaf(c[1], y, fmaf(c[0], x, c[2]));
}

/* 0 <= n/w < lim on the exact quotient: the sign of n - lim*w is exact in
 * one fmaf (the exact value is rounded once and rounding keeps the sign). */
static inline int quotient_in(float n, float w, float lim) {
    if (w > 0) return n >= 0 && fmaf(-lim, w, n) < 0;
    if (w < 0) return n <= 0 && fmaf(-lim, w, n) > 0;
    return 0;
}

/* Sample point of pixel (x, y): returns inImage and the rounded quotients. */
int orc_warp_point(const float *coef, int x, int y, float *out_xy, float mw, float mh) {
    const float fx = (float) x, fy = (float) y;
    const float w = form(coef + 6, fx, fy);
    const float xn = form(coef, fx, fy), yn = form(coef + 3, fx, fy);
    out_xy[0] = xn / w;
    out_xy[1] = yn / w;
    return quotient_in(xn, w, mw) && quotient_in(yn, w, mh);
}

/* -------------------------------------------------------------- sampling */

static inline float pix(const orc_image *im, int64_t lin) {
    const int64_t n = (int64_t) im->width * im->height * im->channels;
    if (lin < 0 || lin >= n) return 0.0f;   /* declared zero pad past the end */
    if (im->pixtype == ORC_PIX_U8) return (float) ((const uint8_t *) im->data)[lin];
    return ((const float *) im->data)[lin];
}

/* CudaImage::inImage (cudaTensor.cuh:548-550). */
static inline int in_image(const orc_image *im, float y, float x) {
    return x >= 0 && y >= 0 && x < (float) im->width && y < (float) im->height;
}

/* toIndex1d<float>({c, x, y}) (cudaTensor.cuh:171-180) with increments
 * {1, C, C*W}: int index = (int)c; index = (int)fmaf(incr, coord, (float)index). */
static inline int64_t lin_index_f(const orc_image *im, float y, float x, int c) {
    int32_t i = c;
    i = (int32_t) fmaf((float) im->channels, x, (float) i);
    i = (int32_t) fmaf((float) (im->channels * im->width), y, (float) i);
    return i;
}

/* valueAt_nearest_neighbor (cudaTensor.cuh:557-559) after the inImage test. */
static inline float nearest_in(const orc_image *im, float y, float x, int c) {
    return pix(im, lin_index_f(im, y, x, c));
}

float orc_sample_nearest(const orc_image *im, float y, float x, int c) {
    if (!in_image(im, y, x)) return 0.0f;
    return nearest_in(im, y, x, c);
}

/* valueAt_bilinear (cudaTensor.cuh:561-575); ceil(x) == W reads the next
 * row's first pixel through the linear index, as the reference does. */
static inline float bilinear_in(const orc_image *im, float y, float x, int c) {
    const float fy = floorf(y), fx = floorf(x), cy = ceilf(y), cx = ceilf(x);
    const float tlc = pix(im, lin_index_f(im, fy, fx, c));
    const float trc = pix(im, lin_index_f(im, fy, cx, c));
    const float blc = pix(im, lin_index_f(im, cy, fx, c));
    const float brc = pix(im, lin_index_f(im, cy, cx, c));
    const float alpha_x = x - fx;
    const float alpha_y = y - fy;
    const float value_top = fmaf(alpha_x, trc, (1.0f - alpha_x) * tlc);
    const float value_bottom = fmaf(alpha_x, brc, (1.0f - alpha_x) * blc);
    return fmaf(alpha_y, value_bottom, (1.0f - alpha_y) * value_top);
}

float orc_sample_bilinear(const orc_image *im, float y, float x, int c) {
    if (!in_image(im, y, x)) return 0.0f;
    return bilinear_in(im, y, x, c);
}

static inline float sample(const orc_image *im, float y, float x, int c, int sampler) {
    return sampler == ORC_NEAREST ? nearest_in(im, y, x, c) : bilinear_in(im, y, x, c);
}

/* ----------------------------------------------------------------- costs */

typedef struct {
    double sum, i1, i2, ic;     /* SAD sum | NCC/SQD sums (see orc_cost) */
    int64_t count;              /* in-image (pixel, channel) samples */
    int64_t isum;               /* exact integer SAD sum (nearest, u8) */
} orc_acc;

/* MI bin of an intensity: floor(bins * (v / 256.0f)), clamped. */
static inline int mi_bin(float v, int bins) {
    float t = floorf((float) bins * (v / 256.0f));
    int b = (int) t;
    if (!(t >= 0.0f)) b = 0;
    if (b > bins - 1) b = bins - 1;
    return b;
}

/* Evaluates one grid point.  hist (bins*bins uint32, row = fixed bin,
 * column = moving bin) is filled for the MI costs when non-NULL or needed.
 * Returns the cost as the product writes it (binary32). */
static float cost_alt(const orc_image *moved, const orc_image *fixed, int cost, const float *params,
                      int bins, uint32_t *hist_out, double *aux_out);

float orc_cost_point(const orc_image *moved, const orc_image *fixed, int cost, int sampler,
                     int mode, const float *params, int bins, uint32_t *hist_out,
                     double *aux_out) {
    if (cost >= ORC_COST_SQD_ALT) return cost_alt(moved, fixed, cost, params, bins, hist_out, aux_out);
    float coef[9];
    const float cx = (float) moved->width / 2, cy = (float) moved->height / 2;
    orc_warp_coeffs(mode, params, cx, cy, coef);
    const int C = fixed->channels;
    const int is_mi = (cost == ORC_COST_MI || cost == ORC_COST_MI_PLAIN);
    uint32_t *hist = NULL;
    if (is_mi) {
        hist = hist_out ? hist_out : (uint32_t *) malloc(sizeof(uint32_t) * bins * bins);
        memset(hist, 0, sizeof(uint32_t) * bins * bins);
    }
    orc_acc a = {0, 0, 0, 0, 0, 0};
    const int exact_int = (sampler == ORC_NEAREST && moved->pixtype == ORC_PIX_U8 &&
                           fixed->pixtype == ORC_PIX_U8);
    const float mw = (float) moved->width, mh = (float) moved->height;
    for (int y = 0; y < fixed->height; y++) {
        for (int x = 0; x < fixed->width; x++) {
            float xy[2];
            const int in = orc_warp_point(coef, x, y, xy, mw, mh);
            /* the MI costs test CudaImage::inImage on calcNewCoordH's binary32
             * quotients, as calcMIstream does (cudaErrorFunction_miStreams.cuh:263-268);
             * SAD/NCC/SQD the declared exact-quotient rule */
            if (is_mi ? !(xy[0] >= 0 && xy[1] >= 0 && xy[0] < mw && xy[1] < mh) : !in) continue;
            for (int c = 0; c < C; c++) {
                const float m = sample(moved, xy[1], xy[0], c, sampler);
                const float f = pix(fixed, ((int64_t) y * fixed->width + x) * C + c);
                a.count++;
                switch (cost) {
                case ORC_COST_SAD: case ORC_COST_SAD_SUM: case ORC_COST_SAD_STREAM:
                    if (exact_int) a.isum += (int64_t) fabsf(m - f);
                    else a.sum += (double) fabsf(m - f);
                    break;
                case ORC_COST_NCC: case ORC_COST_SQD: case ORC_COST_NCC_STREAM: case ORC_COST_SQD_STREAM: {
                    const double md = (double) m / 255.0, fd = (double) f / 255.0;
                    a.i1 += fd * fd;
                    a.i2 += md * md;
                    if (cost == ORC_COST_NCC || cost == ORC_COST_NCC_STREAM) a.ic += fd * md;
                    else a.ic += (md - fd) * (md - fd);
                    break;
                }
                default:
                    hist[mi_bin(f, bins) * bins + mi_bin(m, bins)]++;
                }
            }
        }
    }
    float out = 0.0f;
    switch (cost) {
    case ORC_COST_SAD:
        if (exact_int) out = (float) a.isum / (float) a.count;
        else out = (float) (a.sum / (double) a.count);
        break;
    case ORC_COST_SAD_SUM:
        out = exact_int ? (float) a.isum : (float) a.sum;
        break;
    case ORC_COST_NCC:
        out = a.count > 0 ? (float) (-a.ic / sqrt(a.i1 * a.i2)) : 0.0f;
        break;
    case ORC_COST_SQD:
        out = a.count > 0 ? (float) (a.ic / sqrt(a.i1 * a.i2)) : 123123.0f;
        break;
    case ORC_COST_SAD_STREAM:   /* averageAbsoluteDifference_stream: no overlap -> 0 (cudaErrorFunctionsStreams.cuh:71-78) */
        if (a.count == 0) out = 0.0f;
        else if (exact_int) out = (float) a.isum / (float) a.count;
        else out = (float) (a.sum / (double) a.count);
        break;
    case ORC_COST_NCC_STREAM:   /* calcNCCstream: 0 unless both sums are positive (cudaErrorFunctionsStreams.cuh:158-161) */
        out = (a.i1 > 0 && a.i2 > 0) ? (float) (-a.ic / sqrt(a.i1 * a.i2)) : 0.0f;
        break;
    case ORC_COST_SQD_STREAM:   /* calcSQDstream: no sentinel (cudaErrorFunctionsStreams.cuh:337-350) */
        out = (float) (a.ic / sqrt(a.i1 * a.i2));
        break;
    default: {
        /* calcMIstream (cudaErrorFunction_miStreams.cuh:306-365) in binary64 */
        const double n = (double) a.count;
        double mi = 0, hx = 0, hy = 0;
        double *px = (double *) calloc(bins, sizeof(double));
        double *py = (double *) calloc(bins, sizeof(double));
        for (int i = 0; i < bins; i++)
            for (int j = 0; j < bins; j++) {
                px[i] += hist[i * bins + j];
                py[j] += hist[i * bins + j];
            }
        for (int i = 0; i < bins; i++)
            for (int j = 0; j < bins; j++) {
                const double pxy = hist[i * bins + j] / n, pxi = px[i] / n, pyj = py[j] / n;
                if (pxy > 0 && pxi > 0 && pyj > 0) mi += pxy * (log2(pxy) - (log2(pxi) + log2(pyj)));
            }
        for (int i = 0; i < bins; i++) {
            const double pxi = px[i] / n, pyi = py[i] / n;
            if (pxi > 0) hx -= pxi * log2(pxi);
            if (pyi > 0) hy -= pyi * log2(pyi);
        }
        free(px); free(py);
        double r = (cost == ORC_COST_MI) ? -2.0 * mi / (hx + hy) : -mi;
        if (a.count == 0) r = (cost == ORC_COST_MI) ? NAN : -0.0;
        out = (float) r;
        if (aux_out) { aux_out[5] = mi; aux_out[6] = hx; aux_out[7] = hy; aux_out[8] = r; }
    }
    }
    if (aux_out) {
        aux_out[0] = exact_int ? (double) a.isum : a.sum;
        aux_out[1] = a.i1; aux_out[2] = a.i2; aux_out[3] = a.ic; aux_out[4] = (double) a.count;
    }
    if (is_mi && !hist_out) free(hist);
    return out;
}

/* ------------------------------------------------- line-traversal costs */

/* MI of a joint histogram (row = fixed bin, column = moving bin) of n
 * samples, binary64: mi = sum pxy (log2 pxy - (log2 px + log2 py)) over the
 * cells where all three are positive; hx, hy the marginal entropies. */
static void mi_of_hist(const uint32_t *hist, int bins, double n, double *mi_out, double *hx_out, double *hy_out) {
    double mi = 0, hx = 0, hy = 0;
    double *px = (double *) calloc(bins, sizeof(double));
    double *py = (double *) calloc(bins, sizeof(double));
    for (int i = 0; i < bins; i++)
        for (int j = 0; j < bins; j++) {
            px[i] += hist[i * bins + j];
            py[j] += hist[i * bins + j];
        }
    for (int i = 0; i < bins; i++)
        for (int j = 0; j < bins; j++) {
            const double pxy = hist[i * bins + j] / n, pxi = px[i] / n, pyj = py[j] / n;
            if (pxy > 0 && pxi > 0 && pyj > 0) mi += pxy * (log2(pxy) - (log2(pxi) + log2(pyj)));
        }
    for (int i = 0; i < bins; i++) {
        const double pxi = px[i] / n, pyi = py[i] / n;
        if (pxi > 0) hx -= pxi * log2(pxi);
        if (pyi > 0) hy -= pyi * log2(pyi);
    }
    free(px); free(py);
    *mi_out = mi; *hx_out = hx; *hy_out = hy;
}

/* One fixed row's line through the moving image: the row's samples sit at
 * p0 + lambda v for lambda = ls, ls + step, ... while lambda < le (and the
 * row's own end); in = some image border crossing qualified.  out[7] =
 * {ls, le, vx, vy, px, py, in}. */
typedef struct { float ls, le, vx, vy, px, py; int in; } orc_line;

/* parametricAssignValues (cudaImageFunctions.cuh:315-376).  Each expression
 * is the reference's, in its association order; the factors it repeats are
 * named once here (dt: the shared determinant-like term; a0 / aw: the
 * row's denominators at column 0 and at column f_width; gy / gx: the
 * crossing denominators of the horizontal / vertical borders). */
static void line_centred(const float *h, int mw, int mh, int fw, int fh, int row, orc_line *L) {
    const float h11 = h[0], h12 = h[1], h13 = h[2], h21 = h[3], h22 = h[4], h23 = h[5];
    const float h31 = h[6], h32 = h[7];
    const float r = (float) row, FW = (float) fw, MW = (float) mw, MH = (float) mh, FH = (float) fh;
    const float dt = h11*h22 - h12*h21 - h11*h23*h32 + h12*h23*h31 + h13*h21*h32 - h13*h22*h31;
    const float a0 = h11*h22 - h12*h21 - h11*h32*r + h12*h31*r;
    const float aw = h11*h22 - h12*h21 + FW*h21*h32 - FW*h22*h31 - h11*h32*r + h12*h31*r;
    const float gy = h21 - h31*r, gx = h22 - h32*r;
    L->vx = (gx*(dt))/(a0*aw);
    L->vy = -(gy*(dt))/(a0*aw);
    const float q = h12*h21 - h11*h22 + r*(h11*h32 - h12*h31);
    L->px = (h13*h22 - h12*h23 + r*(h12 - h13*h32))/q;
    L->py = -(h13*h21 - h11*h23 + r*(h11 - h13*h31))/q;
    const float lt = -((h11*h23 - h13*h21 - h11*r + h13*h31*r)*(aw))/(gy*(dt));
    const float xt = -(h23 - r)/gy;
    const float lb = -((aw)*(h11*h23 - h13*h21 - h11*r + h11*h22*MH - h12*h21*MH + h13*h31*r - h11*h32*MH*r
                             + h12*h31*MH*r))/(gy*(dt));
    const float xb = -(h23 - r + h22*MH - h32*MH*r)/gy;
    const float ll = -((h12*h23 - h13*h22 - h12*r + h13*h32*r)*(aw))/(gx*(dt));
    const float yl = -(h23 - r)/gx;
    const float lr = -((aw)*(h12*h23 - h13*h22 - h12*r - h11*h22*MW + h12*h21*MW + h13*h32*r + h11*h32*MW*r
                             - h12*h31*MW*r))/(gx*(dt));
    const float yr = -(h23 - r + h21*MW - h31*MW*r)/gx;
    L->ls = 0.0f;
    L->le = FW;
    L->in = 0;
    if (xt >= 0 && xt < MW && lt >= 0 && lt < FW) { if (L->vy > 0) L->ls = lt; else L->le = lt; L->in = 1; }
    if (xb >= 0 && xb < MW && lb >= 0 && lb < FW) { if (L->vy < 0) L->ls = lb; else L->le = lb; L->in = 1; }
    if (yl >= 0 && yl < MH && ll >= 0 && ll < FH) { if (L->vx > 0) L->ls = ll; else L->le = ll; L->in = 1; }
    if (yr >= 0 && yr < MH && lr >= 0 && lr < FH) { if (L->vx < 0) L->ls = lr; else L->le = lr; L->in = 1; }
}

/* parametricAssignValuesNorm (cudaImageFunctions.cuh:378-441), normalised
 * coordinates (the image spans [-0.5, 0.5]), column -0.5.  ty / tx are the
 * quotients the border crossings share with p0 (p0 = (tx, -ty)). */
static void line_norm(const float *h, float row, orc_line *L) {
    const float h11 = h[0], h12 = h[1], h13 = h[2], h21 = h[3], h22 = h[4], h23 = h[5];
    const float h31 = h[6], h32 = h[7];
    const float col = -0.5f;
    const float dt = h11*h22 - h12*h21 - h11*h23*h32 + h12*h23*h31 + h13*h21*h32 - h13*h22*h31;
    L->vy = -(h21 - h31*row)/dt;
    L->vx = (h22 - h32*row)/dt;
    const float q = h11*h22 - h12*h21 + col*(h21*h32 - h22*h31) - row*(h11*h32 - h12*h31);
    const float tx = (h12*h23 - h13*h22 + col*(h22 - h23*h32) - row*(h12 - h13*h32))/q;
    const float ty = (h11*h23 - h13*h21 + col*(h21 - h23*h31) - row*(h11 - h13*h31))/q;
    L->px = tx;
    L->py = -ty;
    const float gy = h21 - h31*row, gx = h22 - h32*row;
    const float lt = -((ty - 0.5f)*(dt))/gy;
    const float xt = (h22 - 2*h23 + 2*row - h32*row)/(2*gy);
    const float lb = -((ty + 0.5f)*(dt))/gy;
    const float xb = -(h22 + 2*h23 - 2*row - h32*row)/(2*gy);
    const float ll = -((tx + 0.5f)*(dt))/gx;
    const float yl = (h21 - 2*h23 + 2*row - h31*row)/(2*gx);
    const float lr = -((tx - 0.5f)*(dt))/gx;
    const float yr = -(h21 + 2*h23 - 2*row - h31*row)/(2*gx);
    L->ls = 0.0f;
    L->le = 1.0f;
    L->in = 0;
    if (xt > -0.5f && xt < 0.5f && lt >= 0 && lt < 1) { if (L->vy > 0) L->ls = lt; else L->le = lt; L->in = 1; }
    if (xb > -0.5f && xb < 0.5f && lb > 0 && lb < 1) { if (L->vy < 0) L->ls = lb; else L->le = lb; L->in = 1; }
    if (yl > -0.5f && yl < 0.5f && ll > 0 && ll < 1) { if (L->vx > 0) L->ls = ll; else L->le = ll; L->in = 1; }
    if (yr > -0.5f && yr < 0.5f && lr > 0 && lr < 1) { if (L->vx < 0) L->ls = lr; else L->le = lr; L->in = 1; }
}

/* parametersToHomographyNorm (cudaImageFunctions.cuh:181-214) with its
 * integer quotients f_w / m_w, f_h / m_h, f_w / 2, f_h / 2, m_w / 2 and the
 * moving-image centre's y taken from the width (Q11). */
static void homography_norm(const float *p, int mw, int mh, int fw, int fh, float *h) {
    const float c = (float) cos((double) p[0]), s = (float) sin((double) p[0]);
    const float x_ratio = (float) (fw / mw), y_ratio = (float) (fh / mh);
    const float cfx = (float) (fw / 2), cfy = (float) (fh / 2), cmx = (float) (mw / 2), cmy = (float) (mw / 2);
    const float scaleX = p[1] / x_ratio, scaleY = p[2] / y_ratio, shearXY = p[3];
    h[0] = scaleX * c;
    h[1] = scaleY * shearXY * c - scaleY * s;
    h[2] = (p[4] - cfx + cmx) / (float) fw;
    h[3] = scaleX * s;
    h[4] = scaleY * shearXY * s + scaleY * c;
    h[5] = (p[5] - cfy + cmy) / (float) fh;
    h[6] = p[6];
    h[7] = p[7];
}

void orc_alt_homography(int norm, const float *params, int mw, int mh, int fw, int fh, float *h) {
    if (norm) homography_norm(params, mw, mh, fw, fh, h);
    else homography_centred(params, (float) mw / 2, (float) mh / 2, h);
}

void orc_alt_line(int norm, const float *h, int mw, int mh, int fw, int fh, float row, float *out) {
    orc_line L;
    if (norm) line_norm(h, row, &L);
    else line_centred(h, mw, mh, fw, fh, (int) row, &L);
    out[0] = L.ls; out[1] = L.le; out[2] = L.vx; out[3] = L.vy; out[4] = L.px; out[5] = L.py;
    out[6] = (float) L.in;
}

/* Sum of 64 lane partial sums in the xor-butterfly order: at each distance
 * o = 32, 16, ..., 1 every lane i adds lane i ^ o's value. */
static double lane_tree(const double *p) {
    double a[64], b[64];
    memcpy(a, p, sizeof(a));
    for (int o = 32; o >= 1; o >>= 1) {
        for (int i = 0; i < 64; i++) b[i] = a[i] + a[i ^ o];
        memcpy(a, b, sizeof(a));
    }
    return a[0];
}

/* The reference's overlap test of the step-1 Alt functors
 * (cudaErrorFunctions.cuh:296, 372; cudaErrorFunctionsStreams.cuh:272;
 * cudaErrorFunction_miStreams.cuh:199): moved area / (W_m H_m scaleY) > 0.25
 * and fixed area / (W_f H_f) > 0.35, in binary32 with the double compare. */
static int alt_overlap(float mov_area, float fix_area, int mw, int mh, int fw, int fh, float scale_y) {
    return (double) (mov_area / ((float) (mw * mh) * scale_y)) > 0.25 &&
           (double) (fix_area / (float) (fw * fh)) > 0.35;
}

/* calcSQDAlt / calcNCCAlt / calcNCCstreamAlt / calcMIstreamAlt: one line
 * per fixed row (centred H), unit lambda steps from the line's start,
 * accumulated as lambda += 1 in binary32; both images sampled by the gated
 * bilinear valueAt, the fixed one at (row, lambda).  Declared sums: step k of
 * a row goes to lane k mod 64; each lane adds, in row and step order,
 * fma(m, f) (or fma(m - f, m - f) for SQD), fma(m, m), fma(f, f) in binary64
 * (the 1/255 scaling of the reference cancels in every one of these costs);
 * the lanes combine in lane_tree order.  Moved and fixed areas are the
 * reference's binary32 loops: calcSQDAlt / calcNCCAlt add |v| and 1 to one
 * sum at every step of every row (cudaErrorFunctions.cuh:289-290, 365-366);
 * calcNCCstreamAlt adds them per row, once per step
 * (cudaErrorFunctionsStreams.cuh:245-246), and calcMIstreamAlt per row once
 * per step and channel (cudaErrorFunction_miStreams.cuh:110-111), and both
 * then add the row sums in row order (cudaErrorFunctionsStreams.cuh:259-264,
 * cudaErrorFunction_miStreams.cuh:129-135; rows without a line add 0, the MI
 * functor's uninitialised row sums read as 0).  aux[9] is the moved area's
 * exact value, the same adds of |v| summed in binary64.  MI: joint counts of every
 * sample. */
static float cost_alt_step(const orc_image *moved, const orc_image *fixed, int cost, const float *params,
                           int bins, uint32_t *hist, double *aux) {
    const int mw = moved->width, mh = moved->height, fw = fixed->width, fh = fixed->height, C = fixed->channels;
    float h[8];
    homography_centred(params, (float) mw / 2, (float) mh / 2, h);
    double s0[64], s1[64], s2[64];
    for (int i = 0; i < 64; i++) s0[i] = s1[i] = s2[i] = 0.0;
    double mov = 0.0;
    int64_t steps = 0, count = 0;
    const int is_mi = cost == ORC_COST_MI_STREAM_ALT;
    const int per_row = is_mi || cost == ORC_COST_NCC_STREAM_ALT;   /* the stream functors' row sums */
    const int adds = is_mi ? C : 1;                                 /* area adds per step */
    float mov_area = 0.0f, fix_area = 0.0f;
    for (int row = 0; row < fh; row++) {
        orc_line L;
        line_centred(h, mw, mh, fw, fh, row, &L);
        float mov_row = 0.0f, fix_row = 0.0f;
        if (L.in) {
            const float mag = sqrtf(L.vx * L.vx + L.vy * L.vy);
            int64_t k = 0;
            for (float lam = L.ls; lam < L.le && lam < (float) fw; lam += 1.0f, k++) {
                for (int a = 0; a < adds; a++) {
                    if (per_row) {
                        mov_row += mag;
                        fix_row += 1.0f;
                    } else {
                        mov_area += mag;
                        fix_area += 1.0f;
                    }
                }
                const float nx = lam * L.vx + L.px, ny = lam * L.vy + L.py;
                const int lane = (int) (k & 63);
                for (int c = 0; c < C; c++) {
                    const float m = orc_sample_bilinear(moved, ny, nx, c);
                    const float f = orc_sample_bilinear(fixed, (float) row, lam, c);
                    if (is_mi) {
                        hist[mi_bin(f, bins) * bins + mi_bin(m, bins)]++;
                    } else {
                        const double md = (double) m, fd = (double) f;
                        if (cost == ORC_COST_SQD_ALT) s0[lane] = fma(md - fd, md - fd, s0[lane]);
                        else s0[lane] = fma(md, fd, s0[lane]);
                        s1[lane] = fma(fd, fd, s1[lane]);
                        s2[lane] = fma(md, md, s2[lane]);
                    }
                }
                count += C;
            }
            mov += (double) (k * adds) * (double) mag;
            steps += k;
        }
        if (per_row) {
            mov_area += mov_row;
            fix_area += fix_row;
        }
    }
    const int pass = alt_overlap(mov_area, fix_area, mw, mh, fw, fh, params[2]);
    const double S0 = lane_tree(s0), S1 = lane_tree(s1), S2 = lane_tree(s2);
    double mi = 0, hx = 0, hy = 0;
    float out;
    if (is_mi) {
        if (pass) mi_of_hist(hist, bins, (double) count, &mi, &hx, &hy);
        out = pass ? (float) (-2.0 * mi / (hx + hy)) : 123123.0f;
    } else if (cost == ORC_COST_SQD_ALT) {
        out = pass ? (float) (S0 / sqrt(S1 * S2)) : 123123.0f;
    } else {
        out = pass ? (float) (-S0 / sqrt(S1 * S2)) : (cost == ORC_COST_NCC_ALT ? 123123.0f : 0.0f);
    }
    if (aux) {
        aux[0] = (double) steps; aux[1] = S1; aux[2] = S2; aux[3] = S0; aux[4] = (double) count;
        aux[5] = mi; aux[6] = (double) mov_area; aux[7] = (double) fix_area; aux[8] = pass; aux[9] = mov;
    }
    return out;
}

/* Number of rows calcMIAlt visits and their values: row = -0.5f, then
 * row += 1 / (float)(H_f - 1) in binary32 while row < 0.5
 * (cudaErrorFunction_mi.cuh:162).  rows may be NULL; at most cap values
 * are written.  Returns the count. */
static float mi_alt_row_step(int fh) { return 1 / ((float) (fh - 1)); }

int64_t orc_mi_alt_rows(int fh, float *rows, int64_t cap) {
    const float step = mi_alt_row_step(fh);
    int64_t k = 0;
    for (float row = -0.5f; row < 0.5; row += step, k++)
        if (rows && k < cap) rows[k] = row;
    return k;
}

/* calcMIAlt's fixed area after n steps: fix_area += 1.0 / (W_f H_f), a
 * binary64 add rounded to binary32 at every step (cudaErrorFunction_mi.cuh:198). */
float orc_mi_alt_area(int fw, int fh, int64_t n) {
    const double d = 1.0 / (double) (fw * fh);
    float fa = 0.0f;
    for (int64_t i = 0; i < n; i++) fa = (float) ((double) fa + d);
    return fa;
}

/* calcMIAlt (cudaErrorFunction_mi.cuh:117-255): normalised H and lines,
 * rows and lambda stepped in binary32 by 1/(H_f - 1) and 1/(W_f - 1);
 * the sample points scaled back to pixels through binary64 ((x + 0.5) * W);
 * the moving sample is the gated bilinear value, so a point outside the
 * moving image counts in bin 0 (Q3).  Cost -MI * fix_area, MI in binary64
 * from the exact counts. */
static float cost_mi_alt(const orc_image *moved, const orc_image *fixed, const float *params, int bins,
                         uint32_t *hist, double *aux) {
    const int mw = moved->width, mh = moved->height, fw = fixed->width, fh = fixed->height, C = fixed->channels;
    float h[8];
    homography_norm(params, mw, mh, fw, fh, h);
    const float sr = mi_alt_row_step(fh), sc = 1 / ((float) (fw - 1));
    int64_t steps = 0, count = 0;
    for (float row = -0.5f; row < 0.5; row += sr) {
        orc_line L;
        line_norm(h, row, &L);
        if (!L.in) continue;
        for (float lam = L.ls; lam < L.le && lam < 1; lam += sc) {
            float nx = lam * L.vx + L.px, ny = lam * L.vy + L.py;
            nx = (float) (((double) nx + 0.5) * (double) mw);
            ny = (float) (((double) ny + 0.5) * (double) mh);
            const float ri = (float) (((double) row + 0.5) * (double) (fh - 1));
            const float li = lam * (float) (fw - 1);
            for (int c = 0; c < C; c++) {
                const float f = orc_sample_bilinear(fixed, ri, li, c);
                const float m = orc_sample_bilinear(moved, ny, nx, c);
                hist[mi_bin(f, bins) * bins + mi_bin(m, bins)]++;
                count++;
            }
            steps++;
        }
    }
    const float fa = orc_mi_alt_area(fw, fh, steps);
    double mi = 0, hx = 0, hy = 0;
    if (count > 0) mi_of_hist(hist, bins, (double) count, &mi, &hx, &hy);
    if (aux) {
        aux[0] = (double) steps; aux[1] = aux[2] = aux[3] = 0.0; aux[4] = (double) count;
        aux[5] = mi; aux[6] = 0.0; aux[7] = (double) fa; aux[8] = 1.0;
    }
    return (float) (-mi * (double) fa);
}

static float cost_alt(const orc_image *moved, const orc_image *fixed, int cost, const float *params,
                      int bins, uint32_t *hist_out, double *aux_out) {
    const int is_mi = cost == ORC_COST_MI_STREAM_ALT || cost == ORC_COST_MI_ALT;
    uint32_t *hist = NULL;
    if (is_mi) {
        hist = hist_out ? hist_out : (uint32_t *) malloc(sizeof(uint32_t) * bins * bins);
        memset(hist, 0, sizeof(uint32_t) * bins * bins);
    }
    const float v = cost == ORC_COST_MI_ALT ? cost_mi_alt(moved, fixed, params, bins, hist, aux_out)
                                            : cost_alt_step(moved, fixed, cost, params, bins, hist, aux_out);
    if (is_mi && !hist_out) free(hist);
    return v;
}

/* ---------------------------------------------------------- preprocessing */

/* filterProcess (include/cudaTensorKernels.cuh:353-392) with every channel
 * filtered, as CudaImage::filter launches it: output (x, y) = the running sum
 * over filter rows r, columns c of filter[r][c] * in(y - (fh-1)/2 + r,
 * x - (fw-1)/2 + c), an outside tap reading the centre pixel; the sum is not
 * reset between channels; "pixVal += f * p" as fmaf(f, p, pixVal); uint8
 * output truncated toward zero and saturated to [0, 255]. */
void orc_image_filter(const orc_image *in, void *out, const float *filter, int fw, int fh) {
    const int W = in->width, H = in->height, C = in->channels;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float acc = 0.0f;
            for (int c = 0; c < C; c++) {
                for (int r = 0; r < fh; r++)
                    for (int k = 0; k < fw; k++) {
                        const int yy = y - (fh - 1) / 2 + r, xx = x - (fw - 1) / 2 + k;
                        const int inside = xx >= 0 && yy >= 0 && xx < W && yy < H;
                        const int64_t at = inside ? ((int64_t) yy * W + xx) * C + c : ((int64_t) y * W + x) * C + c;
                        acc = fmaf(filter[r * fw + k], pix(in, at), acc);
                    }
                const int64_t o = ((int64_t) y * W + x) * C + c;
                if (in->pixtype == ORC_PIX_U8) {
                    float t = truncf(acc);
                    if (!(t >= 0.0f)) t = 0.0f;   /* negatives and NaN */
                    if (t > 255.0f) t = 255.0f;
                    ((uint8_t *) out)[o] = (uint8_t) t;
                } else {
                    ((float *) out)[o] = acc;
                }
            }
        }
}

/* transformPixel (src/cpu/multispectral_ImageMatcher.cu:411-430), in place:
 * char pixel = byte (signed); int temp = pixel; temp /= mod; temp *= mod;
 * byte = (uint8_t) temp.  The int /= float and int *= float go through
 * binary32 and truncate toward zero. */
void orc_image_quantize(uint8_t *img, int64_t n, float mod) {
    for (int64_t i = 0; i < n; i++) {
        int t = (int) (int8_t) img[i];
        t = (int) ((float) t / mod);
        t = (int) ((float) t * mod);
        img[i] = (uint8_t) t;
    }
}

/* Evaluates grid indices [i0, i1) into result[0 .. i1-i0).  Threads: OpenMP
 * over grid points (nthreads <= 0 keeps the OpenMP default). */
int orc_search(const orc_image *moved, const orc_image *fixed, int cost, int sampler, int mode,
               int D, const float *start, const float *res, const float *count, int bins,
               int64_t i0, int64_t i1, float *result, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    int32_t incr[ORC_MAXD];
    orc_grid_increments(D, count, incr);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = i0; i < i1; i++) {
        int32_t idx[ORC_MAXD];
        float p[ORC_MAXD];
        orc_grid_indices(D, incr, i, idx);
        for (int a = 0; a < D; a++) p[a] = fmaf((float) idx[a], res[a], start[a]);
        result[i - i0] = orc_cost_point(moved, fixed, cost, sampler, mode, p, bins, NULL, NULL);
    }
    return 0;
}

/* ---------------------------------------------------------------- argmin */

/* Rank of an index in find_extrema's tie order (cudaTensor.cuh:215-260,
 * cudaTensorKernels.cuh:133-220): block = idx / 1024 first, then the 9-bit
 * thread id bit-reversed (the LDS tree resolves bits 8..5 and the shuffles
 * bits 4..0; ties keep the lower slot), then which half of the block's 1024
 * elements the thread read first. */
int64_t orc_reference_rank(int64_t idx) {
    const int64_t blk = idx / 1024, t = idx % 512, half = (idx % 1024) / 512;
    int64_t r = 0;
    for (int b = 0; b < 9; b++) if (t & (1 << b)) r |= (int64_t) 1 << (8 - b);
    return blk * 1024 + r * 2 + half;
}

/* Global minimum with strict '<' (NaN and +inf never win; nothing finite or
 * -inf -> index -1, value +inf). */
void orc_argmin(const float *r, int64_t n, int tie_mode, float *value, int64_t *index) {
    float best = INFINITY;
    int64_t bi = -1, brank = INT64_MAX;
    for (int64_t i = 0; i < n; i++) {
        const float v = r[i];
        if (!(v < INFINITY)) continue;
        const int64_t rank = tie_mode == ORC_TIE_REFERENCE ? orc_reference_rank(i) : i;
        if (v < best || (v == best && rank < brank)) { best = v; bi = i; brank = rank; }
    }
    *value = best;
    *index = bi;
}

/* Literal simulation of findExtremaProcess/reduceMin6<float,512> plus the
 * host pass (cudaTensor.cuh:215-260), used to pin orc_reference_rank. */
void orc_argmin_reference_sim(const float *g, int64_t n, float *value, int64_t *index) {
    const int bs = 512;
    const int64_t nblocks = n / bs + 1;
    float *bv = (float *) malloc(sizeof(float) * nblocks);
    int64_t *bix = (int64_t *) malloc(sizeof(int64_t) * nblocks);
    float sv[512];
    int64_t si[512];
    for (int64_t b = 0; b < nblocks; b++) {
        for (int t = 0; t < bs; t++) {
            float mn = INFINITY;
            int64_t mi = -1;
            int64_t i = b * bs * 2 + t;
            const int64_t grid = (int64_t) bs * 2 * nblocks;
            while (i < n) {
                if (g[i] < mn) { mi = i; mn = g[i]; }
                if (i + bs < n && g[i + bs] < mn) { mi = i + bs; mn = g[i + bs]; }
                i += grid;
            }
            sv[t] = mn; si[t] = mi;
        }
        for (int off = 256; off >= 1; off /= 2)
            for (int t = 0; t < off; t++)
                if (sv[t + off] < sv[t]) { sv[t] = sv[t + off]; si[t] = si[t + off]; }
        bv[b] = sv[0]; bix[b] = si[0];
    }
    float best = bv[0];
    int64_t bi = bix[0];
    for (int64_t b = 1; b < nblocks; b++)
        if (bv[b] < best) { best = bv[b]; bi = bix[b]; }
    *value = best; *index = bi;
    free(bv); free(bix);
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* CudaGridSearcher::covariance_by_value's kernel, calculateCovariance_by_value
 * (src/gpu/cudaGridSearch.cuh:317-419), in its own loop order: the minimum
 * scan with strict < from (value 0, index 0) (:331-339), then every other
 * point: f == 0 read as -1e-9 (:346), dx = p - p_min, len = sqrt of the
 * axis-ordered binary32 sum of dx^2 (:349-356), slope = (log|min| - log|f|)
 * * (1 / len) skipped when < 0 (:357-362), g = dx * slope / len (:365-367),
 * covar += g g^T (:369-394).  Binary32 per point with every operation
 * rounded on its own; log declared as the binary64 log rounded to binary32;
 * the sums in binary64 (the reference adds into binary32).  covar: D x D,
 * overwritten.  Returns the minimum's index; *min_value gets its value. */
int64_t orc_covariance(int D, const float *start, const float *res, const float *count, const float *result,
                       int64_t n, double *covar, float *min_value) {
    int64_t mi = 0;
    float mv = 0.0f;
    for (int64_t i = 0; i < n; ++i)
        if (result[i] < mv) {
            mv = result[i];
            mi = i;
        }
    *min_value = mv;
    for (int k = 0; k < D * D; ++k) covar[k] = 0.0;
    float xmin[ORC_MAXD], p[ORC_MAXD], dx[ORC_MAXD], g[ORC_MAXD];
    orc_grid_point(D, start, res, count, mi, xmin);
    const float logmin = (float) log((double) fabsf(mv));
    for (int64_t i = 0; i < n; ++i) {
        if (i == mi) continue;
        orc_grid_point(D, start, res, count, i, p);
        float f = result[i];
        if (f == 0.0f) f = -1e-9f;
        float len2 = 0.0f;
        for (int d = 0; d < D; ++d) {
            dx[d] = p[d] - xmin[d];
            len2 = len2 + dx[d] * dx[d];
        }
        const float len = sqrtf(len2);
        const float slope = (logmin - (float) log((double) fabsf(f))) * (1.0f / len);
        if (slope < 0.0f) continue;
        for (int d = 0; d < D; ++d) g[d] = dx[d] * slope / len;
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) covar[r * D + c] += (double) (g[r] * g[c]);
    }
    return mi;
}

/* ------------------------------------------------------------------ resize
 * stbir_resize_uint8 of stb_image_resize v0.97 (the reference vendors it at
 * include/third_party/stb_image_resize.h and calls it at
 * src/cpu/multispectral_ImageMatcher.cu:531 and :550 to bring the larger image
 * to the smaller one's size): linear colour space, no alpha, clamped edges,
 * Mitchell filter on an axis that shrinks and Catmull-Rom on one that grows
 * (:429-433), channels filtered alike.  Restated from the library's published
 * algorithm with its binary32 operations in its order:
 *   decode     p / 255.0f                                        (:1283-1291)
 *   kernels    (16 + x*x*(21x - 36))/18, (32 + x*(-60 + x*(36 - 7x)))/18;
 *              1 - x*x*(2.5 - 1.5x), 2 - x*(4 + x*(0.5x - 2.5))  (:797-837)
 *   ranges     floor((double) bound +- 0.5) of the float bounds  (:1009-1036)
 *   weights    growing axis: per output pixel, leading exact zeros dropped,
 *              scaled by 1/sum, trailing zeros dropped             (:1038-1090);
 *              shrinking axis: per input pixel kernel(x) * scale, trailing zeros
 *              dropped, then per output pixel scaled by 1/sum over the input
 *              pixels that reach it, then leading zeros and outputs < 0 dropped
 *              (:1092-1201)
 *   sums       each output sample accumulates from 0 in ascending input order
 *              (:1450-1656 horizontally; :1870-2065 and the ring buffer loops
 *              :2067-2204 vertically), then
 *   encode     (uint8) (int) ((double) (sat(v) * 255.0f) + 0.5)  (:1735-1757).
 * Checked bit for bit against the library itself (oracle/_ref, built from the
 * reference's header) by tests/test_oracle_resize.py. */

typedef struct { int n0, n1; } orc_rz_contrib;

static float orc_rz_kernel(int grow, float x) {
    x = fabsf(x);
    if (grow) {   /* Catmull-Rom */
        if (x < 1.0f) return 1.0f - x * x * (2.5f - 1.5f * x);
        if (x < 2.0f) return 2.0f - x * (4.0f + x * (0.5f * x - 2.5f));
        return 0.0f;
    }
    if (x < 1.0f) return (16.0f + x * x * (21.0f * x - 36.0f)) / 18.0f;   /* Mitchell */
    if (x < 2.0f) return (32.0f + x * (-60.0f + x * (36.0f - 7.0f * x))) / 18.0f;
    return 0.0f;
}

/* One axis: fills, for every output index o, the ascending list of input
 * indices (clamped into [0, in_size)) and weights that accumulate into it.
 * first[o]..first[o+1] indexes src/wt (capacity: out_size * cap entries). */
static int orc_rz_axis(int in_size, int out_size, int *first, int *src, float *wt, int cap) {
    const float scale = (float) out_size / (float) in_size;
    const int grow = scale > 1.0f;
    const int width = 4;                                   /* ceil(2 * support) */
    int n = 0;
    if (grow) {
        const float radius = 2.0f * scale;
        for (int o = 0; o < out_size; o++) {
            const float centre = (float) o + 0.5f;
            const float lb = (centre - radius) / scale, ub = (centre + radius) / scale;
            const float in_centre = centre / scale;
            int lo = (int) floor((double) lb + 0.5), hi = (int) floor((double) ub - 0.5);
            float c[16];
            float total = 0.0f;
            int n0 = lo;
            for (int i = 0; i <= hi - lo; i++) {
                c[i] = orc_rz_kernel(1, in_centre - ((float) (i + lo) + 0.5f));
                if (i == 0 && c[i] == 0.0f) {
                    n0 = ++lo;
                    i--;
                    continue;
                }
                total += c[i];
            }
            const float s = 1.0f / total;
            int n1 = hi;
            for (int i = 0; i <= hi - lo; i++) c[i] *= s;
            for (int i = hi - lo; i >= 0; i--) {
                if (c[i] != 0.0f) break;
                n1 = n0 + i - 1;
            }
            first[o] = n;
            for (int k = n0; k <= n1; k++) {
                if (n >= cap) return -1;
                src[n] = k < 0 ? 0 : (k >= in_size ? in_size - 1 : k);
                wt[n++] = c[k - n0];
            }
        }
        first[out_size] = n;
        return n;
    }
    /* shrinking: weights per input pixel (with the margin), normalised per output */
    const int margin = (int) ceil((double) (2.0f * 2.0f / scale)) / 2;
    const int ncontrib = in_size + 2 * margin;
    const float radius = 2.0f / scale;
    orc_rz_contrib *ct = (orc_rz_contrib *) malloc(sizeof(orc_rz_contrib) * (size_t) ncontrib);
    /* the groups are contiguous, as the library lays them out: a fifth weight of
     * pixel j (its output range can span 5) lands in group j+1's first slot
     * and is read back from there; one spare group takes the last pixel's */
    float *cf = (float *) calloc((size_t) (ncontrib + 1) * width, sizeof(float));
    for (int j = 0; j < ncontrib; j++) {
        const float centre = (float) (j - margin) + 0.5f;
        const float lb = (centre - radius) * scale, ub = (centre + radius) * scale;
        const float out_centre = centre * scale;
        const int lo = (int) floor((double) lb + 0.5), hi = (int) floor((double) ub - 0.5);
        ct[j].n0 = lo;
        ct[j].n1 = hi;
        float *c = cf + (size_t) j * width;
        for (int i = 0; i <= hi - lo; i++) c[i] = orc_rz_kernel(0, ((float) (i + lo) + 0.5f) - out_centre) * scale;
        for (int i = hi - lo; i >= 0; i--) {
            if (c[i] != 0.0f) break;
            ct[j].n1 = ct[j].n0 + i - 1;
        }
    }
    for (int o = 0; o < out_size; o++) {
        float total = 0.0f;
        for (int j = 0; j < ncontrib; j++) {
            if (o >= ct[j].n0 && o <= ct[j].n1) total += cf[(size_t) j * width + (o - ct[j].n0)];
            else if (o < ct[j].n0) break;
        }
        const float s = 1.0f / total;
        for (int j = 0; j < ncontrib; j++) {
            if (o >= ct[j].n0 && o <= ct[j].n1) cf[(size_t) j * width + (o - ct[j].n0)] *= s;
            else if (o < ct[j].n0) break;
        }
    }
    for (int j = 0; j < ncontrib; j++) {
        float *c = cf + (size_t) j * width;
        int skip = 0;
        while (skip < width && c[skip] == 0.0f) skip++;
        if (skip == width) {   /* no weight at all: the pixel reaches no output */
            ct[j].n1 = ct[j].n0 - 1;
            continue;
        }
        ct[j].n0 += skip;
        while (ct[j].n0 < 0) {
            ct[j].n0++;
            skip++;
        }
        const int range = ct[j].n1 - ct[j].n0 + 1;
        const int mx = range < width ? range : width;
        for (int i = 0; i < mx && i + skip < width; i++) c[i] = c[i + skip];
        if (ct[j].n1 > out_size - 1) ct[j].n1 = out_size - 1;
    }
    /* gather per output, inputs ascending */
    for (int o = 0; o < out_size; o++) {
        first[o] = n;
        for (int j = 0; j < ncontrib; j++) {
            if (o < ct[j].n0 || o > ct[j].n1) continue;
            if (n >= cap) {
                free(ct);
                free(cf);
                return -1;
            }
            const int k = j - margin;
            src[n] = k < 0 ? 0 : (k >= in_size ? in_size - 1 : k);
            wt[n++] = cf[(size_t) j * width + (o - ct[j].n0)];
        }
    }
    first[out_size] = n;
    free(ct);
    free(cf);
    return n;
}

/* in: ih x iw x C uint8; out: oh x ow x C uint8.  Returns 0, or -1 on bad sizes. */
int orc_image_resize(const uint8_t *in, int iw, int ih, uint8_t *out, int ow, int oh, int C) {
    if (iw < 1 || ih < 1 || ow < 1 || oh < 1 || C < 1) return -1;
    /* entries per axis: at most 4 per input pixel and margin pixel when shrinking,
     * at most 6 per output pixel when growing */
    const int capx = 8 * ow + 4 * (iw + 2 * (int) (4.0 * iw / ow + 2)) + 64;
    const int capy = 8 * oh + 4 * (ih + 2 * (int) (4.0 * ih / oh + 2)) + 64;
    int *fx = (int *) malloc(sizeof(int) * (size_t) (ow + 1)), *fy = (int *) malloc(sizeof(int) * (size_t) (oh + 1));
    int *sx = (int *) malloc(sizeof(int) * (size_t) capx), *sy = (int *) malloc(sizeof(int) * (size_t) capy);
    float *wx = (float *) malloc(sizeof(float) * (size_t) capx), *wy = (float *) malloc(sizeof(float) * (size_t) capy);
    float *hb = (float *) malloc(sizeof(float) * (size_t) ih * ow * C);
    int rc = 0;
    if (orc_rz_axis(iw, ow, fx, sx, wx, capx) < 0 || orc_rz_axis(ih, oh, fy, sy, wy, capy) < 0) {
        rc = -1;
        goto done;
    }
    /* horizontal pass on every input row (rows outside the image are its
     * clamped edge rows, identical values) */
    for (int r = 0; r < ih; r++)
        for (int x = 0; x < ow; x++)
            for (int c = 0; c < C; c++) {
                float acc = 0.0f;
                for (int t = fx[x]; t < fx[x + 1]; t++) {
                    const float d = (float) in[((size_t) r * iw + sx[t]) * C + c] / 255.0f;
                    acc += d * wx[t];
                }
                hb[((size_t) r * ow + x) * C + c] = acc;
            }
    for (int y = 0; y < oh; y++)
        for (int x = 0; x < ow; x++)
            for (int c = 0; c < C; c++) {
                float acc = 0.0f;
                for (int t = fy[y]; t < fy[y + 1]; t++) acc += hb[((size_t) sy[t] * ow + x) * C + c] * wy[t];
                const float v = acc < 0.0f ? 0.0f : (acc > 1.0f ? 1.0f : acc);
                out[((size_t) y * ow + x) * C + c] = (uint8_t) (int) ((double) (v * 255.0f) + 0.5);
            }
done:
    free(fx); free(fy); free(sx); free(sy); free(wx); free(wy); free(hb);
    return rc;
}

/* ----------------------------------------------------------- image writers
 * The output images of the reference's drivers, restated literally (one pass
 * in the reference's loop order, so a later write overwrites an earlier one
 * exactly as in its single-thread kernels):
 *   transformImage    src/gpu/cudaImageFunctions.cuh:40-62 (transformImage_local,
 *                     src/cpu/multispectral_ImageMatcher.cu:181-203, has the same
 *                     body)
 *   fuseAlignedImages src/gpu/cudaImageFunctions.cuh:64-114
 *   fuseAlignedImages_local  src/cpu/multispectral_ImageMatcher.cu:115-179
 * The inverse map is the reference's adjugate expression of H, every product
 * and sum rounded on its own (no contraction), the quotient an IEEE division;
 * the in-image test is inImage on that quotient; samples are valueAt_bilinear
 * (the declared lerps of orc_sample_bilinear); a uint8 output takes the
 * sample truncated toward zero (the float-to-uint8 store of at<float>),
 * saturated to [0, 255]. */

static void orc_inverse_point(const float *H, float x, float y, float *nx, float *ny) {
    const float h11 = H[0], h12 = H[1], h13 = H[2], h21 = H[3], h22 = H[4], h23 = H[5], h31 = H[6], h32 = H[7];
    const float w = (h21 * h32 - h22 * h31) * x + (h12 * h31 - h11 * h32) * y + h11 * h22 - h12 * h21;
    *nx = ((h22 - h23 * h32) * x + (h13 * h32 - h12) * y + h12 * h23 - h13 * h22) / w;
    *ny = ((h23 * h31 - h21) * x + (h11 - h13 * h31) * y + h13 * h21 - h11 * h23) / w;
}

static void orc_store(const orc_image *out, int64_t lin, float v) {
    if (out->pixtype == ORC_PIX_U8) {
        const float t = truncf(v);
        ((uint8_t *) out->data)[lin] = (uint8_t) (!(t >= 0.0f) ? 0.0f : (t > 255.0f ? 255.0f : t));
    } else {
        ((float *) out->data)[lin] = v;
    }
}

/* transformImage: out (rows x cols) pixels whose sample point lies in `in`
 * get its bilinear value; the others keep what out holds. */
void orc_image_warp(const orc_image *in, const orc_image *out, const float *H) {
    const int C = in->channels;
    for (int x = 0; x < out->width; x++)
        for (int y = 0; y < out->height; y++) {
            float nx, ny;
            orc_inverse_point(H, (float) x, (float) y, &nx, &ny);
            if (!in_image(in, ny, nx)) continue;
            for (int c = 0; c < C; c++)
                orc_store(out, ((int64_t) y * out->width + x) * C + c, bilinear_in(in, ny, nx, c));
        }
}

/* fuseAlignedImages (local = 0) / fuseAlignedImages_local (local = 1) into a
 * 3-channel image: channel 0 the moving image under est, 1 the reference under
 * the identity, 2 the moving image under truth, each sampled at `channel`. */
void orc_image_fuse(const orc_image *ref, const orc_image *mov, const float *est, const float *truth, int channel,
                    int local, const orc_image *fused) {
    const float ident[8] = {1, 0, 0, 0, 1, 0, 0, 0};
    const int columns = fused->width, rows = fused->height;
    const int64_t total = (int64_t) columns * rows * 3;
    for (int c = 0; c < 3; c++) {
        const float *H = c == 0 ? est : (c == 1 ? ident : truth);
        const orc_image *src = c == 1 ? ref : mov;
        if (!local) {
            for (int x = 0; x < columns; x++)
                for (int y = 0; y < rows; y++) {
                    float nx, ny;
                    orc_inverse_point(H, (float) x, (float) y, &nx, &ny);
                    if (in_image(src, ny, nx))
                        orc_store(fused, ((int64_t) y * columns + x) * 3 + c, bilinear_in(src, ny, nx, channel));
                }
            continue;
        }
        /* normalised coordinates stepped in binary32 from -0.5 by 1/size; the
         * sample point scaled back by the moving image's size in binary64; the
         * target pixel round(size * (t + 0.5)), its linear index possibly the
         * next row's first pixel (past the end: not written) */
        for (float x = -0.5f; x < 0.5f; x += 1.0f / (float) columns)
            for (float y = -0.5f; y < 0.5f; y += 1.0f / (float) rows) {
                float nx, ny;
                orc_inverse_point(H, x, y, &nx, &ny);
                nx = (float) (((double) nx + 0.5) * (double) mov->width);
                ny = (float) (((double) ny + 0.5) * (double) mov->height);
                const int tx = (int) round((double) columns * ((double) x + 0.5));
                const int ty = (int) round((double) rows * ((double) y + 0.5));
                const int64_t lin = ((int64_t) ty * columns + tx) * 3 + c;
                if (lin < total && in_image(src, ny, nx)) orc_store(fused, lin, bilinear_in(src, ny, nx, channel));
            }
    }
}

/* ------------------------------------------------------- matrix product
 * CudaMatrix::multiply (include/cudaTensor.cuh:474-496, kernel multiplyProcess
 * include/cudaTensorKernels.cuh:396-440): the thread of C[r][c] accumulates
 * Csub += As[ty][k] * Bs[k][tx] over k ascending in 4-wide tiles from 0, nvcc
 * contracting each step into one fma; the zero-padded tile tails add exact
 * zeros, so the value is the plain fma chain over k. */
void orc_matrix_multiply_f32(const float *A, const float *B, float *C, int m, int k, int n) {
    for (int r = 0; r < m; r++)
        for (int c = 0; c < n; c++) {
            float acc = 0.0f;
            for (int j = 0; j < k; j++) acc = fmaf(A[(int64_t) r * k + j], B[(int64_t) j * n + c], acc);
            C[(int64_t) r * n + c] = acc;
        }
}

void orc_matrix_multiply_f64(const double *A, const double *B, double *C, int m, int k, int n) {
    for (int r = 0; r < m; r++)
        for (int c = 0; c < n; c++) {
            double acc = 0.0;
            for (int j = 0; j < k; j++) acc = fma(A[(int64_t) r * k + j], B[(int64_t) j * n + c], acc);
            C[(int64_t) r * n + c] = acc;
        }
}
