/* cgs_search_mex.c -- MATLAB MEX gateway of the grid search itself:
 *
 *   [point, cost, index, costs] = cgs_search_mex(moving, fixed, start, stop, count, functor[, bins[, sampler]])
 *
 * moving, fixed: H x W real matrices (MATLAB's rows are image rows), uint8 or
 * single, one channel; start, stop, count: double vectors of the grid's D
 * values (8, or 2 for the translation functors), as CudaGrid::setStartPoint /
 * setEndPoint / setNumSamples take them (src/gpu/cudaGridSearch.cuh:198-215);
 * functor: the name of a reference functor (calcNCC, calcSQD, calcMIstream,
 * calcMI, averageAbsoluteDifference, averageAbsoluteDifferenceH,
 * sumOfAbsoluteDifferences, calcNCCstream, calcSQDstream,
 * averageAbsoluteDifference_stream, calcSQDAlt, calcNCCAlt, calcNCCstreamAlt,
 * calcMIstreamAlt, calcMIAlt, centredAbsoluteDifference); bins: MI histogram
 * bins (default 64, the reference's); sampler: 'bilinear' (default) or
 * 'nearest'.  Returns the argmin's grid point (1 x D double), its cost, its
 * grid index (0-based, axis 0 fastest, as find_extrema's) and, when asked,
 * the cost of every grid point (N x 1 single, grid order).
 *
 * No reference counterpart: the reference's MATLAB harness runs the search
 * through the multispectral_ImageMatcher CLI (matlab/experiment_00.m:77-98).
 * Built like cgs_mex.c. */
#include <string.h>

#include "mex.h"
#include <hip/hip_runtime_api.h>

#include "cgs.h"

typedef struct {
    const char *name;
    int cost, mode;
} cgs_mex_functor;

static const cgs_mex_functor kFunctors[] = {
    {"calcNCC", CGS_COST_NCC, CGS_PARAM_CENTRED},
    {"calcSQD", CGS_COST_SQD, CGS_PARAM_CENTRED},
    {"calcMIstream", CGS_COST_MI, CGS_PARAM_CENTRED},
    {"calcMI", CGS_COST_MI_PLAIN, CGS_PARAM_CENTRED},
    {"averageAbsoluteDifference", CGS_COST_SAD, CGS_PARAM_NONCENTRED},
    {"averageAbsoluteDifferenceH", CGS_COST_SAD, CGS_PARAM_RAW_H},
    {"sumOfAbsoluteDifferences", CGS_COST_SAD_SUM, CGS_PARAM_TRANSLATION},
    {"calcNCCstream", CGS_COST_NCC_STREAM, CGS_PARAM_CENTRED},
    {"calcSQDstream", CGS_COST_SQD_STREAM, CGS_PARAM_CENTRED},
    {"averageAbsoluteDifference_stream", CGS_COST_SAD_STREAM, CGS_PARAM_TRANSLATION},
    {"calcSQDAlt", CGS_COST_SQD_ALT, CGS_PARAM_CENTRED},
    {"calcNCCAlt", CGS_COST_NCC_ALT, CGS_PARAM_CENTRED},
    {"calcNCCstreamAlt", CGS_COST_NCC_STREAM_ALT, CGS_PARAM_CENTRED},
    {"calcMIstreamAlt", CGS_COST_MI_STREAM_ALT, CGS_PARAM_CENTRED},
    {"calcMIAlt", CGS_COST_MI_ALT, CGS_PARAM_NORM},
    {"centredAbsoluteDifference", CGS_COST_SAD, CGS_PARAM_CENTRED},
};

typedef struct {
    void *mov, *fix;
    float *costs;
} cgs_mex_buffers;

static void cgs_mex_free(cgs_mex_buffers *d) {
    if (d->mov) (void) hipFree(d->mov);
    if (d->fix) (void) hipFree(d->fix);
    if (d->costs) (void) hipFree(d->costs);
    memset(d, 0, sizeof(*d));
}

static void cgs_mex_fail(cgs_mex_buffers *d, const char *id, const char *what, const char *msg) {
    cgs_mex_free(d);
    mexErrMsgIdAndTxt(id, "%s: %s", what, msg);
}

/* An H x W column-major MATLAB image as a row-major device copy. */
static void *cgs_mex_upload(const mxArray *a, int *pixtype, cgs_mex_buffers *d) {
    const mwSize H = mxGetM(a), W = mxGetN(a);
    const size_t esz = mxIsUint8(a) ? 1 : 4;
    *pixtype = mxIsUint8(a) ? CGS_PIX_U8 : CGS_PIX_F32;
    unsigned char *host = (unsigned char *) mxGetData(a);
    unsigned char *rows = (unsigned char *) mxMalloc((size_t) H * W * esz);
    for (mwSize y = 0; y < H; ++y)
        for (mwSize x = 0; x < W; ++x) memcpy(rows + (y * W + x) * esz, host + (x * H + y) * esz, esz);
    void *dev = NULL;
    hipError_t e = hipMalloc(&dev, (size_t) H * W * esz);
    if (e == hipSuccess) e = hipMemcpy(dev, rows, (size_t) H * W * esz, hipMemcpyHostToDevice);
    mxFree(rows);
    if (e != hipSuccess) {
        if (dev) (void) hipFree(dev);
        cgs_mex_fail(d, "cgs:search:gpu", "upload", hipGetErrorString(e));
    }
    return dev;
}

static int cgs_mex_vector(const mxArray *a, float *out, int cap) {
    if (!mxIsDouble(a)) return -1;
    const mwSize n = mxGetNumberOfElements(a);
    if (n < 1 || n > (mwSize) cap) return -1;
    for (mwSize i = 0; i < n; ++i) out[i] = (float) mxGetPr(a)[i];
    return (int) n;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs < 6 || nrhs > 8)
        mexErrMsgIdAndTxt("cgs:search:inputmismatch", "Input arguments must be 6 to 8!");
    if (nlhs < 1 || nlhs > 4) mexErrMsgIdAndTxt("cgs:search:outputmismatch", "Output arguments must be 1 to 4!");
    for (int i = 0; i < 2; ++i)
        if (!(mxIsUint8(prhs[i]) || mxIsSingle(prhs[i])) || mxGetNumberOfElements(prhs[i]) == 0)
            mexErrMsgIdAndTxt("cgs:search:imagetype", "Images must be non-empty uint8 or single matrices!");
    float start[CGS_MAX_DIM], stop[CGS_MAX_DIM], count[CGS_MAX_DIM];
    const int D = cgs_mex_vector(prhs[2], start, CGS_MAX_DIM);
    if (D < 1 || cgs_mex_vector(prhs[3], stop, CGS_MAX_DIM) != D || cgs_mex_vector(prhs[4], count, CGS_MAX_DIM) != D)
        mexErrMsgIdAndTxt("cgs:search:gridmismatch", "start, stop and count must be double vectors of one length!");
    char name[64], samp[16] = "bilinear";
    if (!mxIsChar(prhs[5]) || mxGetString(prhs[5], name, sizeof(name)))
        mexErrMsgIdAndTxt("cgs:search:functor", "The functor must be given by name!");
    const cgs_mex_functor *fn = NULL;
    for (size_t i = 0; i < sizeof(kFunctors) / sizeof(kFunctors[0]); ++i)
        if (!strcmp(kFunctors[i].name, name)) fn = &kFunctors[i];
    if (!fn) mexErrMsgIdAndTxt("cgs:search:functor", "Unknown functor %s!", name);
    cgs_search_opts o;
    cgs_default_opts(&o);
    o.cost = fn->cost;
    o.param_mode = fn->mode;
    o.mi_bins = 64;
    if (nrhs >= 7) {
        if (!mxIsDouble(prhs[6]) || mxGetNumberOfElements(prhs[6]) != 1)
            mexErrMsgIdAndTxt("cgs:search:bins", "bins must be a double scalar!");
        o.mi_bins = (int32_t) mxGetPr(prhs[6])[0];
    }
    if (nrhs == 8 && (!mxIsChar(prhs[7]) || mxGetString(prhs[7], samp, sizeof(samp))))
        mexErrMsgIdAndTxt("cgs:search:sampler", "The sampler must be 'bilinear' or 'nearest'!");
    if (!strcmp(samp, "nearest")) o.sampler = CGS_SAMPLER_NEAREST;
    else if (strcmp(samp, "bilinear")) mexErrMsgIdAndTxt("cgs:search:sampler", "The sampler must be 'bilinear' or 'nearest'!");

    cgs_grid g;
    if (cgs_grid_init(&g, D, start, stop, count) != CGS_OK)
        mexErrMsgIdAndTxt("cgs:search:grid", "cgs_grid_init: %s", cgs_last_error());

    cgs_mex_buffers d = {NULL, NULL, NULL};
    cgs_image mv, fx;
    d.mov = cgs_mex_upload(prhs[0], &mv.pixtype, &d);
    d.fix = cgs_mex_upload(prhs[1], &fx.pixtype, &d);
    mv.data = d.mov;
    mv.width = (int32_t) mxGetN(prhs[0]);
    mv.height = (int32_t) mxGetM(prhs[0]);
    mv.channels = 1;
    fx.data = d.fix;
    fx.width = (int32_t) mxGetN(prhs[1]);
    fx.height = (int32_t) mxGetM(prhs[1]);
    fx.channels = 1;
    const size_t n = (size_t) g.num_elements;
    hipError_t e = hipSuccess;
    if (nlhs == 4 && (e = hipMalloc((void **) &d.costs, n * sizeof(float))) != hipSuccess)
        cgs_mex_fail(&d, "cgs:search:gpu", "hipMalloc", hipGetErrorString(e));
    cgs_argmin am;
    if (cgs_search(&g, &mv, &fx, &o, d.costs, &am, NULL) != CGS_OK)
        cgs_mex_fail(&d, "cgs:search:failed", "cgs_search", cgs_last_error());

    plhs[0] = mxCreateDoubleMatrix(1, (mwSize) D, mxREAL);
    float pt[CGS_MAX_DIM] = {0};
    if (am.index >= 0) (void) cgs_grid_point(&g, am.index, pt);
    for (int a = 0; a < D; ++a) mxGetPr(plhs[0])[a] = pt[a];
    if (nlhs >= 2) plhs[1] = mxCreateDoubleScalar((double) am.value);
    if (nlhs >= 3) plhs[2] = mxCreateDoubleScalar((double) am.index);
    if (nlhs == 4) {
        plhs[3] = mxCreateNumericMatrix((mwSize) n, 1, mxSINGLE_CLASS, mxREAL);
        if ((e = hipMemcpy(mxGetData(plhs[3]), d.costs, n * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)
            cgs_mex_fail(&d, "cgs:search:gpu", "hipMemcpy", hipGetErrorString(e));
    }
    cgs_mex_free(&d);
}
