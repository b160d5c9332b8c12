/* cgs_mex.c -- MATLAB MEX gateway of libcgs: C = cgs_mex(A, B), the element-wise
 * sum of two real double matrices of one size, computed on the GPU.  The
 * counterpart of the reference's cudaGridSearch_mex (mexFunction,
 * src/gpu/cudaGridSearch_mex.cu:48-114, exported by Matlabdef.def), with its
 * argument checks and error identifiers (MATLAB:cudaAdd:inputmismatch,
 * :outputmismatch, :sizemismatch) and its prints of A[0] and B[1]; the sum is
 * cgs_tensor_binary(CGS_OP_ADD, ..., CGS_DT_F64) on device copies.
 *
 * Build with MATLAB (Linux):
 *   mex -I<repo>/include -I/opt/rocm/include -D__HIP_PLATFORM_AMD__ cgs_mex.c \
 *       -L<repo>/cugridsearch_amd -lcgs -L/opt/rocm/lib -lamdhip64
 * and export mexFunction (Matlabdef.def) on Windows. */
#include <string.h>

#include "mex.h"
#include <hip/hip_runtime_api.h>

#include "cgs.h"

/* Device buffers of one call; freed before any error leaves the gateway. */
typedef struct {
    double *a, *b;
} cgs_mex_buffers;

static void cgs_mex_free(cgs_mex_buffers *d) {
    if (d->a) (void) hipFree(d->a);
    if (d->b) (void) hipFree(d->b);
    d->a = d->b = NULL;
}

static void cgs_mex_fail(cgs_mex_buffers *d, const char *what, const char *msg) {
    cgs_mex_free(d);
    mexErrMsgIdAndTxt("cgs:cudaAdd:gpu", "%s: %s", what, msg);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs != 2) mexErrMsgIdAndTxt("MATLAB:cudaAdd:inputmismatch", "Input arguments must be 2!");
    if (nlhs != 1) mexErrMsgIdAndTxt("MATLAB:cudaAdd:outputmismatch", "Output arguments must be 1!");
    if (!mxIsDouble(prhs[0]) || !mxIsDouble(prhs[1]))
        mexErrMsgIdAndTxt("cgs:cudaAdd:inputtype", "Input matrices must be real double!");
    const double *A = mxGetPr(prhs[0]), *B = mxGetPr(prhs[1]);
    const mwSize n_a = mxGetNumberOfElements(prhs[0]), n_b = mxGetNumberOfElements(prhs[1]);
    if (n_a > 0) mexPrintf("%f\n", A[0]);
    if (n_b > 1) mexPrintf("%f\n", B[1]);
    const mwSize Am = mxGetM(prhs[0]), An = mxGetN(prhs[0]);
    if (Am != mxGetM(prhs[1]) || An != mxGetN(prhs[1]))
        mexErrMsgIdAndTxt("MATLAB:cudaAdd:sizemismatch", "Input matrices must have the same size!");

    plhs[0] = mxCreateDoubleMatrix(Am, An, mxREAL);
    double *C = mxGetPr(plhs[0]);
    const size_t n = (size_t) Am * An, bytes = n * sizeof(double);
    if (n == 0) return;
    cgs_mex_buffers d = {NULL, NULL};
    hipError_t e;
    if ((e = hipMalloc((void **) &d.a, bytes)) != hipSuccess || (e = hipMalloc((void **) &d.b, bytes)) != hipSuccess)
        cgs_mex_fail(&d, "hipMalloc", hipGetErrorString(e));
    if ((e = hipMemcpy(d.a, A, bytes, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(d.b, B, bytes, hipMemcpyHostToDevice)) != hipSuccess)
        cgs_mex_fail(&d, "hipMemcpy", hipGetErrorString(e));
    if (cgs_tensor_binary(CGS_OP_ADD, d.a, d.b, (int64_t) n, CGS_DT_F64, NULL) != CGS_OK)
        cgs_mex_fail(&d, "cgs_tensor_binary", cgs_last_error());
    if ((e = hipMemcpy(C, d.a, bytes, hipMemcpyDeviceToHost)) != hipSuccess)
        cgs_mex_fail(&d, "hipMemcpy", hipGetErrorString(e));
    cgs_mex_free(&d);
}
