/* mex.h -- a stand-in for MATLAB's MEX API, enough to build and drive the
 * library's MEX gateways (cugridsearch_amd/mex/) without MATLAB.  Test
 * infrastructure only: the real header comes with MATLAB.  Arrays are
 * column-major M x N blocks of one class; mexErrMsgIdAndTxt leaves the
 * gateway (as MATLAB does) by a longjmp back into fake_mex_call. */
#ifndef FAKE_MEX_H
#define FAKE_MEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef ptrdiff_t mwSignedIndex;
typedef enum { mxUNKNOWN_CLASS = 0, mxCHAR_CLASS, mxDOUBLE_CLASS, mxSINGLE_CLASS, mxUINT8_CLASS, mxINT32_CLASS } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

typedef struct mxArray_tag mxArray;

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c);
mxArray *mxCreateDoubleScalar(double v);
mxArray *mxCreateString(const char *s);
void mxDestroyArray(mxArray *a);
double *mxGetPr(const mxArray *a);
void *mxGetData(const mxArray *a);
mwSize mxGetM(const mxArray *a);
mwSize mxGetN(const mxArray *a);
mwSize mxGetNumberOfElements(const mxArray *a);
mxClassID mxGetClassID(const mxArray *a);
int mxIsDouble(const mxArray *a);
int mxIsSingle(const mxArray *a);
int mxIsUint8(const mxArray *a);
int mxIsChar(const mxArray *a);
int mxGetString(const mxArray *a, char *buf, mwSize len);

void *mxMalloc(mwSize n);
void mxFree(void *p);

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
int mexPrintf(const char *fmt, ...);

/* The harness side: runs a gateway; returns 0, or 1 after an error (id and
 * message copied out). */
typedef void (*fake_mex_gateway)(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
int fake_mex_call(fake_mex_gateway fn, int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[], char *err_id,
                  size_t err_id_len);

#ifdef __cplusplus
}
#endif

#endif
