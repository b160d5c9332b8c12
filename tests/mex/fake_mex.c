/* fake_mex.c -- the MEX API stand-in declared in mex.h (test
 * infrastructure): heap-backed column-major arrays and an error path that
 * longjmps back to the caller of fake_mex_call. */
#include "mex.h"

#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct mxArray_tag {
    mxClassID cls;
    mwSize m, n;
    void *data;
};

static size_t elem_size(mxClassID c) {
    switch (c) {
        case mxDOUBLE_CLASS: return 8;
        case mxSINGLE_CLASS: return 4;
        case mxINT32_CLASS: return 4;
        case mxUINT8_CLASS: return 1;
        case mxCHAR_CLASS: return 2;
        default: return 1;
    }
}

mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c) {
    (void) c;
    mxArray *a = (mxArray *) calloc(1, sizeof(mxArray));
    a->cls = cls;
    a->m = m;
    a->n = n;
    a->data = calloc(m * n ? m * n : 1, elem_size(cls));
    return a;
}
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) { return mxCreateNumericMatrix(m, n, mxDOUBLE_CLASS, c); }
mxArray *mxCreateDoubleScalar(double v) {
    mxArray *a = mxCreateDoubleMatrix(1, 1, mxREAL);
    *(double *) a->data = v;
    return a;
}
mxArray *mxCreateString(const char *s) {
    const size_t n = strlen(s);
    mxArray *a = mxCreateNumericMatrix(1, n, mxCHAR_CLASS, mxREAL);
    for (size_t i = 0; i < n; ++i) ((uint16_t *) a->data)[i] = (uint16_t) (unsigned char) s[i];
    return a;
}
void mxDestroyArray(mxArray *a) {
    if (a) free(a->data);
    free(a);
}
double *mxGetPr(const mxArray *a) { return a && a->cls == mxDOUBLE_CLASS ? (double *) a->data : NULL; }
void *mxGetData(const mxArray *a) { return a ? a->data : NULL; }
mwSize mxGetM(const mxArray *a) { return a ? a->m : 0; }
mwSize mxGetN(const mxArray *a) { return a ? a->n : 0; }
mwSize mxGetNumberOfElements(const mxArray *a) { return a ? a->m * a->n : 0; }
mxClassID mxGetClassID(const mxArray *a) { return a ? a->cls : mxUNKNOWN_CLASS; }
int mxIsDouble(const mxArray *a) { return a && a->cls == mxDOUBLE_CLASS; }
int mxIsSingle(const mxArray *a) { return a && a->cls == mxSINGLE_CLASS; }
int mxIsUint8(const mxArray *a) { return a && a->cls == mxUINT8_CLASS; }
int mxIsChar(const mxArray *a) { return a && a->cls == mxCHAR_CLASS; }
int mxGetString(const mxArray *a, char *buf, mwSize len) {
    if (!a || a->cls != mxCHAR_CLASS || len == 0) return 1;
    const mwSize n = a->m * a->n;
    if (n + 1 > len) return 1;
    for (mwSize i = 0; i < n; ++i) buf[i] = (char) ((uint16_t *) a->data)[i];
    buf[n] = 0;
    return 0;
}

void *mxMalloc(mwSize n) { return malloc(n ? n : 1); }
void mxFree(void *p) { free(p); }

static jmp_buf g_jmp;
static char g_err_id[128];

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "mex error %s: ", id);
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
    snprintf(g_err_id, sizeof(g_err_id), "%s", id);
    longjmp(g_jmp, 1);
}

int mexPrintf(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    const int r = vprintf(fmt, ap);
    va_end(ap);
    return r;
}

int fake_mex_call(fake_mex_gateway fn, int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[], char *err_id,
                  size_t err_id_len) {
    g_err_id[0] = 0;
    if (setjmp(g_jmp)) {
        if (err_id) snprintf(err_id, err_id_len, "%s", g_err_id);
        return 1;
    }
    fn(nlhs, plhs, nrhs, prhs);
    if (err_id && err_id_len) err_id[0] = 0;
    return 0;
}
