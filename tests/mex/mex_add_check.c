/* Drives cugridsearch_amd/mex/cgs_mex.c through the fake MEX API: the
 * reference gateway's argument errors (src/gpu/cudaGridSearch_mex.cu:52-74)
 * and, unless --errors-only (no GPU needed), C = A + B on the GPU against the
 * host sum, element for element.  Prints "mex_add ok" or the failures. */
#include <stdio.h>
#include <string.h>

#include "mex.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

static int g_fail = 0;
#define EXPECT(c)                                                   \
    do {                                                            \
        if (!(c)) {                                                 \
            fprintf(stderr, "%s:%d: EXPECT(%s)\n", __FILE__, __LINE__, #c); \
            g_fail = 1;                                             \
        }                                                           \
    } while (0)

static mxArray *matrix(mwSize m, mwSize n, double base, double step) {
    mxArray *a = mxCreateDoubleMatrix(m, n, mxREAL);
    for (mwSize i = 0; i < m * n; ++i) mxGetPr(a)[i] = base + step * (double) i;
    return a;
}

static void expect_error(int nlhs, int nrhs, const mxArray **in, const char *want) {
    mxArray *out[2] = {NULL, NULL};
    char id[128] = "";
    EXPECT(fake_mex_call(mexFunction, nlhs, out, nrhs, in, id, sizeof(id)) == 1);
    EXPECT(strcmp(id, want) == 0);
}

int main(int argc, char **argv) {
    const int errors_only = argc > 1 && strcmp(argv[1], "--errors-only") == 0;
    mxArray *A = matrix(3, 4, -1.25, 0.5), *B = matrix(3, 4, 1e-3, 1.0 / 3.0), *Bt = matrix(4, 3, 0.0, 1.0);
    mxArray *S = mxCreateNumericMatrix(3, 4, mxSINGLE_CLASS, mxREAL);
    const mxArray *one[1] = {A}, *three[3] = {A, B, B}, *two[2] = {A, B}, *mism[2] = {A, Bt}, *single[2] = {A, S};
    expect_error(1, 1, one, "MATLAB:cudaAdd:inputmismatch");
    expect_error(1, 3, three, "MATLAB:cudaAdd:inputmismatch");
    expect_error(2, 2, two, "MATLAB:cudaAdd:outputmismatch");
    expect_error(0, 2, two, "MATLAB:cudaAdd:outputmismatch");
    expect_error(1, 2, mism, "MATLAB:cudaAdd:sizemismatch");
    expect_error(1, 2, single, "cgs:cudaAdd:inputtype");
    if (!errors_only) {
        mxArray *out[1] = {NULL};
        char id[128];
        EXPECT(fake_mex_call(mexFunction, 1, out, 2, two, id, sizeof(id)) == 0);
        EXPECT(out[0] && mxGetM(out[0]) == 3 && mxGetN(out[0]) == 4);
        for (mwSize i = 0; out[0] && i < 12; ++i) EXPECT(mxGetPr(out[0])[i] == mxGetPr(A)[i] + mxGetPr(B)[i]);
        mxDestroyArray(out[0]);
        mxArray *L = matrix(257, 129, 3.5, -0.0078125), *R = matrix(257, 129, -7.0, 1e-5);
        const mxArray *big[2] = {L, R};
        out[0] = NULL;
        EXPECT(fake_mex_call(mexFunction, 1, out, 2, big, id, sizeof(id)) == 0);
        for (mwSize i = 0; out[0] && i < 257 * 129; ++i) EXPECT(mxGetPr(out[0])[i] == mxGetPr(L)[i] + mxGetPr(R)[i]);
        mxDestroyArray(out[0]);
        mxArray *E = mxCreateDoubleMatrix(0, 0, mxREAL);
        const mxArray *empty[2] = {E, E};
        out[0] = NULL;
        EXPECT(fake_mex_call(mexFunction, 1, out, 2, empty, id, sizeof(id)) == 0);
        EXPECT(out[0] && mxGetNumberOfElements(out[0]) == 0);
    }
    printf("mex_add %s\n", g_fail ? "FAILED" : "ok");
    return g_fail;
}
