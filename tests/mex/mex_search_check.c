/* Drives cugridsearch_amd/mex/cgs_search_mex.c through the fake MEX API.
 *   mex_search_check --errors-only          its argument errors (no GPU needed)
 *   mex_search_check <spec> <mov> <fix> <costs_out>
 * spec: pixtype(u8|f32) W H D start[D] stop[D] count[D] functor bins sampler;
 * the images are row-major raw files, handed over as H x W column-major
 * matrices.  Prints {"point": [...], "cost": c, "index": i} and writes the
 * costs (float32, grid order). */
#include <stdio.h>
#include <stdlib.h>
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

static mxArray *vec(int n, const double *v) {
    mxArray *a = mxCreateDoubleMatrix(1, (mwSize) n, mxREAL);
    memcpy(mxGetPr(a), v, sizeof(double) * n);
    return a;
}

static void expect_error(int nlhs, int nrhs, const mxArray **in, const char *want) {
    mxArray *out[4] = {NULL, NULL, NULL, NULL};
    char id[128] = "";
    EXPECT(fake_mex_call(mexFunction, nlhs, out, nrhs, in, id, sizeof(id)) == 1);
    if (strcmp(id, want)) fprintf(stderr, "got %s, want %s\n", id, want);
    EXPECT(strcmp(id, want) == 0);
}

static int errors(void) {
    mxArray *img = mxCreateNumericMatrix(8, 8, mxUINT8_CLASS, mxREAL), *dimg = mxCreateDoubleMatrix(8, 8, mxREAL);
    const double s[2] = {-1, -1}, e[2] = {1, 1}, c[2] = {3, 3}, c3[3] = {3, 3, 3};
    mxArray *S = vec(2, s), *E = vec(2, e), *C = vec(2, c), *C3 = vec(3, c3);
    mxArray *fn = mxCreateString("sumOfAbsoluteDifferences"), *bad = mxCreateString("noSuchFunctor");
    mxArray *samp = mxCreateString("cubic"), *bins = mxCreateDoubleScalar(32);
    const mxArray *five[5] = {img, img, S, E, C};
    const mxArray *dbl[6] = {dimg, img, S, E, C, fn};
    const mxArray *grid[6] = {img, img, S, E, C3, fn};
    const mxArray *unk[6] = {img, img, S, E, C, bad};
    const mxArray *sam[8] = {img, img, S, E, C, fn, bins, samp};
    const mxArray *ok[6] = {img, img, S, E, C, fn};
    expect_error(1, 5, five, "cgs:search:inputmismatch");
    expect_error(5, 6, ok, "cgs:search:outputmismatch");
    expect_error(1, 6, dbl, "cgs:search:imagetype");
    expect_error(1, 6, grid, "cgs:search:gridmismatch");
    expect_error(1, 6, unk, "cgs:search:functor");
    expect_error(1, 8, sam, "cgs:search:sampler");
    printf("mex_search errors %s\n", g_fail ? "FAILED" : "ok");
    return g_fail;
}

int main(int argc, char **argv) {
    if (argc == 2 && strcmp(argv[1], "--errors-only") == 0) return errors();
    if (argc != 5) {
        fprintf(stderr, "usage: mex_search_check <spec> <mov> <fix> <costs_out>\n");
        return 2;
    }
    FILE *sp = fopen(argv[1], "r");
    char pix[8], fname[64], sname[16];
    int W, H, D;
    double s[8], e[8], c[8], bins;
    if (!sp || fscanf(sp, "%7s %d %d %d", pix, &W, &H, &D) != 4 || D < 1 || D > 8) return 2;
    for (int a = 0; a < D; ++a) if (fscanf(sp, "%lf", &s[a]) != 1) return 2;
    for (int a = 0; a < D; ++a) if (fscanf(sp, "%lf", &e[a]) != 1) return 2;
    for (int a = 0; a < D; ++a) if (fscanf(sp, "%lf", &c[a]) != 1) return 2;
    if (fscanf(sp, "%63s %lf %15s", fname, &bins, sname) != 3) return 2;
    fclose(sp);
    const int u8 = strcmp(pix, "u8") == 0;
    const size_t esz = u8 ? 1 : 4;
    mxArray *img[2];
    for (int k = 0; k < 2; ++k) {
        unsigned char *raw = (unsigned char *) malloc((size_t) W * H * esz);
        FILE *f = fopen(argv[2 + k], "rb");
        if (!f || fread(raw, esz, (size_t) W * H, f) != (size_t) W * H) return 2;
        fclose(f);
        img[k] = mxCreateNumericMatrix((mwSize) H, (mwSize) W, u8 ? mxUINT8_CLASS : mxSINGLE_CLASS, mxREAL);
        unsigned char *col = (unsigned char *) mxGetData(img[k]);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) memcpy(col + ((size_t) x * H + y) * esz, raw + ((size_t) y * W + x) * esz, esz);
        free(raw);
    }
    const mxArray *in[8] = {img[0], img[1], vec(D, s), vec(D, e), vec(D, c), mxCreateString(fname),
                            mxCreateDoubleScalar(bins), mxCreateString(sname)};
    mxArray *out[4] = {NULL, NULL, NULL, NULL};
    char id[128];
    if (fake_mex_call(mexFunction, 4, out, 8, in, id, sizeof(id))) {
        fprintf(stderr, "gateway failed: %s\n", id);
        return 1;
    }
    FILE *fo = fopen(argv[4], "wb");
    fwrite(mxGetData(out[3]), 4, mxGetNumberOfElements(out[3]), fo);
    fclose(fo);
    printf("{\"point\": [");
    for (int a = 0; a < D; ++a) printf("%s%.9g", a ? ", " : "", mxGetPr(out[0])[a]);
    printf("], \"cost\": %.9g, \"index\": %.0f, \"n\": %zu}\n", mxGetPr(out[1])[0], mxGetPr(out[2])[0],
           (size_t) mxGetNumberOfElements(out[3]));
    return 0;
}
