/*
 * cnn_oracle.c -- CPU oracle of the proposal network (forward, loss,
 * backward).  TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the
 * GPU network (cugridsearch_amd/csrc/cgs_cnn.hip); the product never links
 * or calls it.
 *
 * The network has no reference counterpart (the reference has no learned
 * component); this file restates the network's own definition:
 *   input x[n][S][S][2]; six 3x3 convolutions, zero padding 1, channels
 *   2->32->32->64->64->128->128, stride 2 on layers 2, 4, 6, ReLU after each;
 *   global average pooling; FC 128 -> 8; loss = sum (out - target)^2 / (8n).
 * Parameter vector: per conv layer weight [Cout][3][3][Cin] then bias
 * [Cout]; FC weight [8][128], bias [8].
 *
 * All arithmetic is binary64.  emulate != 0 additionally rounds, at the
 * points where the GPU stores bf16, exactly as the GPU does (FP32 first,
 * then round-to-nearest-even to bf16): the conv and FC weights, every conv
 * output (after bias and ReLU) and every inter-layer gradient (the
 * pre-activation gradient of each conv layer).  emulate == 0 is the exact
 * function, whose gradient the tests pin against central finite
 * differences.  The spatial side S (a multiple of 8) is a parameter so that
 * finite-difference checks can run on small inputs; the GPU network is
 * fixed at S = 128.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NL 6
static const int CIN[NL] = {2, 32, 32, 64, 64, 128};
static const int COUT[NL] = {32, 32, 64, 64, 128, 128};
static const int STRIDE[NL] = {1, 2, 1, 2, 1, 2};
#define FEAT 128
#define NOUT 8

static int64_t off_w(int l) {
    int64_t p = 0;
    for (int i = 0; i < l; ++i) p += (int64_t) COUT[i] * 9 * CIN[i] + COUT[i];
    return p;
}
static int64_t off_b(int l) { return off_w(l) + (int64_t) COUT[l] * 9 * CIN[l]; }
static int64_t off_fcw(void) { return off_w(NL); }
static int64_t off_fcb(void) { return off_fcw() + NOUT * FEAT; }

int64_t cnn_param_count(void) { return off_fcb() + NOUT; }

/* binary64 -> binary32 -> bf16 (round to nearest even), back to binary64 */
double cnn_round_bf16(double v) {
    float f = (float) v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return v;
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return (double) f;
}

static void conv_fwd(const double *x, int n, int H, int cin, const double *w, const double *b, int cout, int s,
                     int emulate, double *y) {
    const int Ho = H / s;
#pragma omp parallel for collapse(2) schedule(static)
    for (int bi = 0; bi < n; ++bi)
        for (int oy = 0; oy < Ho; ++oy)
            for (int ox = 0; ox < Ho; ++ox)
                for (int co = 0; co < cout; ++co) {
                    double acc = 0.0;
                    for (int tap = 0; tap < 9; ++tap) {
                        const int iy = oy * s + tap / 3 - 1, ix = ox * s + tap % 3 - 1;
                        if (iy < 0 || iy >= H || ix < 0 || ix >= H) continue;
                        const double *xp = x + (((size_t) bi * H + iy) * H + ix) * cin;
                        const double *wp = w + ((size_t) co * 9 + tap) * cin;
                        for (int ci = 0; ci < cin; ++ci) acc += xp[ci] * wp[ci];
                    }
                    double v = acc + b[co];
                    v = v > 0.0 ? v : 0.0;
                    if (emulate) v = cnn_round_bf16(v);
                    y[(((size_t) bi * Ho + oy) * Ho + ox) * cout + co] = v;
                }
}

/* dz: [n][Ho][Ho][cout] pre-activation gradient of this layer.
 * dw[co][tap][ci] and db[co] are written; dx (may be NULL) receives the
 * input gradient [n][H][H][cin]. */
static void conv_bwd(const double *x, const double *dz, int n, int H, int cin, const double *w, int cout, int s,
                     double *dw, double *db, double *dx) {
    const int Ho = H / s;
#pragma omp parallel for schedule(static)
    for (int co = 0; co < cout; ++co) {
        double gb = 0.0;
        double *g = dw + (size_t) co * 9 * cin;
        for (int k = 0; k < 9 * cin; ++k) g[k] = 0.0;
        for (int bi = 0; bi < n; ++bi)
            for (int oy = 0; oy < Ho; ++oy)
                for (int ox = 0; ox < Ho; ++ox) {
                    const double d = dz[(((size_t) bi * Ho + oy) * Ho + ox) * cout + co];
                    gb += d;
                    if (d == 0.0) continue;
                    for (int tap = 0; tap < 9; ++tap) {
                        const int iy = oy * s + tap / 3 - 1, ix = ox * s + tap % 3 - 1;
                        if (iy < 0 || iy >= H || ix < 0 || ix >= H) continue;
                        const double *xp = x + (((size_t) bi * H + iy) * H + ix) * cin;
                        for (int ci = 0; ci < cin; ++ci) g[tap * cin + ci] += d * xp[ci];
                    }
                }
        db[co] = gb;
    }
    if (!dx) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (int bi = 0; bi < n; ++bi)
        for (int iy = 0; iy < H; ++iy)
            for (int ix = 0; ix < H; ++ix)
                for (int ci = 0; ci < cin; ++ci) {
                    double acc = 0.0;
                    for (int tap = 0; tap < 9; ++tap) {
                        const int ty = iy + 1 - tap / 3, tx = ix + 1 - tap % 3;
                        if (ty < 0 || tx < 0 || ty % s || tx % s) continue;
                        const int oy = ty / s, ox = tx / s;
                        if (oy >= Ho || ox >= Ho) continue;
                        const double *dp = dz + (((size_t) bi * Ho + oy) * Ho + ox) * cout;
                        for (int co = 0; co < cout; ++co) acc += dp[co] * w[((size_t) co * 9 + tap) * cin + ci];
                    }
                    dx[(((size_t) bi * H + iy) * H + ix) * cin + ci] = acc;
                }
}

/* Forward (+ backward when grad != NULL).  params: double[cnn_param_count()];
 * x: double[n][side][side][2]; target: double[n][8] (may be NULL when grad
 * is NULL); out: double[n][8] or NULL.  Returns the loss (0 without target). */
double cnn_loss_grad(const double *params, const double *x, const double *target, int n, int side, int emulate,
                     double *out, double *grad) {
    const int64_t P = cnn_param_count();
    double *prm = (double *) malloc(sizeof(double) * P);
    memcpy(prm, params, sizeof(double) * P);
    if (emulate) {
        for (int l = 0; l < NL; ++l)
            for (int64_t i = off_w(l); i < off_b(l); ++i) prm[i] = cnn_round_bf16(prm[i]);
        for (int64_t i = off_fcw(); i < off_fcb(); ++i) prm[i] = cnn_round_bf16(prm[i]);
        for (int l = 0; l < NL; ++l)
            for (int64_t i = off_b(l); i < off_b(l) + COUT[l]; ++i) prm[i] = (double) (float) prm[i];
        for (int64_t i = off_fcb(); i < P; ++i) prm[i] = (double) (float) prm[i];
    }
    int H[NL + 1];
    H[0] = side;
    for (int l = 0; l < NL; ++l) H[l + 1] = H[l] / STRIDE[l];
    double *act[NL + 1];
    act[0] = (double *) x;
    for (int l = 0; l < NL; ++l) {
        act[l + 1] = (double *) malloc(sizeof(double) * (size_t) n * H[l + 1] * H[l + 1] * COUT[l]);
        conv_fwd(act[l], n, H[l], CIN[l], prm + off_w(l), prm + off_b(l), COUT[l], STRIDE[l], emulate, act[l + 1]);
    }
    const int last = H[NL] * H[NL];
    double *feat = (double *) calloc((size_t) n * FEAT, sizeof(double));
    double *o = (double *) malloc(sizeof(double) * (size_t) n * NOUT);
    for (int bi = 0; bi < n; ++bi) {
        for (int p = 0; p < last; ++p)
            for (int c = 0; c < FEAT; ++c) feat[bi * FEAT + c] += act[NL][((size_t) bi * last + p) * FEAT + c];
        for (int c = 0; c < FEAT; ++c) feat[bi * FEAT + c] /= last;
        for (int j = 0; j < NOUT; ++j) {
            double v = prm[off_fcb() + j];
            for (int c = 0; c < FEAT; ++c) v += prm[off_fcw() + j * FEAT + c] * feat[bi * FEAT + c];
            o[bi * NOUT + j] = v;
        }
    }
    if (out) memcpy(out, o, sizeof(double) * (size_t) n * NOUT);
    double loss = 0.0;
    if (target) {
        for (int i = 0; i < n * NOUT; ++i) loss += (o[i] - target[i]) * (o[i] - target[i]);
        loss /= (double) NOUT * n;
    }
    if (grad && target) {
        double *dout = (double *) malloc(sizeof(double) * (size_t) n * NOUT);
        for (int i = 0; i < n * NOUT; ++i) dout[i] = 2.0 * (o[i] - target[i]) / ((double) NOUT * n);
        for (int j = 0; j < NOUT; ++j) {
            double gb = 0.0;
            for (int bi = 0; bi < n; ++bi) gb += dout[bi * NOUT + j];
            grad[off_fcb() + j] = gb;
            for (int c = 0; c < FEAT; ++c) {
                double g = 0.0;
                for (int bi = 0; bi < n; ++bi) g += dout[bi * NOUT + j] * feat[bi * FEAT + c];
                grad[off_fcw() + j * FEAT + c] = g;
            }
        }
        /* pre-activation gradient of the last conv */
        double *dz = (double *) malloc(sizeof(double) * (size_t) n * last * FEAT);
        for (int bi = 0; bi < n; ++bi)
            for (int c = 0; c < FEAT; ++c) {
                double df = 0.0;
                for (int j = 0; j < NOUT; ++j) df += dout[bi * NOUT + j] * prm[off_fcw() + j * FEAT + c];
                double g = df / last;
                if (emulate) g = cnn_round_bf16(g);
                for (int p = 0; p < last; ++p) {
                    const size_t i = ((size_t) bi * last + p) * FEAT + c;
                    dz[i] = act[NL][i] > 0.0 ? g : 0.0;
                }
            }
        for (int l = NL - 1; l >= 0; --l) {
            double *dx = l > 0 ? (double *) malloc(sizeof(double) * (size_t) n * H[l] * H[l] * CIN[l]) : NULL;
            conv_bwd(act[l], dz, n, H[l], CIN[l], prm + off_w(l), COUT[l], STRIDE[l], grad + off_w(l),
                     grad + off_b(l), dx);
            free(dz);
            dz = dx;
            if (dx) {
                const size_t cnt = (size_t) n * H[l] * H[l] * CIN[l];
                for (size_t i = 0; i < cnt; ++i) {
                    double g = act[l][i] > 0.0 ? dx[i] : 0.0;
                    dx[i] = emulate ? cnn_round_bf16(g) : g;
                }
            }
        }
        free(dout);
    }
    for (int l = 1; l <= NL; ++l) free(act[l]);
    free(feat);
    free(o);
    free(prm);
    return loss;
}
