// cgs_cnn.hip -- the proposal network: forward, backward and Adam on gfx950.
//
// Network (fixed): input 2 x 128 x 128 (moving, fixed; standardised), six
// 3x3 convolutions with zero padding 1 -- channels 32-32-64-64-128-128,
// stride 2 on layers 2, 4, 6, ReLU after each -- global average pooling
// and a fully connected layer to 8 outputs (corner offsets).
//
// Arithmetic: weights are BF16 (rounded from FP32 master weights after
// every optimiser step), activations and the gradients between layers are
// stored as BF16, every convolution accumulates in FP32 on the matrix cores
// (v_mfma_f32_16x16x32_bf16); biases, pooling, the FC layer, the loss and
// the optimiser are FP32.
//
// Layout: activations NHWC (pixel-major, channels contiguous), conv weights
// [Cout][ky][kx][Cin] padded with zeros to a multiple of 32 along the
// 9*Cin reduction, plus a [ky][kx][Cin][Cout] copy for the data gradient.
//
// All three convolution passes are implicit GEMMs run by one skeleton
// (igemm_kernel): a workgroup of 4 waves owns an MT x NT output tile and
// walks the reduction in 32-wide chunks; each chunk's A (MT x 32) and B
// (NT x 32) tiles are gathered into LDS with the reduction index
// contiguous, and each wave multiplies its (MT/2) x (NT/2) quadrant with
// 16x16x32 MFMAs.  The pass decides what the tiles hold:
//   forward      M = output pixels, N = Cout, K = 9*Cin (taps x channels);
//   data grad    M = input pixels,  N = Cin,  K = 9*Cout;
//   weight grad  M = Cout, N = 9*Cin, K = output pixels, split over
//                blockIdx.z into slices whose partial sums a second kernel
//                adds in slice order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "cgs.h"
#include "cgs_host.hpp"

namespace cgs {
namespace cnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kLayers = 6;
constexpr int kIn = 128;            // input side
constexpr int kFeat = 128;          // channels of the last conv = FC inputs
constexpr int kOut = 8;             // FC outputs
constexpr int kCin[kLayers] = {2, 32, 32, 64, 64, 128};
constexpr int kCout[kLayers] = {32, 32, 64, 64, 128, 128};
constexpr int kStride[kLayers] = {1, 2, 1, 2, 1, 2};
constexpr int kHin[kLayers] = {128, 128, 64, 64, 32, 32};
constexpr int hout(int l) { return kHin[l] / kStride[l]; }
constexpr int ktrue(int l) { return 9 * kCin[l]; }
constexpr int kpad(int l) { return (ktrue(l) + 31) / 32 * 32; }
constexpr int kLastSide = hout(kLayers - 1);          // 16
constexpr int kLastPix = kLastSide * kLastSide;       // 256
constexpr int kLDK = 40;            // LDS row pitch in bf16 (32 + 8 pad)

// Offsets of each tensor in the flat FP32 parameter vector:
// per layer weight [Cout][9*Cin] then bias [Cout]; FC weight [8][128], bias [8].
struct Offsets {
    int64_t w[kLayers], b[kLayers], fcw, fcb, total;
};
constexpr Offsets offsets() {
    Offsets o{};
    int64_t p = 0;
    for (int l = 0; l < kLayers; ++l) {
        o.w[l] = p;
        p += (int64_t) kCout[l] * ktrue(l);
        o.b[l] = p;
        p += kCout[l];
    }
    o.fcw = p;
    p += kOut * kFeat;
    o.fcb = p;
    p += kOut;
    o.total = p;
    return o;
}
constexpr Offsets kOff = offsets();

// FP32 -> BF16, round to nearest even (NaN stays a quiet NaN).
__device__ __host__ inline uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t) ((u >> 16) | 0x40);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t) (u >> 16);
}
__device__ __host__ inline float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t) h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
// ReLU outputs are >= 0; "active" means strictly positive (+0 and -0 are not).
__device__ inline bool positive(uint16_t h) { return h != 0 && (h & 0x8000u) == 0; }

// ------------------------------------------------------------ GEMM skeleton
//
// A problem type supplies, for an MT x NT tile and the 32-wide reduction
// chunk at k0, the tile contents in 16-byte pieces (8 bf16): loadA / loadB
// read a thread's pieces from global memory into registers, storeA /
// storeB put them into the LDS tiles (reduction index contiguous per row),
// k_range gives the blockIdx.z slice of the reduction and store() takes one
// accumulator.  The skeleton issues the loads of chunk k+1 before it
// multiplies chunk k, so global-memory latency overlaps the matrix work.

template <int T>
constexpr int pieces() { return (T * 4 + 255) / 256; }   // 16-byte pieces per thread of a T x 32 tile

template <class Prob, int MT, int NT>
__global__ __launch_bounds__(256) void igemm_kernel(Prob pb) {
    static_assert(MT % 32 == 0 && NT % 32 == 0, "tiles are multiples of 32");
    constexpr int WM = MT / 2, WN = NT / 2, FM = WM / 16, FN = WN / 16;
    constexpr int kOperands = (MT + NT) * kLDK, kTile = MT * (NT + 8);   // operand tiles; bf16 output tile
    __shared__ __attribute__((aligned(16))) uint16_t lds[kOperands > kTile ? kOperands : kTile];
    uint16_t *const As = lds, *const Bs = lds + MT * kLDK;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.x * MT, n0 = blockIdx.y * NT;
    int kbeg, kend;
    pb.k_range(blockIdx.z, kbeg, kend);
    f32x4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int lrow = lane & 15, lk = (lane >> 4) * 8;
    uint4 ra[pieces<MT>()], rb[pieces<NT>()];
    if (kbeg < kend) {
        pb.template loadA<MT>(ra, m0, kbeg, tid);
        pb.template loadB<NT>(rb, n0, kbeg, tid);
    }
    for (int k0 = kbeg; k0 < kend; k0 += 32) {
        pb.template storeA<MT>(As, ra, tid);
        pb.template storeB<NT>(Bs, rb, tid);
        __syncthreads();
        if (k0 + 32 < kend) {
            pb.template loadA<MT>(ra, m0, k0 + 32, tid);
            pb.template loadB<NT>(rb, n0, k0 + 32, tid);
        }
        bf16x8 a[FM], b[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
            a[i] = __builtin_bit_cast(bf16x8, *(const uint4 *) &As[(wm * WM + i * 16 + lrow) * kLDK + lk]);
#pragma unroll
        for (int j = 0; j < FN; ++j)
            b[j] = __builtin_bit_cast(bf16x8, *(const uint4 *) &Bs[(wn * WN + j * 16 + lrow) * kLDK + lk]);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // D[i][j] of a 16x16 block sits in lane j + 16*(i/4), element i%4
    if constexpr (Prob::kTileStore) {
        // bf16 results go through LDS so that global stores are 16-byte
        // row pieces: each lane writes its values into an MT x NT tile,
        // then each thread stores whole 8-column pieces
        uint16_t *T = lds;   // the operand tiles are free after the loop (last barrier above)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = wm * WM + i * 16 + (lane >> 4) * 4 + r, col = wn * WN + j * 16 + lrow;
                    T[row * (NT + 8) + col] = pb.value(n0 + col, acc[i][j][r]);
                }
        __syncthreads();
        for (int t = tid; t < MT * NT / 8; t += 256) {
            const int row = t / (NT / 8), c8 = (t % (NT / 8)) * 8;
            pb.store8(m0 + row, n0 + c8, *(const uint4 *) &T[row * (NT + 8) + c8]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    pb.store(m0 + wm * WM + i * 16 + (lane >> 4) * 4 + r, n0 + wn * WN + j * 16 + lrow, acc[i][j][r]);
    }
}

__device__ inline void put8(uint16_t *dst, uint4 v) { *(uint4 *) dst = v; }

__device__ inline uint4 pack8(const uint16_t *e) {
    return make_uint4(e[0] | (uint32_t) e[1] << 16, e[2] | (uint32_t) e[3] << 16, e[4] | (uint32_t) e[5] << 16,
                      e[6] | (uint32_t) e[7] << 16);
}

// Writes the 8 values of a piece down a column of a [row][kLDK] LDS tile
// (rows r .. r+7, column px): the transposing store of the weight gradient.
__device__ inline void put8_column(uint16_t *T, int r, int px, uint4 v) {
    const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 8; ++q) T[(r + q) * kLDK + px] = (uint16_t) (w4[q >> 1] >> ((q & 1) * 16));
}

// Row-major pieces (row r = t / 4, reduction offset kb = 8 (t % 4)) of a T x 32 tile.
template <int T>
__device__ inline void store_rows(uint16_t *S, const uint4 *r, int tid) {
#pragma unroll
    for (int j = 0; j < pieces<T>(); ++j) {
        const int t = tid + 256 * j;
        if (t < T * 4) put8(&S[(t >> 2) * kLDK + (t & 3) * 8], r[j]);
    }
}

// Forward: out[m = (b, oy, ox)][co] = relu(bias + sum_{tap, ci} x[b][oy*S+ky-1][ox*S+kx-1][ci] w[co][tap][ci]).
template <int L>
struct FwdProb {
    static constexpr int Cin = kCin[L], Cout = kCout[L], S = kStride[L], Hi = kHin[L], Ho = hout(L);
    static constexpr int Kt = ktrue(L), Kp = kpad(L);
    const uint16_t *x;      // [n][Hi][Hi][Cin]
    const uint16_t *w;      // [Cout][Kp]
    const float *bias;      // [Cout]
    uint16_t *y;            // [n][Ho][Ho][Cout]
    int M;                  // n * Ho * Ho
    __device__ void k_range(int, int &b, int &e) const { b = 0; e = Kp; }
    __device__ uint16_t tap_value(int b, int oy, int ox, int k) const {
        if (k >= Kt) return 0;
        const int tap = k / Cin, ci = k - tap * Cin;
        const int iy = oy * S + tap / 3 - 1, ix = ox * S + tap % 3 - 1;
        if (iy < 0 || iy >= Hi || ix < 0 || ix >= Hi) return 0;
        return x[(((size_t) b * Hi + iy) * Hi + ix) * Cin + ci];
    }
    template <int MT>
    __device__ void loadA(uint4 *r, int m0, int k0, int tid) const {
#pragma unroll
        for (int j = 0; j < pieces<MT>(); ++j) {
            const int t = tid + 256 * j, m = m0 + (t >> 2), k = k0 + (t & 3) * 8;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (t < MT * 4 && m < M) {
                const int b = m / (Ho * Ho), rem = m - b * Ho * Ho, oy = rem / Ho, ox = rem - oy * Ho;
                if constexpr (Cin % 8 == 0) {
                    const int tap = k / Cin, ci = k - tap * Cin;
                    const int iy = oy * S + tap / 3 - 1, ix = ox * S + tap % 3 - 1;
                    if (iy >= 0 && iy < Hi && ix >= 0 && ix < Hi)
                        v = *(const uint4 *) &x[(((size_t) b * Hi + iy) * Hi + ix) * Cin + ci];
                } else if constexpr (Cin == 2) {
                    // a piece is 4 taps x both channels: one 4-byte load per tap
                    uint32_t e[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int tap = k / 2 + q;
                        const int iy = oy * S + tap / 3 - 1, ix = ox * S + tap % 3 - 1;
                        e[q] = (tap < 9 && iy >= 0 && iy < Hi && ix >= 0 && ix < Hi)
                                   ? *(const uint32_t *) &x[(((size_t) b * Hi + iy) * Hi + ix) * 2]
                                   : 0u;
                    }
                    v = make_uint4(e[0], e[1], e[2], e[3]);
                } else {
                    uint16_t e[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) e[q] = tap_value(b, oy, ox, k + q);
                    v = pack8(e);
                }
            }
            r[j] = v;
        }
    }
    template <int NT>
    __device__ void loadB(uint4 *r, int n0, int k0, int tid) const {
#pragma unroll
        for (int j = 0; j < pieces<NT>(); ++j) {
            const int t = tid + 256 * j;
            r[j] = t < NT * 4 ? *(const uint4 *) &w[(size_t) (n0 + (t >> 2)) * Kp + k0 + (t & 3) * 8]
                              : make_uint4(0, 0, 0, 0);
        }
    }
    template <int MT>
    __device__ void storeA(uint16_t *As, const uint4 *r, int tid) const { store_rows<MT>(As, r, tid); }
    template <int NT>
    __device__ void storeB(uint16_t *Bs, const uint4 *r, int tid) const { store_rows<NT>(Bs, r, tid); }
    static constexpr bool kTileStore = true;
    __device__ uint16_t value(int n, float v) const { return f2bf(fmaxf(v + bias[n], 0.0f)); }
    __device__ void store8(int m, int n, uint4 v) const {
        if (m < M) *(uint4 *) &y[(size_t) m * Cout + n] = v;
    }
};

// Data gradient: dx[(b, iy, ix)][ci] = sum_{tap, co} dz[b][oy][ox][co] w[co][tap][ci]
// over the output pixels with oy*S + ky - 1 = iy; stored as the previous
// layer's pre-activation gradient dz_prev = dx * (x > 0).  For stride 2 the
// input pixels are split by parity class (PY, PX) = (iy & 1, ix & 1): only
// the taps with ky = 1 (PY = 0) or ky in {0, 2} (PY = 1), and likewise kx,
// reach such a pixel, so each class is its own GEMM over 1, 2, 2 or 4 taps
// instead of 9.  For stride 1 the single class (0, 0) holds every pixel and
// every tap.  Rows m enumerate the class's pixels (b, a, c) with
// iy = S a + PY, ix = S c + PX.
template <int L, int PY, int PX>
struct DataProb {
    static constexpr int Cin = kCin[L], Cout = kCout[L], S = kStride[L], Hi = kHin[L], Ho = hout(L);
    static constexpr int Hc = Hi / S;                                 // class pixels per side
    static constexpr int NY = S == 1 ? 3 : (PY == 0 ? 1 : 2), NX = S == 1 ? 3 : (PX == 0 ? 1 : 2);
    static constexpr int Ktaps = NY * NX;
    const uint16_t *dz;     // [n][Ho][Ho][Cout]
    const uint16_t *wt;     // [9][Cin][Cout]
    const uint16_t *x;      // [n][Hi][Hi][Cin] (this layer's input = previous ReLU output)
    uint16_t *dzp;          // [n][Hi][Hi][Cin]
    int M;                  // n * Hc * Hc
    // the class's t-th tap: ky, kx
    __device__ static int tap_ky(int t) { return S == 1 ? t / 3 : (PY == 0 ? 1 : 2 * (t / NX)); }
    __device__ static int tap_kx(int t) { return S == 1 ? t % 3 : (PX == 0 ? 1 : 2 * (t % NX)); }
    __device__ void k_range(int, int &b, int &e) const { b = 0; e = Ktaps * Cout; }
    __device__ void pixel(int m, int &b, int &iy, int &ix) const {
        b = m / (Hc * Hc);
        const int rem = m - b * Hc * Hc, a = rem / Hc, c = rem - a * Hc;
        iy = S * a + PY;
        ix = S * c + PX;
    }
    template <int MT>
    __device__ void loadA(uint4 *r, int m0, int k0, int tid) const {
        const int t9 = k0 / Cout, co0 = k0 - t9 * Cout;   // a chunk never straddles a tap (Cout % 32 == 0)
        const int ky = tap_ky(t9), kx = tap_kx(t9);
#pragma unroll
        for (int j = 0; j < pieces<MT>(); ++j) {
            const int t = tid + 256 * j, m = m0 + (t >> 2), kb = (t & 3) * 8;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (t < MT * 4 && m < M) {
                int b, iy, ix;
                pixel(m, b, iy, ix);
                const int ty = iy + 1 - ky, tx = ix + 1 - kx;   // divisible by S by the class's choice of taps
                if (ty >= 0 && tx >= 0 && ty / S < Ho && tx / S < Ho)
                    v = *(const uint4 *) &dz[(((size_t) b * Ho + ty / S) * Ho + tx / S) * Cout + co0 + kb];
            }
            r[j] = v;
        }
    }
    template <int NT>
    __device__ void loadB(uint4 *r, int n0, int k0, int tid) const {
        const int t9 = k0 / Cout, co0 = k0 - t9 * Cout;
        const int tap = tap_ky(t9) * 3 + tap_kx(t9);
#pragma unroll
        for (int j = 0; j < pieces<NT>(); ++j) {
            const int t = tid + 256 * j;
            r[j] = t < NT * 4 ? *(const uint4 *) &wt[((size_t) tap * Cin + n0 + (t >> 2)) * Cout + co0 + (t & 3) * 8]
                              : make_uint4(0, 0, 0, 0);
        }
    }
    template <int MT>
    __device__ void storeA(uint16_t *As, const uint4 *r, int tid) const { store_rows<MT>(As, r, tid); }
    template <int NT>
    __device__ void storeB(uint16_t *Bs, const uint4 *r, int tid) const { store_rows<NT>(Bs, r, tid); }
    static constexpr bool kTileStore = true;
    __device__ uint16_t value(int, float v) const { return f2bf(v); }
    __device__ void store8(int m, int n, uint4 v) const {
        if (m < M) {
            int b, iy, ix;
            pixel(m, b, iy, ix);
            const size_t i = (((size_t) b * Hi + iy) * Hi + ix) * Cin + n;
            const uint4 xm = *(const uint4 *) &x[i];
            const uint32_t xs[4] = {xm.x, xm.y, xm.z, xm.w};
            uint32_t ds[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t lo = positive((uint16_t) xs[q]) ? (ds[q] & 0xFFFFu) : 0u;
                const uint32_t hi = positive((uint16_t) (xs[q] >> 16)) ? (ds[q] & 0xFFFF0000u) : 0u;
                ds[q] = lo | hi;
            }
            *(uint4 *) &dzp[i] = make_uint4(ds[0], ds[1], ds[2], ds[3]);
        }
    }
};

// Columns of the weight-gradient GEMM: the 9*Cin weight columns, then one
// column of ones whose product with dz is the bias gradient, padded to a
// multiple of the tile width NT.
constexpr int wgrad_cols(int l, int nt) { return (ktrue(l) + 1 + nt - 1) / nt * nt; }

// Weight gradient: dw[co][k = (tap, ci)] = sum_p dz[p][co] x[shift_tap(p)][ci]
// for k < 9 Cin, and db[co] = sum_p dz[p][co] in column 9 Cin (B holds 1
// there); the pixel sum split into slices of `slice` pixels (blockIdx.z),
// partial sums in part[z][co][Kw].
template <int L, int NT>
struct WeightProb {
    static constexpr int Cin = kCin[L], Cout = kCout[L], S = kStride[L], Hi = kHin[L], Ho = hout(L);
    static constexpr int Kt = ktrue(L), Kw = wgrad_cols(L, NT);
    const uint16_t *dz;     // [P][Cout]
    const uint16_t *x;      // [n][Hi][Hi][Cin]
    float *part;            // [slices][Cout][Kw]
    int P;                  // n * Ho * Ho
    int slice;              // pixels per slice, multiple of 32
    __device__ void k_range(int z, int &b, int &e) const {
        b = z * slice;
        e = min(P, b + slice);
    }
    // piece t of a T x 32 tile: pixel px = t % 32 of the chunk, rows o .. o+7 with o = 8 (t / 32)
    template <int MT>
    __device__ void loadA(uint4 *r, int m0, int k0, int tid) const {
#pragma unroll
        for (int j = 0; j < pieces<MT>(); ++j) {
            const int t = tid + 256 * j, p = k0 + t % 32, o = (t / 32) * 8;
            r[j] = (t < MT * 4 && p < P) ? *(const uint4 *) &dz[(size_t) p * Cout + m0 + o] : make_uint4(0, 0, 0, 0);
        }
    }
    template <int NT_>
    __device__ void loadB(uint4 *r, int n0, int k0, int tid) const {
        constexpr uint16_t kOne = 0x3F80;   // bf16 1.0
#pragma unroll
        for (int j = 0; j < pieces<NT_>(); ++j) {
            const int t = tid + 256 * j, p = k0 + t % 32, n = n0 + (t / 32) * 8;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (t < NT_ * 4 && p < P) {
                const int b = p / (Ho * Ho), rem = p - b * Ho * Ho, oy = rem / Ho, ox = rem - oy * Ho;
                if constexpr (Cin % 8 == 0) {
                    if (n < Kt) {
                        const int tap = n / Cin, ci = n - tap * Cin;
                        const int iy = oy * S + tap / 3 - 1, ix = ox * S + tap % 3 - 1;
                        if (iy >= 0 && iy < Hi && ix >= 0 && ix < Hi)
                            v = *(const uint4 *) &x[(((size_t) b * Hi + iy) * Hi + ix) * Cin + ci];
                    } else if (n == Kt) {
                        v.x = kOne;
                    }
                } else {
                    uint16_t e[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int k = n + q;
                        e[q] = k == Kt ? kOne : (uint16_t) 0;
                        if (k < Kt) {
                            const int tap = k / Cin, ci = k - tap * Cin;
                            const int iy = oy * S + tap / 3 - 1, ix = ox * S + tap % 3 - 1;
                            if (iy >= 0 && iy < Hi && ix >= 0 && ix < Hi)
                                e[q] = x[(((size_t) b * Hi + iy) * Hi + ix) * Cin + ci];
                        }
                    }
                    v = pack8(e);
                }
            }
            r[j] = v;
        }
    }
    template <int MT>
    __device__ void storeA(uint16_t *As, const uint4 *r, int tid) const {
#pragma unroll
        for (int j = 0; j < pieces<MT>(); ++j) {
            const int t = tid + 256 * j;
            if (t < MT * 4) put8_column(As, (t / 32) * 8, t % 32, r[j]);
        }
    }
    template <int NT_>
    __device__ void storeB(uint16_t *Bs, const uint4 *r, int tid) const {
#pragma unroll
        for (int j = 0; j < pieces<NT_>(); ++j) {
            const int t = tid + 256 * j;
            if (t < NT_ * 4) put8_column(Bs, (t / 32) * 8, t % 32, r[j]);
        }
    }
    static constexpr bool kTileStore = false;
    __device__ void store(int m, int n, float v) const {
        part[((size_t) blockIdx.z * Cout + m) * Kw + n] = v;
    }
};

// Adds the slices' partial sums: output o = (co, k) for k <= Kt; 64 outputs
// per workgroup, 4 lanes per output each adding every 4th slice, the lanes
// then combined in a fixed order (deterministic).  k < Kt goes to the weight
// gradient [co][Kt], k == Kt to the bias gradient.
__global__ __launch_bounds__(256) void reduce_slices_kernel(const float *part, int slices, int cout, int kw, int kt,
                                                           float *grad_w, float *grad_b) {
    __shared__ float red[256];
    const int tid = threadIdx.x, lane = tid >> 6, o = blockIdx.x * 64 + (tid & 63);
    const int outs = cout * (kt + 1);
    float s = 0.0f;
    int co = 0, k = 0;
    if (o < outs) {
        co = o / (kt + 1);
        k = o - co * (kt + 1);
        for (int z = lane; z < slices; z += 4) s += part[((size_t) z * cout + co) * kw + k];
    }
    red[tid] = s;
    __syncthreads();
    if (tid < 64 && o < outs) {
        const float t = (red[tid] + red[tid + 64]) + (red[tid + 128] + red[tid + 192]);
        if (k < kt) grad_w[(size_t) co * kt + k] = t;
        else grad_b[co] = t;
    }
}

// ------------------------------------------------------------- small passes

// Input preparation: per pair, each image reduced to 128 x 128 (2x2 box
// average for 256 x 256, as is for 128 x 128), standardised to zero mean and
// unit variance (binary64 moments; a flat image maps to 0), rounded to bf16
// and interleaved as channel 0 = moving, 1 = fixed.  One workgroup per
// (pair, image).
template <typename T>
__device__ float prep_pixel(const T *img, int side, int y, int x) {
    if (side == kIn) return (float) img[y * side + x];
    const int r = 2 * y, c = 2 * x;
    const float a = (float) img[r * side + c], b = (float) img[r * side + c + 1];
    const float d = (float) img[(r + 1) * side + c], e = (float) img[(r + 1) * side + c + 1];
    return ((a + b) + (d + e)) * 0.25f;
}

struct PrepImage {
    const void *data;
    int side;
    int u8;
};

__global__ __launch_bounds__(256) void prepare_kernel(const PrepImage *imgs, uint16_t *x) {
    const int pair = blockIdx.x >> 1, ch = blockIdx.x & 1;
    const PrepImage im = imgs[blockIdx.x];
    __shared__ double red[2][256];
    double s = 0.0, ss = 0.0;
    for (int i = threadIdx.x; i < kIn * kIn; i += 256) {
        const int y = i / kIn, xx = i % kIn;
        const float v = im.u8 ? prep_pixel((const uint8_t *) im.data, im.side, y, xx)
                              : prep_pixel((const float *) im.data, im.side, y, xx);
        s += v;
        ss += (double) v * v;
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = ss;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    const double n = (double) kIn * kIn;
    const double mean = red[0][0] / n;
    const double var = red[1][0] / n - mean * mean;
    const double inv = var > 1e-12 ? 1.0 / sqrt(var) : 0.0;
    for (int i = threadIdx.x; i < kIn * kIn; i += 256) {
        const int y = i / kIn, xx = i % kIn;
        const float v = im.u8 ? prep_pixel((const uint8_t *) im.data, im.side, y, xx)
                              : prep_pixel((const float *) im.data, im.side, y, xx);
        x[((size_t) pair * kIn * kIn + i) * 2 + ch] = f2bf((float) (((double) v - mean) * inv));
    }
}

// Global average pooling of the last activation + FC: one workgroup per sample.
__global__ __launch_bounds__(kFeat) void gap_fc_kernel(const uint16_t *y6, const float *fcw, const float *fcb,
                                                      float *feat, float *out) {
    const int b = blockIdx.x, c = threadIdx.x;
    __shared__ float f[kFeat];
    float s = 0.0f;
    for (int p = 0; p < kLastPix; ++p) s += bf2f(y6[((size_t) b * kLastPix + p) * kFeat + c]);
    s *= 1.0f / kLastPix;
    f[c] = s;
    feat[(size_t) b * kFeat + c] = s;
    __syncthreads();
    if (c < kOut) {
        float o = fcb[c];
        for (int q = 0; q < kFeat; ++q) o = fmaf(fcw[c * kFeat + q], f[q], o);
        out[(size_t) b * kOut + c] = o;
    }
}

// Mean squared error over n x 8 outputs and its gradient.  One workgroup.
// loss = sum (out - target)^2 / (8n); dout = 2 (out - target) / (8n).
__global__ __launch_bounds__(256) void loss_kernel(const float *out, const float *target, int n, float *dout,
                                                  float *loss) {
    __shared__ float red[256];
    const float scale = 1.0f / (kOut * (float) n);
    float s = 0.0f;
    for (int i = threadIdx.x; i < n * kOut; i += 256) {
        const float d = out[i] - target[i];
        s += d * d;
        dout[i] = 2.0f * d * scale;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] * scale;
}

// FC gradients: threads [0, 1024) the weight gradient sum_b dout[b][j] feat[b][c],
// [1024, 1032) the bias gradient, then one thread per (b, c) of the pooled
// features' gradient sum_j dout[b][j] fcw[j][c].
__global__ __launch_bounds__(256) void fc_grad_kernel(const float *dout, const float *feat, const float *fcw, int n,
                                                     float *grad_fcw, float *grad_fcb, float *dfeat) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    constexpr int W = kOut * kFeat;
    if (i < W) {
        const int j = i / kFeat, c = i % kFeat;
        float g = 0.0f;
        for (int b = 0; b < n; ++b) g = fmaf(dout[b * kOut + j], feat[(size_t) b * kFeat + c], g);
        grad_fcw[i] = g;
    } else if (i < W + kOut) {
        float g = 0.0f;
        for (int b = 0; b < n; ++b) g += dout[b * kOut + (i - W)];
        grad_fcb[i - W] = g;
    } else if (i < W + kOut + n * kFeat) {
        const int k = i - W - kOut, b = k / kFeat, c = k % kFeat;
        float g = 0.0f;
        for (int j = 0; j < kOut; ++j) g = fmaf(dout[b * kOut + j], fcw[j * kFeat + c], g);
        dfeat[k] = g;
    }
}

// Gradient at the last conv's pre-activation: dfeat / 256 where y6 > 0.
__global__ void gap_backward_kernel(const uint16_t *y6, const float *dfeat, int n, uint16_t *dz6) {
    const size_t total = (size_t) n * kLastPix * kFeat;
    for (size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t) gridDim.x * blockDim.x) {
        const size_t b = i / ((size_t) kLastPix * kFeat);
        const int c = (int) (i % kFeat);
        dz6[i] = positive(y6[i]) ? f2bf(dfeat[b * kFeat + c] * (1.0f / kLastPix)) : (uint16_t) 0;
    }
}

// Adam with bias correction on the flat FP32 parameter vector.
__global__ void adam_kernel(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// Master FP32 weights -> bf16 working copies (padded [Cout][Kp] and the
// transposed [9][Cin][Cout]), FC weights rounded to bf16 (kept as float).
struct PackArgs {
    uint16_t *w[kLayers];
    uint16_t *wt[kLayers];
    float *fcw;
};

__global__ void pack_kernel(const float *master, PackArgs a) {
    const int l = blockIdx.y;
    if (l == kLayers) {
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kOut * kFeat; i += gridDim.x * blockDim.x)
            a.fcw[i] = bf2f(f2bf(master[kOff.fcw + i]));
        return;
    }
    const int cin = kCin[l], cout = kCout[l], kt = 9 * cin, kp = (kt + 31) / 32 * 32;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cout * kp; i += gridDim.x * blockDim.x) {
        const int co = i / kp, k = i % kp;
        const uint16_t h = k < kt ? f2bf(master[kOff.w[l] + (int64_t) co * kt + k]) : (uint16_t) 0;
        a.w[l][i] = h;
        if (k < kt) {
            const int tap = k / cin, ci = k % cin;
            a.wt[l][((size_t) tap * cin + ci) * cout + co] = h;
        }
    }
}

}  // namespace cnn
}  // namespace cgs

using namespace cgs;
using namespace cgs::cnn;

struct cgs_cnn_s {
    int max_batch = 0;
    int cus = 0;
    int adam_t = 0;
    int last_n = 0;                 // batch of the last forward (0: none)
    const uint16_t *last_x = nullptr;
    DevBuf master, grad, adam_m, adam_v, fcw, w[kLayers], wt[kLayers];
    DevBuf act[kLayers];            // outputs of the conv layers
    DevBuf dz[2], part, feat, out, dout, dfeat, loss, loss_host;
    std::vector<float> host_master;
    ~cgs_cnn_s() {
        auto rel = [](DevBuf &b, bool host) {
            if (b.ptr) (void) (host ? hipHostFree(b.ptr) : hipFree(b.ptr));
            b.ptr = nullptr;
        };
        rel(master, false); rel(grad, false); rel(adam_m, false); rel(adam_v, false); rel(fcw, false);
        for (int l = 0; l < kLayers; ++l) { rel(w[l], false); rel(wt[l], false); rel(act[l], false); }
        rel(dz[0], false); rel(dz[1], false); rel(part, false); rel(feat, false);
        rel(out, false); rel(dout, false); rel(dfeat, false); rel(loss, false); rel(loss_host, true);
    }
};

namespace {

// split-K slices of the weight gradient: about 2 waves of workgroups per CU
int wgrad_slices(int l, int n, int cus, int tiles) {
    const int P = n * hout(l) * hout(l);
    int z = std::max(1, (2 * cus + tiles - 1) / tiles);
    z = std::min(z, std::max(1, P / 256));
    return z;
}

int wgrad_slice_len(int l, int n, int z) {
    const int P = n * hout(l) * hout(l);
    return ((P + z - 1) / z + 31) / 32 * 32;
}

constexpr int fwd_mt(int l) { return kCout[l] >= 128 ? 64 : kCout[l] == 64 ? 64 : 128; }
constexpr int fwd_nt(int l) { return kCout[l]; }
constexpr int dat_mt(int l) { return kCin[l] >= 64 ? 64 : 128; }
constexpr int dat_nt(int l) { return kCin[l]; }
constexpr int wgt_mt(int l) { return kCout[l] >= 128 ? 128 : kCout[l]; }
constexpr int wgt_nt(int l) { return l == 0 ? 32 : 64; }

template <int L>
hipError_t launch_fwd(const uint16_t *x, const uint16_t *w, const float *bias, uint16_t *y, int n, hipStream_t s) {
    constexpr int MT = fwd_mt(L), NT = fwd_nt(L);
    FwdProb<L> pb{x, w, bias, y, n * hout(L) * hout(L)};
    hipLaunchKernelGGL((igemm_kernel<FwdProb<L>, MT, NT>), dim3((pb.M + MT - 1) / MT, kCout[L] / NT, 1), dim3(256),
                       0, s, pb);
    return hipGetLastError();
}

template <int L, int PY, int PX>
hipError_t launch_data_class(const uint16_t *dz, const uint16_t *wt, const uint16_t *x, uint16_t *dzp, int n,
                             hipStream_t s) {
    constexpr int MT = dat_mt(L), NT = dat_nt(L);
    using Prob = DataProb<L, PY, PX>;
    Prob pb{dz, wt, x, dzp, n * Prob::Hc * Prob::Hc};
    hipLaunchKernelGGL((igemm_kernel<Prob, MT, NT>), dim3((pb.M + MT - 1) / MT, kCin[L] / NT, 1), dim3(256), 0, s, pb);
    return hipGetLastError();
}

template <int L>
hipError_t launch_data(const uint16_t *dz, const uint16_t *wt, const uint16_t *x, uint16_t *dzp, int n,
                       hipStream_t s) {
    hipError_t e = launch_data_class<L, 0, 0>(dz, wt, x, dzp, n, s);
    if constexpr (kStride[L] == 2) {
        if (e == hipSuccess) e = launch_data_class<L, 0, 1>(dz, wt, x, dzp, n, s);
        if (e == hipSuccess) e = launch_data_class<L, 1, 0>(dz, wt, x, dzp, n, s);
        if (e == hipSuccess) e = launch_data_class<L, 1, 1>(dz, wt, x, dzp, n, s);
    }
    return e;
}

template <int L>
hipError_t launch_weight(const uint16_t *dz, const uint16_t *x, float *part, float *grad_w, float *grad_b, int n,
                         int cus, hipStream_t s) {
    constexpr int MT = wgt_mt(L), NT = wgt_nt(L), Kw = wgrad_cols(L, NT);
    const int tiles = (kCout[L] / MT) * (Kw / NT);
    const int z = wgrad_slices(L, n, cus, tiles);
    const int P = n * hout(L) * hout(L);
    WeightProb<L, NT> pb{dz, x, part, P, wgrad_slice_len(L, n, z)};
    hipLaunchKernelGGL((igemm_kernel<WeightProb<L, NT>, MT, NT>), dim3(kCout[L] / MT, Kw / NT, z), dim3(256), 0, s,
                       pb);
    const int outs = kCout[L] * (ktrue(L) + 1);
    hipLaunchKernelGGL(reduce_slices_kernel, dim3((outs + 63) / 64), dim3(256), 0, s, (const float *) part, z,
                       kCout[L], Kw, ktrue(L), grad_w, grad_b);
    return hipGetLastError();
}

size_t act_elems(int l, int n) { return (size_t) n * hout(l) * hout(l) * kCout[l]; }

size_t part_floats(int n, int cus) {
    size_t mx = 0;
    auto one = [&](int l, int mt, int nt) {
        const int kw = (ktrue(l) + 1 + nt - 1) / nt * nt;
        const int tiles = (kCout[l] / mt) * (kw / nt);
        mx = std::max(mx, (size_t) wgrad_slices(l, n, cus, tiles) * kCout[l] * kw);
    };
    one(0, wgt_mt(0), wgt_nt(0)); one(1, wgt_mt(1), wgt_nt(1)); one(2, wgt_mt(2), wgt_nt(2));
    one(3, wgt_mt(3), wgt_nt(3)); one(4, wgt_mt(4), wgt_nt(4)); one(5, wgt_mt(5), wgt_nt(5));
    return mx;
}

cgs_status pack(cgs_cnn *c, hipStream_t s) {
    PackArgs a;
    for (int l = 0; l < kLayers; ++l) {
        a.w[l] = (uint16_t *) c->w[l].ptr;
        a.wt[l] = (uint16_t *) c->wt[l].ptr;
    }
    a.fcw = (float *) c->fcw.ptr;
    hipLaunchKernelGGL(pack_kernel, dim3(64, kLayers + 1), dim3(256), 0, s, (const float *) c->master.ptr, a);
    HIP_TRY(hipGetLastError());
    return CGS_OK;
}

cgs_status check_batch(const cgs_cnn *c, int32_t n) {
    if (!c) return set_error(CGS_ERR_INVALID_ARGUMENT, "null network");
    if (n < 1 || n > c->max_batch) return set_error(CGS_ERR_TOO_LARGE, "batch outside 1..max_batch");
    return CGS_OK;
}

}  // namespace

extern "C" {

int64_t cgs_cnn_param_count(void) { return kOff.total; }

cgs_status cgs_cnn_create(int32_t max_batch, cgs_cnn **out) {
    clear_error();
    if (!out || max_batch < 1 || max_batch > 4096)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "cnn_create: max_batch must be 1..4096");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return set_error(CGS_ERR_NO_DEVICE, "no HIP device");
    cgs_cnn *c = new cgs_cnn_s();
    c->max_batch = max_batch;
    HIP_TRY(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t np = (size_t) kOff.total;
    cgs_status st = CGS_OK;
    auto need = [&](DevBuf &b, size_t bytes, bool host = false) {
        if (st == CGS_OK) st = b.ensure(bytes, host);
    };
    need(c->master, np * 4); need(c->grad, np * 4); need(c->adam_m, np * 4); need(c->adam_v, np * 4);
    need(c->fcw, kOut * kFeat * 4);
    size_t dzmax = 0;
    for (int l = 0; l < kLayers; ++l) {
        need(c->w[l], (size_t) kCout[l] * kpad(l) * 2);
        need(c->wt[l], (size_t) 9 * kCin[l] * kCout[l] * 2);
        need(c->act[l], act_elems(l, max_batch) * 2);
        dzmax = std::max(dzmax, act_elems(l, max_batch));
    }
    need(c->dz[0], dzmax * 2); need(c->dz[1], dzmax * 2);
    need(c->part, part_floats(max_batch, c->cus) * 4);
    need(c->feat, (size_t) max_batch * kFeat * 4);
    need(c->out, (size_t) max_batch * kOut * 4);
    need(c->dout, (size_t) max_batch * kOut * 4);
    need(c->dfeat, (size_t) max_batch * kFeat * 4);
    need(c->loss, 16); need(c->loss_host, 16, true);
    if (st != CGS_OK) { delete c; return st; }
    if (hipMemset(c->master.ptr, 0, np * 4) != hipSuccess || hipMemset(c->adam_m.ptr, 0, np * 4) != hipSuccess ||
        hipMemset(c->adam_v.ptr, 0, np * 4) != hipSuccess) {
        delete c;
        return set_error(CGS_ERR_HIP, "cnn_create: clearing parameters failed");
    }
    if ((st = pack(c, nullptr)) != CGS_OK || hipDeviceSynchronize() != hipSuccess) {
        delete c;
        return st != CGS_OK ? st : set_error(CGS_ERR_HIP, "cnn_create: packing failed");
    }
    *out = c;
    return CGS_OK;
}

cgs_status cgs_cnn_destroy(cgs_cnn *c) {
    delete c;
    return CGS_OK;
}

cgs_status cgs_cnn_set_params(cgs_cnn *c, const float *params) {
    clear_error();
    if (!c || !params) return set_error(CGS_ERR_INVALID_ARGUMENT, "cnn_set_params: null argument");
    HIP_TRY(hipMemcpy(c->master.ptr, params, (size_t) kOff.total * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(c->adam_m.ptr, 0, (size_t) kOff.total * 4));
    HIP_TRY(hipMemset(c->adam_v.ptr, 0, (size_t) kOff.total * 4));
    c->adam_t = 0;
    c->last_n = 0;
    cgs_status st = pack(c, nullptr);
    if (st != CGS_OK) return st;
    HIP_TRY(hipDeviceSynchronize());
    return CGS_OK;
}

cgs_status cgs_cnn_get_params(const cgs_cnn *c, float *params) {
    clear_error();
    if (!c || !params) return set_error(CGS_ERR_INVALID_ARGUMENT, "cnn_get_params: null argument");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(params, c->master.ptr, (size_t) kOff.total * 4, hipMemcpyDeviceToHost));
    return CGS_OK;
}

cgs_status cgs_cnn_get_grads(const cgs_cnn *c, float *grads) {
    clear_error();
    if (!c || !grads) return set_error(CGS_ERR_INVALID_ARGUMENT, "cnn_get_grads: null argument");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(grads, c->grad.ptr, (size_t) kOff.total * 4, hipMemcpyDeviceToHost));
    return CGS_OK;
}

cgs_status cgs_cnn_prepare(const cgs_image *moved, const cgs_image *fixed, int32_t n, uint16_t *input, void *stream) {
    clear_error();
    if (!moved || !fixed || !input || n < 1) return set_error(CGS_ERR_INVALID_ARGUMENT, "cnn_prepare: bad arguments");
    std::vector<PrepImage> h(2 * (size_t) n);
    for (int32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 2; ++k) {
            const cgs_image &im = k == 0 ? moved[i] : fixed[i];
            cgs_status st = validate_image(&im, k == 0 ? "moved" : "fixed");
            if (st != CGS_OK) return st;
            if (im.channels != 1 || im.width != im.height || (im.width != kIn && im.width != 2 * kIn))
                return set_error(CGS_ERR_UNSUPPORTED, "cnn_prepare: images must be one-channel 128x128 or 256x256");
            h[2 * i + k] = PrepImage{im.data, im.width, im.pixtype == CGS_PIX_U8 ? 1 : 0};
        }
    }
    Workspace *ws = nullptr;
    cgs_status st = workspace_for_current_device(&ws);
    if (st != CGS_OK) return st;
    std::lock_guard<std::mutex> lock(ws->mu);
    hipStream_t s = (hipStream_t) stream;
    if ((st = ws->prep.ensure(h.size() * sizeof(PrepImage))) != CGS_OK ||
        (st = ws->prep_host.ensure(h.size() * sizeof(PrepImage), true)) != CGS_OK)
        return st;
    // the pinned staging copy is reused by the next call: wait for this one
    std::memcpy(ws->prep_host.ptr, h.data(), h.size() * sizeof(PrepImage));
    HIP_TRY(hipMemcpyAsync(ws->prep.ptr, ws->prep_host.ptr, h.size() * sizeof(PrepImage), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(prepare_kernel, dim3(2 * n), dim3(256), 0, s, (const PrepImage *) ws->prep.ptr, input);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return CGS_OK;
}

cgs_status cgs_cnn_forward(cgs_cnn *c, const uint16_t *input, int32_t n, float *offsets, void *stream) {
    clear_error();
    cgs_status st = check_batch(c, n);
    if (st != CGS_OK) return st;
    if (!input) return set_error(CGS_ERR_INVALID_ARGUMENT, "cnn_forward: null input");
    hipStream_t s = (hipStream_t) stream;
    auto A = [&](int l) { return (uint16_t *) c->act[l].ptr; };
    auto W = [&](int l) { return (const uint16_t *) c->w[l].ptr; };
    const float *mp = (const float *) c->master.ptr;
    HIP_TRY(launch_fwd<0>(input, W(0), mp + kOff.b[0], A(0), n, s));
    HIP_TRY(launch_fwd<1>(A(0), W(1), mp + kOff.b[1], A(1), n, s));
    HIP_TRY(launch_fwd<2>(A(1), W(2), mp + kOff.b[2], A(2), n, s));
    HIP_TRY(launch_fwd<3>(A(2), W(3), mp + kOff.b[3], A(3), n, s));
    HIP_TRY(launch_fwd<4>(A(3), W(4), mp + kOff.b[4], A(4), n, s));
    HIP_TRY(launch_fwd<5>(A(4), W(5), mp + kOff.b[5], A(5), n, s));
    hipLaunchKernelGGL(gap_fc_kernel, dim3(n), dim3(kFeat), 0, s, (const uint16_t *) A(5), (const float *) c->fcw.ptr,
                       mp + kOff.fcb, (float *) c->feat.ptr, (float *) c->out.ptr);
    HIP_TRY(hipGetLastError());
    if (offsets)
        HIP_TRY(hipMemcpyAsync(offsets, c->out.ptr, (size_t) n * kOut * 4, hipMemcpyDeviceToDevice, s));
    c->last_n = n;
    c->last_x = input;
    return CGS_OK;
}

cgs_status cgs_cnn_backward(cgs_cnn *c, const float *target, float *loss_out, void *stream) {
    clear_error();
    if (!c || !target) return set_error(CGS_ERR_INVALID_ARGUMENT, "cnn_backward: null argument");
    if (c->last_n < 1) return set_error(CGS_ERR_INVALID_ARGUMENT, "cnn_backward: no forward pass to differentiate");
    const int n = c->last_n;
    hipStream_t s = (hipStream_t) stream;
    float *g = (float *) c->grad.ptr;
    auto A = [&](int l) { return (const uint16_t *) c->act[l].ptr; };
    auto WT = [&](int l) { return (const uint16_t *) c->wt[l].ptr; };
    uint16_t *dz[2] = {(uint16_t *) c->dz[0].ptr, (uint16_t *) c->dz[1].ptr};
    float *part = (float *) c->part.ptr;
    hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(256), 0, s, (const float *) c->out.ptr, target, n,
                       (float *) c->dout.ptr, (float *) c->loss.ptr);
    const int fc_threads = kOut * kFeat + kOut + n * kFeat;
    hipLaunchKernelGGL(fc_grad_kernel, dim3((fc_threads + 255) / 256), dim3(256), 0, s, (const float *) c->dout.ptr,
                       (const float *) c->feat.ptr, (const float *) c->fcw.ptr, n, g + kOff.fcw, g + kOff.fcb,
                       (float *) c->dfeat.ptr);
    hipLaunchKernelGGL(gap_backward_kernel, dim3(512), dim3(256), 0, s, A(5), (const float *) c->dfeat.ptr, n, dz[1]);
    HIP_TRY(hipGetLastError());
    // dz of layer l lives in dz[l & 1]
    HIP_TRY(launch_weight<5>(dz[1], A(4), part, g + kOff.w[5], g + kOff.b[5], n, c->cus, s));
    HIP_TRY(launch_data<5>(dz[1], WT(5), A(4), dz[0], n, s));
    HIP_TRY(launch_weight<4>(dz[0], A(3), part, g + kOff.w[4], g + kOff.b[4], n, c->cus, s));
    HIP_TRY(launch_data<4>(dz[0], WT(4), A(3), dz[1], n, s));
    HIP_TRY(launch_weight<3>(dz[1], A(2), part, g + kOff.w[3], g + kOff.b[3], n, c->cus, s));
    HIP_TRY(launch_data<3>(dz[1], WT(3), A(2), dz[0], n, s));
    HIP_TRY(launch_weight<2>(dz[0], A(1), part, g + kOff.w[2], g + kOff.b[2], n, c->cus, s));
    HIP_TRY(launch_data<2>(dz[0], WT(2), A(1), dz[1], n, s));
    HIP_TRY(launch_weight<1>(dz[1], A(0), part, g + kOff.w[1], g + kOff.b[1], n, c->cus, s));
    HIP_TRY(launch_data<1>(dz[1], WT(1), A(0), dz[0], n, s));
    HIP_TRY(launch_weight<0>(dz[0], c->last_x, part, g + kOff.w[0], g + kOff.b[0], n, c->cus, s));
    if (loss_out) {
        HIP_TRY(hipMemcpyAsync(c->loss_host.ptr, c->loss.ptr, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        *loss_out = *(const float *) c->loss_host.ptr;
    }
    return CGS_OK;
}

cgs_status cgs_cnn_adam_step(cgs_cnn *c, float lr, float beta1, float beta2, float eps, void *stream) {
    clear_error();
    if (!c) return set_error(CGS_ERR_INVALID_ARGUMENT, "null network");
    if (!(lr >= 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return set_error(CGS_ERR_INVALID_ARGUMENT, "cnn_adam_step: bad hyper-parameters");
    hipStream_t s = (hipStream_t) stream;
    c->adam_t += 1;
    const float c1 = (float) (1.0 - std::pow((double) beta1, c->adam_t));
    const float c2 = (float) (1.0 - std::pow((double) beta2, c->adam_t));
    hipLaunchKernelGGL(adam_kernel, dim3(512), dim3(256), 0, s, (float *) c->master.ptr, (const float *) c->grad.ptr,
                       (float *) c->adam_m.ptr, (float *) c->adam_v.ptr, kOff.total, lr, beta1, beta2, eps, c1, c2);
    HIP_TRY(hipGetLastError());
    return pack(c, s);
}

}  // extern "C"
