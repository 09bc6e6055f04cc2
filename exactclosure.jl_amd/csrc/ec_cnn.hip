// Learned closure: a 4-layer 3D CNN, 3 -> 32 -> 32 -> 32 -> 3 channels,
// 3x3x3 kernels, periodic padding, GELU after the first three layers.  It maps
// the (scaled) filtered LES velocity to the (scaled) exact closure forcing.
// The reference has no learned part; this network is specified by the
// workload only (BF16 weights and activations, FP32 accumulation, Adam).
//
// Every convolution is an implicit GEMM on gfx950 MFMA
// (v_mfma_f32_32x32x16_bf16):
//   forward   Y[p][co] = sum_t sum_ci X[p + o_t][ci] W[t][co][ci]       M = points, N = co, K = ci per tap
//   dgrad     dX[q][ci] = sum_t sum_co dY[q + o_t][co] W[26 - t][co][ci] the forward kernel with flipped weights
//   wgrad     dW[t][co][ci] = sum_p dY[p][co] X[p + o_t][ci]          M = co, N = ci, K = points
// with o_t = (dx, dy, dz), t = 9 (dz + 1) + 3 (dy + 1) + (dx + 1).
// Activations are channels-last BF16, [sample][z][y][x][channel] with x
// fastest among the spatial axes (the column-major order of the fields).
// Weight gradients are reduced deterministically (per-block partials, then a
// fixed-order sum).
#include "ec_internal.h"

#include <cmath>

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

namespace {

constexpr int NL = 4;
constexpr int CH = 32;                       // hidden channels = MFMA N and M width
constexpr int CIN[NL] = {3, CH, CH, CH};     // real input channels
constexpr int COUT[NL] = {CH, CH, CH, 3};    // real output channels
constexpr int CINP[NL] = {16, CH, CH, CH};   // padded input channels (MFMA K granularity)
constexpr int CX0 = 16;                      // channels of the packed input
constexpr int CD4 = 16;                      // channels of the output gradient

__device__ __forceinline__ int wrapn(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__device__ __forceinline__ uint16_t bf16_bits(float x) {
    const __bf16 b = (__bf16)x;  // round to nearest even
    return __builtin_bit_cast(uint16_t, b);
}
__device__ __forceinline__ float bf16_float(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }

// ---------------------------------------------------------------------------
// forward / dgrad convolution
// tile 8 x 8 x 8 points (512 threads = 8 waves, wave w owns z = w, two
// M-blocks of 8 x 4 points); halo tile 10 x 10 x 10 in LDS, all weights in LDS
// (146 KB for 32 channels: one block of eight waves per CU, the weights loaded
// once per 512 points)
constexpr int FX = 8, FY = 8, FZ = 8, FHX = 10, FHY = 10, FHZ = 10, FHP = FHX * FHY * FHZ;
constexpr int FT = 64 * FZ;  // threads per block: one wave per z-plane of the tile

enum { CONV_HIDDEN = 0, CONV_OUT = 1, CONV_DGRAD = 2 };

template <int CP, int MODE>
__global__ void __launch_bounds__(FT) k_conv(const uint16_t *__restrict__ X, const uint16_t *__restrict__ W,
                                              const float *__restrict__ bias, int n, int tiles_per_sample,
                                              uint16_t *__restrict__ out_a, uint16_t *__restrict__ out_b,
                                              float *__restrict__ out_f, const uint16_t *__restrict__ pre_in) {
    constexpr int SP = CP + 8;  // LDS row stride in bf16: 16-byte rows, odd multiple of 16 bytes
    extern __shared__ uint16_t smem_cnn[];
    uint16_t *xs = smem_cnn;            // [FHP][SP]
    uint16_t *ws = smem_cnn + FHP * SP; // [27 * 32][SP]
    const int tid = threadIdx.x;
    const int total = gridDim.x;
    const int L = (int)ec_xcd_tile(blockIdx.x, total);
    const int s = L / tiles_per_sample, tl = L - s * tiles_per_sample;
    const int gx = n / FX, gy = n / FY;
    const int x0 = (tl % gx) * FX, y0 = ((tl / gx) % gy) * FY, z0 = (tl / (gx * gy)) * FZ;
    const size_t n3 = (size_t)n * n * n, sbase = (size_t)s * n3;

    constexpr int CH8 = CP / 8;  // 16-byte chunks per point
    for (int i = tid; i < 27 * CH * CH8; i += FT) {
        const int row = i / CH8, part = i - row * CH8;
        *(uint4 *)(ws + row * SP + part * 8) = ((const uint4 *)W)[i];
    }
    for (int i = tid; i < FHP * CH8; i += FT) {
        const int pt = i / CH8, part = i - pt * CH8;
        const int hx = pt % FHX, hy = (pt / FHX) % FHY, hz = pt / (FHX * FHY);
        const size_t g = (size_t)wrapn(x0 + hx - 1, n) + (size_t)n * ((size_t)wrapn(y0 + hy - 1, n) +
                                                                     (size_t)n * (size_t)wrapn(z0 + hz - 1, n));
        *(uint4 *)(xs + pt * SP + part * 8) = *(const uint4 *)(X + (sbase + g) * CP + part * 8);
    }
    __syncthreads();

    const int w = tid >> 6, l = tid & 63, m = l & 31, kh = l >> 5;
    f32x16 acc0 = {}, acc1 = {};
    for (int t = 0; t < 27; t++) {
        const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
        const int hx = (m & 7) + 1 + dx, hz = w + 1 + dz;
        const int p0 = ((hz * FHY + ((m >> 3) + 1 + dy)) * FHX + hx) * SP;      // M-block 0: y = m / 8
        const int p1 = ((hz * FHY + ((m >> 3) + 5 + dy)) * FHX + hx) * SP;      // M-block 1: y = 4 + m / 8
#pragma unroll
        for (int kc = 0; kc < CP / 16; kc++) {
            const int ko = kc * 16 + kh * 8;
            const bf16x8 b = *(const bf16x8 *)(ws + (t * CH + m) * SP + ko);
            const bf16x8 a0 = *(const bf16x8 *)(xs + p0 + ko);
            const bf16x8 a1 = *(const bf16x8 *)(xs + p1 + ko);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, acc1, 0, 0, 0);
        }
    }

    // D layout: lane l, register r -> point row i = 8 (r / 4) + 4 kh + r % 4, channel m
    const int co = m;
    const float bco = (MODE == CONV_DGRAD) ? 0.0f : bias[co];
    if constexpr (MODE == CONV_OUT) {
#pragma unroll
        for (int mb = 0; mb < 2; mb++) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int i = 8 * (r >> 2) + 4 * kh + (r & 3);
                const int x = x0 + (i & 7), y = y0 + 4 * mb + (i >> 3), z = z0 + w;
                const size_t p = sbase + (size_t)x + (size_t)n * ((size_t)y + (size_t)n * (size_t)z);
                if (co < 3) out_f[p * 4 + co] = (mb ? acc1[r] : acc0[r]) + bco;
            }
        }
        return;
    }
    // BF16 outputs go through LDS: the wave's 64 points x 32 channels (4 KB,
    // a wave-private slice of the halo tile, free once every wave is past the
    // taps) are written in the D layout and stored as 16-byte chunks, four per
    // lane (64 contiguous bytes per point, eight points per row of the tile)
    // instead of 32 two-byte stores per lane and output.
    static_assert(FZ * 64 * CH <= FHP * (16 + 8), "staging slices fit the halo tile");
    uint16_t *stg = xs + w * (64 * CH);
    auto chunk_addr = [&](int c) {  // chunk c of the wave: point q = c / 4, 8 channels from 8 (c % 4)
        const int q = c >> 2;
        const int x = x0 + (q & 7), y = y0 + (q >> 3), z = z0 + w;
        return (sbase + (size_t)x + (size_t)n * ((size_t)y + (size_t)n * (size_t)z)) * CH + (size_t)(c & 3) * 8;
    };
    auto flush = [&](uint16_t *dst) {  // stg -> global, after a barrier
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int c = l + 64 * k;
            *(uint4 *)(dst + chunk_addr(c)) = *(const uint4 *)(stg + c * 8);
        }
    };
    __syncthreads();  // every wave is past the taps: the halo tile is free
    if constexpr (MODE == CONV_DGRAD) {
        // the saved pre-activations of the same points, staged the same way
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int c = l + 64 * k;
            *(uint4 *)(stg + c * 8) = *(const uint4 *)(pre_in + chunk_addr(c));
        }
        __syncthreads();
    }
    float v[2][16];
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int q = 32 * mb + 8 * (r >> 2) + 4 * kh + (r & 3);
            v[mb][r] = (mb ? acc1[r] : acc0[r]) + bco;
            if constexpr (MODE == CONV_DGRAD) v[mb][r] *= gelu_grad(bf16_float(stg[q * CH + co]));
        }
    if constexpr (MODE == CONV_DGRAD) __syncthreads();
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
        for (int r = 0; r < 16; r++) stg[(32 * mb + 8 * (r >> 2) + 4 * kh + (r & 3)) * CH + co] = bf16_bits(v[mb][r]);
    __syncthreads();
    flush(out_a);  // pre-activation (hidden) or input gradient (dgrad)
    if constexpr (MODE == CONV_HIDDEN) {
        __syncthreads();
#pragma unroll
        for (int mb = 0; mb < 2; mb++)
#pragma unroll
            for (int r = 0; r < 16; r++)
                stg[(32 * mb + 8 * (r >> 2) + 4 * kh + (r & 3)) * CH + co] = bf16_bits(gelu(v[mb][r]));
        __syncthreads();
        flush(out_b);  // activation
    }
}

// ---------------------------------------------------------------------------
// weight gradient: tile 16 x 8 x 4 points; X halo (18 x 10 x 6, rows padded to
// 24) and dY stored channel-major in LDS so eight consecutive x-points of one
// channel are one 16-byte read.  Wave w accumulates taps w, w + 4, ...
constexpr int WX = 16, WY = 8, WZ = 4, WP = WX * WY * WZ, WHX = 24, WHY = 10, WHZ = 6, WHP = WHX * WHY * WHZ;
// channel strides in LDS: one 16-byte pad per channel row, so that the 32
// lanes of an MFMA operand read (one channel each) fall on distinct banks
// (unpadded, the dY rows are 1024 bytes apart and every lane hits the same banks)
constexpr int WPS = WP + 8, WHPS = WHP + 8;

// eight bf16 starting at element offset q8 * 8 + rr of a 16-byte aligned row
template <int RR>
__device__ __forceinline__ bf16x8 ld8_shift(const uint16_t *row, int q8) {
    const uint4 a = *(const uint4 *)(row + q8 * 8);
    if constexpr (RR == 0) {
        return __builtin_bit_cast(bf16x8, a);
    } else {
        const uint4 b = *(const uint4 *)(row + q8 * 8 + 8);
        uint4 o;
        if constexpr (RR == 1) {
            o.x = __builtin_amdgcn_alignbyte(a.y, a.x, 2);
            o.y = __builtin_amdgcn_alignbyte(a.z, a.y, 2);
            o.z = __builtin_amdgcn_alignbyte(a.w, a.z, 2);
            o.w = __builtin_amdgcn_alignbyte(b.x, a.w, 2);
        } else {  // RR == 2: one dword
            o.x = a.y, o.y = a.z, o.z = a.w, o.w = b.x;
        }
        return __builtin_bit_cast(bf16x8, o);
    }
}

template <int CP, int CY, int DX>
__device__ __forceinline__ void wgrad_tap(const uint16_t *xs, const uint16_t *ys, int m, int kh, int dy, int dz,
                                          f32x16 &acc) {
    const bool aok = m < CY, bok = m < CP;
    const uint16_t *yrow = ys + (aok ? m : 0) * WPS;
    const uint16_t *xch = xs + (bok ? m : 0) * WHPS;
    const bf16x8 zero = {};
#pragma unroll 4
    for (int kc = 0; kc < WY * WZ; kc++) {
        const int y = kc % WY, z = kc / WY;
        bf16x8 a = *(const bf16x8 *)(yrow + (z * WY + y) * WX + kh * 8);
        const uint16_t *xr = xch + ((z + 1 + dz) * WHY + (y + 1 + dy)) * WHX;
        bf16x8 b = ld8_shift<(DX + 1) & 7>(xr, kh + ((DX + 1) >> 3));
        if (!aok) a = zero;
        if (!bok) b = zero;
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
}

template <int CP, int CY>
__global__ void __launch_bounds__(256) k_wgrad(const uint16_t *__restrict__ X, const uint16_t *__restrict__ dY,
                                               int n, int tiles_per_sample, int tiles_total,
                                               float *__restrict__ wpart) {
    extern __shared__ uint16_t smem_cnn[];
    uint16_t *xs = smem_cnn;             // [CP][WHZ][WHY][WHX]
    uint16_t *ys = smem_cnn + CP * WHPS;  // [CY][WZ][WY][WX], channel stride WPS
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, m = l & 31, kh = l >> 5;
    const int gx = n / WX, gy = n / WY;
    const size_t n3 = (size_t)n * n * n;
    f32x16 acc[7];
#pragma unroll
    for (int j = 0; j < 7; j++) acc[j] = f32x16{};

    for (int tile = blockIdx.x; tile < tiles_total; tile += gridDim.x) {
        const int s = tile / tiles_per_sample, tl = tile - s * tiles_per_sample;
        const int x0 = (tl % gx) * WX, y0 = ((tl / gx) % gy) * WY, z0 = (tl / (gx * gy)) * WZ;
        const size_t sbase = (size_t)s * n3;
        __syncthreads();  // the previous tile's reads are done
        for (int i = tid; i < WHZ * WHY * 18 * (CP / 8); i += 256) {
            const int pt = i / (CP / 8), part = i - pt * (CP / 8);
            const int hx = pt % 18, hy = (pt / 18) % WHY, hz = pt / (18 * WHY);
            const size_t g = (size_t)wrapn(x0 + hx - 1, n) + (size_t)n * ((size_t)wrapn(y0 + hy - 1, n) +
                                                                         (size_t)n * (size_t)wrapn(z0 + hz - 1, n));
            const uint4 v = *(const uint4 *)(X + (sbase + g) * CP + part * 8);
            const uint16_t *e = (const uint16_t *)&v;
            const int o = (hz * WHY + hy) * WHX + hx;
#pragma unroll
            for (int c = 0; c < 8; c++) xs[(part * 8 + c) * WHPS + o] = e[c];
        }
        for (int i = tid; i < WP * (CY / 8); i += 256) {
            const int pt = i / (CY / 8), part = i - pt * (CY / 8);
            const int x = pt % WX, y = (pt / WX) % WY, z = pt / (WX * WY);
            const size_t g = (size_t)(x0 + x) + (size_t)n * ((size_t)(y0 + y) + (size_t)n * (size_t)(z0 + z));
            const uint4 v = *(const uint4 *)(dY + (sbase + g) * CY + part * 8);
            const uint16_t *e = (const uint16_t *)&v;
#pragma unroll
            for (int c = 0; c < 8; c++) ys[(part * 8 + c) * WPS + pt] = e[c];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 7; j++) {
            const int t = w + 4 * j;
            if (t < 27) {
                const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
                if (dx == -1) wgrad_tap<CP, CY, -1>(xs, ys, m, kh, dy, dz, acc[j]);
                else if (dx == 0) wgrad_tap<CP, CY, 0>(xs, ys, m, kh, dy, dz, acc[j]);
                else wgrad_tap<CP, CY, 1>(xs, ys, m, kh, dy, dz, acc[j]);
            }
        }
    }
    // partial dW of this block: [27][32 co][32 ci]
    float *dst = wpart + (size_t)blockIdx.x * 27 * CH * CH;
#pragma unroll
    for (int j = 0; j < 7; j++) {
        const int t = w + 4 * j;
        if (t < 27) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int co = 8 * (r >> 2) + 4 * kh + (r & 3);
                dst[(t * CH + co) * CH + m] = acc[j][r];
            }
        }
    }
}

// bias gradient partials: db[c] = sum_p dY[p][c]; a thread sums one 16-byte
// chunk (8 channels) of every point it visits
template <int CY>
__global__ void __launch_bounds__(256) k_bias_grad(const uint16_t *__restrict__ dY, size_t npts,
                                                   float *__restrict__ bpart) {
    constexpr int NQ = CY / 8, PP = 256 / NQ;  // chunks per point, points per block pass
    __shared__ float red[PP][CY + 1];
    const int tid = threadIdx.x, q = tid % NQ, lp = tid / NQ;
    float acc[8] = {};
    for (size_t p = (size_t)blockIdx.x * PP + lp; p < npts; p += (size_t)gridDim.x * PP) {
        const uint4 v = *(const uint4 *)(dY + p * CY + q * 8);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            acc[2 * k] += __uint_as_float(w[k] << 16);
            acc[2 * k + 1] += __uint_as_float(w[k] & 0xffff0000u);
        }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) red[lp][q * 8 + k] = acc[k];
    __syncthreads();
    if (tid < CY) {
        float s = 0.0f;
        for (int k = 0; k < PP; k++) s += red[k][tid];
        bpart[blockIdx.x * CH + tid] = s;
    }
}

// fixed-order sum of the partials into the FP32 gradient of one layer
__global__ void k_reduce_grad(const float *__restrict__ wpart, int nw, const float *__restrict__ bpart, int nbk,
                              int cin, int cout, float *__restrict__ gW, float *__restrict__ gb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nwgt = 27 * cout * cin;
    if (i < nwgt) {
        const int t = i / (cout * cin), r = i - t * cout * cin, co = r / cin, ci = r - co * cin;
        const size_t o = ((size_t)t * CH + co) * CH + ci;
        float s = 0.0f;
        for (int b = 0; b < nw; b++) s += wpart[(size_t)b * 27 * CH * CH + o];
        gW[i] = s;
    } else if (i < nwgt + cout) {
        const int co = i - nwgt;
        float s = 0.0f;
        for (int b = 0; b < nbk; b++) s += bpart[b * CH + co];
        gb[co] = s;
    }
}

// FP32 master parameters -> BF16 forward and flipped (dgrad) weights, FP32 padded bias
__global__ void k_make_weights(const float *__restrict__ Wm, const float *__restrict__ bm, int cin, int cout,
                               int cinp, int coutp, uint16_t *__restrict__ wf, uint16_t *__restrict__ wd,
                               float *__restrict__ b32) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // forward: [27][32 co][cinp ci]
    if (i < 27 * CH * cinp) {
        const int t = i / (CH * cinp), r = i - t * CH * cinp, co = r / cinp, ci = r - co * cinp;
        const float v = (co < cout && ci < cin) ? Wm[(t * cout + co) * cin + ci] : 0.0f;
        wf[i] = bf16_bits(v);
    }
    // dgrad: [27][32 ci][coutp co] = W[26 - t][co][ci]
    if (wd && i < 27 * CH * coutp) {
        const int t = i / (CH * coutp), r = i - t * CH * coutp, ci = r / coutp, co = r - ci * coutp;
        const float v = (co < cout && ci < cin) ? Wm[((26 - t) * cout + co) * cin + ci] : 0.0f;
        wd[i] = bf16_bits(v);
    }
    if (i < CH) b32[i] = i < cout ? bm[i] : 0.0f;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, int np, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// three FP64 column-major fields of one sample -> channels 0..2 of the packed BF16 input
__global__ void k_pack(const double *__restrict__ u1, const double *__restrict__ u2, const double *__restrict__ u3,
                       size_t n3, double scale, uint16_t *__restrict__ x) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n3) return;
    uint16_t v[CX0] = {};
    v[0] = bf16_bits((float)(u1[i] * scale));
    v[1] = bf16_bits((float)(u2[i] * scale));
    v[2] = bf16_bits((float)(u3[i] * scale));
    uint4 *d = (uint4 *)(x + i * CX0);
    d[0] = *(const uint4 *)&v[0];
    d[1] = *(const uint4 *)&v[8];
}

// prediction (scaled units) of one sample -> three FP64 fields times inv_scale
__global__ void k_unpack(const float *__restrict__ out, size_t n3, double inv_scale, double *__restrict__ m1,
                         double *__restrict__ m2, double *__restrict__ m3) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n3) return;
    m1[i] = out[i * 4 + 0] * inv_scale;
    m2[i] = out[i * 4 + 1] * inv_scale;
    m3[i] = out[i * 4 + 2] * inv_scale;
}

// mean-squared-error gradient of one sample (BF16, 16 channels) and FP64 partial sums
__global__ void __launch_bounds__(256) k_loss(const float *__restrict__ out, const double *__restrict__ t1,
                                              const double *__restrict__ t2, const double *__restrict__ t3,
                                              size_t n3, double scale, double gscale, uint16_t *__restrict__ dout,
                                              double *__restrict__ part) {
    __shared__ double red[256];
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (i < n3) {
        const double t[3] = {t1[i] * scale, t2[i] * scale, t3[i] * scale};
        uint16_t v[CD4] = {};
        for (int c = 0; c < 3; c++) {
            const double d = (double)out[i * 4 + c] - t[c];
            acc += d * d;
            v[c] = bf16_bits((float)(2.0 * d * gscale));
        }
        uint4 *dst = (uint4 *)(dout + i * CD4);
        dst[0] = *(const uint4 *)&v[0];
        dst[1] = *(const uint4 *)&v[8];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// deterministic host-side layer bookkeeping
struct Layer {
    int cin, cout, cinp, coutp, woff, boff;
};

uint64_t sm64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

}  // namespace

struct ec_cnn {
    Layer L[NL];
    int np = 0, step = 0;
    float *param = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t *wf[NL] = {}, *wd[NL] = {};
    float *b32[NL] = {};
    // activations, sized for cap_pts points (all samples of a batch)
    size_t cap_pts = 0;
    int n = 0, batch = 0;
    uint16_t *x0 = nullptr, *pre[3] = {}, *act[3] = {}, *dout = nullptr, *dpre[2] = {};
    float *out = nullptr;
    float *wpart = nullptr, *bpart = nullptr;
    double *lpart = nullptr, *lpart_h = nullptr;
    int wblocks = 0, bblocks = 0;
};

static int cnn_free_acts(ec_cnn *c) {
    void *bufs[] = {c->x0, c->pre[0], c->pre[1], c->pre[2], c->act[0], c->act[1], c->act[2],
                    c->dout, c->dpre[0], c->dpre[1], c->out};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    c->x0 = c->dout = nullptr;
    for (int i = 0; i < 3; i++) c->pre[i] = c->act[i] = nullptr;
    c->dpre[0] = c->dpre[1] = nullptr;
    c->out = nullptr;
    c->cap_pts = 0;
    return EC_OK;
}

static int cnn_reserve(ec_ctx *ctx, ec_cnn *c, size_t pts) {
    if (pts <= c->cap_pts) return EC_OK;
    EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    cnn_free_acts(c);
    EC_HIP(ctx, hipMalloc(&c->x0, pts * CX0 * 2));
    for (int i = 0; i < 3; i++) {
        EC_HIP(ctx, hipMalloc(&c->pre[i], pts * CH * 2));
        EC_HIP(ctx, hipMalloc(&c->act[i], pts * CH * 2));
    }
    EC_HIP(ctx, hipMalloc(&c->dout, pts * CD4 * 2));
    EC_HIP(ctx, hipMalloc(&c->dpre[0], pts * CH * 2));
    EC_HIP(ctx, hipMalloc(&c->dpre[1], pts * CH * 2));
    EC_HIP(ctx, hipMalloc(&c->out, pts * 4 * sizeof(float)));
    c->cap_pts = pts;
    return EC_OK;
}

static int cnn_upload_weights(ec_ctx *ctx, ec_cnn *c) {
    for (int l = 0; l < NL; l++) {
        const Layer &y = c->L[l];
        const int cnt = 27 * CH * std::max(y.cinp, y.coutp);
        k_make_weights<<<(cnt + 255) / 256, 256, 0, ctx->stream>>>(c->param + y.woff, c->param + y.boff, y.cin,
                                                                    y.cout, y.cinp, y.coutp, c->wf[l],
                                                                    l > 0 ? c->wd[l] : nullptr, c->b32[l]);
        EC_CHECK_LAUNCH(ctx, "k_make_weights");
    }
    return EC_OK;
}

extern "C" int ec_cnn_create(ec_ctx *ctx, uint64_t seed, ec_cnn **net) {
    EC_REQUIRE(ctx, ctx && net, "ec_cnn_create: null argument");
    *net = nullptr;
    ec_cnn *c = new ec_cnn();
    int o = 0;
    for (int l = 0; l < NL; l++) {
        Layer &y = c->L[l];
        y.cin = CIN[l], y.cout = COUT[l], y.cinp = CINP[l], y.coutp = l == NL - 1 ? CD4 : CH;
        y.woff = o;
        o += 27 * y.cout * y.cin;
        y.boff = o;
        o += y.cout;
    }
    c->np = o;
    // default 3D convolution initialisation: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and bias
    std::vector<float> h(o);
    for (int l = 0; l < NL; l++) {
        const Layer &y = c->L[l];
        const double bound = 1.0 / std::sqrt(27.0 * y.cin);
        for (int i = y.woff; i < y.boff + y.cout; i++) {
            const double u = (double)(sm64(seed ^ sm64((uint64_t)i)) >> 11) * 0x1.0p-53;
            h[i] = (float)((2.0 * u - 1.0) * bound);
        }
    }
    auto fail = [&](int rc) {
        ec_cnn_destroy(ctx, c);
        return rc;
    };
    if (hipMalloc(&c->param, o * 4) || hipMalloc(&c->grad, o * 4) || hipMalloc(&c->m1, o * 4) ||
        hipMalloc(&c->m2, o * 4))
        return fail(ec_fail(ctx, EC_E_OOM, "ec_cnn_create: out of device memory"));
    for (int l = 0; l < NL; l++) {
        const Layer &y = c->L[l];
        if (hipMalloc(&c->wf[l], 27 * CH * y.cinp * 2) || (l > 0 && hipMalloc(&c->wd[l], 27 * CH * y.coutp * 2)) ||
            hipMalloc(&c->b32[l], CH * 4))
            return fail(ec_fail(ctx, EC_E_OOM, "ec_cnn_create: out of device memory"));
    }
    c->wblocks = 512;
    c->bblocks = 1024;
    if (hipMalloc(&c->wpart, (size_t)c->wblocks * 27 * CH * CH * 4) || hipMalloc(&c->bpart, c->bblocks * CH * 4) ||
        hipHostMalloc(&c->lpart_h, sizeof(double) * 65536))
        return fail(ec_fail(ctx, EC_E_OOM, "ec_cnn_create: out of memory"));
    c->lpart = nullptr;
    if (hipMemcpyAsync(c->param, h.data(), o * 4, hipMemcpyHostToDevice, ctx->stream) ||
        hipMemsetAsync(c->m1, 0, o * 4, ctx->stream) || hipMemsetAsync(c->m2, 0, o * 4, ctx->stream) ||
        hipMemsetAsync(c->grad, 0, o * 4, ctx->stream) || hipStreamSynchronize(ctx->stream))
        return fail(ec_fail(ctx, EC_E_HIP, "ec_cnn_create: upload failed"));
    int rc = cnn_upload_weights(ctx, c);
    if (rc) return fail(rc);
    *net = c;
    return EC_OK;
}

extern "C" int ec_cnn_destroy(ec_ctx *ctx, ec_cnn *c) {
    if (!c) return EC_OK;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    cnn_free_acts(c);
    void *bufs[] = {c->param, c->grad, c->m1, c->m2, c->wpart, c->bpart, c->lpart};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (int l = 0; l < NL; l++) {
        if (c->wf[l]) (void)hipFree(c->wf[l]);
        if (c->wd[l]) (void)hipFree(c->wd[l]);
        if (c->b32[l]) (void)hipFree(c->b32[l]);
    }
    if (c->lpart_h) (void)hipHostFree(c->lpart_h);
    delete c;
    return EC_OK;
}

extern "C" int ec_cnn_num_params(ec_cnn *c) { return c ? c->np : 0; }

extern "C" int ec_cnn_params(ec_ctx *ctx, ec_cnn *c, int which, int get, float *host, int count) {
    EC_REQUIRE(ctx, ctx && c && host, "ec_cnn_params: null argument");
    EC_REQUIRE(ctx, count == c->np, "ec_cnn_params: count %d, the network has %d parameters", count, c->np);
    EC_REQUIRE(ctx, which == 0 || which == 1, "ec_cnn_params: which %d (0 parameters, 1 gradients)", which);
    float *d = which == 0 ? c->param : c->grad;
    if (get) {
        EC_HIP(ctx, hipMemcpyAsync(host, d, count * 4, hipMemcpyDeviceToHost, ctx->stream));
        EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        EC_HIP(ctx, hipMemcpyAsync(d, host, count * 4, hipMemcpyHostToDevice, ctx->stream));
        EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (which == 0) return cnn_upload_weights(ctx, c);
    }
    return EC_OK;
}

template <int CP, int MODE>
static int launch_conv(ec_ctx *ctx, const uint16_t *X, const uint16_t *W, const float *bias, int n, int batch,
                       uint16_t *oa, uint16_t *ob, float *of, const uint16_t *pre_in) {
    const int tps = (n / FX) * (n / FY) * (n / FZ);
    const size_t lds = (size_t)(FHP + 27 * CH) * (CP + 8) * 2;
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_conv<CP, MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    k_conv<CP, MODE><<<tps * batch, FT, lds, ctx->stream>>>(X, W, bias, n, tps, oa, ob, of, pre_in);
    EC_CHECK_LAUNCH(ctx, "k_conv");
    return EC_OK;
}

template <int CP, int CY>
static int launch_wgrad(ec_ctx *ctx, ec_cnn *c, const uint16_t *X, const uint16_t *dY, int n, int batch, float *gW,
                        float *gb, int cin, int cout) {
    const int tps = (n / WX) * (n / WY) * (n / WZ), total = tps * batch;
    const int blocks = std::min(total, c->wblocks);
    const size_t lds = (size_t)(CP * WHPS + CY * WPS) * 2;
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_wgrad<CP, CY>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    k_wgrad<CP, CY><<<blocks, 256, lds, ctx->stream>>>(X, dY, n, tps, total, c->wpart);
    EC_CHECK_LAUNCH(ctx, "k_wgrad");
    const size_t npts = (size_t)batch * ec_cube(n);
    const int bb = (int)std::min<size_t>((size_t)c->bblocks, (npts + 255) / 256);
    k_bias_grad<CY><<<bb, 256, 0, ctx->stream>>>(dY, npts, c->bpart);
    EC_CHECK_LAUNCH(ctx, "k_bias_grad");
    const int cnt = 27 * cout * cin + cout;
    k_reduce_grad<<<(cnt + 255) / 256, 256, 0, ctx->stream>>>(c->wpart, blocks, c->bpart, bb, cin, cout, gW, gb);
    EC_CHECK_LAUNCH(ctx, "k_reduce_grad");
    return EC_OK;
}

static int cnn_check_shape(ec_ctx *ctx, int n, int batch) {
    EC_REQUIRE(ctx, n >= 16 && n % 16 == 0, "CNN grids need n a multiple of 16 (got %d)", n);
    EC_REQUIRE(ctx, batch >= 1, "CNN batch %d", batch);
    EC_REQUIRE(ctx, (size_t)batch * ec_cube(n) / 256 < (size_t)0x7fffffff, "CNN batch too large");
    return EC_OK;
}

extern "C" int ec_cnn_forward(ec_ctx *ctx, ec_cnn *c, int n, int batch, const void *const *u, double in_scale,
                              void *const *out, double out_scale) {
    EC_REQUIRE(ctx, ctx && c && u, "ec_cnn_forward: null argument");
    int rc = cnn_check_shape(ctx, n, batch);
    if (rc) return rc;
    for (int i = 0; i < 3 * batch; i++) EC_REQUIRE(ctx, u[i], "ec_cnn_forward: null input field %d", i);
    const size_t n3 = ec_cube(n), pts = (size_t)batch * n3;
    if ((rc = cnn_reserve(ctx, c, pts))) return rc;
    c->n = n;
    c->batch = batch;
    ec_prof_scope pr(ctx, "cnn_forward", 0.0);
    for (int s = 0; s < batch; s++) {
        k_pack<<<(unsigned)((n3 + 255) / 256), 256, 0, ctx->stream>>>(
            (const double *)u[3 * s], (const double *)u[3 * s + 1], (const double *)u[3 * s + 2], n3, in_scale,
            c->x0 + s * n3 * CX0);
        EC_CHECK_LAUNCH(ctx, "k_pack");
    }
    if ((rc = launch_conv<16, CONV_HIDDEN>(ctx, c->x0, c->wf[0], c->b32[0], n, batch, c->pre[0], c->act[0], nullptr,
                                           nullptr)))
        return rc;
    for (int l = 1; l < 3; l++)
        if ((rc = launch_conv<CH, CONV_HIDDEN>(ctx, c->act[l - 1], c->wf[l], c->b32[l], n, batch, c->pre[l],
                                               c->act[l], nullptr, nullptr)))
            return rc;
    if ((rc = launch_conv<CH, CONV_OUT>(ctx, c->act[2], c->wf[3], c->b32[3], n, batch, nullptr, nullptr, c->out,
                                        nullptr)))
        return rc;
    if (out) {
        for (int s = 0; s < batch; s++) {
            EC_REQUIRE(ctx, out[3 * s] && out[3 * s + 1] && out[3 * s + 2], "ec_cnn_forward: null output");
            k_unpack<<<(unsigned)((n3 + 255) / 256), 256, 0, ctx->stream>>>(
                c->out + s * n3 * 4, n3, 1.0 / out_scale, (double *)out[3 * s], (double *)out[3 * s + 1],
                (double *)out[3 * s + 2]);
            EC_CHECK_LAUNCH(ctx, "k_unpack");
        }
    }
    return EC_OK;
}

extern "C" int ec_cnn_backward(ec_ctx *ctx, ec_cnn *c, const void *const *target, double tgt_scale, double *loss) {
    EC_REQUIRE(ctx, ctx && c && target, "ec_cnn_backward: null argument");
    EC_REQUIRE(ctx, c->n > 0 && c->cap_pts > 0, "ec_cnn_backward: call ec_cnn_forward first");
    const int n = c->n, batch = c->batch;
    const size_t n3 = ec_cube(n), pts = (size_t)batch * n3;
    for (int i = 0; i < 3 * batch; i++) EC_REQUIRE(ctx, target[i], "ec_cnn_backward: null target %d", i);
    ec_prof_scope pr(ctx, "cnn_backward", 0.0);
    const unsigned lb = (unsigned)((n3 + 255) / 256);
    EC_REQUIRE(ctx, (size_t)lb * batch <= 65536, "ec_cnn_backward: loss partial buffer too small");
    if (!c->lpart) EC_HIP(ctx, hipMalloc(&c->lpart, sizeof(double) * 65536));
    const double gscale = 1.0 / (3.0 * (double)pts);
    for (int s = 0; s < batch; s++) {
        k_loss<<<lb, 256, 0, ctx->stream>>>(c->out + s * n3 * 4, (const double *)target[3 * s],
                                            (const double *)target[3 * s + 1], (const double *)target[3 * s + 2],
                                            n3, tgt_scale, gscale, c->dout + s * n3 * CD4, c->lpart + s * lb);
        EC_CHECK_LAUNCH(ctx, "k_loss");
    }
    int rc;
    const Layer *L = c->L;
    float *g = c->grad;
    // layer 4: dW from (act3, dout); dpre3 = dgrad(dout) * gelu'(pre3)
    if ((rc = launch_wgrad<CH, CD4>(ctx, c, c->act[2], c->dout, n, batch, g + L[3].woff, g + L[3].boff, L[3].cin,
                                    L[3].cout)))
        return rc;
    if ((rc = launch_conv<CD4, CONV_DGRAD>(ctx, c->dout, c->wd[3], nullptr, n, batch, c->dpre[0], nullptr, nullptr,
                                           c->pre[2])))
        return rc;
    // layer 3
    if ((rc = launch_wgrad<CH, CH>(ctx, c, c->act[1], c->dpre[0], n, batch, g + L[2].woff, g + L[2].boff, L[2].cin,
                                   L[2].cout)))
        return rc;
    if ((rc = launch_conv<CH, CONV_DGRAD>(ctx, c->dpre[0], c->wd[2], nullptr, n, batch, c->dpre[1], nullptr, nullptr,
                                          c->pre[1])))
        return rc;
    // layer 2
    if ((rc = launch_wgrad<CH, CH>(ctx, c, c->act[0], c->dpre[1], n, batch, g + L[1].woff, g + L[1].boff, L[1].cin,
                                   L[1].cout)))
        return rc;
    if ((rc = launch_conv<CH, CONV_DGRAD>(ctx, c->dpre[1], c->wd[1], nullptr, n, batch, c->dpre[0], nullptr, nullptr,
                                          c->pre[0])))
        return rc;
    // layer 1 (no input gradient)
    if ((rc = launch_wgrad<CX0, CH>(ctx, c, c->x0, c->dpre[0], n, batch, g + L[0].woff, g + L[0].boff, L[0].cin,
                                    L[0].cout)))
        return rc;
    if (loss) {
        const size_t nl = (size_t)lb * batch;
        EC_HIP(ctx, hipMemcpyAsync(c->lpart_h, c->lpart, nl * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double s = 0.0;
        for (size_t i = 0; i < nl; i++) s += c->lpart_h[i];
        *loss = s / (3.0 * (double)pts);
    }
    return EC_OK;
}

extern "C" int ec_cnn_adam(ec_ctx *ctx, ec_cnn *c, double lr, double beta1, double beta2, double eps) {
    EC_REQUIRE(ctx, ctx && c, "ec_cnn_adam: null argument");
    EC_REQUIRE(ctx, lr > 0 && beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 && eps > 0, "ec_cnn_adam: bad hyperparameters");
    c->step++;
    const float c1 = (float)(1.0 - std::pow(beta1, c->step)), c2 = (float)(1.0 - std::pow(beta2, c->step));
    k_adam<<<(c->np + 255) / 256, 256, 0, ctx->stream>>>(c->param, c->grad, c->m1, c->m2, c->np, (float)lr,
                                                         (float)beta1, (float)beta2, (float)eps, c1, c2);
    EC_CHECK_LAUNCH(ctx, "k_adam");
    return cnn_upload_weights(ctx, c);
}
