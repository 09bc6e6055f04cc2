// Hand-written 3D real FFT and the FFT Poisson projection (poissonsolver /
// project!, src/N.jl:202-307) for gfx950.
//
// Layout follows FFTW's rfft of a column-major array: the half spectrum is
// (nh = n/2+1, n, n) complex with kx fastest.  Five passes per projection:
//   X  r2c  : two real x-lines packed as one complex line, LDS Stockham FFT,
//             untangled into two half spectra (contiguous reads and writes)
//   Y  fwd  : tiles of TK consecutive kx columns, whole y extent staged in LDS
//   Z  fwd + Poisson scale + zero mode + Z inverse, one LDS residency
//   Y  inv
//   X  c2r  : Hermitian rebuild of two lines, one complex inverse, split
// Twiddles and sin^2 eigenvalue tables are computed on the host in FP64.
#include "ec_internal.h"

#include <algorithm>
#include <cmath>

namespace {

struct RadixPlan {
    int nrad;
    int r[16];
};

constexpr int FFT_E = 16;  // maximum complex elements per thread per line

// ---- radix butterflies: in-place DFT of size R with kernel exp(sign*2*pi*i/R) ----
template <typename T>
__device__ __forceinline__ cplx<T> mul_i(cplx<T> z, int sign) {  // (sign*i) * z
    return sign > 0 ? cplx<T>{-z.y, z.x} : cplx<T>{z.y, -z.x};
}

template <typename T, int R>
struct Bfly;

template <typename T>
struct Bfly<T, 2> {
    __device__ static void run(cplx<T> *v, int) {
        cplx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename T>
struct Bfly<T, 4> {
    __device__ static void run(cplx<T> *v, int sign) {
        cplx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        cplx<T> b0 = cadd(v[1], v[3]), b1 = mul_i(csub(v[1], v[3]), sign);
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[3] = csub(a1, b1);
    }
};

template <typename T>
struct Bfly<T, 8> {
    __device__ static void run(cplx<T> *v, int sign) {
        cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
        Bfly<T, 4>::run(e, sign);
        Bfly<T, 4>::run(o, sign);
        const T c = (T)0.70710678118654752440;
        // o[k] *= exp(sign*2*pi*i*k/8)
        o[1] = cplx<T>{c * (o[1].x - sign * o[1].y), c * (o[1].y + sign * o[1].x)};
        o[2] = mul_i(o[2], sign);
        o[3] = cplx<T>{-c * (o[3].x + sign * o[3].y), c * (sign * o[3].x - o[3].y)};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
};

template <typename T>
struct Bfly<T, 16> {
    __device__ static void run(cplx<T> *v, int sign) {
        cplx<T> e[8], o[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            e[k] = v[2 * k];
            o[k] = v[2 * k + 1];
        }
        Bfly<T, 8>::run(e, sign);
        Bfly<T, 8>::run(o, sign);
        // o[k] *= exp(sign*2*pi*i*k/16): (cos, sin) of k*pi/8
        const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, c2 = (T)0.70710678118654752440;
        const T cs[8] = {T(1), c1, c2, s1, T(0), -s1, -c2, -c1};
        const T sn[8] = {T(0), s1, c2, c1, T(1), c1, c2, s1};
#pragma unroll
        for (int k = 1; k < 8; k++) {
            const T sk = sign > 0 ? sn[k] : -sn[k];
            o[k] = cplx<T>{o[k].x * cs[k] - o[k].y * sk, o[k].y * cs[k] + o[k].x * sk};
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            v[k] = cadd(e[k], o[k]);
            v[k + 8] = csub(e[k], o[k]);
        }
    }
};

template <typename T, int R>
struct BflyGeneric {
    __device__ static void run(cplx<T> *v, int sign) {
        // cos / sin of 2*pi*m/R for m < R
        double cs[R], sn[R];
#pragma unroll
        for (int m = 0; m < R; m++) {
            cs[m] = R == 3 ? (m == 0 ? 1.0 : -0.5)
                           : (R == 5 ? (m == 0 ? 1.0 : (m == 1 || m == 4) ? 0.30901699437494742410 : -0.80901699437494742410)
                                     : 0.0);
            sn[m] = R == 3 ? (m == 0 ? 0.0 : (m == 1 ? 0.86602540378443864676 : -0.86602540378443864676))
                           : (R == 5 ? (m == 0 ? 0.0
                                               : m == 1 ? 0.95105651629515357212
                                               : m == 2 ? 0.58778525229247312917
                                               : m == 3 ? -0.58778525229247312917
                                                        : -0.95105651629515357212)
                                     : 0.0);
        }
        cplx<T> out[R];
#pragma unroll
        for (int q = 0; q < R; q++) {
            cplx<T> acc = v[0];
#pragma unroll
            for (int r = 1; r < R; r++) {
                const int m = (r * q) % R;
                const cplx<T> w{(T)cs[m], (T)(sign * sn[m])};
                acc = cadd(acc, cmul(v[r], w));
            }
            out[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < R; q++) v[q] = out[q];
    }
};
template <typename T>
struct Bfly<T, 3> : BflyGeneric<T, 3> {};
template <typename T>
struct Bfly<T, 5> : BflyGeneric<T, 5> {};

// One Stockham stage over a line held in LDS at lds[e * ls], e in [0, n).
template <typename T, int R, int MAXB>
__device__ __forceinline__ void fft_stage(cplx<T> *lds, int ls, int n, int t, int nthr, int Ns, int sign,
                                          const cplx<T> *__restrict__ tw) {
    const int nR = n / R;
    const int tstep = n / (Ns * R);
    cplx<T> v[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int b = t + q * nthr;
        if (b < nR) {
            const int jm = b % Ns;
#pragma unroll
            for (int r = 0; r < R; r++) {
                cplx<T> x = lds[(b + r * nR) * ls];
                if (r > 0 && jm > 0) {
                    cplx<T> w = tw[r * jm * tstep];
                    if (sign > 0) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[q][r] = x;
            }
            Bfly<T, R>::run(v[q], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        const int b = t + q * nthr;
        if (b < nR) {
            const int jm = b % Ns;
            const int idxD = (b - jm) * R + jm;
#pragma unroll
            for (int r = 0; r < R; r++) lds[(idxD + r * Ns) * ls] = v[q][r];
        }
    }
    __syncthreads();
}

// Full line FFT; every thread of the block must call it (it contains barriers).
template <typename T>
__device__ void lds_fft(cplx<T> *lds, int ls, int n, int t, int nthr, int sign, const cplx<T> *__restrict__ tw,
                        const RadixPlan &rp) {
    int Ns = 1;
    for (int s = 0; s < rp.nrad; s++) {
        const int R = rp.r[s];
        switch (R) {
            case 8: fft_stage<T, 8, (FFT_E + 7) / 8>(lds, ls, n, t, nthr, Ns, sign, tw); break;
            case 4: fft_stage<T, 4, (FFT_E + 3) / 4>(lds, ls, n, t, nthr, Ns, sign, tw); break;
            case 2: fft_stage<T, 2, (FFT_E + 1) / 2>(lds, ls, n, t, nthr, Ns, sign, tw); break;
            case 3: fft_stage<T, 3, (FFT_E + 2) / 3>(lds, ls, n, t, nthr, Ns, sign, tw); break;
            default: fft_stage<T, 5, (FFT_E + 4) / 5>(lds, ls, n, t, nthr, Ns, sign, tw); break;
        }
        Ns *= R;
    }
}

// ---- X pass, forward: pairs of real lines -> half spectra ----
template <typename T>
__global__ void k_fft_x_r2c(const T *__restrict__ in, cplx<T> *__restrict__ out, int n, int nh, int pitch,
                            long nlines, RadixPlan rp, const cplx<T> *__restrict__ tw, int nthr) {
    extern __shared__ char smem_raw[];
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
    const int pair = threadIdx.x / nthr, t = threadIdx.x % nthr;
    const int P = blockDim.x / nthr;
    const long L0 = 2 * ((long)blockIdx.x * P + pair);
    const bool v0 = L0 < nlines, v1 = L0 + 1 < nlines;
    cplx<T> *line = lds + (size_t)pair * n;
    const T *a = in + L0 * n, *b = in + (L0 + 1) * n;
    for (int e = t; e < n; e += nthr) line[e] = cplx<T>{v0 ? a[e] : T(0), v1 ? b[e] : T(0)};
    __syncthreads();
    lds_fft(line, 1, n, t, nthr, -1, tw, rp);
    cplx<T> *oa = out + L0 * pitch, *ob = out + (L0 + 1) * pitch;
    for (int k = t; k < nh; k += nthr) {
        const cplx<T> z = line[k];
        const cplx<T> zc = line[k == 0 ? 0 : n - k];
        // A = (Z_k + conj Z_{n-k}) / 2, B = (Z_k - conj Z_{n-k}) / (2i)
        const cplx<T> A{(z.x + zc.x) / 2, (z.y - zc.y) / 2};
        const cplx<T> B{(z.y + zc.y) / 2, (zc.x - z.x) / 2};
        if (v0) oa[k] = A;
        if (v1) ob[k] = B;
    }
}

// ---- X pass, inverse: two half spectra -> two real lines (c2r semantics) ----
template <typename T>
__global__ void k_fft_x_c2r(const cplx<T> *__restrict__ in, T *__restrict__ out, int n, int nh, int pitch,
                            long nlines, RadixPlan rp, const cplx<T> *__restrict__ tw, int nthr, T scale) {
    extern __shared__ char smem_raw[];
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
    const int pair = threadIdx.x / nthr, t = threadIdx.x % nthr;
    const int P = blockDim.x / nthr;
    const long L0 = 2 * ((long)blockIdx.x * P + pair);
    const bool v0 = L0 < nlines, v1 = L0 + 1 < nlines;
    cplx<T> *line = lds + (size_t)pair * n;
    const cplx<T> *ia = in + L0 * pitch, *ib = in + (L0 + 1) * pitch;
    const int nyq = (n % 2 == 0) ? n / 2 : -1;
    for (int e = t; e < n; e += nthr) {
        cplx<T> A{T(0), T(0)}, B{T(0), T(0)};
        if (e < nh) {
            if (v0) A = ia[e];
            if (v1) B = ib[e];
            if (e == 0 || e == nyq) {  // self-conjugate bins: c2r ignores their imaginary part
                A.y = T(0);
                B.y = T(0);
            }
        } else {
            if (v0) A = ia[n - e];
            if (v1) B = ib[n - e];
            A.y = -A.y;
            B.y = -B.y;
        }
        line[e] = cplx<T>{A.x - B.y, A.y + B.x};
    }
    __syncthreads();
    lds_fft(line, 1, n, t, nthr, +1, tw, rp);
    T *oa = out + L0 * n, *ob = out + (L0 + 1) * n;
    for (int e = t; e < n; e += nthr) {
        const cplx<T> z = line[e];
        if (v0) oa[e] = z.x * scale;
        if (v1) ob[e] = z.y * scale;
    }
}

// ---- Y / Z passes over tiles of TK consecutive kx columns ----
// MODE 0: one transform with `sign`.  MODE 1 (Z pass only): forward, Poisson
// scale, zero mode, inverse.
// The Poisson scale 1 / (-4/h^2 sum sin^2) of a spectral bin: FP32 takes the
// hardware reciprocal (1 ulp; the FP32 transforms round at ~1e-7 of the
// field anyway), FP64 the correctly rounded division.
__device__ __forceinline__ float poisson_factor(float scale, float denom) {
    return scale * __builtin_amdgcn_rcpf(denom);
}
__device__ __forceinline__ double poisson_factor(double scale, double denom) { return scale / denom; }

template <typename T, int MODE>
__global__ void k_fft_cols(cplx<T> *__restrict__ data, int n, int ncols, long elem_stride, long outer_stride,
                           int TK, RadixPlan rp, const cplx<T> *__restrict__ tw, int sign,
                           const T *__restrict__ eig, T poisson_a, T scale) {
    extern __shared__ char smem_raw[];
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
    const int col0 = blockIdx.x * TK;
    const int outer = blockIdx.y;
    const int c = threadIdx.x % TK, t = threadIdx.x / TK;
    const int nthr = blockDim.x / TK;
    cplx<T> *base = data + col0 + (long)outer * outer_stride;
    const int total = n * TK;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int cc = idx % TK, e = idx / TK;
        lds[idx] = (col0 + cc < ncols) ? base[cc + e * elem_stride] : cplx<T>{T(0), T(0)};
    }
    __syncthreads();
    if (MODE == 0) {
        lds_fft(lds + c, TK, n, t, nthr, sign, tw, rp);
    } else {
        lds_fft(lds + c, TK, n, t, nthr, -1, tw, rp);
        // Z pass: column = kx, outer = ky, element = kz.
        const int kx = col0 + c;
        const T sxy = (kx < ncols ? eig[kx] : T(0)) + eig[outer];
        for (int e = t; e < n; e += nthr) {
            const T denom = poisson_a * (sxy + eig[e]);
            const bool zero_mode = (kx == 0 && outer == 0 && e == 0);
            const T f = zero_mode ? T(0) : poisson_factor(scale, denom);
            cplx<T> z = lds[e * TK + c];
            z.x *= f;
            z.y *= f;
            lds[e * TK + c] = z;
        }
        __syncthreads();
        lds_fft(lds + c, TK, n, t, nthr, +1, tw, rp);
    }
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int cc = idx % TK, e = idx / TK;
        if (col0 + cc < ncols) base[cc + e * elem_stride] = lds[idx];
    }
}

// ---- energy-shell histogram and rescale for randomfield (src/N.jl:375-449) ----
template <typename T>
__global__ void k_shell_hist(const cplx<T> *__restrict__ spec, int n, int nh, int pitch, int kmax, int nshell,
                             double inv_n3,
                             double *__restrict__ hist) {
    extern __shared__ double sh[];
    for (int s = threadIdx.x; s < nshell; s += blockDim.x) sh[s] = 0.0;
    __syncthreads();
    const size_t total = (size_t)nh * n * n;
    for (size_t I = blockIdx.x * (size_t)blockDim.x + threadIdx.x; I < total; I += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(I % nh);
        const int iy = (int)((I / nh) % n), iz = (int)(I / ((size_t)nh * n));
        const long ky = iy <= (n - 1) / 2 ? iy : iy - n;  // wavenumber_int (src/N.jl:32)
        const long kz = iz <= (n - 1) / 2 ? iz : iz - n;
        const long kk = (long)kx * kx + ky * ky + kz * kz;
        long s = (long)sqrt((double)kk);
        while (s * s > kk) s--;
        while ((s + 1) * (s + 1) <= kk) s++;
        const size_t J = (size_t)kx + (size_t)pitch * ((size_t)iy + (size_t)n * iz);
        const double re = (double)spec[J].x * inv_n3, im = (double)spec[J].y * inv_n3;
        const double w = (kx >= 1 && kx <= kmax - 1) ? 2.0 : 1.0;
        if (s < nshell) atomicAdd(&sh[s], w * (re * re + im * im) / 2);
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nshell; s += blockDim.x)
        if (sh[s] != 0.0) atomicAdd(&hist[s], sh[s]);
}

template <typename T>
__global__ void k_shell_scale(cplx<T> *__restrict__ spec, int n, int nh, int pitch, int nshell,
                              const double *__restrict__ factor,
                              double extra) {
    const size_t total = (size_t)nh * n * n;
    for (size_t I = blockIdx.x * (size_t)blockDim.x + threadIdx.x; I < total; I += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(I % nh);
        const int iy = (int)((I / nh) % n), iz = (int)(I / ((size_t)nh * n));
        const long ky = iy <= (n - 1) / 2 ? iy : iy - n;
        const long kz = iz <= (n - 1) / 2 ? iz : iz - n;
        const long kk = (long)kx * kx + ky * ky + kz * kz;
        long s = (long)sqrt((double)kk);
        while (s * s > kk) s--;
        while ((s + 1) * (s + 1) <= kk) s++;
        const double f = (s < nshell ? factor[s] : 0.0) * extra;
        const size_t J = (size_t)kx + (size_t)pitch * ((size_t)iy + (size_t)n * iz);
        spec[J].x = (T)((double)spec[J].x * f);
        spec[J].y = (T)((double)spec[J].y * f);
    }
}


// ---------------------------------------------------------------------------
// Power-of-two sizes: the whole radix plan, every stride and the tile width
// are compile-time constants, so stages unroll with shift/mask index math and
// no plan array lives in memory.  n = 2^m is split into radix-8 stages and a
// final radix-4 or radix-2 stage; each thread owns 16 elements of a line.
constexpr int ct_radix(int m, int maxr = 16) {
    return (maxr >= 16 && m % 16 == 0) ? 16 : (m % 8 == 0 ? 8 : (m % 4 == 0 ? 4 : 2));
}

// LDS position of line element e: one padding slot after every 8 elements,
// which spreads the strided Stockham scatters over all banks.
#ifndef EC_FFT_PAD_SHIFT
#define EC_FFT_PAD_SHIFT 5
#endif
// LDS padding: one spare element every 2^EC_FFT_PAD_SHIFT elements
__device__ __forceinline__ constexpr int pad8(int e) { return e + (e >> EC_FFT_PAD_SHIFT); }
constexpr int padded_len(int n) { return n + (n >> EC_FFT_PAD_SHIFT); }
// pad8(x + r S) = pad8(x) + pad_off<S>(r) for the accesses of the power-of-two
// stages: with P = 2^EC_FFT_PAD_SHIFT and S, P powers of two, either S is a
// multiple of P, or S divides P and x + r S stays inside the period that
// r S starts in (reads: x = b < S; writes: x = idxD = S R m + jm, jm < S,
// with S R a multiple or a divisor of P; tile rows: x = e0 < S).  So every
// LDS access is a per-thread base plus a compile-time offset, with no
// per-element index arithmetic.
template <int S>
__device__ __forceinline__ constexpr int pad_off(int r) {
    static_assert((S & (S - 1)) == 0, "power-of-two stride");
    return r * S + ((r * S) >> EC_FFT_PAD_SHIFT);
}

// Ordering between the stages of a line: a block barrier, or (WAVE) when one
// wave owns the line, only the completion of its own LDS operations (a wave's
// LDS operations execute in order; the clobber keeps the compiler from moving
// LDS accesses across).
template <bool WAVE>
__device__ __forceinline__ void line_sync() {
    if constexpr (WAVE) asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    else __syncthreads();
}

// A hook on the values the first stage reads: pre(e, x) for element e (the
// Poisson pass scales the spectrum as its inverse transform loads it).
struct NoPre {
    template <typename X>
    __device__ __forceinline__ X operator()(int, X x) const { return x; }
};

// MAXR caps the radix (a 1024-thread block keeps radix 8 to stay within 128 VGPRs)
template <typename T, int N, int NS, int NTHR, int LS, int SIGN, int MAXR = 16, bool WAVE = false, class PRE = NoPre>
__device__ __forceinline__ void ct_stages(cplx<T> *lds, int t, const cplx<T> *__restrict__ tw, const PRE &pre = PRE()) {
    static_assert((N & (N - 1)) == 0, "power-of-two length");
    if constexpr (NS < N) {
        constexpr int R = ct_radix(N / NS, MAXR);
        constexpr int NR = N / R;
        constexpr int B = (NR + NTHR - 1) / NTHR;
        constexpr int TSTEP = N / (NS * R);
        cplx<T> v[B][R];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int b = t + q * NTHR;
            if (NR % NTHR == 0 || b < NR) {
                const int jm = b & (NS - 1);
                const int pb = pad8(b) * LS;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    cplx<T> x = lds[pb + pad_off<NR>(r) * LS];
                    if constexpr (NS == 1) x = pre(b + r * NR, x);
                    if (NS > 1 && r > 0) {
                        cplx<T> w = tw[r * jm * TSTEP];
                        if (SIGN > 0) w.y = -w.y;
                        x = cmul(x, w);
                    }
                    v[q][r] = x;
                }
                Bfly<T, R>::run(v[q], SIGN);
            }
        }
        line_sync<WAVE>();
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int b = t + q * NTHR;
            if (NR % NTHR == 0 || b < NR) {
                const int jm = b & (NS - 1);
                const int pd = pad8((b - jm) * R + jm) * LS;
#pragma unroll
                for (int r = 0; r < R; r++) lds[pd + pad_off<NS>(r) * LS] = v[q][r];
            }
        }
        line_sync<WAVE>();
        ct_stages<T, N, NS * R, NTHR, LS, SIGN, MAXR, WAVE, NoPre>(lds, t, tw);
    }
}

// X passes.  NTHR = N / 8 threads transform a line pair (8 elements each,
// radix-8 stages); a block of xblock(NTHR) threads (one wave, or the two waves
// of a 1024-point pair) transforms W = xblock / NTHR pairs at a time and loops
// over G such groups, loading the next group into registers
// while the current one is transformed.  Each thread meets the same twiddles
// in every group, so it keeps, per butterfly and stage, the powers w^(2^i) of
// its base twiddle w in registers (loaded once) and forms the other powers as
// products of these (at most three complex products).  No global load is left
// inside the stages, so the prefetch stays in flight through the transform.
constexpr int xblock(int nthr) { return nthr > 64 ? nthr : 64; }

template <typename T, int N, int NS, int NTHR, int MAXR, bool MORE = (NS < N)>
struct RegStages {
    __device__ __forceinline__ void load(int, const cplx<T> *__restrict__) {}
    template <int SIGN>
    __device__ __forceinline__ void run(cplx<T> *, int) {}
};

template <typename T, int N, int NS, int NTHR, int MAXR>
struct RegStages<T, N, NS, NTHR, MAXR, true> {
    static constexpr int R = ct_radix(N / NS, MAXR);
    static constexpr int NR = N / R;
    static constexpr int B = (NR + NTHR - 1) / NTHR;
    static constexpr int TSTEP = N / (NS * R);
    static constexpr int NB = NS > 1 ? (R == 16 ? 4 : (R == 8 ? 3 : (R == 4 ? 2 : 1))) : 1;
    cplx<T> w[B][NB];  // w[q][i] = tw[(jm * TSTEP) << i] of butterfly q
    RegStages<T, N, NS * R, NTHR, MAXR> next;

    __device__ __forceinline__ void load(int t, const cplx<T> *__restrict__ tw) {
        if constexpr (NS > 1) {
#pragma unroll
            for (int q = 0; q < B; q++) {
                const int b = t + q * NTHR;
                const int jm = (NR % NTHR == 0 || b < NR) ? (b & (NS - 1)) : 0;
#pragma unroll
                for (int i = 0; i < NB; i++) w[q][i] = tw[(jm * TSTEP) << i];
            }
        }
        next.load(t, tw);
    }

    template <int SIGN>
    __device__ __forceinline__ void run(cplx<T> *lds, int t) {
        cplx<T> v[B][R];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int b = t + q * NTHR;
            if (NR % NTHR == 0 || b < NR) {
                cplx<T> pw[R];
                if constexpr (NS > 1) {
                    // opaque copies: the products are formed here, not hoisted
                    // out of the caller's group loop (they would hold registers)
#pragma unroll
                    for (int i = 0; i < NB; i++) {
                        cplx<T> x = w[q][i];
                        asm volatile("" : "+v"(x.x), "+v"(x.y));
                        pw[1 << i] = x;
                    }
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        const int low = r & -r;
                        if (r != low) pw[r] = cmul(pw[r - low], pw[low]);
                    }
                }
                const int pb = pad8(b);
#pragma unroll
                for (int r = 0; r < R; r++) {
                    cplx<T> x = lds[pb + pad_off<NR>(r)];
                    if (NS > 1 && r > 0) {
                        cplx<T> ww = pw[r];
                        if (SIGN > 0) ww.y = -ww.y;
                        x = cmul(x, ww);
                    }
                    v[q][r] = x;
                }
                Bfly<T, R>::run(v[q], SIGN);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int b = t + q * NTHR;
            if (NR % NTHR == 0 || b < NR) {
                const int jm = b & (NS - 1);
                const int pd = pad8((b - jm) * R + jm);
#pragma unroll
                for (int r = 0; r < R; r++) lds[pd + pad_off<NS>(r)] = v[q][r];
            }
        }
        __syncthreads();
        next.template run<SIGN>(lds, t);
    }
};

template <typename T, int N, int NTHR, int G>
__global__ void __launch_bounds__(xblock(NTHR)) k_fft_x_r2c_ct(const T *__restrict__ in, cplx<T> *__restrict__ out,
                                                               long nlines, int pitch, const cplx<T> *__restrict__ tw) {
    extern __shared__ char smem_raw[];
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
    constexpr int NH = N / 2 + 1, W = xblock(NTHR) / NTHR, PER = N / NTHR;
    const int pair = threadIdx.x / NTHR, t = threadIdx.x % NTHR;
    cplx<T> *line = lds + (size_t)pair * padded_len(N);
    const long npairs = (nlines + 1) / 2;
    RegStages<T, N, 1, NTHR, 8> st;
    st.load(t, tw);
    T ra[PER], rb[PER];
    auto fetch = [&](long pr) {
        const long L0 = 2 * pr;
        const bool v0 = pr < npairs, v1 = L0 + 1 < nlines;
        const T *a = in + L0 * N, *b = a + N;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = t + q * NTHR;
            ra[q] = v0 ? ec_nt_load(a + e) : T(0);
            rb[q] = v1 ? ec_nt_load(b + e) : T(0);
        }
    };
    const long g0 = (long)blockIdx.x * G * W;  // first pair of the block
    fetch(g0 + pair);
    for (int g = 0; g < G; g++) {
        if (g0 + (long)g * W >= npairs) break;  // uniform over the block
        const long pr = g0 + (long)g * W + pair;
#pragma unroll
        for (int q = 0; q < PER; q++) line[pad8(t) + pad_off<NTHR>(q)] = cplx<T>{ra[q], rb[q]};
        if (g + 1 < G) fetch(pr + W);
        __syncthreads();
        st.template run<-1>(line, t);
        const long L0 = 2 * pr;
        const bool v0 = pr < npairs, v1 = L0 + 1 < nlines;
        cplx<T> *oa = out + L0 * pitch, *ob = oa + pitch;
#pragma unroll
        for (int k = t; k < NH; k += NTHR) {
            const cplx<T> z = line[pad8(k)];
            const cplx<T> zc = line[pad8((N - k) & (N - 1))];
            const cplx<T> A{(z.x + zc.x) / 2, (z.y - zc.y) / 2};
            const cplx<T> Bv{(z.y + zc.y) / 2, (zc.x - z.x) / 2};
            if (v0) ec_nt_store(oa + k, A);
            if (v1) ec_nt_store(ob + k, Bv);
        }
        __syncthreads();
    }
}

template <typename T, int N, int NTHR, int G>
__global__ void __launch_bounds__(xblock(NTHR)) k_fft_x_c2r_ct(const cplx<T> *__restrict__ in, T *__restrict__ out,
                                                               long nlines, int pitch, const cplx<T> *__restrict__ tw,
                                                               T scale) {
    extern __shared__ char smem_raw[];
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
    constexpr int NH = N / 2 + 1, W = xblock(NTHR) / NTHR, PER = N / NTHR, KQ = (NH + NTHR - 1) / NTHR;
    const int pair = threadIdx.x / NTHR, t = threadIdx.x % NTHR;
    cplx<T> *line = lds + (size_t)pair * padded_len(N);
    const long npairs = (nlines + 1) / 2;
    RegStages<T, N, 1, NTHR, 8> st;
    st.load(t, tw);
    // only the stored half spectrum is loaded: bin k < NH of both lines
    cplx<T> ra[KQ], rb[KQ];
    auto fetch = [&](long pr) {
        const long L0 = 2 * pr;
        const bool v0 = pr < npairs, v1 = L0 + 1 < nlines;
        const cplx<T> *ia = in + L0 * pitch, *ib = ia + pitch;
#pragma unroll
        for (int q = 0; q < KQ; q++) {
            const int k = t + q * NTHR;
            const bool kin = q < KQ - 1 || k < NH;
            ra[q] = v0 && kin ? ec_nt_load(ia + k) : cplx<T>{T(0), T(0)};
            rb[q] = v1 && kin ? ec_nt_load(ib + k) : cplx<T>{T(0), T(0)};
        }
    };
    const long g0 = (long)blockIdx.x * G * W;
    fetch(g0 + pair);
    for (int g = 0; g < G; g++) {
        if (g0 + (long)g * W >= npairs) break;  // uniform over the block
        const long pr = g0 + (long)g * W + pair;
        // line = A + iB with the Hermitian halves rebuilt: bin k and its
        // mirror N - k (conj(A) + i conj(B)); the self-conjugate bins 0 and
        // N/2 drop their imaginary parts
#pragma unroll
        for (int q = 0; q < KQ; q++) {
            const int k = t + q * NTHR;
            if (q < KQ - 1 || k < NH) {
                cplx<T> A = ra[q], B = rb[q];
                if (k == 0 || k == N / 2) {
                    A.y = T(0);
                    B.y = T(0);
                }
                line[pad8(k)] = cplx<T>{A.x - B.y, A.y + B.x};
                if (k > 0 && k < N / 2) line[pad8(N - k)] = cplx<T>{A.x + B.y, B.x - A.y};
            }
        }
        if (g + 1 < G) fetch(pr + W);
        __syncthreads();
        st.template run<+1>(line, t);
        const long L0 = 2 * pr;
        const bool v0 = pr < npairs, v1 = L0 + 1 < nlines;
        T *oa = out + L0 * N, *ob = oa + N;
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int e = t + q * NTHR;
            const cplx<T> z = line[pad8(t) + pad_off<NTHR>(q)];
            if (v0) ec_nt_store(oa + e, z.x * scale);
            if (v1) ec_nt_store(ob + e, z.y * scale);
        }
        __syncthreads();
    }
}

// Y / Z passes: TK consecutive kx columns, the whole column staged in LDS
// (element e of column c at lds[e * TK + c]: consecutive lanes take
// consecutive columns, so both global rows and LDS rows are contiguous).
// The fused Poisson pass (MODE 1) also stages the twiddle table and the
// eigenvalues sin^2(pi k / N) in LDS behind the tile (lds_cols_ct): read from
// global memory they are small scattered loads, issued between the stages and
// between the two transforms, whose latency under the streaming traffic of
// the other blocks left the pass at half the speed of a single transform.
// The Poisson scale of bin (kx, ky, kz) = 1 / (a (s_x + s_y + s_z)), s the
// eigenvalues sin^2(pi k / N) (kx, ky fixed per column: sxy = s_x + s_y), and 0
// for the zero mode; applied by the inverse transform as it loads bin kz = e.
template <typename T>
struct PoissonScale {
    T a, scale, sxy;
    const T *eig;   // LDS
    bool origin;    // the column holds the zero mode (kx = ky = 0)
    __device__ __forceinline__ cplx<T> operator()(int e, cplx<T> z) const {
        const T denom = a * (sxy + eig[e]);
        const T f = (origin && e == 0) ? T(0) : poisson_factor(scale, denom);
        return cplx<T>{z.x * f, z.y * f};
    }
};
__device__ __forceinline__ bool kx_valid(int kx, int ncols) { return kx < ncols; }

constexpr size_t lds_cols_ct(int n, int tk, int mode, size_t tsize) {
    return (size_t)padded_len(n) * tk * 2 * tsize + (mode == 1 ? (size_t)n * 3 * tsize : 0);
}

template <typename T, int N, int TK, int NTHR, int MODE, int SIGN>
__global__ void __launch_bounds__(TK *NTHR) k_fft_cols_ct(cplx<T> *__restrict__ data, int ncols, int ntiles,
                                                         long elem_stride, long outer_stride,
                                                         const cplx<T> *__restrict__ tw, const T *__restrict__ eig,
                                                         T poisson_a, T scale) {
    extern __shared__ char smem_raw[];
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
    // 1D grid of ntiles column tiles x N outer indices, column tiles fastest, in
    // XCD order: neighbouring tiles share 128-byte lines in one L2
    const int L = (int)ec_xcd_tile(blockIdx.x, gridDim.x);
    const int col0 = (L % ntiles) * TK;
    const int outer = L / ntiles;
    const int c = threadIdx.x % TK, t = threadIdx.x / TK;
    cplx<T> *base = data + col0 + (long)outer * outer_stride;
    constexpr int TOTAL = N * TK;
    constexpr int TBQ = (N + TK * NTHR - 1) / (TK * NTHR);  // table entries per thread
    cplx<T> *twl = lds + (size_t)padded_len(N) * TK;         // MODE 1: N twiddles, then N eigenvalues
    T *eigl = reinterpret_cast<T *>(twl + N);
    // element q of a thread is (column cc, row e0 + q NTHR): a uniform row
    // base plus one loop-invariant 32-bit byte offset per lane, and in LDS a
    // per-thread base plus a compile-time offset
    static_assert(TOTAL % (TK * NTHR) == 0, "whole rows per thread");
    constexpr int PER = TOTAL / (TK * NTHR);
    const int cc = threadIdx.x % TK, e0 = threadIdx.x / TK;
    const bool cok = col0 + cc < ncols;
    const unsigned voff = (unsigned)((cc + (long)e0 * elem_stride) * (long)sizeof(cplx<T>));
    const long qstep = (long)NTHR * elem_stride * (long)sizeof(cplx<T>);
    const int lrow = pad8(e0) * TK + cc;
    // all loads of the tile (and of the tables) in flight at once
    {
        cplx<T> tmp[PER];
        cplx<T> twr[TBQ];
        T eir[TBQ];
        if constexpr (MODE == 1) {
#pragma unroll
            for (int q = 0; q < TBQ; q++) {
                const int i = threadIdx.x + q * TK * NTHR;
                if (i < N) {
                    twr[q] = tw[i];
                    eir[q] = eig[i];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < PER; q++)
            tmp[q] = cok ? *(const cplx<T> *)((const char *)base + q * qstep + voff) : cplx<T>{T(0), T(0)};
        if constexpr (MODE == 1) {
#pragma unroll
            for (int q = 0; q < TBQ; q++) {
                const int i = threadIdx.x + q * TK * NTHR;
                if (i < N) {
                    twl[i] = twr[q];
                    eigl[i] = eir[q];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < PER; q++) lds[lrow + pad_off<NTHR>(q) * TK] = tmp[q];
    }
    __syncthreads();
    if (MODE == 0) {
        ct_stages<T, N, 1, NTHR, TK, SIGN>(lds + c, t, tw);
    } else {
        ct_stages<T, N, 1, NTHR, TK, -1, (TK * NTHR > 512 ? 8 : 16)>(lds + c, t, twl);
        // the Poisson scale and zero mode, applied as the inverse loads each bin
        const PoissonScale<T> ps{poisson_a, scale, (kx_valid(col0 + c, ncols) ? eigl[col0 + c] : T(0)) + eigl[outer],
                                 eigl, col0 + c == 0 && outer == 0};
        ct_stages<T, N, 1, NTHR, TK, +1, (TK * NTHR > 512 ? 8 : 16)>(lds + c, t, twl, ps);
    }
    if (cok) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            cplx<T> *dst = (cplx<T> *)((char *)base + q * qstep + voff);
            const cplx<T> z = lds[lrow + pad_off<NTHR>(q) * TK];
            // rows shorter than a 128-byte line: plain stores, so that L2 joins
            // the halves the neighbouring tile writes before the line goes out
            if constexpr (TK * sizeof(cplx<T>) < 128) *dst = z;
            else ec_nt_store(dst, z);
        }
    }
}

// The z pass of the FP32 Poisson solve as the tridiagonal solve it is.  After
// the x and y transforms, column (kx, ky) of the half spectrum holds D_z, and
// the z transform, the division by the eigenvalues a (s_x + s_y + s_z) of the
// discrete Laplacian and the inverse z transform solve, exactly, the cyclic
// system (P_{z+1} - 2 P_z + P_{z-1}) / h^2 + c P_z = D_z with
// c = a (s_x + s_y) = -4 (s_x + s_y) / h^2 (the z modes are its eigenvectors).
// With beta = 2 - c h^2 = 2 + 4 s, s = s_x + s_y, and r the root of
// r^2 - beta r + 1 = 0 in (0, 1], r = 1 / (1 + 2 s + 2 sqrt(s (1 + s))), the
// operator factors as S+ + S- - beta = -(1/r) (1 - r S+)(1 - r S-), so
//     Q_z = F_z + r Q_{z+1},   P_z = Q_z + r P_{z-1}   (cyclic in z),
// with F = -r h^2 D / N^2 (N^2: the inverse x and y transforms are not
// normalised).  The tile of TK columns is staged in LDS as in the transform
// pass (whole row segments per lane group); thread (column c, block t) then
// runs each recursion over its V = N/64 consecutive planes from zero, and the
// value entering each block comes from a scan over the 64 blocks of the
// column (factor R = r^V per block): within a wave over its 64/TK blocks by
// shuffles, across the TK waves through LDS, cyclic through 1 / (1 - r^N).
// Running sums, carries and scans are FP64 (the recursions damp rounding by r;
// 1 / (1 - r) reaches ~N / (2 pi) for the lowest modes); the planes stay FP32
// between the passes, each rounded once at its own scale.  The zero column
// (kx = ky = 0, r = 1, whose zero mode the transforms drop) keeps its input
// for the transform pass, run on that one column afterwards.  About a fifth of
// the two transforms' arithmetic.
template <int N, int TK>
__global__ void __launch_bounds__(64 * TK) __attribute__((amdgpu_waves_per_eu(4))) k_z_poisson_tri(
    cplx<float> *__restrict__ data, int ncols, int ntiles, long elem_stride, long outer_stride, double h2n) {
    using T = float;
    typedef cplx<double> cd;
    constexpr int V = N / 64, BW = 64 / TK, NW = TK;  // planes per block, blocks per wave and column, waves
    constexpr int LB = BW == 16 ? 4 : (BW == 8 ? 3 : (BW == 4 ? 2 : 1));
    static_assert(N % 64 == 0 && V >= 1 && V <= 16 && (1 << LB) == BW, "64 blocks of V planes per column");
    extern __shared__ char smem_raw[];
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
    cd *gsc = reinterpret_cast<cd *>(lds + (size_t)padded_len(N) * TK);  // [2][NW waves][TK columns]
    const int L = (int)ec_xcd_tile(blockIdx.x, gridDim.x);
    const int col0 = (L % ntiles) * TK, outer = L / ntiles;
    cplx<T> *base = data + col0 + (long)outer * outer_stride;
    // tile in and out as in k_fft_cols_ct: element q of a thread is column cc,
    // row e0 + 64 q (whole row segments per TK lanes)
    const int cc = threadIdx.x % TK, e0 = threadIdx.x / TK;
    const bool cok = col0 + cc < ncols;
    const unsigned voff = (unsigned)((cc + (long)e0 * elem_stride) * (long)sizeof(cplx<T>));
    const long qstep = 64L * elem_stride * (long)sizeof(cplx<T>);
    const int lrow = pad8(e0) * TK + cc;
    {
        cplx<T> tmp[V];
#pragma unroll
        for (int q = 0; q < V; q++)
            tmp[q] = cok ? *(const cplx<T> *)((const char *)base + q * qstep + voff) : cplx<T>{T(0), T(0)};
#pragma unroll
        for (int q = 0; q < V; q++) lds[lrow + pad_off<64>(q) * TK] = tmp[q];
    }
    __syncthreads();
    // the same thread now owns column c = cc, planes V t .. V t + V - 1 with
    // t = e0 = BW w + j (w the wave, j = lane / TK)
    const int c = cc, t = e0, w = t / BW, j = t % BW, lane = threadIdx.x & 63;
    const int kx = col0 + c;
    const bool solve = kx < ncols && !(kx == 0 && outer == 0);
    cplx<T> *col = lds + pad8(V * t) * TK + c;  // plane V t + v at col[v * TK]
    const double sx = sinpi((double)kx / N), sy = sinpi((double)outer / N);
    const double sxy = solve ? sx * sx + sy * sy : 1.0;
    const double r = 1.0 / (1.0 + 2.0 * sxy + 2.0 * sqrt(sxy * (1.0 + sxy)));
    double Rk[5];  // R^(2^k), R = r^V
    Rk[0] = r;
#pragma unroll
    for (int v = 1; v < V; v++) Rk[0] *= r;
#pragma unroll
    for (int k = 1; k < 5; k++) Rk[k] = Rk[k - 1] * Rk[k - 1];
    const double RW = Rk[LB];  // R^BW: one wave's blocks
    double RN = RW;            // R^64 = r^N
#pragma unroll
    for (int k = LB; k < 6; k++) RN *= RN;
    const double cyc = 1.0 / (1.0 - RN);
    auto rpow = [&](int e) {  // R^e, 0 <= e <= 16
        double p = 1.0;
#pragma unroll
        for (int k = 0; k < 5; k++)
            if (e & (1 << k)) p *= Rk[k];
        return p;
    };
    auto shfl = [&](const cd &z, int src) { return cd{__shfl(z.x, src, 64), __shfl(z.y, src, 64)}; };
    auto fma_c = [](const cd &acc, double a, const cd &z) { return cd{fma(a, z.x, acc.x), fma(a, z.y, acc.y)}; };
    // the planes stay FP32 in registers between the passes; every running sum,
    // carry and scan is FP64 (each stored value is rounded once, at its own
    // scale)
    cplx<T> x[V];
#pragma unroll
    for (int v = 0; v < V; v++) x[v] = col[v * TK];
    // Q_z = F_z + r Q_{z+1}: from the block's end with zero, E = the block's first value
    cd acc{0.0, 0.0};
    {
        const double f = -r * h2n;
#pragma unroll
        for (int v = V - 1; v >= 0; v--) {
            acc = cd{fma(f, (double)x[v].x, r * acc.x), fma(f, (double)x[v].y, r * acc.y)};
            x[v] = cplx<T>{(T)acc.x, (T)acc.y};
        }
    }
    // suffix over the wave's BW blocks: X_j = sum_{i <= BW-1-j} R^i E_{j+i}
    cd X = acc;
#pragma unroll
    for (int k = 0; k < LB; k++) {
        const int s = 1 << k;
        const cd o = shfl(X, (lane + TK * s) & 63);
        if (j + s < BW) X = fma_c(X, Rk[k], o);
    }
    if (j == 0) gsc[w * TK + c] = X;  // G_w: the wave's blocks as one transfer
    const cd Xn = shfl(X, (lane + TK) & 63);
    __syncthreads();
    // Q at the end of this wave's blocks: sum_i (R^BW)^i G_{w+1+i} / (1 - R^64)
    cd Cw{0.0, 0.0};
    {
        double f = 1.0;
#pragma unroll
        for (int i = 0; i < NW; i++) {
            Cw = fma_c(Cw, f, gsc[((w + 1 + i) % NW) * TK + c]);
            f *= RW;
        }
        Cw = cd{Cw.x * cyc, Cw.y * cyc};
    }
    // Q_z = l_z + r^(V - v) C with C the Q entering this block (at plane
    // V (t + 1)); then P_z = Q_z + r P_{z-1} from the block's start with zero,
    // E' = the block's last value
    {
        const cd C = j == BW - 1 ? Cw : fma_c(Xn, rpow(BW - 1 - j), Cw);
        const double ri = 1.0 / r;
        double p = Rk[0];  // r^(V - v)
        acc = cd{0.0, 0.0};
#pragma unroll
        for (int v = 0; v < V; v++) {
            const cd q = fma_c(cd{(double)x[v].x, (double)x[v].y}, p, C);
            p *= ri;
            asm volatile("" : "+v"(p));  // formed here, not all ahead (they would hold registers)
            acc = fma_c(q, r, acc);
            x[v] = cplx<T>{(T)acc.x, (T)acc.y};
        }
    }
    cd Y = acc;
#pragma unroll
    for (int k = 0; k < LB; k++) {
        const int s = 1 << k;
        const cd o = shfl(Y, (lane - TK * s) & 63);
        if (j - s >= 0) Y = fma_c(Y, Rk[k], o);
    }
    if (j == BW - 1) gsc[NW * TK + w * TK + c] = Y;  // H_w
    const cd Yp = shfl(Y, (lane - TK) & 63);
    __syncthreads();
    cd Cp{0.0, 0.0};
    {
        double f = 1.0;
#pragma unroll
        for (int i = 0; i < NW; i++) {
            Cp = fma_c(Cp, f, gsc[NW * TK + ((w + NW - 1 - i) % NW) * TK + c]);
            f *= RW;
        }
        Cp = cd{Cp.x * cyc, Cp.y * cyc};
    }
    if (solve) {
        // P entering this block (P at plane V t - 1)
        const cd C = j == 0 ? Cp : fma_c(Yp, rpow(j), Cp);
        double p = r;  // r^(v + 1)
#pragma unroll
        for (int v = 0; v < V; v++) {
            const cd z = fma_c(cd{(double)x[v].x, (double)x[v].y}, p, C);
            p *= r;
            asm volatile("" : "+v"(p));
            col[v * TK] = cplx<T>{(T)z.x, (T)z.y};
        }
    }
    __syncthreads();
    // the zero column keeps its input for the transform pass that solves it
    // (the store addresses are formed again here: held from the loads they
    // would occupy 2 V registers through the solve)
    unsigned soff = voff;
    asm volatile("" : "+v"(soff));
    if (cok && !(col0 + cc == 0 && outer == 0)) {
#pragma unroll
        for (int q = 0; q < V; q++)
            *(cplx<T> *)((char *)base + q * qstep + soff) = lds[lrow + pad_off<64>(q) * TK];
    }
}

// Plain (non-Poisson) column transforms of large FP32 grids with one wave per
// column: the tile of TK columns is staged column-major (column stride CS,
// CS = 64 / TK mod 32, so the row-wise transposing writes share banks at most
// two ways), and the Stockham stages of a column order only through its own
// wave.  Block barriers remain only after the tile load and before the tile
// store.  (The fused Poisson pass keeps k_fft_cols_ct: with two transforms per
// residency this layout needs more than 128 VGPRs and loses a block per CU.)
constexpr int col_stride(int n, int tk) {
    const int pl = padded_len(n), want = (64 / tk) % 32;
    return pl + ((want - pl % 32) % 32 + 32) % 32;
}

template <typename T, int N, int TK, int SIGN>
__global__ void __launch_bounds__(TK * 64) k_fft_cols_w(cplx<T> *__restrict__ data, int ncols, int ntiles,
                                                       long elem_stride, long outer_stride,
                                                       const cplx<T> *__restrict__ tw) {
    constexpr int CS = col_stride(N, TK), TOTAL = N * TK, PER = TOTAL / (TK * 64);
    extern __shared__ char smem_raw[];
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
    const int L = (int)ec_xcd_tile(blockIdx.x, gridDim.x);
    const int col0 = (L % ntiles) * TK;
    const int outer = L / ntiles;
    const int c = threadIdx.x / 64, t = threadIdx.x % 64;
    cplx<T> *base = data + col0 + (long)outer * outer_stride;
    // element q of a thread: column cc, row e0 + 64 q (uniform row base, one
    // 32-bit lane offset; LDS base plus a compile-time offset)
    const int cc = threadIdx.x % TK, e0 = threadIdx.x / TK;
    const bool cok = col0 + cc < ncols;
    const unsigned voff = (unsigned)((cc + (long)e0 * elem_stride) * (long)sizeof(cplx<T>));
    const long qstep = 64L * elem_stride * (long)sizeof(cplx<T>);
    const int lrow = cc * CS + pad8(e0);
    {
        cplx<T> tmp[PER];
#pragma unroll
        for (int q = 0; q < PER; q++)
            tmp[q] = cok ? *(const cplx<T> *)((const char *)base + q * qstep + voff) : cplx<T>{T(0), T(0)};
#pragma unroll
        for (int q = 0; q < PER; q++) lds[lrow + pad_off<64>(q)] = tmp[q];
    }
    __syncthreads();
    ct_stages<T, N, 1, 64, 1, SIGN, 16, true>(lds + c * CS, t, tw);
    __syncthreads();
    if (cok) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            cplx<T> *dst = (cplx<T> *)((char *)base + q * qstep + voff);
            const cplx<T> z = lds[lrow + pad_off<64>(q)];
            if constexpr (TK * sizeof(cplx<T>) < 128) *dst = z;
            else ec_nt_store(dst, z);
        }
    }
}

// Persistent form of the plain column pass above (same tiles, threads and
// stages, so the same bits).  A block that owns one tile loads it, transforms
// it and stores it one phase after the other, and with one or two blocks per
// CU the CU's memory traffic stops while its tiles are transformed.  Here the
// grid holds as many blocks as fit on the device at once; each walks the
// tiles L = blockIdx.x, + gridDim.x, ... and, as soon as tile L is in LDS,
// issues the global loads of its next tile into registers, which arrive
// while tile L is transformed and written back.  The twiddles are staged in
// LDS once per block: a global table load among the stages would wait, in the
// in-order vector-memory counter, for the whole prefetch issued before it.
// (The fused Poisson pass stays with one block per tile: its two transforms
// and a 16-value prefetch per lane need more than the 128 VGPRs of a
// 1024-thread block.)
template <typename T, int N, int TK, int SIGN>
__global__ void __launch_bounds__(TK * 64) k_fft_cols_wp(cplx<T> *__restrict__ data, int ncols, int ntiles, int ntot,
                                                        long elem_stride, long outer_stride,
                                                        const cplx<T> *__restrict__ tw) {
    constexpr int CS = col_stride(N, TK), TOTAL = N * TK, NT = TK * 64, PER = TOTAL / NT;
    extern __shared__ char smem_raw[];
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *twl = lds + (size_t)CS * TK;  // the N twiddles
    for (int i = threadIdx.x; i < N; i += NT) twl[i] = tw[i];
    const int c = threadIdx.x / 64, t = threadIdx.x % 64;
    const int cc = threadIdx.x % TK, e0 = threadIdx.x / TK;
    const unsigned voff = (unsigned)((cc + (long)e0 * elem_stride) * (long)sizeof(cplx<T>));
    const long qstep = 64L * elem_stride * (long)sizeof(cplx<T>);
    const int lrow = cc * CS + pad8(e0);
    cplx<T> buf[PER];
    auto load = [&](int L) {
        const int col0 = (L % ntiles) * TK, outer = L / ntiles;
        const char *base = (const char *)(data + col0 + (long)outer * outer_stride);
        const bool ok = col0 + cc < ncols;
#pragma unroll
        for (int q = 0; q < PER; q++)
            buf[q] = ok ? *(const cplx<T> *)(base + q * qstep + voff) : cplx<T>{T(0), T(0)};
    };
    int L = blockIdx.x;
    if (L < ntot) load(L);
    for (; L < ntot; L += gridDim.x) {
#pragma unroll
        for (int q = 0; q < PER; q++) lds[lrow + pad_off<64>(q)] = buf[q];
        __syncthreads();
        if (L + (int)gridDim.x < ntot) load(L + gridDim.x);
        const int col0 = (L % ntiles) * TK, outer = L / ntiles;
        ct_stages<T, N, 1, 64, 1, SIGN, 16, true>(lds + c * CS, t, twl);
        __syncthreads();
        if (col0 + cc < ncols) {
            char *base = (char *)(data + col0 + (long)outer * outer_stride);
#pragma unroll
            for (int q = 0; q < PER; q++) {
                cplx<T> *dst = (cplx<T> *)(base + q * qstep + voff);
                const cplx<T> z = lds[lrow + pad_off<64>(q)];
                if constexpr (TK * sizeof(cplx<T>) < 128) *dst = z;
                else ec_nt_store(dst, z);
            }
        }
        __syncthreads();
    }
}

// blocks of a persistent launch: as many as fit on the device at once (by
// LDS per block), at most one per tile
inline long persistent_grid(ec_ctx *ctx, size_t lds_bytes, long ntot) {
    const long per_cu = std::max<long>(1, (long)(160 * 1024 / std::max<size_t>(lds_bytes, 1)));
    return std::min<long>(ntot, (long)std::max(1, ctx->ncu) * per_cu);
}

// groups per block of the x passes: enough blocks to fill the device several times
inline int x_groups(long npairs, int W) {
    if (npairs / (16L * W) >= 8192) return 16;
    if (npairs / (4L * W) >= 2048) return 4;
    return 1;
}

template <typename T, int N, int G>
int r2c_x_launch(ec_ctx *ctx, ec_poisson *ps, const void *in, void *out, int pitch) {
    constexpr int NTHR = N / 8, W = xblock(NTHR) / NTHR;
    const size_t lds = (size_t)W * padded_len(N) * sizeof(cplx<T>);
    const long nlines = (long)N * N, npairs = (nlines + 1) / 2;
    const long blocks = (npairs + (long)G * W - 1) / ((long)G * W);
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_x_r2c_ct<T, N, NTHR, G>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_fft_x_r2c_ct<T, N, NTHR, G><<<(unsigned)blocks, xblock(NTHR), lds, ctx->stream>>>((const T *)in, (cplx<T> *)out, nlines,
                                                                             pitch, (const cplx<T> *)ps->twiddle);
    EC_CHECK_LAUNCH(ctx, "k_fft_x_r2c_ct");
    return EC_OK;
}

template <typename T, int N>
int r2c_x_ct(ec_ctx *ctx, ec_poisson *ps, const void *in, void *out, int pitch) {
    constexpr int W = xblock(N / 8) / (N / 8);
    switch (x_groups(((long)N * N + 1) / 2, W)) {
        case 16: return r2c_x_launch<T, N, 16>(ctx, ps, in, out, pitch);
        case 4: return r2c_x_launch<T, N, 4>(ctx, ps, in, out, pitch);
        default: return r2c_x_launch<T, N, 1>(ctx, ps, in, out, pitch);
    }
}

template <typename T, int N, int G>
int c2r_x_launch(ec_ctx *ctx, ec_poisson *ps, const void *in, void *out, double scale, int pitch) {
    constexpr int NTHR = N / 8, W = xblock(NTHR) / NTHR;
    const size_t lds = (size_t)W * padded_len(N) * sizeof(cplx<T>);
    const long nlines = (long)N * N, npairs = (nlines + 1) / 2;
    const long blocks = (npairs + (long)G * W - 1) / ((long)G * W);
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_x_c2r_ct<T, N, NTHR, G>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_fft_x_c2r_ct<T, N, NTHR, G><<<(unsigned)blocks, xblock(NTHR), lds, ctx->stream>>>(
        (const cplx<T> *)in, (T *)out, nlines, pitch, (const cplx<T> *)ps->twiddle, (T)scale);
    EC_CHECK_LAUNCH(ctx, "k_fft_x_c2r_ct");
    return EC_OK;
}

template <typename T, int N>
int c2r_x_ct(ec_ctx *ctx, ec_poisson *ps, const void *in, void *out, double scale, int pitch) {
    constexpr int W = xblock(N / 8) / (N / 8);
    switch (x_groups(((long)N * N + 1) / 2, W)) {
        case 16: return c2r_x_launch<T, N, 16>(ctx, ps, in, out, scale, pitch);
        case 4: return c2r_x_launch<T, N, 4>(ctx, ps, in, out, scale, pitch);
        default: return c2r_x_launch<T, N, 1>(ctx, ps, in, out, scale, pitch);
    }
}

template <typename T, int N, int SIGN>
int cols_w(ec_ctx *ctx, ec_poisson *ps, void *spec, int axis, int pitch) {
    constexpr int TK = 16;
    const int nh = ps->nh;
    const long elem_stride = axis == 1 ? (long)pitch : (long)pitch * N;
    const long outer_stride = axis == 1 ? (long)pitch * N : (long)pitch;
    const int ntiles = (nh + TK - 1) / TK;
    const size_t lds = (size_t)col_stride(N, TK) * TK * sizeof(cplx<T>);
    const long ntot = (long)ntiles * N;
    if (ctx->fft_cols_persistent && ntot <= 0x7fffffffL) {
        const size_t ldsp = lds + (size_t)N * sizeof(cplx<T>);  // + the twiddle table
        if (ldsp <= 160 * 1024) {
            EC_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_cols_wp<T, N, TK, SIGN>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsp));
            k_fft_cols_wp<T, N, TK, SIGN><<<(unsigned)persistent_grid(ctx, ldsp, ntot), TK * 64, ldsp, ctx->stream>>>(
                (cplx<T> *)spec, nh, ntiles, (int)ntot, elem_stride, outer_stride, (const cplx<T> *)ps->twiddle);
            EC_CHECK_LAUNCH(ctx, "k_fft_cols_wp");
            return EC_OK;
        }
    }
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_cols_w<T, N, TK, SIGN>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_fft_cols_w<T, N, TK, SIGN><<<(unsigned)(ntiles * N), TK * 64, lds, ctx->stream>>>(
        (cplx<T> *)spec, nh, ntiles, elem_stride, outer_stride, (const cplx<T> *)ps->twiddle);
    EC_CHECK_LAUNCH(ctx, "k_fft_cols_w");
    return EC_OK;
}

// the FP32 z Poisson pass as tridiagonal solves (k_z_poisson_tri), then the
// zero column by the transform pass on that one column
// (8-column tiles: 64-byte row segments; 4-column tiles, four blocks per CU,
// made the 1024^3 step 1% slower)
template <int N>
int z_poisson_tri(ec_ctx *ctx, ec_poisson *ps, void *spec, int pitch) {
    constexpr int TK = 8, TT = 8;  // TK: the transform pass's tile, TT: the tridiagonal tile
    const int nh = ps->nh, ntt = (nh + TT - 1) / TT;
    const long elem_stride = (long)pitch * N, outer_stride = pitch;
    constexpr size_t lds = (size_t)padded_len(N) * TT * sizeof(cplx<float>) + 2 * 64 * sizeof(cplx<double>);
    static_assert(lds <= 80 * 1024, "two blocks per CU");
    const double h = ps->l / N;
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_z_poisson_tri<N, TT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    k_z_poisson_tri<N, TT><<<(unsigned)(ntt * N), 64 * TT, lds, ctx->stream>>>(
        (cplx<float> *)spec, nh, ntt, elem_stride, outer_stride, h * h / ((double)N * N));
    EC_CHECK_LAUNCH(ctx, "k_z_poisson_tri");
    // column (0, 0): one block of the transform pass with one valid column
    constexpr int NTHR = N / 16;
    constexpr size_t ldsz = lds_cols_ct(N, TK, 1, sizeof(float));
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_cols_ct<float, N, TK, NTHR, 1, -1>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsz));
    k_fft_cols_ct<float, N, TK, NTHR, 1, -1><<<1, TK * NTHR, ldsz, ctx->stream>>>(
        (cplx<float> *)spec, 1, 1, elem_stride, outer_stride, (const cplx<float> *)ps->twiddle,
        (const float *)ps->eig, (float)(-4.0 / (h * h)), (float)(1.0 / ((double)N * N * N)));
    EC_CHECK_LAUNCH(ctx, "k_fft_cols_ct");
    return EC_OK;
}

template <typename T, int N, int MODE, int SIGN>
int cols_ct(ec_ctx *ctx, ec_poisson *ps, void *spec, int axis, int pitch) {
    if constexpr (sizeof(T) == 4 && N >= 512 && MODE == 0) return cols_w<T, N, SIGN>(ctx, ps, spec, axis, pitch);
    // (below 512 the transform pass is as fast, and takes one launch instead of two)
    if constexpr (sizeof(T) == 4 && N >= 512 && MODE == 1)
        if (ctx->z_poisson_tri && axis == 2) return z_poisson_tri<N>(ctx, ps, spec, pitch);
    constexpr int NTHR = N / 16;
    // one 128-byte line of complex values per row; the fused Poisson pass (two
    // transforms) takes 8 columns so that 512-thread blocks can use radix 16
    constexpr int TK0 = (sizeof(T) == 4 && MODE == 0) ? 16 : 8;
    constexpr int TK = (TK0 * NTHR > 1024) ? 1024 / NTHR : TK0;
    const int nh = ps->nh;
    const long elem_stride = axis == 1 ? (long)pitch : (long)pitch * N;
    const long outer_stride = axis == 1 ? (long)pitch * N : (long)pitch;
    const int ntiles = (nh + TK - 1) / TK;
    dim3 grid((unsigned)(ntiles * N));
    constexpr size_t lds = lds_cols_ct(N, TK, MODE, sizeof(T));
    static_assert(lds <= 160 * 1024, "column tile and tables exceed the LDS of a CU");
    const double h = ps->l / N;
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_cols_ct<T, N, TK, NTHR, MODE, SIGN>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_fft_cols_ct<T, N, TK, NTHR, MODE, SIGN><<<grid, TK * NTHR, lds, ctx->stream>>>(
        (cplx<T> *)spec, nh, ntiles, elem_stride, outer_stride, (const cplx<T> *)ps->twiddle, (const T *)ps->eig,
        (T)(-4.0 / (h * h)), (T)(1.0 / ((double)N * N * N)));
    EC_CHECK_LAUNCH(ctx, "k_fft_cols_ct");
    return EC_OK;
}

template <typename T, int N>
int cols_ct_any(ec_ctx *ctx, ec_poisson *ps, void *spec, int axis, int sign, int mode, int pitch) {
    if (mode) return cols_ct<T, N, 1, -1>(ctx, ps, spec, axis, pitch);
    return sign < 0 ? cols_ct<T, N, 0, -1>(ctx, ps, spec, axis, pitch)
                    : cols_ct<T, N, 0, +1>(ctx, ps, spec, axis, pitch);
}

#define EC_FFT_CT_SWITCH(n, CALL) \
    switch (n) {                  \
        case 32: CALL(32);        \
        case 64: CALL(64);        \
        case 128: CALL(128);      \
        case 256: CALL(256);      \
        case 512: CALL(512);      \
        case 1024: CALL(1024);    \
        default: break;           \
    }

bool make_plan(int n, RadixPlan &rp) {
    rp.nrad = 0;
    int m = n;
    while (m % 8 == 0) { rp.r[rp.nrad++] = 8; m /= 8; }
    while (m % 4 == 0) { rp.r[rp.nrad++] = 4; m /= 4; }
    while (m % 2 == 0) { rp.r[rp.nrad++] = 2; m /= 2; }
    while (m % 3 == 0) { rp.r[rp.nrad++] = 3; m /= 3; }
    while (m % 5 == 0) { rp.r[rp.nrad++] = 5; m /= 5; }
    return m == 1 && rp.nrad <= 16;
}

struct XCfg {
    int nthr, P;
    size_t lds;
};

template <typename T>
XCfg xcfg(int n) {
    XCfg c;
    c.nthr = (n + FFT_E - 1) / FFT_E;
    if (c.nthr < 1) c.nthr = 1;
    c.P = 256 / c.nthr;
    if (c.P < 1) c.P = 1;
    while (c.P > 1 && (size_t)c.P * n * sizeof(cplx<T>) > 65536) c.P /= 2;
    c.lds = (size_t)c.P * n * sizeof(cplx<T>);
    return c;
}

struct CCfg {
    int TK, nthr;
    size_t lds;
};

template <typename T>
CCfg ccfg(int n) {
    CCfg c;
    c.TK = sizeof(T) == 4 ? 16 : 8;
    while (c.TK > 1 && (size_t)n * c.TK * sizeof(cplx<T>) > 128 * 1024) c.TK /= 2;
    c.nthr = (n + FFT_E - 1) / FFT_E;
    if (c.nthr < 1) c.nthr = 1;
    while (c.TK * c.nthr > 1024 && c.TK > 1) c.TK /= 2;
    c.lds = (size_t)n * c.TK * sizeof(cplx<T>);
    return c;
}

template <typename T>
int set_lds_limits(ec_ctx *ctx) {
    const int maxlds = 160 * 1024;
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_cols<T, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, maxlds));
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_cols<T, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, maxlds));
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_x_r2c<T>, hipFuncAttributeMaxDynamicSharedMemorySize, maxlds));
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_x_c2r<T>, hipFuncAttributeMaxDynamicSharedMemorySize, maxlds));
    return EC_OK;
}

template <typename T>
int r2c_x(ec_ctx *ctx, ec_poisson *ps, const void *in, void *out, int pitch) {
    ec_prof_scope pr(ctx, "fft_x_r2c", (double)sizeof(T) * (ec_cube(ps->n) + 2.0 * ps->nh * ps->n * ps->n));
#define EC_CALL(N) return r2c_x_ct<T, N>(ctx, ps, in, out, pitch)
    EC_FFT_CT_SWITCH(ps->n, EC_CALL)
#undef EC_CALL
    RadixPlan rp;
    make_plan(ps->n, rp);
    const int n = ps->n, nh = ps->nh;
    const long nlines = (long)n * n;
    XCfg c = xcfg<T>(n);
    const long npairs = (nlines + 1) / 2;
    const int blocks = (int)((npairs + c.P - 1) / c.P);
    k_fft_x_r2c<T><<<blocks, c.P * c.nthr, c.lds, ctx->stream>>>((const T *)in, (cplx<T> *)out, n, nh, pitch, nlines,
                                                                  rp, (const cplx<T> *)ps->twiddle, c.nthr);
    EC_CHECK_LAUNCH(ctx, "k_fft_x_r2c");
    return EC_OK;
}

template <typename T>
int c2r_x(ec_ctx *ctx, ec_poisson *ps, const void *in, void *out, double scale, int pitch) {
    ec_prof_scope pr(ctx, "fft_x_c2r", (double)sizeof(T) * (ec_cube(ps->n) + 2.0 * ps->nh * ps->n * ps->n));
#define EC_CALL(N) return c2r_x_ct<T, N>(ctx, ps, in, out, scale, pitch)
    EC_FFT_CT_SWITCH(ps->n, EC_CALL)
#undef EC_CALL
    RadixPlan rp;
    make_plan(ps->n, rp);
    const int n = ps->n, nh = ps->nh;
    const long nlines = (long)n * n;
    XCfg c = xcfg<T>(n);
    const long npairs = (nlines + 1) / 2;
    const int blocks = (int)((npairs + c.P - 1) / c.P);
    k_fft_x_c2r<T><<<blocks, c.P * c.nthr, c.lds, ctx->stream>>>((const cplx<T> *)in, (T *)out, n, nh, pitch, nlines,
                                                                  rp, (const cplx<T> *)ps->twiddle, c.nthr, (T)scale);
    EC_CHECK_LAUNCH(ctx, "k_fft_x_c2r");
    return EC_OK;
}

// axis 1 (y) or 2 (z); mode 1 = Poisson (z only)
template <typename T>
int cols(ec_ctx *ctx, ec_poisson *ps, void *spec, int axis, int sign, int mode, int pitch) {
    const char *name = mode ? "fft_z_poisson" : (axis == 1 ? "fft_y" : "fft_z");
    ec_prof_scope pr(ctx, name, 2.0 * sizeof(T) * 2.0 * ps->nh * ps->n * ps->n);
#define EC_CALL(N) return cols_ct_any<T, N>(ctx, ps, spec, axis, sign, mode, pitch)
    EC_FFT_CT_SWITCH(ps->n, EC_CALL)
#undef EC_CALL
    RadixPlan rp;
    make_plan(ps->n, rp);
    const int n = ps->n, nh = ps->nh;
    CCfg c = ccfg<T>(n);
    const long elem_stride = axis == 1 ? (long)pitch : (long)pitch * n;
    const long outer_stride = axis == 1 ? (long)pitch * n : (long)pitch;
    dim3 grid((nh + c.TK - 1) / c.TK, n);
    const double h = ps->l / n;
    const T a = (T)(-4.0 / (h * h));
    const T scale = (T)(1.0 / ((double)n * n * n));
    if (mode)
        k_fft_cols<T, 1><<<grid, c.TK * c.nthr, c.lds, ctx->stream>>>((cplx<T> *)spec, n, nh, elem_stride, outer_stride,
                                                                       c.TK, rp, (const cplx<T> *)ps->twiddle, sign,
                                                                       (const T *)ps->eig, a, scale);
    else
        k_fft_cols<T, 0><<<grid, c.TK * c.nthr, c.lds, ctx->stream>>>((cplx<T> *)spec, n, nh, elem_stride, outer_stride,
                                                                       c.TK, rp, (const cplx<T> *)ps->twiddle, sign,
                                                                       (const T *)ps->eig, a, scale);
    EC_CHECK_LAUNCH(ctx, "k_fft_cols");
    return EC_OK;
}

// p <- L^-1 p in place (p = the solver's pressure buffer holding a divergence)
template <typename T>
int poisson_fft_t(ec_ctx *ctx, ec_poisson *ps) {
    int rc;
    const int sp = ps->pitch;
    if ((rc = r2c_x<T>(ctx, ps, ps->p, ps->spec, sp))) return rc;
    if ((rc = cols<T>(ctx, ps, ps->spec, 1, -1, 0, sp))) return rc;
    if ((rc = cols<T>(ctx, ps, ps->spec, 2, -1, 1, sp))) return rc;
    if ((rc = cols<T>(ctx, ps, ps->spec, 1, +1, 0, sp))) return rc;
    return c2r_x<T>(ctx, ps, ps->spec, ps->p, 1.0, sp);
}

template <typename T>
int pressure_t(ec_ctx *ctx, ec_poisson *ps, void *const u[3]) {
    const int n = ps->n;
    int rc = ec_ns_divergence(ctx, ps->dtype, n, ps->l, u, ps->p);
    if (rc) return rc;
    const int sp = ps->pitch;
    if ((rc = r2c_x<T>(ctx, ps, ps->p, ps->spec, sp))) return rc;
    if ((rc = cols<T>(ctx, ps, ps->spec, 1, -1, 0, sp))) return rc;
    if ((rc = cols<T>(ctx, ps, ps->spec, 2, -1, 1, sp))) return rc;
    if ((rc = cols<T>(ctx, ps, ps->spec, 1, +1, 0, sp))) return rc;
    return c2r_x<T>(ctx, ps, ps->spec, ps->p, 1.0, sp);
}

template <typename T>
int project_t(ec_ctx *ctx, ec_poisson *ps, void *const u[3]) {
    const int n = ps->n;
    int rc = ec_ns_divergence(ctx, ps->dtype, n, ps->l, u, ps->p);
    if (rc) return rc;
    const int sp = ps->pitch;
    if ((rc = r2c_x<T>(ctx, ps, ps->p, ps->spec, sp))) return rc;
    if ((rc = cols<T>(ctx, ps, ps->spec, 1, -1, 0, sp))) return rc;
    if ((rc = cols<T>(ctx, ps, ps->spec, 2, -1, 1, sp))) return rc;
    if ((rc = cols<T>(ctx, ps, ps->spec, 1, +1, 0, sp))) return rc;
    if ((rc = c2r_x<T>(ctx, ps, ps->spec, ps->p, 1.0, sp))) return rc;
    return ec_ns_pressuregradient(ctx, ps->dtype, n, ps->l, u, ps->p);
}

// dst = P src (src unchanged): the divergence of src, the solve, dst = src - grad p
template <typename T>
int project_into_t(ec_ctx *ctx, ec_poisson *ps, const void *const src[3], void *const dst[3]) {
    const int n = ps->n;
    void *const s3[3] = {const_cast<void *>(src[0]), const_cast<void *>(src[1]), const_cast<void *>(src[2])};
    int rc = ec_ns_divergence(ctx, ps->dtype, n, ps->l, s3, ps->p);
    if (rc) return rc;
    if ((rc = poisson_fft_t<T>(ctx, ps))) return rc;
    return ec_ns_pressuregradient_into(ctx, ps->dtype, n, ps->l, src, ps->p, dst);
}

}  // namespace

extern "C" int ec_poisson_set_z_tridiagonal(ec_ctx *ctx, int on) {
    EC_REQUIRE(ctx, ctx && (on == 0 || on == 1), "ec_poisson_set_z_tridiagonal: bad argument");
    ctx->z_poisson_tri = on;
    return EC_OK;
}

extern "C" int ec_fft_set_cols_persistent(ec_ctx *ctx, int on) {
    EC_REQUIRE(ctx, ctx && (on == 0 || on == 1), "ec_fft_set_cols_persistent: bad argument");
    ctx->fft_cols_persistent = on;
    return EC_OK;
}

extern "C" int ec_poisson_create(ec_ctx *ctx, int dtype, int n, double l, ec_poisson **out) {
    EC_REQUIRE(ctx, ctx && out, "ec_poisson_create: null argument");
    *out = nullptr;
    EC_REQUIRE(ctx, dtype == EC_F32 || dtype == EC_F64, "ec_poisson_create: bad dtype %d", dtype);
    EC_REQUIRE(ctx, n >= 2, "ec_poisson_create: n=%d", n);
    RadixPlan rp;
    if (!make_plan(n, rp))
        return ec_fail(ctx, EC_E_UNSUPPORTED, "FFT length %d: only factors 2, 3, 5 are supported", n);
    EC_REQUIRE(ctx, (n + FFT_E - 1) / FFT_E <= 1024, "FFT length %d too large", n);
    {
        // every butterfly of every stage must fall on a thread slot
        const int nthr = std::max(1, (n + FFT_E - 1) / FFT_E);
        for (int s = 0; s < rp.nrad; s++) {
            const int R = rp.r[s];
            const int need = (n / R + nthr - 1) / nthr, have = (FFT_E + R - 1) / R;
            if (need > have) return ec_fail(ctx, EC_E_UNSUPPORTED, "FFT length %d: radix-%d stage not covered", n, R);
        }
    }
    const size_t bt = dtype == EC_F64 ? 8 : 4;
    if ((size_t)n * 1 * 2 * bt > 128 * 1024)
        return ec_fail(ctx, EC_E_UNSUPPORTED, "FFT length %d exceeds the LDS column budget", n);
    ec_poisson *ps = new ec_poisson();
    ps->n = n;
    ps->dtype = dtype;
    ps->l = l;
    ps->nh = n / 2 + 1;
    // rows of whole 128-byte lines for the column passes (FP32 complex: 16 per line)
    ps->pitch = (ps->nh + 15) / 16 * 16;
    const size_t nspec = (size_t)ps->pitch * n * n;
    hipError_t e = hipMalloc(&ps->spec, nspec * 2 * bt);
    if (e == hipSuccess) e = hipMalloc(&ps->p, ec_cube(n) * bt);
    if (e == hipSuccess) e = hipMalloc(&ps->twiddle, (size_t)n * 2 * bt);
    if (e == hipSuccess) e = hipMalloc(&ps->eig, (size_t)n * bt);
    if (e != hipSuccess) {
        hipFree(ps->spec);
        hipFree(ps->p);
        hipFree(ps->twiddle);
        hipFree(ps->eig);
        delete ps;
        return ec_fail(ctx, EC_E_OOM, "ec_poisson_create(n=%d): %s", n, hipGetErrorString(e));
    }
    std::vector<double> tw(2 * (size_t)n), eg((size_t)n);
    for (int t = 0; t < n; t++) {
        // exp(-2 pi i t / n) from the folded angle for accuracy
        const double a = 2.0 * M_PI * (double)t / (double)n;
        tw[2 * t] = cos(a);
        tw[2 * t + 1] = -sin(a);
        const double s = sin(M_PI * (double)t / (double)n);
        eg[t] = s * s;
    }
    int rc = EC_OK;
    if (dtype == EC_F64) {
        if (hipMemcpy(ps->twiddle, tw.data(), tw.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(ps->eig, eg.data(), eg.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
            rc = EC_E_HIP;
        if (rc == EC_OK) rc = set_lds_limits<double>(ctx);
    } else {
        std::vector<float> twf(tw.begin(), tw.end()), egf(eg.begin(), eg.end());
        if (hipMemcpy(ps->twiddle, twf.data(), twf.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(ps->eig, egf.data(), egf.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            rc = EC_E_HIP;
        if (rc == EC_OK) rc = set_lds_limits<float>(ctx);
    }
    if (rc != EC_OK) {
        ec_poisson_destroy(ctx, ps);
        return rc == EC_E_HIP ? ec_fail(ctx, rc, "ec_poisson_create: table upload failed") : rc;
    }
    *out = ps;
    return EC_OK;
}

extern "C" int ec_poisson_destroy(ec_ctx *ctx, ec_poisson *ps) {
    if (!ps) return EC_OK;
    if (ctx) hipStreamSynchronize(ctx->stream);
    hipFree(ps->spec);
    hipFree(ps->p);
    hipFree(ps->twiddle);
    hipFree(ps->eig);
    delete ps;
    return EC_OK;
}

extern "C" void *ec_poisson_pressure(ec_poisson *ps) { return ps ? ps->p : nullptr; }

extern "C" int ec_ns_project(ec_ctx *ctx, ec_poisson *ps, void *const u[3]) {
    EC_REQUIRE(ctx, ctx && ps && u, "ec_ns_project: null argument");
    return ps->dtype == EC_F64 ? project_t<double>(ctx, ps, u) : project_t<float>(ctx, ps, u);
}

extern "C" int ec_ns_project_into(ec_ctx *ctx, ec_poisson *ps, const void *const src[3], void *const dst[3]) {
    EC_REQUIRE(ctx, ctx && ps && src && dst, "ec_ns_project_into: null argument");
    return ps->dtype == EC_F64 ? project_into_t<double>(ctx, ps, src, dst) : project_into_t<float>(ctx, ps, src, dst);
}

extern "C" int ec_ns_pressure_solve(ec_ctx *ctx, ec_poisson *ps, void *const u[3]) {
    EC_REQUIRE(ctx, ctx && ps && u, "ec_ns_pressure_solve: null argument");
    return ps->dtype == EC_F64 ? pressure_t<double>(ctx, ps, u) : pressure_t<float>(ctx, ps, u);
}

extern "C" int ec_poisson_solve(ec_ctx *ctx, ec_poisson *ps) {
    EC_REQUIRE(ctx, ctx && ps, "ec_poisson_solve: null argument");
    return ps->dtype == EC_F64 ? poisson_fft_t<double>(ctx, ps) : poisson_fft_t<float>(ctx, ps);
}

extern "C" int ec_fft_r2c(ec_ctx *ctx, ec_poisson *ps, const void *in, void *out) {
    EC_REQUIRE(ctx, ctx && ps && in && out, "ec_fft_r2c: null argument");
    EC_REQUIRE(ctx, (uintptr_t)out % (ps->dtype == EC_F64 ? 16 : 8) == 0,
               "ec_fft_r2c: spectrum not aligned to its complex element size");
    // a caller's buffer is the packed (nh, n, n) spectrum; the solver's own is pitched
    const int sp = out == ps->spec ? ps->pitch : ps->nh;
    int rc;
    if (ps->dtype == EC_F64) {
        if ((rc = r2c_x<double>(ctx, ps, in, out, sp))) return rc;
        if ((rc = cols<double>(ctx, ps, out, 1, -1, 0, sp))) return rc;
        return cols<double>(ctx, ps, out, 2, -1, 0, sp);
    }
    if ((rc = r2c_x<float>(ctx, ps, in, out, sp))) return rc;
    if ((rc = cols<float>(ctx, ps, out, 1, -1, 0, sp))) return rc;
    return cols<float>(ctx, ps, out, 2, -1, 0, sp);
}

// Inverse, normalised by 1/n^3 (ldiv!).  Works in place on a copy held in
// the solver's spectrum buffer unless `in` already is that buffer.
extern "C" int ec_fft_c2r(ec_ctx *ctx, ec_poisson *ps, const void *in, void *out) {
    EC_REQUIRE(ctx, ctx && ps && in && out, "ec_fft_c2r: null argument");
    const size_t bc = (ps->dtype == EC_F64 ? 8 : 4) * 2;  // bytes per complex
    // a caller's packed (nh, n, n) spectrum goes into the pitched solver buffer
    if (in != ps->spec)
        EC_HIP(ctx, hipMemcpy2DAsync(ps->spec, ps->pitch * bc, in, ps->nh * bc, ps->nh * bc, (size_t)ps->n * ps->n,
                                     hipMemcpyDeviceToDevice, ctx->stream));
    const double scale = 1.0 / ((double)ps->n * ps->n * ps->n);
    const int sp = ps->pitch;
    int rc;
    if (ps->dtype == EC_F64) {
        if ((rc = cols<double>(ctx, ps, ps->spec, 2, +1, 0, sp))) return rc;
        if ((rc = cols<double>(ctx, ps, ps->spec, 1, +1, 0, sp))) return rc;
        return c2r_x<double>(ctx, ps, ps->spec, out, scale, sp);
    }
    if ((rc = cols<float>(ctx, ps, ps->spec, 2, +1, 0, sp))) return rc;
    if ((rc = cols<float>(ctx, ps, ps->spec, 1, +1, 0, sp))) return rc;
    return c2r_x<float>(ctx, ps, ps->spec, out, scale, sp);
}

static double profile_eval(int profile, double param, int k) {
    if (profile == 0) return pow((double)k, 4) * exp(-2 * pow((double)k / param, 2));
    return k > 0 ? pow((double)k, param) : 0.0;
}

template <typename T>
static int randomfield_t(ec_ctx *ctx, ec_poisson *ps, int profile, double param, double totalenergy, uint64_t seed,
                         void *const u[3]) {
    const int n = ps->n, nh = ps->nh, kmax = n / 2;
    const size_t N = ec_cube(n);
    int rc;
    for (int c = 0; c < 3; c++)
        if ((rc = ec_randn(ctx, ps->dtype, N, seed, (uint64_t)c * N, u[c]))) return rc;
    if ((rc = ec_ns_project(ctx, ps, u))) return rc;
    const int kdiag = (int)floor(sqrt(3.0) * kmax);
    const int nshell = kdiag + 1;
    double *dhist = nullptr;
    EC_HIP(ctx, hipMallocAsync((void **)&dhist, sizeof(double) * 2 * nshell, ctx->stream));
    // the shell sums and factors live until the end of the body; every exit frees them
    auto body = [&]() -> int {
        EC_HIP(ctx, hipMemsetAsync(dhist, 0, sizeof(double) * 2 * nshell, ctx->stream));
        const double inv_n3 = 1.0 / ((double)n * n * n);
        const size_t nspec = (size_t)nh * n * n;
        const int blocks = (int)std::min<size_t>((nspec + 255) / 256, 4096);
        int rc2;
        for (int c = 0; c < 3; c++) {
            if ((rc2 = ec_fft_r2c(ctx, ps, u[c], ps->spec))) return rc2;
            k_shell_hist<T><<<blocks, 256, sizeof(double) * nshell, ctx->stream>>>(
                (const cplx<T> *)ps->spec, n, nh, ps->pitch, kmax, nshell, inv_n3, dhist);
            EC_CHECK_LAUNCH(ctx, "k_shell_hist");
        }
        std::vector<double> hist(nshell), factor(nshell);
        EC_HIP(ctx, hipMemcpyAsync(hist.data(), dhist, sizeof(double) * nshell, hipMemcpyDeviceToHost, ctx->stream));
        EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double totalprofile = 0.0;
        for (int k = 0; k <= kdiag; k++) totalprofile += profile_eval(profile, param, k);
        for (int k = 0; k <= kdiag; k++) {
            const double E0 = totalenergy * profile_eval(profile, param, k) / totalprofile;
            factor[k] = sqrt(E0 / hist[k]);
        }
        EC_HIP(ctx, hipMemcpyAsync(dhist + nshell, factor.data(), sizeof(double) * nshell, hipMemcpyHostToDevice,
                                   ctx->stream));
        for (int c = 0; c < 3; c++) {
            if ((rc2 = ec_fft_r2c(ctx, ps, u[c], ps->spec))) return rc2;
            // rfft / n^3, times the shell factor; the inverse then runs unnormalised
            k_shell_scale<T><<<blocks, 256, 0, ctx->stream>>>((cplx<T> *)ps->spec, n, nh, ps->pitch, nshell,
                                                              dhist + nshell, inv_n3);
            EC_CHECK_LAUNCH(ctx, "k_shell_scale");
            if ((rc2 = cols<T>(ctx, ps, ps->spec, 2, +1, 0, ps->pitch))) return rc2;
            if ((rc2 = cols<T>(ctx, ps, ps->spec, 1, +1, 0, ps->pitch))) return rc2;
            if ((rc2 = c2r_x<T>(ctx, ps, ps->spec, u[c], 1.0, ps->pitch))) return rc2;
        }
        // the host factor vector must outlive the copy above
        EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return EC_OK;
    };
    rc = body();
    const hipError_t fe = hipFreeAsync(dhist, ctx->stream);
    if (rc != EC_OK) return rc;
    if (fe != hipSuccess) return ec_fail(ctx, EC_E_HIP, "ec_ns_randomfield: free failed");
    return ec_ns_project(ctx, ps, u);
}

extern "C" int ec_ns_randomfield(ec_ctx *ctx, ec_poisson *ps, int profile, double param, double totalenergy,
                                 uint64_t seed, void *const u[3]) {
    EC_REQUIRE(ctx, ctx && ps && u, "ec_ns_randomfield: null argument");
    EC_REQUIRE(ctx, profile == 0 || profile == 1, "ec_ns_randomfield: bad profile %d", profile);
    return ps->dtype == EC_F64 ? randomfield_t<double>(ctx, ps, profile, param, totalenergy, seed, u)
                               : randomfield_t<float>(ctx, ps, profile, param, totalenergy, seed, u);
}

// Kinetic-energy spectrum over integer shells (spectrum, src/N.jl:576-590, with
// the shells of getshells, src/N.jl:478-545): out[s] = sum over half-spectrum
// bins with s^2 <= |k|^2 < (s+1)^2 of w(kx) |u^_a(k) / n^3|^2 / 2, summed over
// the three components; w = 2 for 1 <= kx <= n/2 - 1 (their conjugate bins are
// not stored), 1 otherwise.  The same shell histogram as the initial-field
// rescale, FP64 sums; the transforms use and overwrite the plan's spectrum.
template <typename T>
static int spectrum_t(ec_ctx *ctx, ec_poisson *ps, void *const u[3], int nshell, double *out) {
    const int n = ps->n, nh = ps->nh, kmax = n / 2;
    const double inv_n3 = 1.0 / ((double)n * n * n);
    const size_t nspec = (size_t)nh * n * n;
    const int blocks = (int)std::min<size_t>((nspec + 255) / 256, 4096);
    double *dhist = nullptr;
    EC_HIP(ctx, hipMallocAsync((void **)&dhist, sizeof(double) * nshell, ctx->stream));
    int rc = EC_OK;
    if (hipMemsetAsync(dhist, 0, sizeof(double) * nshell, ctx->stream) != hipSuccess)
        rc = ec_fail(ctx, EC_E_HIP, "ec_ns_spectrum: clearing the shell sums failed");
    for (int c = 0; c < 3 && rc == EC_OK; c++) {
        rc = ec_fft_r2c(ctx, ps, u[c], ps->spec);
        if (rc == EC_OK) {
            k_shell_hist<T><<<blocks, 256, sizeof(double) * nshell, ctx->stream>>>(
                (const cplx<T> *)ps->spec, n, nh, ps->pitch, kmax, nshell, inv_n3, dhist);
            if (hipGetLastError() != hipSuccess) rc = ec_fail(ctx, EC_E_HIP, "ec_ns_spectrum: k_shell_hist launch");
        }
    }
    if (rc == EC_OK &&
        (hipMemcpyAsync(out, dhist, sizeof(double) * nshell, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
         hipStreamSynchronize(ctx->stream) != hipSuccess))
        rc = ec_fail(ctx, EC_E_HIP, "ec_ns_spectrum: reading the shell sums failed");
    const hipError_t fe = hipFreeAsync(dhist, ctx->stream);
    if (rc == EC_OK && fe != hipSuccess) rc = ec_fail(ctx, EC_E_HIP, "ec_ns_spectrum: free failed");
    return rc;
}

extern "C" int ec_ns_spectrum(ec_ctx *ctx, ec_poisson *ps, void *const u[3], int nshell, double *out) {
    EC_REQUIRE(ctx, ctx && ps && u && u[0] && u[1] && u[2] && out, "ec_ns_spectrum: null argument");
    EC_REQUIRE(ctx, nshell >= 1 && nshell <= 8192, "ec_ns_spectrum: nshell %d outside [1, 8192]", nshell);
    return ps->dtype == EC_F64 ? spectrum_t<double>(ctx, ps, u, nshell, out)
                               : spectrum_t<float>(ctx, ps, u, nshell, out);
}
