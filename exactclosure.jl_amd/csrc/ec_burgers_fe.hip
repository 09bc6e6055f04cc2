// The reference's Burgers closure comparison on the GPU: dns_aided_les
// (src/Burgers.jl:350-445) for many samples in one launch, and its initial
// field (randomfield with bump_profile, src/Burgers.jl:192-202).
//
// Per sample (one 1024-thread block) and time step, with the CFL step
// dt = cfl_factor * min(h / max|u|, h^2 / visc) cut at tstop:
//   s = s_h(u)                      DNS stress at the centres
//   fsu = C_c(F * s), vsu = C_c(D * s), vu = C_f(D * u), sVU = C_c(s_h(D * u)),
//   svu = s_H(vu)                   (F the LES kernel, D the double kernel)
//   LES stresses  nomodel  s_H(u)
//                 class_m  s_H(u) + vsu - sVU
//                 class_p  s_H(u) + vsu - svu
//                 swapfil  s_H(u) + fsu - svu
//   every LES and the DNS take one forward-Euler step u += dt delta_coll(sigma).
// The DNS row sits in LDS (FP64, at most 16384 points); the coarse arrays sit
// in a per-sample global work area.  All arithmetic is FP64.
#include "ec_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int FE_NT = 1024;
constexpr int FE_PT = 16;  // DNS points per thread: nh <= 16384

__device__ __forceinline__ uint64_t hmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

__device__ __forceinline__ int wrapm(int m, int n) {  // m in (-n, 2n)
    m += m < 0 ? n : 0;
    return m >= n ? m - n : m;
}

// stress(u, visc) at centre i from faces i - 1 and i (src/Burgers.jl:97-102),
// with the reference's operation order (division by the spacing)
__device__ __forceinline__ double bst(double ua, double ub, double sp, double visc) {
    const double e = ua + ub;
    return -(e * e) / 8 + visc * (ub - ua) / sp;
}

// Block reductions (sum or max) of one double per thread; every thread gets the result.
template <int OP>
__device__ double block_reduce(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = OP == 0 ? v + w : fmax(v, w);
    }
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[wv] = v;
    __syncthreads();
    double r = OP == 0 ? 0.0 : -1.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r = OP == 0 ? r + red[i] : fmax(r, red[i]);
    return r;
}

// randomfield (src/Burgers.jl:195-202) with bump_profile: coefficients
// c_k = (k / kp)^4 exp(-2 (k / kp)^2 + 2 pi i phi_k), k = 0 .. n/2, the field
// irfft(c n, n) = sum_k c_k e^{2 pi i k m / n} over the Hermitian spectrum,
// i.e. 2 sum_k |c_k| cos(2 pi (k m / n + phi_k)), scaled to energy E with
// sum u^2 / 2 h.  Modes past 6 kp (|c_k| below 1e-27 of the peak) and the
// Nyquist mode are left out.  phi_k from a counter-based hash of (seed, k).
__global__ void __launch_bounds__(FE_NT) k_bump_field(int n, double L, double kp, double energy, int kc, uint64_t seed0,
                                                      double *__restrict__ u) {
    __shared__ double red[FE_NT / 64];
    const int q = blockIdx.x;
    const uint64_t seed = seed0 + (uint64_t)q;
    double *uq = u + (size_t)q * n;
    double ss = 0.0;
    for (int m = threadIdx.x; m < n; m += FE_NT) {
        double a = 0.0;
        for (int k = 1; k <= kc; k++) {
            const double x = k / kp;
            const double amp = x * x * x * x * exp(-2.0 * x * x);
            const double ph = (double)(hmix(seed ^ hmix((uint64_t)k)) >> 11) * 0x1.0p-53;
            const long mk = ((long)k * m) % n;
            a += 2.0 * amp * cos(2.0 * M_PI * ((double)mk / n + ph));
        }
        uq[m] = a;
        ss += a * a;
    }
    const double etot = block_reduce<0>(ss, red) / 2 * (L / n);
    const double sc = sqrt(energy / etot);
    for (int m = threadIdx.x; m < n; m += FE_NT) uq[m] *= sc;
}

struct FeArgs {
    int nh, nH, Rl, K;
    double L, visc, tstop, cfl;
    const double *wl, *wd;  // device weights
    const double *ustart;   // nsamples x nh
    double *work;           // per sample: 9 nH doubles
    double *out;            // per sample: 2 nh + 5 nH doubles
    int *steps;             // per sample step count (nullable)
};

__global__ void __launch_bounds__(FE_NT) k_burgers_dnsaid_fe(FeArgs A) {
    extern __shared__ double xs[];  // the DNS row
    __shared__ double red[FE_NT / 64];
    const int q = blockIdx.x, tid = threadIdx.x;
    const int nh = A.nh, nH = A.nH, comp = nh / nH, half = comp / 2, Rl = A.Rl, K = A.K;
    const double h = A.L / nh, H = A.L / nH, visc = A.visc;
    double *w = A.work + (size_t)q * 9 * nH;
    double *les[4] = {w, w + nH, w + 2 * nH, w + 3 * nH};  // nomodel, class_m, class_p, swapfil
    double *vu = w + 4 * nH, *sVU = w + 5 * nH, *vsu = w + 6 * nH, *fsu = w + 7 * nH;
    double *sig = w + 8 * nH;  // one LES stress at a time
    const double *wl = A.wl, *wd = A.wd;
    for (int i = tid; i < nh; i += FE_NT) xs[i] = A.ustart[(size_t)q * nh + i];
    __syncthreads();
    // lazyfilter(u, D) at fine face i: sum_k wd[k + K] u[i + k]
    auto Du = [&](int i) {
        double acc = 0.0;
        for (int k = -K; k <= K; k++) acc += wd[k + K] * xs[wrapm(i + k, nh)];
        return acc;
    };
    // 0-based fine indices of coarse point J: face J comp + comp - 1, centre that - half
    // initial double-filtered state of every LES (src/Burgers.jl:391-407)
    for (int J = tid; J < nH; J += FE_NT) {
        const double v = Du(J * comp + comp - 1);
        for (int m = 0; m < 4; m++) les[m][J] = v;
    }
    __syncthreads();
    double t = 0.0;
    int nstep = 0;
    while (t < A.tstop) {
        double um = 0.0;
        for (int i = tid; i < nh; i += FE_NT) um = fmax(um, fabs(xs[i]));
        um = block_reduce<1>(um, red);
        double dt = A.cfl * fmin(h / um, h * h / visc);
        dt = fmin(dt, A.tstop - t);
        t += dt;
        nstep++;
        // coarse quantities of the current DNS state
        for (int J = tid; J < nH; J += FE_NT) {
            const int f = J * comp + comp - 1, c = f - half;
            double a = 0.0, b = 0.0;
            for (int r = -Rl; r <= Rl; r++) {
                const int i = wrapm(c + r, nh);
                a += wl[r + Rl] * bst(xs[wrapm(i - 1, nh)], xs[i], h, visc);
            }
            for (int k = -K; k <= K; k++) {
                const int i = wrapm(c + k, nh);
                b += wd[k + K] * bst(xs[wrapm(i - 1, nh)], xs[i], h, visc);
            }
            fsu[J] = a;
            vsu[J] = b;
            vu[J] = Du(f);
            sVU[J] = bst(Du(wrapm(c - 1, nh)), Du(c), h, visc);
        }
        // the DNS step (src/Burgers.jl:437-438), new values kept in registers
        double un[FE_PT];
#pragma unroll
        for (int p = 0; p < FE_PT; p++) {
            const int i = tid + p * FE_NT;
            if (i < nh) {
                const double s0 = bst(xs[wrapm(i - 1, nh)], xs[i], h, visc);
                const double s1 = bst(xs[i], xs[wrapm(i + 1, nh)], h, visc);
                un[p] = xs[i] + dt * ((s1 - s0) / h);
            }
        }
        __syncthreads();
        // the four LES steps (stepfromstress!, src/Burgers.jl:338-347)
        for (int m = 0; m < 4; m++) {
            double *ul = les[m];
            for (int J = tid; J < nH; J += FE_NT) {
                const int Jm = J == 0 ? nH - 1 : J - 1;
                double sg = bst(ul[Jm], ul[J], H, visc);
                if (m == 1) {
                    sg = sg + vsu[J] - sVU[J];
                } else if (m > 1) {
                    const double svu = bst(vu[Jm], vu[J], H, visc);
                    sg = sg + (m == 2 ? vsu[J] : fsu[J]) - svu;
                }
                sig[J] = sg;
            }
            __syncthreads();
            for (int J = tid; J < nH; J += FE_NT) {
                const int Jp = J + 1 == nH ? 0 : J + 1;
                ul[J] += dt * ((sig[Jp] - sig[J]) / H);
            }
            __syncthreads();
        }
#pragma unroll
        for (int p = 0; p < FE_PT; p++) {
            const int i = tid + p * FE_NT;
            if (i < nh) xs[i] = un[p];
        }
        __syncthreads();
    }
    // outputs: dns_ref, dns_ref_fil = D * u, dns_fil = its coarse faces, the LES
    double *o = A.out + (size_t)q * (2 * (size_t)nh + 5 * (size_t)nH);
    for (int i = tid; i < nh; i += FE_NT) {
        o[i] = xs[i];
        o[nh + i] = Du(i);
    }
    for (int J = tid; J < nH; J += FE_NT) {
        o[2 * nh + J] = Du(J * comp + comp - 1);
        for (int m = 0; m < 4; m++) o[2 * nh + (m + 1) * nH + J] = les[m][J];
    }
    if (A.steps && tid == 0) A.steps[q] = nstep;
}

}  // namespace

extern "C" int ec_burgers_randomfield(ec_ctx *ctx, int n, double L, double kpeak, double totalenergy, int nsamples,
                                      uint64_t seed0, void *u) {
    EC_REQUIRE(ctx, ctx && u, "ec_burgers_randomfield: null argument");
    EC_REQUIRE(ctx, n >= 4 && nsamples >= 1 && kpeak > 0 && totalenergy > 0 && L > 0,
               "ec_burgers_randomfield: bad argument");
    const int kc = std::min(n / 2 - 1, (int)std::ceil(6 * kpeak));
    k_bump_field<<<nsamples, FE_NT, 0, ctx->stream>>>(n, L, kpeak, totalenergy, kc, seed0, (double *)u);
    EC_CHECK_LAUNCH(ctx, "k_bump_field");
    return EC_OK;
}

extern "C" int ec_burgers_dns_aided_les(ec_ctx *ctx, int nh, int nH, double L, double visc, double tstop,
                                        double cfl_factor, int nsamples, const void *ustart, int Rl, const double *wl,
                                        int K, const double *wd, void *out, int *steps) {
    EC_REQUIRE(ctx, ctx && ustart && wl && wd && out, "ec_burgers_dns_aided_les: null argument");
    EC_REQUIRE(ctx, nH >= 2 && nh % nH == 0, "ec_burgers_dns_aided_les: nH=%d must divide nh=%d", nH, nh);
    EC_REQUIRE(ctx, nh <= FE_NT * FE_PT, "ec_burgers_dns_aided_les: nh=%d beyond %d points", nh, FE_NT * FE_PT);
    EC_REQUIRE(ctx, Rl >= 0 && K >= 0 && 2 * K + 1 <= nh && 2 * Rl + 1 <= nh,
               "ec_burgers_dns_aided_les: kernels wider than the grid");
    EC_REQUIRE(ctx, nsamples >= 1 && tstop >= 0 && cfl_factor > 0 && visc > 0, "ec_burgers_dns_aided_les: bad argument");
    // weights and the per-sample coarse work area; the call returns when done
    const size_t nw = (size_t)(2 * Rl + 1) + (2 * K + 1), nwork = (size_t)nsamples * 9 * nH;
    double *buf = nullptr;
    int *dsteps = nullptr;
    EC_HIP(ctx, hipMalloc(&buf, sizeof(double) * (nw + nwork)));
    std::vector<double> w(nw);
    std::copy(wl, wl + 2 * Rl + 1, w.begin());
    std::copy(wd, wd + 2 * K + 1, w.begin() + 2 * Rl + 1);
    hipError_t e = hipMemcpyAsync(buf, w.data(), sizeof(double) * nw, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && steps) e = hipMalloc(&dsteps, sizeof(int) * nsamples);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(buf);
        return ec_fail(ctx, EC_E_HIP, "ec_burgers_dns_aided_les: setup failed: %s", hipGetErrorString(e));
    }
    FeArgs A;
    A.nh = nh;
    A.nH = nH;
    A.Rl = Rl;
    A.K = K;
    A.L = L;
    A.visc = visc;
    A.tstop = tstop;
    A.cfl = cfl_factor;
    A.wl = buf;
    A.wd = buf + 2 * Rl + 1;
    A.ustart = (const double *)ustart;
    A.work = buf + nw;
    A.out = (double *)out;
    A.steps = dsteps;
    const size_t lds = sizeof(double) * (size_t)nh;
    int rc = EC_OK;
    if (hipFuncSetAttribute((const void *)k_burgers_dnsaid_fe, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess)
        rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_dns_aided_les: LDS attribute");
    if (!rc) {
        k_burgers_dnsaid_fe<<<nsamples, FE_NT, lds, ctx->stream>>>(A);
        if (hipGetLastError() != hipSuccess) rc = ec_fail(ctx, EC_E_HIP, "launch of k_burgers_dnsaid_fe failed");
    }
    if (!rc && steps &&
        hipMemcpyAsync(steps, dsteps, sizeof(int) * nsamples, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_dns_aided_les: step count copy");
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc)
        rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_dns_aided_les: kernel failed");
    (void)hipFree(buf);
    if (dsteps) (void)hipFree(dsteps);
    return rc;
}
