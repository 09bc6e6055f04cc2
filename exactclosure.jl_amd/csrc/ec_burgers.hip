// Batched 1D viscous Burgers DNS-aided LES (src/Burgers.jl:89-142, 350-445).
//
// The reference advances one instance at a time with ~12 tiny launches per
// step (src/Burgers.jl:307-324, 395-439), so it is launch-latency bound.
// Here one persistent launch advances every instance of a heterogeneous
// batch: workgroups pull instances from an atomic work queue ordered by
// size (largest first), keep the stage state of the DNS and of the two LES
// variants in LDS, keep the RK4 base state and accumulator in registers,
// and evaluate the exact-closure commutator tau = Pc(s) - s_H(V u) in FP64 at
// every stage.  The launch can be captured once into a hipGraph and replayed.
#include "ec_internal.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

namespace {

struct BInst {
    int n, c, off, offl;
};

__device__ __forceinline__ uint64_t sm64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Initial condition: sum_{k=1}^{kmax} (2/k) cos(2 pi (k i / n + phase_k)), the
// explicit form of irfft of modes with amplitude 1/k and uniform phase
// (randomfield, src/Burgers.jl:195-202).
template <typename T>
__global__ void k_burgers_ic(const BInst *__restrict__ inst, int count, int kmax, uint64_t seed0, T *__restrict__ U) {
    const int b = blockIdx.y;
    if (b >= count) return;
    const BInst I = inst[b];
    const uint64_t seed = seed0 + (uint64_t)b;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < I.n; i += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int k = 1; k <= kmax; k++) {
            const double ph = (double)(sm64(seed ^ sm64((uint64_t)k)) >> 11) * 0x1.0p-53;
            const long m = ((long)k * i) % I.n;
            s += 2.0 / k * cos(2.0 * M_PI * ((double)m / I.n + ph));
        }
        U[I.off + i] = (T)s;
    }
}

// Volume-average filter of the DNS face values at coarse face J (odd c:
// top-hat of c faces; even c: trapezoid over c+1 faces).  `at` maps a point
// index in [0, n) to its storage index.  A window starts at f - h >= c / 2 - 1
// >= 0 and ends at f + h < 2 n, so only its top end wraps, once.
template <typename T, typename At>
__device__ __forceinline__ double vavg(const T *u, int n, int c, int J, At at) {
    const int f = c * (J + 1) - 1;
    const int h = c / 2;
    double s = 0.0;
    if (c & 1) {
#pragma unroll 8
        for (int t = -h; t <= h; t++) {
            int m = f + t;
            m = m >= n ? m - n : m;
            s += (double)u[at(m)];
        }
        return s / c;
    }
#pragma unroll 8
    for (int t = -h + 1; t <= h - 1; t++) {
        int m = f + t;
        m = m >= n ? m - n : m;
        s += (double)u[at(m)];
    }
    int a = f - h, b = f + h;
    b = b >= n ? b - n : b;
    return (s + 0.5 * ((double)u[at(a)] + (double)u[at(b)])) / c;
}

// The DNS stage state in LDS skips one word after every 2^kPadShift points:
// point m is stored at pad_at(m) = m + (m >> kPadShift).  The face averages
// and the commutator read it at a stride of c points, which would put the
// lanes of a wave on a few LDS banks for power-of-two c; with the skip they
// fall on distinct banks.  pad_at(-1) = -2 and pad_at(n) hold ghost copies
// of points n - 1 and 0.
constexpr int kPadShift = 5;
__device__ __forceinline__ int pad_at(int m) { return m + (m >> kPadShift); }

template <typename T>
__global__ void k_burgers_les_init(const BInst *__restrict__ inst, int count, const T *__restrict__ U,
                                   T *__restrict__ UL, T *__restrict__ UM) {
    const int b = blockIdx.y;
    if (b >= count) return;
    const BInst I = inst[b];
    const int nH = I.n / I.c;
    for (int J = blockIdx.x * blockDim.x + threadIdx.x; J < nH; J += gridDim.x * blockDim.x) {
        const T v = (T)vavg(U + I.off, I.n, I.c, J, [](int m) { return m; });
        UL[I.offl + J] = v;
        UM[I.offl + J] = v;
    }
}

// x / y correctly rounded, from r = RN(1 / y) and two corrections with exact
// FMA residuals: x r is within two ulps of x / y, the first correction makes
// it faithful, and then q + (x - y q) r rounds to RN(x / y) (Markstein,
// IBM J. Res. Dev. 34 (1990) 111-119).  For finite x and normal y, r, x / y.
__device__ __forceinline__ double div_by(double x, double y, double r) {
    double q = x * r;
    double e = fma(-q, y, x);
    q = fma(e, r, q);
    e = fma(-q, y, x);
    return fma(e, r, q);
}

template <typename T>
__device__ __forceinline__ T bstress(T a, T b, T visc, T ih) {  // stress (src/Burgers.jl:89-94)
    return -((a + b) * (a + b)) / 8 + visc * (b - a) * ih;
}

__global__ void k_reset_counter(int *counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0;
}

// PTL: LES points per thread, PT / 2 (any compression c >= 2) or, where the
// batch's largest LES grid allows, PT / 4 (c >= 4), which frees registers.
template <typename T, int PT, int NT, int PTL>
__global__ void __launch_bounds__(NT) k_burgers_persistent(const BInst *__restrict__ inst, const int *__restrict__ order,
                                                            int count, int *__restrict__ counter, T *__restrict__ U,
                                                            T *__restrict__ UL, T *__restrict__ UM,
                                                            double *__restrict__ dts, double L, double visc,
                                                            double cfl, int nsteps) {
    extern __shared__ char smem_raw[];
    __shared__ int s_idx;
    __shared__ double s_red[16];
    // the LES face fluxes of lane 0 of each wave, for lane 63 of the wave before
    constexpr int NWV = NT / 64;
    __shared__ T s_xs[NWV * PTL], s_xm[NWV * PTL];
    const int tid = threadIdx.x;
    constexpr int nthr = NT;  // the block size (launched with NT threads): point offsets are compile-time
    for (;;) {
        if (tid == 0) s_idx = atomicAdd(counter, 1);
        __syncthreads();
        const int w = s_idx;
        __syncthreads();
        if (w >= count) break;
        const int id = order[w];
        const BInst I = inst[id];
        const int n = I.n, c = I.c, nH = n / c;
        // coarse face averages V u (FP64) first, so every array stays aligned;
        // then the DNS stage state, point m at us[pad_at(m)], with ghost
        // copies of points n - 1 and 0 at pad_at(-1) and pad_at(n), so the
        // stencil reads need no wrap-around selects
        double *vbs = reinterpret_cast<double *>(smem_raw);
        T *us = reinterpret_cast<T *>(vbs + nH) + 2;
        T *ul = us + pad_at(n) + 1;
        T *um = ul + nH;
        // this thread's points tid + q NT sit at us[ps + q * PSTR]; their left
        // and right neighbours one word further, or two where the skipped
        // word lies between (every 2^kPadShift-th lane)
        static_assert(NT % (1 << kPadShift) == 0, "block size a multiple of the pad period");
        constexpr int PSTR = NT + (NT >> kPadShift);
        constexpr int PMASK = (1 << kPadShift) - 1;
        const int ps = pad_at(tid);
        T *const usq = us + ps;
        const T *const usl = usq - ((tid & PMASK) == 0 ? 2 : 1);
        const T *const usr = usq + ((tid & PMASK) == PMASK ? 2 : 1);
        T *const gl = us + pad_at(-1), *const gr = us + pad_at(n);
        const double h = L / n, H = L / nH;
        const double rh = 1.0 / h, rH = 1.0 / H;
        const T ih = (T)(1.0 / h), iH = (T)(1.0 / H), tv = (T)visc;

        T u0[PT], acc[PT];
        T l0[PTL], lacc[PTL], m0[PTL], macc[PTL];
        double umax = 0.0;
#pragma unroll
        for (int q = 0; q < PT; q++) {
            const int i = tid + q * nthr;
            if (i < n) {
                u0[q] = U[I.off + i];
                usq[q * PSTR] = u0[q];
                if (i == n - 1) *gl = u0[q];
                if (i == 0) *gr = u0[q];
                umax = fmax(umax, fabs((double)u0[q]));
            }
        }
#pragma unroll
        for (int q = 0; q < PTL; q++) {
            const int J = tid + q * nthr;
            if (J < nH) {
                l0[q] = UL[I.offl + J];
                m0[q] = UM[I.offl + J];
                ul[J] = l0[q];
                um[J] = m0[q];
            }
        }
        // time step from the initial state: cfl * min(h / max|u|, h^2 / visc)
        // (cfl, src/Burgers.jl:97-101), fixed for the instance
        for (int o = 32; o > 0; o >>= 1) umax = fmax(umax, __shfl_xor(umax, o, 64));
        if ((tid & 63) == 0) s_red[tid >> 6] = umax;
        __syncthreads();
        double dt = dts[id];
        if (dt <= 0.0) {
            double m = 0.0;
            for (int k = 0; k < (nthr >> 6); k++) m = fmax(m, s_red[k]);
            dt = cfl * fmin(h / m, h * h / visc);
        }
        __syncthreads();
        if (tid == 0) dts[id] = dt;
        const T tdt = (T)dt;

        for (int step = 0; step < nsteps; step++) {
#pragma unroll
            for (int st = 0; st < 4; st++) {
                // --- coarse face averages of the stage state, once per face (the
                // commutator of coarse point J reads faces J - 1 and J)
                for (int J = tid; J < nH; J += nthr) vbs[J] = vavg(us, n, c, J, pad_at);
                __syncthreads();
                // --- closure commutator in FP64 at coarse point Jx: tau = Pc(s) - s_H(V u)
                auto tau_at = [&](int Jx) {
                    const double vb = vbs[Jx], va = vbs[Jx == 0 ? nH - 1 : Jx - 1];
                    const int m0i = c * (Jx + 1) - 1 - c / 2;
                    // m lies in [0, n) (c / 2 - 1 <= m0i, m0i + 1 <= n - c / 2);
                    // point m - 1 = -1 is the ghost copy of point n - 1
                    auto sfine = [&](int m) {
                        const double a = (double)us[pad_at(m - 1)], b = (double)us[pad_at(m)];
                        return -((a + b) * (a + b)) / 8 + div_by(visc * (b - a), h, rh);
                    };
                    const double fsu = (c & 1) ? sfine(m0i) : 0.5 * (sfine(m0i) + sfine(m0i + 1));
                    const double svu = -((va + vb) * (va + vb)) / 8 + div_by(visc * (vb - va), H, rH);
                    return (T)(fsu - svu);
                };
                // --- LES face fluxes, once per coarse face J (between coarse points
                // J - 1 and J): exact closure fs = s_H(ul) + tau, no model fm = s_H(um).
                // The right-hand side at J is (flux(J + 1) - flux(J)) / H.
                T fs[PTL], fm[PTL];
#pragma unroll
                for (int q = 0; q < PTL; q++) {
                    const int J = tid + q * nthr;
                    fs[q] = T(0);
                    fm[q] = T(0);
                    if (J < nH) {
                        const int Jm = J == 0 ? nH - 1 : J - 1;
                        fs[q] = bstress(ul[Jm], ul[J], tv, iH) + tau_at(J);
                        fm[q] = bstress(um[Jm], um[J], tv, iH);
                        if ((tid & 63) == 0) {
                            s_xs[(tid >> 6) * PTL + q] = fs[q];
                            s_xm[(tid >> 6) * PTL + q] = fm[q];
                        }
                    }
                }
                // --- DNS right-hand side, after the LES ones so that it is not live
                // beside the commutator's temporaries (rhs!, src/Burgers.jl:103-109)
                T kd[PT];
#pragma unroll
                for (int q = 0; q < PT; q++) {
                    const int i = tid + q * nthr;
                    if (i < n) {
                        const T a = usl[q * PSTR], b = usq[q * PSTR], cc = usr[q * PSTR];
                        kd[q] = (bstress(b, cc, tv, ih) - bstress(a, b, tv, ih)) * ih;
                    }
                }
                __syncthreads();
                // --- LES right-hand sides: flux(J + 1) is lane + 1's, except for the
                // last lane of a wave (lane 0 of the next wave, or thread 0 at the
                // next q) and for J = nH - 1 (face nH is face 0: thread 0, q = 0)
                T kl[PTL], km[PTL];
#pragma unroll
                for (int q = 0; q < PTL; q++) {
                    const int J = tid + q * nthr;
                    T ns = __shfl_down(fs[q], 1, 64), nm = __shfl_down(fm[q], 1, 64);
                    if (J < nH) {
                        if (J + 1 == nH) {
                            ns = s_xs[0];
                            nm = s_xm[0];
                        } else if ((tid & 63) == 63) {
                            const int x = tid + 1 == NT ? q + 1 : ((tid + 1) >> 6) * PTL + q;
                            ns = s_xs[x];
                            nm = s_xm[x];
                        }
                        kl[q] = (ns - fs[q]) * iH;
                        km[q] = (nm - fm[q]) * iH;
                    }
                }
                // --- RK4 bookkeeping (rk4_timestep!, src/Burgers.jl:120-142)
                const T bw = (st == 0 || st == 3) ? T(1) : T(2);
                const T a_next = st == 2 ? tdt : tdt / 2;
#pragma unroll
                for (int q = 0; q < PT; q++) {
                    const int i = tid + q * nthr;
                    if (i < n) {
                        acc[q] = st == 0 ? kd[q] : acc[q] + bw * kd[q];
                        T v;
                        if (st < 3) {
                            v = u0[q] + a_next * kd[q];
                        } else {
                            u0[q] = u0[q] + tdt / 6 * acc[q];
                            v = u0[q];
                        }
                        usq[q * PSTR] = v;
                        if (i == n - 1) *gl = v;
                        if (i == 0) *gr = v;
                    }
                }
#pragma unroll
                for (int q = 0; q < PTL; q++) {
                    const int J = tid + q * nthr;
                    if (J < nH) {
                        lacc[q] = st == 0 ? kl[q] : lacc[q] + bw * kl[q];
                        macc[q] = st == 0 ? km[q] : macc[q] + bw * km[q];
                        if (st < 3) {
                            ul[J] = l0[q] + a_next * kl[q];
                            um[J] = m0[q] + a_next * km[q];
                        } else {
                            l0[q] = l0[q] + tdt / 6 * lacc[q];
                            m0[q] = m0[q] + tdt / 6 * macc[q];
                            ul[J] = l0[q];
                            um[J] = m0[q];
                        }
                    }
                }
                __syncthreads();
            }
        }
#pragma unroll
        for (int q = 0; q < PT; q++) {
            const int i = tid + q * nthr;
            if (i < n) U[I.off + i] = u0[q];
        }
#pragma unroll
        for (int q = 0; q < PTL; q++) {
            const int J = tid + q * nthr;
            if (J < nH) {
                UL[I.offl + J] = l0[q];
                UM[I.offl + J] = m0[q];
            }
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void k_burgers_rhs(int n, T ih, T visc, const T *__restrict__ u, T *__restrict__ du) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T a = u[i == 0 ? n - 1 : i - 1], b = u[i], c = u[i + 1 == n ? 0 : i + 1];
    du[i] = (bstress(b, c, visc, ih) - bstress(a, b, visc, ih)) * ih;
}

}  // namespace

struct ec_burgers_batch {
    int dtype = EC_F32;
    int count = 0;
    double L = 0, visc = 0, cfl = 0;
    int threads = 512;
    int nmax = 0, nHmax = 0;
    std::vector<BInst> host_inst;
    BInst *inst = nullptr;
    int *order = nullptr;
    int *counter = nullptr;
    void *U = nullptr, *UL = nullptr, *UM = nullptr;
    double *dts = nullptr;
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;  // (nsteps, threads) -> graph
    hipStream_t cap_stream = nullptr;                      // private stream used for capture
};

static int pt_for(int nmax, int threads) {
    int pt = 1;
    while (pt * threads < nmax) pt *= 2;
    return pt;
}

// the largest points-per-thread the persistent kernel is built with for a block size
static int pt_max(int threads) { return std::min(64, 32768 / threads); }

template <typename T>
static size_t batch_lds(const ec_burgers_batch *b) {
    // V u at the coarse faces (FP64), then the DNS stage state (the skipped
    // words, a ghost point at either end) and the two LES stage states
    const size_t dns = (size_t)b->nmax + ((size_t)b->nmax >> kPadShift) + 3;
    return sizeof(double) * (size_t)b->nHmax + sizeof(T) * (dns + 2 * (size_t)b->nHmax);
}

template <typename T, int PT, int NT, int PTL>
static int launch_ptl(ec_ctx *ctx, ec_burgers_batch *b, int nsteps, int blocks, size_t lds, hipStream_t st) {
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_burgers_persistent<T, PT, NT, PTL>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_burgers_persistent<T, PT, NT, PTL><<<blocks, NT, lds, st>>>(
        b->inst, b->order, b->count, b->counter, (T *)b->U, (T *)b->UL, (T *)b->UM, b->dts, b->L, b->visc, b->cfl,
        nsteps);
    EC_CHECK_LAUNCH(ctx, "k_burgers_persistent");
    return EC_OK;
}

template <typename T, int PT, int NT>
static int launch_pt(ec_ctx *ctx, ec_burgers_batch *b, int nsteps, int blocks, size_t lds, hipStream_t st) {
    constexpr int WIDE = PT > 1 ? PT / 2 : 1, NARROW = PT > 3 ? PT / 4 : WIDE;
    if (NARROW < WIDE && b->nHmax <= NARROW * NT) return launch_ptl<T, PT, NT, NARROW>(ctx, b, nsteps, blocks, lds, st);
    return launch_ptl<T, PT, NT, WIDE>(ctx, b, nsteps, blocks, lds, st);
}

// Points per thread PT (a power of two with PT * NT >= the largest instance);
// PT * NT <= 32768 covers instances up to 16384 points at every block size.
template <typename T, int NT>
static int dispatch_pt(ec_ctx *ctx, ec_burgers_batch *b, int nsteps, int blocks, size_t lds, hipStream_t st) {
    const int pt = pt_for(b->nmax, NT);
    if (pt > pt_max(NT)) return ec_fail(ctx, EC_E_UNSUPPORTED, "instance size %d too large for %d threads", b->nmax, NT);
    switch (pt) {
        case 1: return launch_pt<T, 1, NT>(ctx, b, nsteps, blocks, lds, st);
        case 2: return launch_pt<T, 2, NT>(ctx, b, nsteps, blocks, lds, st);
        case 4: return launch_pt<T, 4, NT>(ctx, b, nsteps, blocks, lds, st);
        case 8: return launch_pt<T, 8, NT>(ctx, b, nsteps, blocks, lds, st);
        case 16: return launch_pt<T, 16, NT>(ctx, b, nsteps, blocks, lds, st);
        case 32: return launch_pt<T, (32 * NT <= 32768 ? 32 : 1), NT>(ctx, b, nsteps, blocks, lds, st);
        default: return launch_pt<T, (64 * NT <= 32768 ? 64 : 1), NT>(ctx, b, nsteps, blocks, lds, st);
    }
}

template <typename T>
static int launch_run(ec_ctx *ctx, ec_burgers_batch *b, int nsteps, hipStream_t st) {
    const size_t lds = batch_lds<T>(b);
    int dev_cus = 0;
    EC_HIP(ctx, hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const int per_cu = std::max(1, (int)((160 * 1024) / std::max<size_t>(lds + 1024, 1)));
    const int blocks = std::min(b->count, dev_cus * std::min(per_cu, std::max(1, 2048 / b->threads)));
    k_reset_counter<<<1, 64, 0, st>>>(b->counter);
    EC_CHECK_LAUNCH(ctx, "k_reset_counter");
    switch (b->threads) {
        case 128: return dispatch_pt<T, 128>(ctx, b, nsteps, blocks, lds, st);
        case 256: return dispatch_pt<T, 256>(ctx, b, nsteps, blocks, lds, st);
        case 512: return dispatch_pt<T, 512>(ctx, b, nsteps, blocks, lds, st);
        default: return dispatch_pt<T, 1024>(ctx, b, nsteps, blocks, lds, st);
    }
}

extern "C" int ec_burgers_batch_create(ec_ctx *ctx, int dtype, int count, const int *n, const int *c, double L,
                                       double visc, double cfl, int kmax, uint64_t seed, ec_burgers_batch **out) {
    EC_REQUIRE(ctx, ctx && out && n && c && count > 0, "ec_burgers_batch_create: bad argument");
    *out = nullptr;
    EC_REQUIRE(ctx, dtype == EC_F32 || dtype == EC_F64, "bad dtype");
    ec_burgers_batch *b = new ec_burgers_batch();
    b->dtype = dtype;
    b->count = count;
    b->L = L;
    b->visc = visc;
    b->cfl = cfl;
    size_t tot = 0, totl = 0;
    for (int i = 0; i < count; i++) {
        if (n[i] < 4 || c[i] < 2 || n[i] % c[i] != 0 || n[i] / c[i] < 2) {
            delete b;
            return ec_fail(ctx, EC_E_INVALID, "instance %d: n=%d, c=%d (need c >= 2 dividing n, n/c >= 2)", i, n[i], c[i]);
        }
        b->host_inst.push_back(BInst{n[i], c[i], (int)tot, (int)totl});
        tot += n[i];
        totl += n[i] / c[i];
        b->nmax = std::max(b->nmax, n[i]);
        b->nHmax = std::max(b->nHmax, n[i] / c[i]);
    }
    const size_t bt = dtype == EC_F64 ? 8 : 4;
    const size_t lds = dtype == EC_F64 ? batch_lds<double>(b) : batch_lds<float>(b);
    if (lds > 160 * 1024 - 256) {
        delete b;
        return ec_fail(ctx, EC_E_UNSUPPORTED, "batch needs %zu B of LDS per instance (limit 160 KiB)", lds);
    }
    // largest instances first (longest-processing-time order for the queue)
    std::vector<int> order(count);
    for (int i = 0; i < count; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int bb) { return n[a] > n[bb]; });
    hipError_t e = hipMalloc(&b->inst, sizeof(BInst) * count);
    if (e == hipSuccess) e = hipMalloc(&b->order, sizeof(int) * count);
    if (e == hipSuccess) e = hipMalloc(&b->counter, sizeof(int) * 64);
    if (e == hipSuccess) e = hipMalloc(&b->U, bt * tot);
    if (e == hipSuccess) e = hipMalloc(&b->UL, bt * totl);
    if (e == hipSuccess) e = hipMalloc(&b->UM, bt * totl);
    if (e == hipSuccess) e = hipMalloc(&b->dts, sizeof(double) * count);
    if (e == hipSuccess) e = hipMemcpy(b->inst, b->host_inst.data(), sizeof(BInst) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->order, order.data(), sizeof(int) * count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b->dts, 0, sizeof(double) * count);
    if (e != hipSuccess) {
        ec_burgers_batch_destroy(ctx, b);
        return ec_fail(ctx, EC_E_OOM, "ec_burgers_batch_create: %s", hipGetErrorString(e));
    }
    dim3 grid(std::max(1, (b->nmax + 255) / 256), count);
    if (dtype == EC_F64) {
        k_burgers_ic<double><<<grid, 256, 0, ctx->stream>>>(b->inst, count, kmax, seed, (double *)b->U);
        k_burgers_les_init<double><<<grid, 256, 0, ctx->stream>>>(b->inst, count, (const double *)b->U, (double *)b->UL,
                                                                  (double *)b->UM);
    } else {
        k_burgers_ic<float><<<grid, 256, 0, ctx->stream>>>(b->inst, count, kmax, seed, (float *)b->U);
        k_burgers_les_init<float><<<grid, 256, 0, ctx->stream>>>(b->inst, count, (const float *)b->U, (float *)b->UL,
                                                                 (float *)b->UM);
    }
    EC_CHECK_LAUNCH(ctx, "k_burgers_ic");
    *out = b;
    return EC_OK;
}

extern "C" int ec_burgers_batch_destroy(ec_ctx *ctx, ec_burgers_batch *b) {
    if (!b) return EC_OK;
    if (ctx) hipStreamSynchronize(ctx->stream);
    for (auto &kv : b->graphs) hipGraphExecDestroy(kv.second);
    if (b->cap_stream) hipStreamDestroy(b->cap_stream);
    hipFree(b->inst);
    hipFree(b->order);
    hipFree(b->counter);
    hipFree(b->U);
    hipFree(b->UL);
    hipFree(b->UM);
    hipFree(b->dts);
    delete b;
    return EC_OK;
}

extern "C" int ec_burgers_batch_set_threads(ec_ctx *ctx, ec_burgers_batch *b, int threads) {
    EC_REQUIRE(ctx, ctx && b, "null argument");
    EC_REQUIRE(ctx, threads == 128 || threads == 256 || threads == 512 || threads == 1024, "threads=%d", threads);
    EC_REQUIRE(ctx, pt_for(b->nmax, threads) <= pt_max(threads),
               "threads=%d: n=%d needs more than %d points per thread", threads, b->nmax, pt_max(threads));
    b->threads = threads;
    return EC_OK;
}

extern "C" int ec_burgers_batch_run(ec_ctx *ctx, ec_burgers_batch *b, int nsteps, int use_graph) {
    EC_REQUIRE(ctx, ctx && b, "null argument");
    EC_REQUIRE(ctx, nsteps >= 0, "nsteps=%d", nsteps);
    if (nsteps == 0) return EC_OK;
    auto run = [&](hipStream_t st) {
        return b->dtype == EC_F64 ? launch_run<double>(ctx, b, nsteps, st) : launch_run<float>(ctx, b, nsteps, st);
    };
    if (!use_graph) return run(ctx->stream);
    auto key = std::make_pair(nsteps, b->threads);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        // Captured on a private stream (the legacy default stream cannot be
        // captured); the instantiated graph is then launched on the context
        // stream, ordered with the caller's other work.
        if (!b->cap_stream) EC_HIP(ctx, hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        hipGraph_t g;
        EC_HIP(ctx, hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = run(b->cap_stream);
        hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
        if (e != hipSuccess) return rc ? rc : ec_fail(ctx, EC_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
        if (rc) {
            hipGraphDestroy(g);
            return rc;
        }
        hipGraphExec_t ge;
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return ec_fail(ctx, EC_E_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
        it = b->graphs.emplace(key, ge).first;
    }
    EC_HIP(ctx, hipGraphLaunch(it->second, ctx->stream));
    return EC_OK;
}

extern "C" int ec_burgers_batch_get(ec_ctx *ctx, ec_burgers_batch *b, int instance, int which, double *out, int cap) {
    EC_REQUIRE(ctx, ctx && b && out, "null argument");
    EC_REQUIRE(ctx, instance >= 0 && instance < b->count, "instance %d out of range", instance);
    EC_REQUIRE(ctx, which >= 0 && which <= 2, "which=%d", which);
    const BInst &I = b->host_inst[instance];
    const int len = which == 0 ? I.n : I.n / I.c;
    EC_REQUIRE(ctx, cap >= len, "output capacity %d < %d", cap, len);
    const size_t bt = b->dtype == EC_F64 ? 8 : 4;
    const char *base = (const char *)(which == 0 ? b->U : (which == 1 ? b->UL : b->UM));
    const size_t off = which == 0 ? I.off : I.offl;
    std::vector<char> tmp(bt * len);
    EC_HIP(ctx, hipMemcpyAsync(tmp.data(), base + bt * off, bt * len, hipMemcpyDeviceToHost, ctx->stream));
    EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < len; i++)
        out[i] = b->dtype == EC_F64 ? ((const double *)tmp.data())[i] : (double)((const float *)tmp.data())[i];
    return EC_OK;
}

extern "C" int ec_burgers_batch_dt(ec_ctx *ctx, ec_burgers_batch *b, int instance, double *dt) {
    EC_REQUIRE(ctx, ctx && b && dt, "ec_burgers_batch_dt: null argument");
    EC_REQUIRE(ctx, instance >= 0 && instance < b->count, "instance %d out of range", instance);
    // ordered after the persistent kernel that writes dts on the context stream
    EC_HIP(ctx, hipMemcpyAsync(dt, b->dts + instance, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EC_OK;
}

extern "C" int ec_burgers_rhs(ec_ctx *ctx, int dtype, int n, double L, double visc, const void *u, void *du) {
    EC_REQUIRE(ctx, ctx && u && du && n >= 2, "ec_burgers_rhs: bad argument");
    const double ih = n / L;
    if (dtype == EC_F64)
        k_burgers_rhs<double><<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, ih, visc, (const double *)u, (double *)du);
    else
        k_burgers_rhs<float><<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, (float)ih, (float)visc, (const float *)u,
                                                                       (float *)du);
    EC_CHECK_LAUNCH(ctx, "k_burgers_rhs");
    return EC_OK;
}

// ---- closure fractions (compute_fractions, src/Burgers.jl:1022-1099) ----
// For DNS snapshots u (n points, samples stored one after another) with the
// LES kernel F = (Rl, wl) and the double kernel D = (K, wd) (the finite-volume
// top-hat convolved with F, build_doublekernel), per coarse point J:
//   fsu = C_c(F * s), vsu = C_c(D * s), vu = C_f(D * u), svu = s_H(vu),
//   sVU = C_c(s_h(D * u)),  s = s_h(u) the DNS stress,
//   classic = vsu - sVU, flux = sVU - svu, div = fsu - vsu,
//   f_x = |x| / (|classic| + |flux| + |div| + eps),
// with C_c sampling fine centre (J+1) comp - comp/2 - 1 and C_f fine face
// (J+1) comp - 1 (0-based forms of lazycoarsegrain, src/Burgers.jl:174-180).
// All FP64.  The fine pass forms s and D * u; the coarse pass one block per
// sample, the three fraction sums reduced in a fixed order.
__device__ __forceinline__ int bwrap(int i, int n) { return ((i % n) + n) % n; }

__device__ __forceinline__ double bstress(double ua, double ub, double ih, double visc) {
    return -((ua + ub) * (ua + ub)) / 8 + visc * (ub - ua) * ih;
}

__global__ void k_bfrac_fine(int n, int ns, double ih, double visc, const double *__restrict__ u, int K,
                             const double *__restrict__ wd, double *__restrict__ s, double *__restrict__ vu_f) {
    const long total = (long)n * ns;
    for (long I = blockIdx.x * (long)blockDim.x + threadIdx.x; I < total; I += (long)gridDim.x * blockDim.x) {
        const int i = (int)(I % n);
        const double *us = u + (I - i);
        s[I] = bstress(us[bwrap(i - 1, n)], us[i], ih, visc);
        double acc = 0.0;
        for (int k = -K; k <= K; k++) acc += wd[k + K] * us[bwrap(i + k, n)];
        vu_f[I] = acc;
    }
}

// The same fine pass with the block's stretch of the sample row staged in LDS:
// block (x, q) writes points [x TILE, x TILE + TILE) of sample q from LDS copies
// of u[i - K - 1 .. i + K]; each output sums its taps in the same order as
// k_bfrac_fine, so the results are the same, with one global load per staged
// value instead of one per tap.
constexpr int BF_TILE = 1024;
constexpr int BF_LDS = 8192;  // doubles of staged row per block (64 KB)

__global__ void __launch_bounds__(256) k_bfrac_fine_t(int n, double ih, double visc, const double *__restrict__ u,
                                                      int K, const double *__restrict__ wd, double *__restrict__ s,
                                                      double *__restrict__ vu_f) {
    extern __shared__ double xs[];
    const int i0 = blockIdx.x * BF_TILE, cnt = min(BF_TILE, n - i0);
    const long row = (long)blockIdx.y * n;
    const double *us = u + row;
    const int len = cnt + 2 * K + 1;  // xs[j] = u[i0 - K - 1 + j]
    for (int j = threadIdx.x; j < len; j += 256) xs[j] = us[bwrap(i0 - K - 1 + j, n)];
    __syncthreads();
    for (int t = threadIdx.x; t < cnt; t += 256) {
        const int o = t + K + 1;  // xs index of point i0 + t
        s[row + i0 + t] = bstress(xs[o - 1], xs[o], ih, visc);
        double acc = 0.0;
        for (int k = -K; k <= K; k++) acc += wd[k + K] * xs[o + k];
        vu_f[row + i0 + t] = acc;
    }
}

// s = s_h(u) and vu_f = D * u for every fine point of every sample
static void bfrac_fine(ec_ctx *ctx, int n, int ns, double ih, double visc, const double *u, int K,
                       const double *wd, double *s, double *vf) {
    if (BF_TILE + 2 * K + 1 <= BF_LDS && ns <= 65535) {
        const dim3 grid((unsigned)((n + BF_TILE - 1) / BF_TILE), (unsigned)ns);
        const size_t lds = sizeof(double) * (size_t)(std::min(BF_TILE, n) + 2 * K + 1);
        // a failed attribute call shows up as a failed launch, which the callers check
        (void)hipFuncSetAttribute((const void *)k_bfrac_fine_t, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        k_bfrac_fine_t<<<grid, 256, lds, ctx->stream>>>(n, ih, visc, u, K, wd, s, vf);
    } else {
        const long total = (long)n * ns;
        const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
        k_bfrac_fine<<<blocks, 256, 0, ctx->stream>>>(n, ns, ih, visc, u, K, wd, s, vf);
    }
}

// Coarse samples of the fine fields: G lanes (a power of two dividing 64) share
// one coarse point, each summing every G-th tap of the two filters, so the
// lanes of a group read neighbouring fine points (the points of one wave are
// comp fine cells apart); the group's partial sums meet in a fixed xor-shuffle
// order.  Every lane of a group ends with the full sums.
__device__ __forceinline__ int bwrap1(int m, int n) {  // m in (-n, 2n)
    m += m < 0 ? n : 0;
    return m >= n ? m - n : m;
}

__device__ __forceinline__ double grp_filter(const double *__restrict__ x, int n, int c, int R,
                                             const double *__restrict__ w, int g, int G) {
    double acc = 0.0;
    for (int k = -R + g; k <= R; k += G) acc += w[k + R] * x[bwrap1(c + k, n)];
    for (int o = G >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    return acc;
}

// The closure parts at coarse point J of one sample row (fractions and
// dissipation): fsu = (F * s)(centre), vsu = (D * s)(centre), svu = s_H(vu),
// sVU = s_h(D * u)(centre), vb - va = vu_J - vu_{J-1}.
struct BParts {
    double fsu, vsu, svu, sVU, dvu;
};

__device__ __forceinline__ BParts bparts(const double *__restrict__ ss, const double *__restrict__ vf, int n, int comp,
                                         int a, int J, double ih, double iH, double visc, int Rl,
                                         const double *__restrict__ wl, int K, const double *__restrict__ wd, int g,
                                         int G) {
    const int ic = (J + 1) * comp - a - 1;  // fine centre of coarse cell J
    BParts p;
    p.fsu = grp_filter(ss, n, ic, Rl, wl, g, G);
    p.vsu = grp_filter(ss, n, ic, K, wd, g, G);
    const double vb = vf[(J + 1) * comp - 1], va = vf[bwrap1(J * comp - 1, n)];  // coarse faces J, J-1
    p.svu = -((va + vb) * (va + vb)) / 8 + visc * (vb - va) * iH;
    p.sVU = bstress(vf[bwrap1(ic - 1, n)], vf[ic], ih, visc);
    p.dvu = vb - va;
    return p;
}

// Fractions: one block per sample, groups of G lanes over the coarse points,
// per-thread sums over the points a group's leader owns, then a fixed-order
// block reduction.
__global__ void __launch_bounds__(256) k_bfrac_coarse(int n, int nH, int G, double ih, double iH, double visc,
                                                      const double *__restrict__ s, const double *__restrict__ vu_f,
                                                      int Rl, const double *__restrict__ wl, int K,
                                                      const double *__restrict__ wd, double *__restrict__ sums) {
    const int comp = n / nH, a = comp / 2;
    const double *ss = s + (long)blockIdx.x * n, *vf = vu_f + (long)blockIdx.x * n;
    const int lane = threadIdx.x & 63, g = lane % G, per_wave = 64 / G;
    const int J0 = (threadIdx.x >> 6) * per_wave + lane / G;  // wave-uniform loop over point batches
    double fc = 0.0, ff = 0.0, fd = 0.0;
    for (int Jb = J0 - lane / G; Jb < nH; Jb += 4 * per_wave) {
        const int J = Jb + lane / G;
        const BParts p = bparts(ss, vf, n, comp, a, min(J, nH - 1), ih, iH, visc, Rl, wl, K, wd, g, G);
        if (J < nH && g == 0) {
            const double c = p.vsu - p.sVU, f = p.sVU - p.svu, d = p.fsu - p.vsu;
            const double tot = fabs(c) + fabs(f) + fabs(d) + 2.220446049250313e-16;
            fc += fabs(c) / tot;
            ff += fabs(f) / tot;
            fd += fabs(d) / tot;
        }
    }
    __shared__ double red[3][256];
    red[0][threadIdx.x] = fc;
    red[1][threadIdx.x] = ff;
    red[2][threadIdx.x] = fd;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int x = 0; x < 3; x++) red[x][threadIdx.x] += red[x][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int x = 0; x < 3; x++) sums[3 * blockIdx.x + x] = red[x][0];
}

// lanes per coarse point: about a quarter of the wider kernel's taps, at most 64
static int coarse_group(int Rl, int K) {
    const int taps = 2 * std::max(Rl, K) + 1;
    int G = 1;
    while (G < 64 && 4 * G <= taps) G *= 2;
    return G;
}

extern "C" int ec_burgers_fractions(ec_ctx *ctx, int n, int nH, double L, double visc, int nsamples,
                                    const void *u, int Rl, const double *wl, int K, const double *wd,
                                    double out[3]) {
    EC_REQUIRE(ctx, ctx && u && wl && wd && out, "ec_burgers_fractions: null argument");
    EC_REQUIRE(ctx, n >= 2 && nH >= 1 && n % nH == 0 && nsamples >= 1 && L > 0,
               "ec_burgers_fractions: need n %% nH == 0 (n=%d, nH=%d), nsamples=%d >= 1", n, nH, nsamples);
    EC_REQUIRE(ctx, Rl >= 0 && K >= 0 && 2 * Rl + 1 <= n && 2 * K + 1 <= n,
               "ec_burgers_fractions: kernel radii Rl=%d, K=%d must fit the grid n=%d", Rl, K, n);
    const long total = (long)n * nsamples;
    double *buf = nullptr;
    const size_t bytes = sizeof(double) * ((size_t)2 * total + (2 * Rl + 1) + (2 * K + 1) + 3 * (size_t)nsamples);
    EC_HIP(ctx, hipMallocAsync((void **)&buf, bytes, ctx->stream));
    double *s = buf, *vf = buf + total, *dwl = vf + total, *dwd = dwl + (2 * Rl + 1), *sums = dwd + (2 * K + 1);
    std::vector<double> hs(nsamples * 3);
    int rc = EC_OK;
    if (hipMemcpyAsync(dwl, wl, sizeof(double) * (2 * Rl + 1), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(dwd, wd, sizeof(double) * (2 * K + 1), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_fractions: kernel upload failed");
    if (rc == EC_OK) {
        bfrac_fine(ctx, n, nsamples, n / L, visc, (const double *)u, K, dwd, s, vf);
        k_bfrac_coarse<<<nsamples, 256, 0, ctx->stream>>>(n, nH, coarse_group(Rl, K), n / L, nH / L, visc, s, vf,
                                                          Rl, dwl, K, dwd, sums);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(hs.data(), sums, sizeof(double) * 3 * nsamples, hipMemcpyDeviceToHost, ctx->stream) !=
                hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_fractions: kernels failed");
    }
    const hipError_t fe = hipFreeAsync(buf, ctx->stream);
    if (rc == EC_OK && fe != hipSuccess) rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_fractions: free failed");
    if (rc != EC_OK) return rc;
    // per sample: sum / nH; then the mean over samples (src/Burgers.jl:1090-1096)
    for (int x = 0; x < 3; x++) {
        double acc = 0.0;
        for (int q = 0; q < nsamples; q++) acc += hs[3 * q + x] / nH;
        out[x] = acc / nsamples;
    }
    return EC_OK;
}

// ---- SFS dissipation coefficients (compute_dissipation, src/Burgers.jl:447-516) ----
// Per coarse point J of sample q, from the filtered quantities of the
// fractions pass: sigma_class_m = vsu - sVU, sigma_class_p = vsu - svu and
// sigma_swapfil = fsu - svu, each times delta_stag(vu)_J = (vu_J - vu_{J-1}) / H
// (dissipation, src/Burgers.jl:204-208).  compute_dissipation calls
// three-argument forms of stress, delta_stag and the coarse-grain convolutions
// that src/Burgers.jl does not define; they are read as the two-argument stress
// and delta_stag (src/Burgers.jl:86-94) and the coarse-graining of
// compute_fractions, whose quantities they name.  G lanes per coarse point
// (bparts), grid-stride over all points of all samples;
// out[(m * ns + q) * nH + J] with m = 0 class_m, 1 class_p, 2 swapfil.
__global__ void __launch_bounds__(256) k_bdiss_coarse(int n, int nH, int ns, int G, double ih, double iH,
                                                      double visc, const double *__restrict__ s,
                                                      const double *__restrict__ vu_f, int Rl,
                                                      const double *__restrict__ wl, int K,
                                                      const double *__restrict__ wd, double *__restrict__ out) {
    const int comp = n / nH, a = comp / 2;
    const long total = (long)nH * ns;
    const int lane = threadIdx.x & 63, g = lane % G, sub = lane / G;
    // point batches of 64 / G per wave; the loop bound is the same for every lane of a wave
    const long wave0 = ((long)blockIdx.x * 256 + (threadIdx.x & ~63)) / G;
    const long stride = (long)gridDim.x * 256 / G;
    for (long P0 = wave0; P0 < total; P0 += stride) {
        const long I = min(P0 + sub, total - 1);
        const int J = (int)(I % nH);
        const long q = I / nH;
        const BParts p = bparts(s + q * n, vu_f + q * n, n, comp, a, J, ih, iH, visc, Rl, wl, K, wd, g, G);
        if (P0 + sub < total && g == 0) {
            const double dv = p.dvu * iH;
            out[I] = (p.vsu - p.sVU) * dv;
            out[total + I] = (p.vsu - p.svu) * dv;
            out[2 * total + I] = (p.fsu - p.svu) * dv;
        }
    }
}

extern "C" int ec_burgers_dissipation(ec_ctx *ctx, int n, int nH, double L, double visc, int nsamples,
                                      const void *u, int Rl, const double *wl, int K, const double *wd, void *out) {
    EC_REQUIRE(ctx, ctx && u && wl && wd && out, "ec_burgers_dissipation: null argument");
    EC_REQUIRE(ctx, n >= 2 && nH >= 1 && n % nH == 0 && nsamples >= 1 && L > 0,
               "ec_burgers_dissipation: need n %% nH == 0 (n=%d, nH=%d), nsamples=%d >= 1", n, nH, nsamples);
    EC_REQUIRE(ctx, Rl >= 0 && K >= 0 && 2 * Rl + 1 <= n && 2 * K + 1 <= n,
               "ec_burgers_dissipation: kernel radii Rl=%d, K=%d must fit the grid n=%d", Rl, K, n);
    const long total = (long)n * nsamples, ctotal = (long)nH * nsamples;
    double *buf = nullptr;
    const size_t bytes = sizeof(double) * ((size_t)2 * total + (2 * Rl + 1) + (2 * K + 1));
    EC_HIP(ctx, hipMallocAsync((void **)&buf, bytes, ctx->stream));
    double *s = buf, *vf = buf + total, *dwl = vf + total, *dwd = dwl + (2 * Rl + 1);
    int rc = EC_OK;
    if (hipMemcpyAsync(dwl, wl, sizeof(double) * (2 * Rl + 1), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(dwd, wd, sizeof(double) * (2 * K + 1), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_dissipation: kernel upload failed");
    if (rc == EC_OK) {
        const int G = coarse_group(Rl, K);
        const int cblocks = (int)std::min<long>((ctotal * G + 255) / 256, 65536);
        bfrac_fine(ctx, n, nsamples, n / L, visc, (const double *)u, K, dwd, s, vf);
        k_bdiss_coarse<<<cblocks, 256, 0, ctx->stream>>>(n, nH, nsamples, G, n / L, nH / L, visc, s, vf, Rl, dwl,
                                                         K, dwd, (double *)out);
        // the weights were copied from host memory the caller may free on return
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_dissipation: kernels failed");
    }
    const hipError_t fe = hipFreeAsync(buf, ctx->stream);
    if (rc == EC_OK && fe != hipSuccess) rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_dissipation: free failed");
    return rc;
}

// ---- Smagorinsky study (smagorinsky_fields, src/Burgers.jl:547-680; the
// least-squares coefficients of smagorinsky.jl:24-36; solve_smagorinsky,
// src/Burgers.jl:682-772) ----
// Fit: per fine point i and sample, with U = D * u and s = s_h(u),
//   classic   T = (D * s)[i] - s_h(U)[i],  S = -c |dU| dU,  dU = (U[i] - U[i-1]) / h
//   discrete  T = (F * s)[i] - s_H(U[i-a-1], U[i+a]),  S likewise with dU = (U[i+a] - U[i-a-1]) / H
// (c = Delta^2 + H^2, a = comp / 2), summed into <S,T> and <S,S> of each kind;
// theta^2 = -<S,T> / <S,S>.  Blocks write partial sums, reduced on the host in
// a fixed order.
__global__ void __launch_bounds__(256) k_bsmag_fit(int n, int ns, int a, double ih, double iH, double visc,
                                                   double cs, const double *__restrict__ s,
                                                   const double *__restrict__ U, int Rl,
                                                   const double *__restrict__ wl, int K,
                                                   const double *__restrict__ wd, double *__restrict__ part) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const long total = (long)n * ns;
    for (long I = blockIdx.x * (long)blockDim.x + threadIdx.x; I < total; I += (long)gridDim.x * blockDim.x) {
        const int i = (int)(I % n);
        const double *ss = s + (I - i), *UU = U + (I - i);
        double ds = 0.0, fs = 0.0;
        for (int k = -K; k <= K; k++) ds += wd[k + K] * ss[bwrap(i + k, n)];
        for (int r = -Rl; r <= Rl; r++) fs += wl[r + Rl] * ss[bwrap(i + r, n)];
        const double um = UU[bwrap(i - 1, n)], u0 = UU[i];
        const double th = ds - bstress(um, u0, ih, visc);
        const double dh = (u0 - um) * ih, sh = -cs * fabs(dh) * dh;
        const double ul = UU[bwrap(i - a - 1, n)], ur = UU[bwrap(i + a, n)];
        const double tH = fs - bstress(ul, ur, iH, visc);
        const double dH = (ur - ul) * iH, sH = -cs * fabs(dH) * dH;
        acc[0] += sh * th;
        acc[1] += sh * sh;
        acc[2] += sH * tH;
        acc[3] += sH * sH;
    }
    __shared__ double red[4][256];
    for (int x = 0; x < 4; x++) red[x][threadIdx.x] = acc[x];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int x = 0; x < 4; x++) red[x][threadIdx.x] += red[x][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int x = 0; x < 4; x++) part[4 * blockIdx.x + x] = red[x][0];
}

// The fit with the block's stretch of s and U staged in LDS (rows
// [i0 - K - 1, i0 + cnt + K] of one sample; a <= K and Rl <= K, as the double
// kernel is the LES kernel widened by the volume filter).  Per-point sums in
// the order of k_bsmag_fit; one partial per block.
__global__ void __launch_bounds__(256) k_bsmag_fit_t(int n, int a, double ih, double iH, double visc, double cs,
                                                     const double *__restrict__ s, const double *__restrict__ U,
                                                     int Rl, const double *__restrict__ wl, int K,
                                                     const double *__restrict__ wd, double *__restrict__ part) {
    extern __shared__ double xs[];
    const int i0 = blockIdx.x * BF_TILE, cnt = min(BF_TILE, n - i0);
    const long row = (long)blockIdx.y * n;
    const int len = cnt + 2 * K + 2;  // xs[j] = s[i0 - K - 1 + j], Us[j] = U[i0 - K - 1 + j]
    double *Us = xs + len;
    for (int j = threadIdx.x; j < len; j += 256) {
        const long g = row + bwrap(i0 - K - 1 + j, n);
        xs[j] = s[g];
        Us[j] = U[g];
    }
    __syncthreads();
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = threadIdx.x; t < cnt; t += 256) {
        const int o = t + K + 1;
        double ds = 0.0, fs = 0.0;
        for (int k = -K; k <= K; k++) ds += wd[k + K] * xs[o + k];
        for (int r = -Rl; r <= Rl; r++) fs += wl[r + Rl] * xs[o + r];
        const double um = Us[o - 1], u0 = Us[o];
        const double th = ds - bstress(um, u0, ih, visc);
        const double dh = (u0 - um) * ih, sh = -cs * fabs(dh) * dh;
        const double ul = Us[o - a - 1], ur = Us[o + a];
        const double tH = fs - bstress(ul, ur, iH, visc);
        const double dH = (ur - ul) * iH, sH = -cs * fabs(dH) * dH;
        acc[0] += sh * th;
        acc[1] += sh * sh;
        acc[2] += sH * tH;
        acc[3] += sH * sH;
    }
    __shared__ double red[4][256];
    for (int x = 0; x < 4; x++) red[x][threadIdx.x] = acc[x];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int x = 0; x < 4; x++) red[x][threadIdx.x] += red[x][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const long b = (long)blockIdx.y * gridDim.x + blockIdx.x;
        for (int x = 0; x < 4; x++) part[4 * b + x] = red[x][0];
    }
}

extern "C" int ec_burgers_smag_fit(ec_ctx *ctx, int n, int nH, double L, double visc, double delta, int nsamples,
                                   const void *u, int Rl, const double *wl, int K, const double *wd,
                                   double out[4]) {
    EC_REQUIRE(ctx, ctx && u && wl && wd && out, "ec_burgers_smag_fit: null argument");
    EC_REQUIRE(ctx, n >= 2 && nH >= 1 && n % nH == 0 && nsamples >= 1 && L > 0 && delta >= 0,
               "ec_burgers_smag_fit: need n %% nH == 0 (n=%d, nH=%d), nsamples=%d >= 1", n, nH, nsamples);
    EC_REQUIRE(ctx, Rl >= 0 && K >= 0 && 2 * Rl + 1 <= n && 2 * K + 1 <= n,
               "ec_burgers_smag_fit: kernel radii Rl=%d, K=%d must fit the grid n=%d", Rl, K, n);
    const long total = (long)n * nsamples;
    const int a = (n / nH) / 2;
    // LDS-staged pass while the staged rows fit (a <= K and Rl <= K hold for the
    // reference's kernels; others take the direct pass)
    const bool tiled = 2 * (BF_TILE + 2 * K + 2) <= BF_LDS && a <= K && Rl <= K && nsamples <= 65535;
    const int ntile = (n + BF_TILE - 1) / BF_TILE;
    const long blocks = tiled ? (long)ntile * nsamples : std::min<long>((total + 255) / 256, 4096);
    double *buf = nullptr;
    const size_t bytes = sizeof(double) * ((size_t)2 * total + (2 * Rl + 1) + (2 * K + 1) + 4 * (size_t)blocks);
    EC_HIP(ctx, hipMallocAsync((void **)&buf, bytes, ctx->stream));
    double *s = buf, *U = buf + total, *dwl = U + total, *dwd = dwl + (2 * Rl + 1), *part = dwd + (2 * K + 1);
    std::vector<double> hp(4 * (size_t)blocks);
    const double H = L / nH;
    int rc = EC_OK;
    if (hipMemcpyAsync(dwl, wl, sizeof(double) * (2 * Rl + 1), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(dwd, wd, sizeof(double) * (2 * K + 1), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_smag_fit: kernel upload failed");
    if (rc == EC_OK) {
        bfrac_fine(ctx, n, nsamples, n / L, visc, (const double *)u, K, dwd, s, U);
        if (tiled) {
            const size_t lds = sizeof(double) * 2 * (size_t)(std::min(BF_TILE, n) + 2 * K + 2);
            if (hipFuncSetAttribute((const void *)k_bsmag_fit_t, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds) != hipSuccess)
                rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_smag_fit: LDS attribute");
            else
                k_bsmag_fit_t<<<dim3((unsigned)ntile, (unsigned)nsamples), 256, lds, ctx->stream>>>(
                    n, a, n / L, 1.0 / H, visc, delta * delta + H * H, s, U, Rl, dwl, K, dwd, part);
        } else {
            k_bsmag_fit<<<(unsigned)blocks, 256, 0, ctx->stream>>>(n, nsamples, a, n / L, 1.0 / H, visc,
                                                                   delta * delta + H * H, s, U, Rl, dwl, K, dwd,
                                                                   part);
        }
        if (rc == EC_OK &&
            (hipGetLastError() != hipSuccess ||
             hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, ctx->stream) !=
                 hipSuccess ||
             hipStreamSynchronize(ctx->stream) != hipSuccess))
            rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_smag_fit: kernels failed");
    }
    const hipError_t fe = hipFreeAsync(buf, ctx->stream);
    if (rc == EC_OK && fe != hipSuccess) rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_smag_fit: free failed");
    if (rc != EC_OK) return rc;
    for (int x = 0; x < 4; x++) {
        double acc = 0.0;
        for (long b = 0; b < blocks; b++) acc += hp[4 * b + x];
        out[x] = acc;
    }
    return EC_OK;
}

// Solve: one block per sample; ubar = C_f(D * u) (coarse faces), then RK4 of
// du = delta_coll(s_H(ubar) + C |delta_stag ubar| delta_stag ubar) to tstop
// with dt = min(cfl min(H / max|ubar|, H^2 / visc), tstop - t).  The stages
// run as the reference writes them: each axpy! lands on the running state
// (ubar += dt/2 k1; ubar += dt/2 k2; ubar += dt k3), so stage 3 sees
// ubar0 + dt/2 (k1 + k2); the step ends at ubar0 + dt/6 (k1 + 2 k2 + 2 k3 + k4).
// Only ubar lives in LDS (nH doubles).  Thread t owns the points t + m 256,
// m < PT, and keeps their step-start state, stage sum and slope in registers;
// a slope is the difference of the two face stresses, each formed from ubar
// where it is needed, so a stage takes two block barriers (slopes from the
// old state, then the update).  max|ubar| comes from wave shuffles and one
// value per wave in LDS, written in the step's last update.
constexpr long SMAG_MAX_STEPS = 1000000;  // exit bound for pathological inputs
constexpr int SMAG_NT = 256;

template <int PT>
__global__ void __launch_bounds__(SMAG_NT) k_bsmag_solve(int n, int nH, double iH, double H, double visc, double C,
                                                         double tstop, double cflf, const double *__restrict__ u,
                                                         int K, const double *__restrict__ wd,
                                                         double *__restrict__ ubar_out) {
    extern __shared__ double ub[];
    __shared__ double red[2][SMAG_NT / 64];
    const int comp = n / nH, tid = threadIdx.x;
    const double *us = u + (long)blockIdx.x * n;
    for (int J = tid; J < nH; J += SMAG_NT) {
        const int f = (J + 1) * comp - 1;  // fine face of coarse face J
        double v = 0.0;
        for (int k = -K; k <= K; k++) v += wd[k + K] * us[bwrap(f + k, n)];
        ub[J] = v;
    }
    __syncthreads();
    auto wave_max_to = [&](double m, int slot) {
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        if ((tid & 63) == 0) red[slot][tid >> 6] = m;
    };
    auto sigma = [&](double ua, double ubb) {
        const double d = (ubb - ua) * iH;
        return -((ua + ubb) * (ua + ubb)) / 8 + visc * d + C * fabs(d) * d;
    };
    auto slopes = [&](double (&kk)[PT]) {  // kk = delta_coll(sigma) at the owned points
#pragma unroll
        for (int m = 0; m < PT; m++) {
            const int J = tid + m * SMAG_NT;
            if (J < nH) {
                const double um = ub[J == 0 ? nH - 1 : J - 1], u0 = ub[J], up = ub[J == nH - 1 ? 0 : J + 1];
                kk[m] = (sigma(u0, up) - sigma(um, u0)) * iH;
            }
        }
    };
    double m0 = 0.0;
    for (int J = tid; J < nH; J += SMAG_NT) m0 = fmax(m0, fabs(ub[J]));
    wave_max_to(m0, 0);
    __syncthreads();
    double u0r[PT], acc[PT], kk[PT];
    double t = 0.0;
    int slot = 0;
    for (long step = 0; step < SMAG_MAX_STEPS && t < tstop; step++) {
        double mx = red[slot][0];
        for (int w = 1; w < SMAG_NT / 64; w++) mx = fmax(mx, red[slot][w]);
        const double dt = fmin(cflf * fmin(H / mx, H * H / visc), tstop - t);
        if (!(dt > 0)) break;  // max|ubar| infinite: no step can advance t
        t += dt;
#pragma unroll
        for (int m = 0; m < PT; m++) {
            const int J = tid + m * SMAG_NT;
            if (J < nH) u0r[m] = ub[J];
        }
        slopes(kk);  // k1
        __syncthreads();
#pragma unroll
        for (int m = 0; m < PT; m++) {
            const int J = tid + m * SMAG_NT;
            if (J < nH) {
                acc[m] = kk[m];
                ub[J] += dt / 2 * kk[m];
            }
        }
        __syncthreads();
        slopes(kk);  // k2
        __syncthreads();
#pragma unroll
        for (int m = 0; m < PT; m++) {
            const int J = tid + m * SMAG_NT;
            if (J < nH) {
                acc[m] += 2 * kk[m];
                ub[J] += dt / 2 * kk[m];
            }
        }
        __syncthreads();
        slopes(kk);  // k3
        __syncthreads();
#pragma unroll
        for (int m = 0; m < PT; m++) {
            const int J = tid + m * SMAG_NT;
            if (J < nH) {
                acc[m] += 2 * kk[m];
                ub[J] += dt * kk[m];
            }
        }
        __syncthreads();
        slopes(kk);  // k4
        __syncthreads();
        double mnew = 0.0;
#pragma unroll
        for (int m = 0; m < PT; m++) {
            const int J = tid + m * SMAG_NT;
            if (J < nH) {
                const double v = u0r[m] + dt / 6 * (acc[m] + kk[m]);
                ub[J] = v;
                mnew = fmax(mnew, fabs(v));
            }
        }
        slot ^= 1;
        wave_max_to(mnew, slot);
        __syncthreads();
    }
    for (int J = tid; J < nH; J += SMAG_NT) ubar_out[(long)blockIdx.x * nH + J] = ub[J];
}

extern "C" int ec_burgers_smag_solve(ec_ctx *ctx, int n, int nH, double L, double visc, double C, double tstop,
                                     double cfl, int nsamples, const void *ustart, int K, const double *wd,
                                     void *ubar_out) {
    EC_REQUIRE(ctx, ctx && ustart && wd && ubar_out, "ec_burgers_smag_solve: null argument");
    EC_REQUIRE(ctx, n >= 2 && nH >= 2 && n % nH == 0 && nsamples >= 1 && L > 0 && visc > 0 && cfl > 0 &&
                        tstop >= 0 && C >= 0,
               "ec_burgers_smag_solve: bad argument (n=%d, nH=%d, nsamples=%d, visc=%g, cfl=%g, C=%g)", n, nH,
               nsamples, visc, cfl, C);
    EC_REQUIRE(ctx, nH <= 16 * SMAG_NT, "ec_burgers_smag_solve: nH=%d above %d (points per thread)", nH,
               16 * SMAG_NT);
    const size_t lds = sizeof(double) * (size_t)nH;
    const int pt = nH <= 4 * SMAG_NT ? 4 : nH <= 8 * SMAG_NT ? 8 : nH <= 12 * SMAG_NT ? 12 : 16;
    const void *kern = pt == 4    ? (const void *)k_bsmag_solve<4>
                       : pt == 8  ? (const void *)k_bsmag_solve<8>
                       : pt == 12 ? (const void *)k_bsmag_solve<12>
                                  : (const void *)k_bsmag_solve<16>;
    EC_REQUIRE(ctx, K >= 0 && 2 * K + 1 <= n, "ec_burgers_smag_solve: kernel radius K=%d must fit n=%d", K, n);
    double *dwd = nullptr;
    EC_HIP(ctx, hipMallocAsync((void **)&dwd, sizeof(double) * (2 * K + 1), ctx->stream));
    int rc = EC_OK;
    if (hipMemcpyAsync(dwd, wd, sizeof(double) * (2 * K + 1), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_smag_solve: kernel upload failed");
    if (rc == EC_OK &&
        hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
            hipSuccess)
        rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_smag_solve: LDS attribute");
    if (rc == EC_OK) {
#define EC_SMAG_LAUNCH(P)                                                                                  \
    k_bsmag_solve<P><<<nsamples, SMAG_NT, lds, ctx->stream>>>(n, nH, nH / L, L / nH, visc, C, tstop, cfl,     \
                                                              (const double *)ustart, K, dwd, (double *)ubar_out)
        if (pt == 4)
            EC_SMAG_LAUNCH(4);
        else if (pt == 8)
            EC_SMAG_LAUNCH(8);
        else if (pt == 12)
            EC_SMAG_LAUNCH(12);
        else
            EC_SMAG_LAUNCH(16);
#undef EC_SMAG_LAUNCH
        if (hipGetLastError() != hipSuccess) rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_smag_solve: launch failed");
    }
    const hipError_t fe = hipFreeAsync(dwd, ctx->stream);
    if (rc == EC_OK && fe != hipSuccess) rc = ec_fail(ctx, EC_E_HIP, "ec_burgers_smag_solve: free failed");
    return rc;
}
