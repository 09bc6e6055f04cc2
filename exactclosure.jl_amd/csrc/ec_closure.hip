// Separable filters sampled at LES points and the exact-closure fluxes
// (tophat/gaussian/composekernel/kernelproduct, lazyfilter, lazycoarsegrain,
// get_r, get_rDHi, get_sigDH, apply_DHpi: src/N.jl:592-698, 999-1135).
//
// Every quantity is  out[J] = sum_t w1[t1] w2[t2] w3[t3] src[c(J_d+1)-1+off_d+t_d]
// where src is evaluated on the fly from the DNS velocity and pressure
// (stresses are formed in FP64 from the stored values).  All accumulation and
// every output is FP64: the closure commutator S r(u) - r_H(ubar) cancels
// heavily and is kept in double on the GPU.
#include "ec_internal.h"

#include <deque>
#include <map>
#include <mutex>

namespace {

constexpr int MAXQ = 24;

struct QDesc {
    int source;
    int len[3];
    int off[3];
    const double *w[3];
    double *out;
};

struct QList {
    int nq;
    QDesc q[MAXQ];
};

__device__ __forceinline__ int wrapn(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

template <typename T>
struct Src {
    const T *u[3];
    const T *p;
    int n;
    double ih, visc;
    __device__ __forceinline__ double at(const T *f, int i, int j, int k) const {
        return (double)f[(size_t)i + (size_t)n * ((size_t)j + (size_t)n * k)];
    }
    // stress(u, visc, a, b) of src/N.jl:106-118 at cell (i, j, k)
    __device__ double stress(int a, int b, int i, int j, int k) const {
        int I[3] = {i, j, k};
        if (a == b) {
            int L[3] = {i, j, k};
            L[a] = I[a] == 0 ? n - 1 : I[a] - 1;
            const double r = at(u[a], i, j, k), l = at(u[a], L[0], L[1], L[2]);
            const double e = (r + l) / 2, d = (r - l) * ih;
            return e * e - visc * (d + d);
        }
        int Rb[3] = {i, j, k}, Ra[3] = {i, j, k};
        Rb[b] = I[b] + 1 == n ? 0 : I[b] + 1;
        Ra[a] = I[a] + 1 == n ? 0 : I[a] + 1;
        const double ar = at(u[a], Rb[0], Rb[1], Rb[2]), al = at(u[a], i, j, k);
        const double br = at(u[b], Ra[0], Ra[1], Ra[2]), bl = at(u[b], i, j, k);
        return ((ar + al) / 2) * ((br + bl) / 2) - visc * ((ar - al) * ih + (br - bl) * ih);
    }
    __device__ double eval(int s, int i, int j, int k) const {
        switch (s) {
            case EC_SRC_U1: return at(u[0], i, j, k);
            case EC_SRC_U2: return at(u[1], i, j, k);
            case EC_SRC_U3: return at(u[2], i, j, k);
            case EC_SRC_S11: return stress(0, 0, i, j, k);
            case EC_SRC_S22: return stress(1, 1, i, j, k);
            case EC_SRC_S33: return stress(2, 2, i, j, k);
            case EC_SRC_S12: return stress(0, 1, i, j, k);
            case EC_SRC_S23: return stress(1, 2, i, j, k);
            case EC_SRC_S31: return stress(2, 0, i, j, k);
            case EC_SRC_R11: return stress(0, 0, i, j, k) + at(p, i, j, k);
            case EC_SRC_R22: return stress(1, 1, i, j, k) + at(p, i, j, k);
            case EC_SRC_R33: return stress(2, 2, i, j, k) + at(p, i, j, k);
            case EC_SRC_K1: return at(u[0], i, j, k) - (at(p, i + 1 == n ? 0 : i + 1, j, k) - at(p, i, j, k)) * ih;
            case EC_SRC_K2: return at(u[1], i, j, k) - (at(p, i, j + 1 == n ? 0 : j + 1, k) - at(p, i, j, k)) * ih;
            case EC_SRC_K3: return at(u[2], i, j, k) - (at(p, i, j, k + 1 == n ? 0 : k + 1) - at(p, i, j, k)) * ih;
            default: return at(p, i, j, k);
        }
    }
};

// One thread per (coarse point, quantity).
template <typename T>
__global__ void k_filter_coarse(Src<T> src, int nbar, int c, QList ql) {
    const size_t nb3 = (size_t)nbar * nbar * nbar;
    const size_t J = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const int qi = blockIdx.y;
    if (J >= nb3 || qi >= ql.nq) return;
    const QDesc q = ql.q[qi];
    const int n = src.n;
    const int J1 = (int)(J % nbar), J2 = (int)((J / nbar) % nbar), J3 = (int)(J / ((size_t)nbar * nbar));
    const int b1 = c * (J1 + 1) - 1 + q.off[0];
    const int b2 = c * (J2 + 1) - 1 + q.off[1];
    const int b3 = c * (J3 + 1) - 1 + q.off[2];
    double s = 0.0;
    for (int t3 = 0; t3 < q.len[2]; t3++) {
        const int k = wrapn(b3 + t3, n);
        for (int t2 = 0; t2 < q.len[1]; t2++) {
            const int j = wrapn(b2 + t2, n);
            double sx = 0.0;
            for (int t1 = 0; t1 < q.len[0]; t1++) sx += q.w[0][t1] * src.eval(q.source, wrapn(b1 + t1, n), j, k);
            s += q.w[2][t3] * q.w[1][t2] * sx;
        }
    }
    q.out[J] = s;
}

// Separable form of k_filter_coarse for wide kernels (Gaussian-composed
// filters: (2K+1)^3 taps per coarse point in the dense form): three passes,
// each a 1D filter sampled at the coarse points of one axis,
//   t1[J1, j, k]  = sum_t w1[t] src(c(J1+1)-1+off1+t, j, k)        (source on the fly)
//   t2[J1, J2, k] = sum_t w2[t] t1[J1, c(J2+1)-1+off2+t, k]
//   out[J]        = sum_t w3[t] t2[J1, J2, c(J3+1)-1+off3+t]
// so a coarse point costs len1 + len2 + len3 multiply-adds per fine point
// it covers instead of len1 len2 len3.  All sums FP64.
template <typename T>
__global__ void k_sep_x(Src<T> src, int source, int nbar, int c, int len, int off, const double *__restrict__ w,
                        double *__restrict__ t1) {
    const int n = src.n;
    const size_t total = (size_t)nbar * n * n;
    const size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (id >= total) return;
    const int J1 = (int)(id % nbar);
    const size_t jk = id / nbar;
    const int j = (int)(jk % n), k = (int)(jk / n);
    const int b = c * (J1 + 1) - 1 + off;
    double s = 0.0;
    for (int t = 0; t < len; t++) s += w[t] * src.eval(source, wrapn(b + t, n), j, k);
    t1[id] = s;
}

__global__ void k_sep_y(int n, int nbar, int c, int len, int off, const double *__restrict__ w,
                        const double *__restrict__ t1, double *__restrict__ t2) {
    const size_t total = (size_t)nbar * nbar * n;
    const size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (id >= total) return;
    const int J1 = (int)(id % nbar), J2 = (int)((id / nbar) % nbar);
    const size_t k = id / ((size_t)nbar * nbar);
    const int b = c * (J2 + 1) - 1 + off;
    double s = 0.0;
    for (int t = 0; t < len; t++) s += w[t] * t1[(size_t)J1 + (size_t)nbar * ((size_t)wrapn(b + t, n) + (size_t)n * k)];
    t2[id] = s;
}

__global__ void k_sep_z(int n, int nbar, int c, int len, int off, const double *__restrict__ w,
                        const double *__restrict__ t2, double *__restrict__ out) {
    const size_t total = (size_t)nbar * nbar * nbar;
    const size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (id >= total) return;
    const size_t J12 = id % ((size_t)nbar * nbar);
    const int J3 = (int)(id / ((size_t)nbar * nbar));
    const int b = c * (J3 + 1) - 1 + off;
    double s = 0.0;
    for (int t = 0; t < len; t++) s += w[t] * t2[J12 + (size_t)nbar * nbar * (size_t)wrapn(b + t, n)];
    out[id] = s;
}

// applyfilter! (src/N.jl:639-653) with a kernelproduct of three centred 1D
// kernels, as three 1D passes along x, y, z (periodic):
//   out[i] = sum_{r=-R}^{R} w[R + r] in[i + r e_axis].
template <typename TI, typename TO>
__global__ void k_apply_axis(int n, int axis, int R, const double *__restrict__ w, const TI *__restrict__ in,
                             TO *__restrict__ out) {
    const size_t n1 = (size_t)n, total = n1 * n1 * n1;
    const size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (id >= total) return;
    const int i = (int)(id % n1), j = (int)((id / n1) % n1), k = (int)(id / (n1 * n1));
    const int x0 = axis == 0 ? i : (axis == 1 ? j : k);
    const size_t stride = axis == 0 ? 1 : (axis == 1 ? n1 : n1 * n1);
    const size_t base = id - (size_t)x0 * stride;
    double s = 0.0;
    for (int r = -R; r <= R; r++) s += w[R + r] * (double)in[base + (size_t)wrapn(x0 + r, n) * stride];
    out[id] = (TO)s;
}

// tau_ab = flux_ab - sigma_H(ubar)_ab on the LES grid, nonsym order.
__global__ void k_commutator(int nb, double ih, double visc, const double *__restrict__ u1,
                             const double *__restrict__ u2, const double *__restrict__ u3, QList fl, QList tl) {
    const size_t nb3 = (size_t)nb * nb * nb;
    const size_t I = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (I >= nb3) return;
    Src<double> s;
    s.u[0] = u1;
    s.u[1] = u2;
    s.u[2] = u3;
    s.p = nullptr;
    s.n = nb;
    s.ih = ih;
    s.visc = visc;
    const int i = (int)(I % nb), j = (int)((I / nb) % nb), k = (int)(I / ((size_t)nb * nb));
    const double s11 = s.stress(0, 0, i, j, k), s22 = s.stress(1, 1, i, j, k), s33 = s.stress(2, 2, i, j, k);
    const double s12 = s.stress(0, 1, i, j, k), s23 = s.stress(1, 2, i, j, k), s31 = s.stress(2, 0, i, j, k);
    // 11, 21, 31, 12, 22, 32, 13, 23, 33
    const double sig[9] = {s11, s12, s31, s12, s22, s23, s31, s23, s33};
#pragma unroll
    for (int m = 0; m < 9; m++) tl.q[m].out[I] = fl.q[m].out[I] - sig[m];
}

}  // namespace

// Device copies of 1D weight vectors, uploaded once per distinct vector and
// kept for the life of the process (filter sets repeat every time step).
static const double *weights_on_device(ec_ctx *ctx, const double *w, int len, int *rc) {
    static std::map<std::pair<int, std::vector<double>>, double *> cache;
    std::vector<double> key(w, w + len);
    auto k = std::make_pair(ctx->device, key);
    auto it = cache.find(k);
    if (it != cache.end()) return it->second;
    double *d = nullptr;
    if (hipMalloc(&d, sizeof(double) * (len > 0 ? len : 1)) != hipSuccess ||
        hipMemcpy(d, w, sizeof(double) * len, hipMemcpyHostToDevice) != hipSuccess) {
        *rc = ec_fail(ctx, EC_E_HIP, "weight upload failed");
        return nullptr;
    }
    cache[k] = d;
    return d;
}

// Scratch of the separable passes (t1, t2 and the three weight vectors) in
// doubles for one quantity.
static size_t sep_doubles(int n, int nbar, const int len[3]) {
    return (size_t)nbar * n * n + (size_t)nbar * nbar * n + (size_t)len[0] + len[1] + len[2];
}

// The three separable passes of one quantity; ws holds sep_doubles(...)
// doubles, the weights already in place at its end (w[a] device pointers).
template <typename T>
static int sep_launch(ec_ctx *ctx, const Src<T> &s, int nbar, int source, const int len[3], const int off[3],
                      const double *const w[3], double *ws, double *out) {
    const int n = s.n, c = n / nbar;
    double *t1 = ws, *t2 = ws + (size_t)nbar * n * n;
    const size_t m1 = (size_t)nbar * n * n, m2 = (size_t)nbar * nbar * n, m3 = (size_t)nbar * nbar * nbar;
    k_sep_x<T><<<(unsigned)((m1 + 255) / 256), 256, 0, ctx->stream>>>(s, source, nbar, c, len[0], off[0], w[0], t1);
    EC_CHECK_LAUNCH(ctx, "k_sep_x");
    k_sep_y<<<(unsigned)((m2 + 255) / 256), 256, 0, ctx->stream>>>(n, nbar, c, len[1], off[1], w[1], t1, t2);
    EC_CHECK_LAUNCH(ctx, "k_sep_y");
    k_sep_z<<<(unsigned)((m3 + 255) / 256), 256, 0, ctx->stream>>>(n, nbar, c, len[2], off[2], w[2], t2, out);
    EC_CHECK_LAUNCH(ctx, "k_sep_z");
    return EC_OK;
}

// Context-owned workspace of at least `doubles` doubles (grown on demand;
// freed with the context).
static double *workspace(ec_ctx *ctx, size_t doubles, int *rc) {
    if (ctx->ws_doubles < doubles) {
        if (ctx->ws) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(ctx->ws);
            ctx->ws = nullptr;
            ctx->ws_doubles = 0;
        }
        if (hipMalloc(&ctx->ws, sizeof(double) * doubles) != hipSuccess) {
            *rc = ec_fail(ctx, EC_E_OOM, "workspace of %zu doubles", doubles);
            return nullptr;
        }
        ctx->ws_doubles = doubles;
    }
    return (double *)ctx->ws;
}

// Quantities whose dense tap count exceeds this take the separable passes.
constexpr long EC_DENSE_TAPS = 64;

template <typename T>
static int filter_launch(ec_ctx *ctx, int n, int nbar, double l, double visc, void *const u[3], const void *p,
                         int nq, const ec_filter_quantity *q) {
    Src<T> s;
    for (int a = 0; a < 3; a++) s.u[a] = (const T *)u[a];
    s.p = (const T *)p;
    s.n = n;
    s.ih = n / l;
    s.visc = visc;
    const int c = n / nbar;
    const size_t nb3 = (size_t)nbar * nbar * nbar;
    // wide quantities: separable passes, one quantity at a time
    std::vector<int> dense;
    for (int m = 0; m < nq; m++) {
        const ec_filter_quantity &src = q[m];
        for (int a = 0; a < 3; a++) {
            EC_REQUIRE(ctx, src.len[a] >= 1 && src.len[a] <= 4 * n && src.w[a], "filter quantity %d: bad axis %d", m, a);
            EC_REQUIRE(ctx, src.off[a] >= -4 * n && src.off[a] <= 4 * n, "filter quantity %d: offset out of range", m);
        }
        EC_REQUIRE(ctx, src.out, "filter quantity %d: null output", m);
        EC_REQUIRE(ctx, src.source >= 0 && src.source <= EC_SRC_K3, "filter quantity %d: bad source", m);
        EC_REQUIRE(ctx, (src.source < EC_SRC_R11) || p, "pressure source without pressure field");
        if ((long)src.len[0] * src.len[1] * src.len[2] <= EC_DENSE_TAPS) {
            dense.push_back(m);
            continue;
        }
        int rc = EC_OK;
        double *ws = workspace(ctx, sep_doubles(n, nbar, src.len), &rc);
        if (!ws) return rc;
        const double *w[3];
        for (int a = 0; a < 3; a++) {
            w[a] = weights_on_device(ctx, src.w[a], src.len[a], &rc);
            if (rc) return rc;
        }
        ec_prof_scope pr(ctx, "filter_sep", (double)sizeof(T) * 4.0 * ec_cube(n) + 8.0 * nb3);
        if ((rc = sep_launch<T>(ctx, s, nbar, src.source, src.len, src.off, w, ws, (double *)src.out))) return rc;
    }
    const int ndense = (int)dense.size();
    for (int q0 = 0; q0 < ndense; q0 += MAXQ) {
        QList ql;
        ql.nq = std::min(MAXQ, ndense - q0);
        for (int m = 0; m < ql.nq; m++) {
            const ec_filter_quantity &src = q[dense[q0 + m]];
            QDesc &d = ql.q[m];
            d.source = src.source;
            for (int a = 0; a < 3; a++) {
                d.len[a] = src.len[a];
                d.off[a] = src.off[a];
                int rc = EC_OK;
                d.w[a] = weights_on_device(ctx, src.w[a], src.len[a], &rc);
                if (rc) return rc;
            }
            d.out = (double *)src.out;
        }
        dim3 grid((unsigned)((nb3 + 255) / 256), ql.nq);
        ec_prof_scope pr(ctx, "filter_coarse", (double)sizeof(T) * 4.0 * ec_cube(n) + 8.0 * nb3 * ql.nq);
        k_filter_coarse<T><<<grid, 256, 0, ctx->stream>>>(s, nbar, c, ql);
        EC_CHECK_LAUNCH(ctx, "k_filter_coarse");
    }
    return EC_OK;
}

// ---- the filter without allocation (ec_filter_coarse_noalloc) ----
// The separable passes of one quantity in a fixed device array: t1, t2 and
// the three weight vectors of sep_doubles(n, nbar, len).  Each GPU holds its
// own copy for the life of the process; every call on that GPU shares it.
constexpr size_t EC_NOALLOC_DOUBLES = (size_t)4 << 20;  // 32 MiB
constexpr int EC_NOALLOC_TAPS = 64;                       // per axis
static __device__ __align__(256) double scratch[EC_NOALLOC_DOUBLES];

namespace {
struct WArg {
    int len[3];
    double w[3][EC_NOALLOC_TAPS];
};
// the weights reach the device as kernel arguments (no host copy)
__global__ void k_put_weights(WArg a, double *__restrict__ dst) {
    int o = 0;
    for (int d = 0; d < 3; d++) {
        for (int t = threadIdx.x; t < a.len[d]; t += blockDim.x) dst[o + t] = a.w[d][t];
        o += a.len[d];
    }
}
}  // namespace

// The device address of this GPU's scratch, looked up once per device with
// that device current (the host-side symbol itself is not a device address).
static double *scratch_on_device(ec_ctx *ctx, int *rc) {
    static std::mutex mu;
    static double *addr[64] = {nullptr};
    if (ctx->device < 0 || ctx->device >= 64) {
        *rc = ec_fail(ctx, EC_E_UNSUPPORTED, "device %d beyond the scratch table", ctx->device);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[ctx->device]) {
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess || (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess)) {
            *rc = ec_fail(ctx, EC_E_HIP, "cannot make device %d current", ctx->device);
            return nullptr;
        }
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != ctx->device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr) {
            *rc = ec_fail(ctx, EC_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return nullptr;
        }
        addr[ctx->device] = (double *)ptr;
    }
    return addr[ctx->device];
}

template <typename T>
static int noalloc_launch(ec_ctx *ctx, int n, int nbar, double l, double visc, void *const u[3], const void *p,
                          int nq, const ec_filter_quantity *q) {
    Src<T> s;
    for (int a = 0; a < 3; a++) s.u[a] = (const T *)u[a];
    s.p = (const T *)p;
    s.n = n;
    s.ih = n / l;
    s.visc = visc;
    int rc = EC_OK;
    double *ws = scratch_on_device(ctx, &rc);
    if (!ws) return rc;
    for (int m = 0; m < nq; m++) {
        const ec_filter_quantity &src = q[m];
        WArg wa;
        for (int a = 0; a < 3; a++) {
            wa.len[a] = src.len[a];
            for (int t = 0; t < src.len[a]; t++) wa.w[a][t] = src.w[a][t];
        }
        double *wdev = ws + (size_t)nbar * n * n + (size_t)nbar * nbar * n;
        k_put_weights<<<1, 64, 0, ctx->stream>>>(wa, wdev);
        EC_CHECK_LAUNCH(ctx, "k_put_weights");
        const double *w[3] = {wdev, wdev + src.len[0], wdev + src.len[0] + src.len[1]};
        if ((rc = sep_launch<T>(ctx, s, nbar, src.source, src.len, src.off, w, ws, (double *)src.out))) return rc;
    }
    return EC_OK;
}

extern "C" int ec_filter_coarse_noalloc(ec_ctx *ctx, int dtype, int n, int nbar, double l, double visc,
                                        void *const u[3], const void *p, int nq, const ec_filter_quantity *q) {
    EC_REQUIRE(ctx, ctx && u && (nq == 0 || q), "ec_filter_coarse_noalloc: null argument");
    EC_REQUIRE(ctx, nbar >= 1 && n >= 2 && n % nbar == 0, "ec_filter_coarse_noalloc: n=%d not a multiple of nbar=%d",
               n, nbar);
    for (int m = 0; m < nq; m++) {
        const ec_filter_quantity &src = q[m];
        for (int a = 0; a < 3; a++) {
            EC_REQUIRE(ctx, src.len[a] >= 1 && src.len[a] <= 4 * n && src.w[a], "filter quantity %d: bad axis %d", m, a);
            EC_REQUIRE(ctx, src.len[a] <= EC_NOALLOC_TAPS,
                       "ec_filter_coarse_noalloc: quantity %d axis %d has %d taps, more than the %d it holds", m, a,
                       src.len[a], EC_NOALLOC_TAPS);
            EC_REQUIRE(ctx, src.off[a] >= -4 * n && src.off[a] <= 4 * n, "filter quantity %d: offset out of range", m);
        }
        EC_REQUIRE(ctx, src.out, "filter quantity %d: null output", m);
        EC_REQUIRE(ctx, src.source >= 0 && src.source <= EC_SRC_K3, "filter quantity %d: bad source", m);
        EC_REQUIRE(ctx, (src.source < EC_SRC_R11) || p, "pressure source without pressure field");
        const size_t need = sep_doubles(n, nbar, src.len);
        EC_REQUIRE(ctx, need <= EC_NOALLOC_DOUBLES,
                   "ec_filter_coarse_noalloc: n=%d, nbar=%d needs %zu doubles of scratch, more than its %zu "
                   "(nbar n^2 + nbar^2 n + taps)", n, nbar, need, EC_NOALLOC_DOUBLES);
    }
    if (nq == 0) return EC_OK;
    return dtype == EC_F64 ? noalloc_launch<double>(ctx, n, nbar, l, visc, u, p, nq, q)
                           : noalloc_launch<float>(ctx, n, nbar, l, visc, u, p, nq, q);
}

// applyfilter!(out, in, kernelproduct(w1, w2, w3)) as three 1D passes; the
// intermediates are FP64 in a second context workspace (2 n^3 doubles).
template <typename T>
static int apply_launch(ec_ctx *ctx, int n, const int len[3], const double *const w[3], const T *in, T *out) {
    const size_t N = ec_cube(n);
    if (ctx->ws2_doubles < 2 * N) {
        if (ctx->ws2) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(ctx->ws2);
            ctx->ws2 = nullptr;
            ctx->ws2_doubles = 0;
        }
        if (hipMalloc(&ctx->ws2, sizeof(double) * 2 * N) != hipSuccess)
            return ec_fail(ctx, EC_E_OOM, "ec_filter_apply: workspace of %zu doubles", 2 * N);
        ctx->ws2_doubles = 2 * N;
    }
    double *ta = (double *)ctx->ws2, *tb = ta + N;
    const double *wd[3];
    for (int a = 0; a < 3; a++) {
        int rc = EC_OK;
        wd[a] = weights_on_device(ctx, w[a], len[a], &rc);
        if (rc) return rc;
    }
    const unsigned blocks = (unsigned)((N + 255) / 256);
    ec_prof_scope pr(ctx, "filter_apply", (double)(2 * sizeof(T) + 32) * N);
    k_apply_axis<T, double><<<blocks, 256, 0, ctx->stream>>>(n, 0, len[0] / 2, wd[0], in, ta);
    EC_CHECK_LAUNCH(ctx, "k_apply_axis");
    k_apply_axis<double, double><<<blocks, 256, 0, ctx->stream>>>(n, 1, len[1] / 2, wd[1], ta, tb);
    EC_CHECK_LAUNCH(ctx, "k_apply_axis");
    k_apply_axis<double, T><<<blocks, 256, 0, ctx->stream>>>(n, 2, len[2] / 2, wd[2], tb, out);
    EC_CHECK_LAUNCH(ctx, "k_apply_axis");
    return EC_OK;
}

extern "C" int ec_filter_apply(ec_ctx *ctx, int dtype, int n, const int len[3], const double *const w[3],
                               const void *in, void *out) {
    EC_REQUIRE(ctx, ctx && len && w && in && out, "ec_filter_apply: null argument");
    EC_REQUIRE(ctx, n >= 2, "ec_filter_apply: n=%d", n);
    EC_REQUIRE(ctx, in != out, "ec_filter_apply: in place is not supported");
    for (int a = 0; a < 3; a++) {
        EC_REQUIRE(ctx, w[a] && len[a] >= 1 && (len[a] & 1), "ec_filter_apply: axis %d needs an odd, centred kernel", a);
        EC_REQUIRE(ctx, len[a] / 2 <= 4 * n, "ec_filter_apply: axis %d kernel radius beyond 4 n", a);
    }
    return dtype == EC_F64 ? apply_launch<double>(ctx, n, len, w, (const double *)in, (double *)out)
                           : apply_launch<float>(ctx, n, len, w, (const float *)in, (float *)out);
}

extern "C" int ec_filter_coarse(ec_ctx *ctx, int dtype, int n, int nbar, double l, double visc, void *const u[3],
                                const void *p, int nq, const ec_filter_quantity *q) {
    EC_REQUIRE(ctx, ctx && u && (nq == 0 || q), "ec_filter_coarse: null argument");
    EC_REQUIRE(ctx, nbar >= 1 && n >= 2 && n % nbar == 0, "ec_filter_coarse: n=%d not a multiple of nbar=%d", n, nbar);
    if (nq == 0) return EC_OK;
    return dtype == EC_F64 ? filter_launch<double>(ctx, n, nbar, l, visc, u, p, nq, q)
                           : filter_launch<float>(ctx, n, nbar, l, visc, u, p, nq, q);
}

// ---- closure sets: the 1D pieces of Appendix-style volume / face averages ----
// Offsets are relative to the coarse face index f = c(J+1)-1.
namespace {
struct Piece {
    std::vector<double> w;
    int off;
};
Piece box(int c) { return {std::vector<double>(c, 1.0 / c), -(c - 1)}; }
Piece face(int c) {
    if (c % 2) return {std::vector<double>(c, 1.0 / c), -(c / 2)};
    std::vector<double> w(c + 1, 1.0 / c);
    w[0] = w[c] = 0.5 / c;
    return {w, -(c / 2)};
}
Piece centre(int c) {
    if (c % 2) return {{1.0}, -(c / 2)};
    return {{0.5, 0.5}, -(c / 2)};
}
Piece point() { return {{1.0}, 0}; }

// r_ab = sigma_ab + p delta_ab as a filter source
const int rsrc[3][3] = {{EC_SRC_R11, EC_SRC_S12, EC_SRC_S31},
                        {EC_SRC_S12, EC_SRC_R22, EC_SRC_S23},
                        {EC_SRC_S31, EC_SRC_S23, EC_SRC_R33}};

// A list of filter quantities whose weight vectors stay put while it grows.
struct Quantities {
    std::deque<Piece> keep;
    std::vector<ec_filter_quantity> q;
    void add(int source, const Piece (&ax)[3], void *out) {
        if (!out) return;
        ec_filter_quantity f;
        f.source = source;
        for (int d = 0; d < 3; d++) {
            keep.push_back(ax[d]);
            f.len[d] = (int)keep.back().w.size();
            f.off[d] = keep.back().off;
            f.w[d] = keep.back().w.data();
        }
        f.out = out;
        q.push_back(f);
    }
};
}  // namespace

extern "C" int ec_ns_closure(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3], const void *p,
                             int nsets, const ec_closure_set *sets) {
    EC_REQUIRE(ctx, ctx && u && p && (nsets == 0 || sets), "ec_ns_closure: null argument");
    // one plane-marching launch for all sets when they fit the tiled scheme
    const int trc = ec_closure_tiled(ctx, dtype, n, l, visc, u, p, nsets, sets);
    if (trc != EC_E_UNSUPPORTED) return trc;
    for (int s = 0; s < nsets; s++) {
        const ec_closure_set &S = sets[s];
        EC_REQUIRE(ctx, S.nbar >= 1 && n % S.nbar == 0, "closure set %d: n=%d not a multiple of nbar=%d", s, n, S.nbar);
        EC_REQUIRE(ctx, S.kind == 0 || S.kind == 1, "closure set %d: kind %d", s, S.kind);
        const int c = n / S.nbar;
        Quantities Q;
        for (int a = 0; a < 3; a++) {
            Piece ax[3];
            for (int d = 0; d < 3; d++) ax[d] = d == a ? (S.kind == 0 ? face(c) : point()) : box(c);
            Q.add(EC_SRC_U1 + a, ax, S.ubar[a]);
        }
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) {
                Piece ax[3];
                for (int d = 0; d < 3; d++) {
                    if (a == b) ax[d] = d == a ? centre(c) : box(c);
                    else if (S.kind == 0) ax[d] = d == b ? point() : (d == a ? face(c) : box(c));
                    else ax[d] = (d == a || d == b) ? point() : box(c);
                }
                Q.add(rsrc[a][b], ax, S.flux[a + 3 * b]);
            }
        int rc = ec_filter_coarse(ctx, dtype, n, S.nbar, l, visc, u, p, (int)Q.q.size(), Q.q.data());
        if (rc) return rc;
    }
    return EC_OK;
}

extern "C" int ec_ns_commutator(ec_ctx *ctx, int nbar, double l, double visc, void *const ubar[3],
                                void *const flux9[9], void *const tau9[9]) {
    EC_REQUIRE(ctx, ctx && ubar && flux9 && tau9, "ec_ns_commutator: null argument");
    EC_REQUIRE(ctx, nbar >= 2, "ec_ns_commutator: nbar=%d", nbar);
    QList fl, tl;
    fl.nq = tl.nq = 9;
    for (int m = 0; m < 9; m++) {
        fl.q[m].out = (double *)flux9[m];
        tl.q[m].out = (double *)tau9[m];
    }
    const size_t nb3 = (size_t)nbar * nbar * nbar;
    k_commutator<<<(unsigned)((nb3 + 255) / 256), 256, 0, ctx->stream>>>(
        nbar, nbar / l, visc, (const double *)ubar[0], (const double *)ubar[1], (const double *)ubar[2], fl, tl);
    EC_CHECK_LAUNCH(ctx, "k_commutator");
    return EC_OK;
}

// ---- sub-filter stresses and their dissipation (src/NavierStokes.jl:470-593, process.jl:107-193) ----

extern "C" int ec_ns_sfs_filter(ec_ctx *ctx, int dtype, int n, int nbar, double l, double visc, void *const u[3],
                                const void *p, int kind, void *const fu[3], void *const fr9[9], void *const sr9[9]) {
    EC_REQUIRE(ctx, ctx && u && p && fu && fr9 && sr9, "ec_ns_sfs_filter: null argument");
    EC_REQUIRE(ctx, nbar >= 1 && n >= 2 && n % nbar == 0, "ec_ns_sfs_filter: n=%d not a multiple of nbar=%d", n,
               nbar);
    EC_REQUIRE(ctx, kind == 0 || kind == 1, "ec_ns_sfs_filter: kind %d (0 volume, 1 surface)", kind);
    const int c = n / nbar;
    Quantities Q;
    // the LES velocity filter: volume average (c^2 box times the face window
    // along a) or face average (c^2 box in the face plane)
    for (int a = 0; a < 3; a++) {
        Piece ax[3];
        for (int d = 0; d < 3; d++) ax[d] = d == a ? (kind == 0 ? face(c) : point()) : box(c);
        Q.add(EC_SRC_U1 + a, ax, fu[a]);
    }
    for (int b = 0; b < 3; b++)
        for (int a = 0; a < 3; a++) {
            Piece cl[3], sw[3];
            for (int d = 0; d < 3; d++) {
                if (a == b) {
                    // volume: the c^3 box of the coarse cell; surface: the c^2
                    // box of the planes normal to a through the coarse centre
                    cl[d] = kind == 0 ? box(c) : (d == a ? centre(c) : box(c));
                    sw[d] = d == a ? centre(c) : box(c);
                } else if (kind == 0) {
                    // classic: the volume around the coarse edge (a, b);
                    // swap: the face normal to b, the derivative direction
                    cl[d] = (d == a || d == b) ? face(c) : box(c);
                    sw[d] = d == b ? point() : (d == a ? face(c) : box(c));
                } else {
                    // classic: the face normal to a (the filter of u_a);
                    // swap: the line along the third direction
                    cl[d] = d == a ? point() : (d == b ? face(c) : box(c));
                    sw[d] = (d == a || d == b) ? point() : box(c);
                }
            }
            Q.add(rsrc[a][b], cl, fr9[a + 3 * b]);
            Q.add(rsrc[a][b], sw, sr9[a + 3 * b]);
        }
    return ec_filter_coarse(ctx, dtype, n, nbar, l, visc, u, p, (int)Q.q.size(), Q.q.data());
}

namespace {
struct Tensor9 {
    const double *s[9];
};

// d = sigma : grad u on the staggered grid, at the cell centres: the diagonal
// products sigma_aa delta_a u_a sit at the centre; each off-diagonal product
// sigma_ab delta_b u_a sits at its edge (a, b) and is averaged over the four
// edges around the centre, so sum_I d[I] = sum_I u . (-div^ns sigma) exactly.
__global__ void k_tensordissipation(int nb, double ih, Tensor9 S, const double *__restrict__ u1,
                                    const double *__restrict__ u2, const double *__restrict__ u3,
                                    double *__restrict__ d) {
    const size_t nb3 = (size_t)nb * nb * nb;
    const size_t I = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (I >= nb3) return;
    const double *u[3] = {u1, u2, u3};
    const int x[3] = {(int)(I % nb), (int)((I / nb) % nb), (int)(I / ((size_t)nb * nb))};
    auto at = [nb](const int (&y)[3]) { return (size_t)y[0] + (size_t)nb * ((size_t)y[1] + (size_t)nb * y[2]); };
    double acc = 0.0;
    for (int a = 0; a < 3; a++) {
        int m[3] = {x[0], x[1], x[2]};
        m[a] = ec_dn(m[a], nb);
        acc += S.s[4 * a][I] * ((u[a][I] - u[a][at(m)]) * ih);
        for (int b = 0; b < 3; b++) {
            if (b == a) continue;
            double e = 0.0;
            for (int sb = 0; sb < 2; sb++)
                for (int sa = 0; sa < 2; sa++) {
                    int J[3] = {x[0], x[1], x[2]};
                    if (sa) J[a] = ec_dn(J[a], nb);
                    if (sb) J[b] = ec_dn(J[b], nb);
                    int Jb[3] = {J[0], J[1], J[2]};
                    Jb[b] = ec_up(Jb[b], nb);
                    const size_t j = at(J);
                    e += S.s[a + 3 * b][j] * ((u[a][at(Jb)] - u[a][j]) * ih);
                }
            acc += 0.25 * e;
        }
    }
    d[I] = acc;
}
}  // namespace

extern "C" int ec_ns_tensordissipation(ec_ctx *ctx, int nbar, double l, void *const sigma9[9], void *const u[3],
                                       void *d) {
    EC_REQUIRE(ctx, ctx && sigma9 && u && d, "ec_ns_tensordissipation: null argument");
    EC_REQUIRE(ctx, nbar >= 2, "ec_ns_tensordissipation: nbar=%d", nbar);
    Tensor9 S;
    for (int m = 0; m < 9; m++) {
        EC_REQUIRE(ctx, sigma9[m], "ec_ns_tensordissipation: sigma component %d is null", m);
        S.s[m] = (const double *)sigma9[m];
    }
    EC_REQUIRE(ctx, u[0] && u[1] && u[2], "ec_ns_tensordissipation: null velocity component");
    const size_t nb3 = (size_t)nbar * nbar * nbar;
    k_tensordissipation<<<(unsigned)((nb3 + 255) / 256), 256, 0, ctx->stream>>>(
        nbar, nbar / l, S, (const double *)u[0], (const double *)u[1], (const double *)u[2], (double *)d);
    EC_CHECK_LAUNCH(ctx, "k_tensordissipation");
    return EC_OK;
}
