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

#include <map>

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
    for (int q0 = 0; q0 < nq; q0 += MAXQ) {
        QList ql;
        ql.nq = std::min(MAXQ, nq - q0);
        for (int m = 0; m < ql.nq; m++) {
            const ec_filter_quantity &src = q[q0 + m];
            QDesc &d = ql.q[m];
            d.source = src.source;
            for (int a = 0; a < 3; a++) {
                EC_REQUIRE(ctx, src.len[a] >= 1 && src.len[a] <= n && src.w[a], "filter quantity %d: bad axis %d", q0 + m, a);
                EC_REQUIRE(ctx, src.off[a] > -n && src.off[a] < n, "filter quantity %d: offset out of range", q0 + m);
                d.len[a] = src.len[a];
                d.off[a] = src.off[a];
                int rc = EC_OK;
                d.w[a] = weights_on_device(ctx, src.w[a], src.len[a], &rc);
                if (rc) return rc;
            }
            EC_REQUIRE(ctx, src.out, "filter quantity %d: null output", q0 + m);
            EC_REQUIRE(ctx, src.source >= 0 && src.source <= EC_SRC_K3, "filter quantity %d: bad source", q0 + m);
            EC_REQUIRE(ctx, (src.source < EC_SRC_R11) || p, "pressure source without pressure field");
            d.out = (double *)src.out;
        }
        dim3 grid((unsigned)((nb3 + 255) / 256), ql.nq);
        ec_prof_scope pr(ctx, "filter_coarse", (double)sizeof(T) * 4.0 * ec_cube(n) + 8.0 * nb3 * ql.nq);
        k_filter_coarse<T><<<grid, 256, 0, ctx->stream>>>(s, nbar, c, ql);
        EC_CHECK_LAUNCH(ctx, "k_filter_coarse");
    }
    return EC_OK;
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
}  // namespace

extern "C" int ec_ns_closure(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3], const void *p,
                             int nsets, const ec_closure_set *sets) {
    EC_REQUIRE(ctx, ctx && u && p && (nsets == 0 || sets), "ec_ns_closure: null argument");
    // one plane-marching launch for all sets when they fit the tiled scheme
    const int trc = ec_closure_tiled(ctx, dtype, n, l, visc, u, p, nsets, sets);
    if (trc != EC_E_UNSUPPORTED) return trc;
    static const int rsrc[3][3] = {{EC_SRC_R11, EC_SRC_S12, EC_SRC_S31},
                                   {EC_SRC_S12, EC_SRC_R22, EC_SRC_S23},
                                   {EC_SRC_S31, EC_SRC_S23, EC_SRC_R33}};
    for (int s = 0; s < nsets; s++) {
        const ec_closure_set &S = sets[s];
        EC_REQUIRE(ctx, S.nbar >= 1 && n % S.nbar == 0, "closure set %d: n=%d not a multiple of nbar=%d", s, n, S.nbar);
        EC_REQUIRE(ctx, S.kind == 0 || S.kind == 1, "closure set %d: kind %d", s, S.kind);
        const int c = n / S.nbar;
        std::vector<Piece> keep;
        keep.reserve(64);
        std::vector<ec_filter_quantity> q;
        auto add = [&](int source, Piece a, Piece b, Piece cc, void *out) {
            keep.push_back(a);
            keep.push_back(b);
            keep.push_back(cc);
            ec_filter_quantity f;
            f.source = source;
            const Piece *pp[3] = {&keep[keep.size() - 3], &keep[keep.size() - 2], &keep[keep.size() - 1]};
            for (int d = 0; d < 3; d++) {
                f.len[d] = (int)pp[d]->w.size();
                f.off[d] = pp[d]->off;
                f.w[d] = pp[d]->w.data();
            }
            f.out = out;
            q.push_back(f);
        };
        for (int a = 0; a < 3; a++) {
            if (!S.ubar[a]) continue;
            Piece ax[3];
            for (int d = 0; d < 3; d++) ax[d] = d == a ? (S.kind == 0 ? face(c) : point()) : box(c);
            add(EC_SRC_U1 + a, ax[0], ax[1], ax[2], S.ubar[a]);
        }
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) {
                void *out = S.flux[a + 3 * b];
                if (!out) continue;
                Piece ax[3];
                for (int d = 0; d < 3; d++) {
                    if (a == b) ax[d] = d == a ? centre(c) : box(c);
                    else if (S.kind == 0) ax[d] = d == b ? point() : (d == a ? face(c) : box(c));
                    else ax[d] = (d == a || d == b) ? point() : box(c);
                }
                add(rsrc[a][b], ax[0], ax[1], ax[2], out);
            }
        int rc = ec_filter_coarse(ctx, dtype, n, S.nbar, l, visc, u, p, (int)q.size(), q.data());
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
