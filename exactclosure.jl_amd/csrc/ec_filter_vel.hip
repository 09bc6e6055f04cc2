// Velocity-type filters of several LES grids in one launch, for the RK4
// hot path of the DNS-aided LES:
//
//   out_s[a] = V_s(v_a - delta_a p)     (p optional)
//
// V_s is the filter that defines the LES velocity of set s: volume average
// (face top-hat / trapezoid along a, box over the other two axes) or face
// average (face point along a, box over the others), the pieces of
// apply_DHpi (src/N.jl:1228-1237) for any compression.  Applied to the
// projected DNS tendency k = du - grad p this gives V(k) = -div_H^ns (S r),
// the filtered flux divergence of get_rDHi (src/N.jl:1038-1073), so the exact
// closure forcing of the LES is available without forming the nine stresses.
//
// Structure: the plane-marching block of the one-launch closure kernel, with
// the compressions and kinds of the sets as template parameters so every
// piece sum is an unrolled, compile-time-length loop.  All sums are FP64.
#include "ec_internal.h"

#include <algorithm>
#include <numeric>

namespace {

// periodic wrap of an index within one period of [0, n)
__device__ __forceinline__ int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// piece kinds: 0 box (cells cI..cI+c-1), 1 face (top-hat / trapezoid around
// the coarse face f = c(I+1)-1), 2 point at f
template <int C, int PIECE, typename V>
__device__ __forceinline__ double psum(const V *p, int stride, int f) {
    if constexpr (PIECE == 0) {
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int t = 0; t < C; t++) {
            if (t & 1) a1 += (double)p[(f - (C - 1) + t) * stride];
            else a0 += (double)p[(f - (C - 1) + t) * stride];
        }
        return (a0 + a1) * (1.0 / C);
    } else if constexpr (PIECE == 1) {
        constexpr int H = C / 2;
        double a0 = 0.0, a1 = 0.0;
        if constexpr (C & 1) {
#pragma unroll
            for (int t = -H; t <= H; t++) {
                if (t & 1) a1 += (double)p[(f + t) * stride];
                else a0 += (double)p[(f + t) * stride];
            }
            return (a0 + a1) * (1.0 / C);
        } else {
#pragma unroll
            for (int t = -H + 1; t <= H - 1; t++) {
                if (t & 1) a1 += (double)p[(f + t) * stride];
                else a0 += (double)p[(f + t) * stride];
            }
            return ((a0 + a1) + 0.5 * ((double)p[(f - H) * stride] + (double)p[(f + H) * stride])) * (1.0 / C);
        }
    } else {
        return (double)p[f * stride];
    }
}

// piece of component a along axis d for a set of kind K (0 volume, 1 face)
template <int K>
__host__ __device__ constexpr int piece_of(int a, int d) {
    return d == a ? (K == 0 ? 1 : 2) : 0;
}

// high extent of a piece beyond the coarse face (face: c/2, else 0)
template <int C>
constexpr int hext() { return C / 2; }

struct FSet {
    double *out[3];
    int nb;
};

struct FArgs {
    FSet s[4];
};

template <int C, int K>
struct SetCfg {
    static constexpr int c = C, kind = K;
};

// ---- per-set phases, instantiated per (C, K) ----
template <typename T, int C, int K>
__device__ __forceinline__ void xpartials(const T *S, int SSTR, int SXP, int TX, int TY, double *XP, int XOFF,
                                          int tid, int nthr) {
    // items ordered (component, Ix, row): rows fastest so a wave shares one piece
    constexpr int R_EXTRA = C / 2;
    const int rows = TY + R_EXTRA;
    const int nix = TX / C;
    const int per_a = nix * rows;
    for (int it = tid; it < 3 * per_a; it += nthr) {
        const int a = it / per_a, r0 = it - a * per_a;
        const int Ix = r0 / rows, row = r0 - Ix * rows;
        const T *srow = S + a * SSTR + row * SXP;
        const int f = C * (Ix + 1) - 1;
        double v;
        if (a == 0) v = psum<C, piece_of<K>(0, 0)>(srow, 1, f);
        else v = psum<C, 0>(srow, 1, f);  // components 2, 3: box along x
        XP[XOFF + (a * nix + Ix) * 32 + row] = v;
    }
}

template <typename T, int C, int K>
__device__ __forceinline__ void slots(const double *XP, int XOFF, int TX, int TY, int ZC, int x0, int y0, int z0,
                                      int k, int tid, int nthr, int soff, const FSet &fs, double *acc0,
                                      double *acc1, int m_base) {
    // slot (component, Iy, Ix) for this set, slots soff.. handled by threads
    const int nix = TX / C, niy = TY / C;
    const int ns = 3 * nix * niy;
    const int sl = tid - soff;
    (void)m_base;
    if (sl < 0 || sl >= ns) return;
    const int a = sl / (nix * niy), c0 = sl - a * nix * niy;
    const int Iy = c0 / nix, Ix = c0 - Iy * nix;
    const double *xcol = XP + XOFF + (a * nix + Ix) * 32;
    const int fy = C * (Iy + 1) - 1;
    double P;
    if (a == 1) P = psum<C, piece_of<K>(1, 1)>(xcol, 1, fy);
    else P = psum<C, 0>(xcol, 1, fy);
    // z piece: component 3 takes the face piece, the others the box
    const int Kmin = z0 / C, Kmax = (z0 + ZC) / C;
    const int K1 = k / C;
#pragma unroll
    for (int dd = 0; dd < 2; dd++) {
        const int Kc = K1 - dd;
        if (Kc < Kmin || Kc >= Kmax) continue;
        const int t = k - (C * (Kc + 1) - 1);
        double w = 0.0;
        if (a == 2) {
            if constexpr (K == 0) {
                if (t >= -(C / 2) && t <= C / 2)
                    w = ((C & 1) == 0 && (t == -(C / 2) || t == C / 2)) ? 0.5 / C : 1.0 / C;
            } else {
                w = t == 0 ? 1.0 : 0.0;
            }
        } else {
            w = (t >= -(C - 1) && t <= 0) ? 1.0 / C : 0.0;
        }
        if (Kc & 1) *acc1 += w * P;
        else *acc0 += w * P;
        if (t == C / 2) {
            const size_t nb = (size_t)fs.nb;
            const size_t o = (size_t)(x0 / C + Ix) + nb * ((size_t)(y0 / C + Iy) + nb * (size_t)Kc);
            double *out = fs.out[a];
            if (Kc & 1) {
                out[o] = *acc1;
                *acc1 = 0.0;
            } else {
                out[o] = *acc0;
                *acc0 = 0.0;
            }
        }
    }
}

// Up to four sets; C = 0 marks an absent set.
template <typename T, int C0, int K0, int C1, int K1, int C2, int K2, int C3, int K3>
__global__ void __launch_bounds__(256) k_filter_vel(const T *__restrict__ v1, const T *__restrict__ v2,
                                                   const T *__restrict__ v3, const T *__restrict__ p, int n, int TX,
                                                   int TY, int ZC, double ih, FArgs A) {
    constexpr int CMAX = std::max(std::max(C0, C1), std::max(C2, C3));
    constexpr int HMAX = CMAX / 2;
    extern __shared__ double smem_d[];
    const int SX = TX + HMAX, SY = TY + HMAX;
    const int SXP = SX | 1;
    const int SSTR = (SXP * SY) | 1;
    // XP: per set, 3 components x (TX/C) columns x 32 rows (rows <= TY + C/2 <= 32)
    constexpr int NCOL0 = C0 ? 3 : 0, NCOL1 = C1 ? 3 : 0, NCOL2 = C2 ? 3 : 0, NCOL3 = C3 ? 3 : 0;
    const int xo0 = 0;
    const int xo1 = xo0 + (C0 ? NCOL0 * (TX / (C0 ? C0 : 1)) * 32 : 0);
    const int xo2 = xo1 + (C1 ? NCOL1 * (TX / (C1 ? C1 : 1)) * 32 : 0);
    const int xo3 = xo2 + (C2 ? NCOL2 * (TX / (C2 ? C2 : 1)) * 32 : 0);
    const int xend = xo3 + (C3 ? NCOL3 * (TX / (C3 ? C3 : 1)) * 32 : 0);
    double *XP = smem_d;
    T *S = reinterpret_cast<T *>(XP + xend);          // 3 source planes
    T *Pp = S + 3 * SSTR;                             // p at planes k, k+1: (SX+1) x (SY+1) each
    const int PW = SX + 1, PH = SY + 1;
    const size_t nn = (size_t)n, n2 = nn * nn;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, z0 = blockIdx.z * ZC;
    const int tid = threadIdx.x, nthr = blockDim.x;

    double acc0 = 0.0, acc1 = 0.0;
    // slot ownership: one slot per thread, sets laid out consecutively
    const int ns0 = C0 ? 3 * (TX / (C0 ? C0 : 1)) * (TY / (C0 ? C0 : 1)) : 0;
    const int ns1 = C1 ? 3 * (TX / (C1 ? C1 : 1)) * (TY / (C1 ? C1 : 1)) : 0;
    const int ns2 = C2 ? 3 * (TX / (C2 ? C2 : 1)) * (TY / (C2 ? C2 : 1)) : 0;
    const int so1 = ns0, so2 = so1 + ns1, so3 = so2 + ns2;

    // fetch layout: 32 lanes along x, rows ly0, ly0 + 8, ... (NPASS passes)
    constexpr int LW = 32, RP = 256 / LW, NPASS = 4;  // rows <= 32 (checked on the host)
    const int lx = tid % LW, ly0 = tid / LW;
    const bool in_v = lx < SX, in_p = lx < PW;
    const size_t Xo = (size_t)wrap1(x0 + lx, n);
    auto rowoff = [&](int ly) { return Xo + nn * (size_t)wrap1(y0 + ly, n); };

    // v at plane z0, p at planes z0 and z0+1
    {
        const size_t k0 = (size_t)wrap1(z0, n) * n2, k1 = (size_t)wrap1(z0 + 1, n) * n2;
#pragma unroll
        for (int i = 0; i < NPASS; i++) {
            const int ly = ly0 + i * RP;
            if (in_v && ly < SY) {
                const size_t o = rowoff(ly) + k0;
                const int so = ly * SXP + lx;
                S[so] = v1[o];
                S[SSTR + so] = v2[o];
                S[2 * SSTR + so] = v3[o];
            }
            if (in_p && ly < PH) {
                const size_t o = rowoff(ly);
                Pp[ly * PW + lx] = p ? p[o + k0] : T(0);
                Pp[PW * PH + ly * PW + lx] = p ? p[o + k1] : T(0);
            }
        }
    }
    __syncthreads();
    const int kend = z0 + ZC - 1 + HMAX;
    for (int k = z0; k <= kend; ++k) {
        const int cur = (k - z0) & 1, nxt = cur ^ 1;
        // prefetch v(k+1) and p(k+2); they land after the slot phase
        T pv[3][NPASS], pp[NPASS];
        {
            const size_t kv = (size_t)wrap1(k + 1, n) * n2, kq = (size_t)wrap1(k + 2, n) * n2;
#pragma unroll
            for (int i = 0; i < NPASS; i++) {
                const int ly = ly0 + i * RP;
                const size_t o = rowoff(ly < PH ? ly : 0);
                if (in_v && ly < SY) {
                    pv[0][i] = v1[o + kv];
                    pv[1][i] = v2[o + kv];
                    pv[2][i] = v3[o + kv];
                }
                if (p && in_p && ly < PH) pp[i] = p[o + kq];
            }
        }
        // sources: k_a = v_a - delta_a p
        if (p) {
            const T *pc = Pp + cur * PW * PH, *pn = Pp + nxt * PW * PH;
            for (int e = tid; e < SX * SY; e += nthr) {
                const int ly = (int)(((float)e + 0.5f) * (1.0f / SX)), lx2 = e - ly * SX;
                const int so = ly * SXP + lx2, pi = ly * PW + lx2;
                const double p0 = (double)pc[pi];
                S[so] = (T)((double)S[so] - ((double)pc[pi + 1] - p0) * ih);
                S[SSTR + so] = (T)((double)S[SSTR + so] - ((double)pc[pi + PW] - p0) * ih);
                S[2 * SSTR + so] = (T)((double)S[2 * SSTR + so] - ((double)pn[pi] - p0) * ih);
            }
        }
        __syncthreads();
        if constexpr (C0 > 0) xpartials<T, C0, K0>(S, SSTR, SXP, TX, TY, XP, xo0, tid, nthr);
        if constexpr (C1 > 0) xpartials<T, C1, K1>(S, SSTR, SXP, TX, TY, XP, xo1, tid, nthr);
        if constexpr (C2 > 0) xpartials<T, C2, K2>(S, SSTR, SXP, TX, TY, XP, xo2, tid, nthr);
        if constexpr (C3 > 0) xpartials<T, C3, K3>(S, SSTR, SXP, TX, TY, XP, xo3, tid, nthr);
        __syncthreads();
        if constexpr (C0 > 0) slots<T, C0, K0>(XP, xo0, TX, TY, ZC, x0, y0, z0, k, tid, nthr, 0, A.s[0], &acc0, &acc1, 0);
        if constexpr (C1 > 0) slots<T, C1, K1>(XP, xo1, TX, TY, ZC, x0, y0, z0, k, tid, nthr, so1, A.s[1], &acc0, &acc1, 0);
        if constexpr (C2 > 0) slots<T, C2, K2>(XP, xo2, TX, TY, ZC, x0, y0, z0, k, tid, nthr, so2, A.s[2], &acc0, &acc1, 0);
        if constexpr (C3 > 0) slots<T, C3, K3>(XP, xo3, TX, TY, ZC, x0, y0, z0, k, tid, nthr, so3, A.s[3], &acc0, &acc1, 0);
#pragma unroll
        for (int i = 0; i < NPASS; i++) {
            const int ly = ly0 + i * RP;
            if (in_v && ly < SY) {
                const int so = ly * SXP + lx;
                S[so] = pv[0][i];
                S[SSTR + so] = pv[1][i];
                S[2 * SSTR + so] = pv[2][i];
            }
            if (in_p && ly < PH) Pp[cur * PW * PH + ly * PW + lx] = p ? pp[i] : T(0);
        }
        __syncthreads();
    }
}

int largest_tile(int n, int L, int target) {
    int best = 0;
    for (int t = L; t <= n; t += L)
        if (n % t == 0 && t <= target) best = t;
    if (best == 0)
        for (int t = L; t <= n; t += L)
            if (n % t == 0) return t;
    return best;
}

template <typename T, int C0, int K0, int C1, int K1, int C2, int K2, int C3, int K3>
int launch_fv(ec_ctx *ctx, int n, double l, void *const v[3], const void *p, const FArgs &A, int tx_target,
              int tz_target) {
    constexpr int CMAX = std::max(std::max(C0, C1), std::max(C2, C3));
    int L = 1;
    for (int c : {C0, C1, C2, C3})
        if (c) L = std::lcm(L, c);
    if (n % L) return EC_E_UNSUPPORTED;
    const int TX = largest_tile(n, L, tx_target), TY = TX, ZC = largest_tile(n, L, tz_target);
    // rows and columns (with the p halo) fit one 32 x 32 fetch; every plane index
    // read (up to n + CMAX/2 + 1) lies within one period, as wrap1 assumes
    if (!TX || !ZC || TY + CMAX / 2 + 1 > 32 || TX + CMAX / 2 + 1 > 32 || CMAX / 2 + 1 >= n) return EC_E_UNSUPPORTED;
    const int SX = TX + CMAX / 2, SY = TY + CMAX / 2, SXP = SX | 1, SSTR = (SXP * SY) | 1;
    int slots = 0, xp = 0;
    for (int c : {C0, C1, C2, C3})
        if (c) {
            slots += 3 * (TX / c) * (TY / c);
            xp += 3 * (TX / c) * 32;
        }
    const int threads = 256;
    if (slots > threads) return EC_E_UNSUPPORTED;
    const size_t lds = (size_t)xp * 8 + sizeof(T) * (3 * (size_t)SSTR + 2 * (size_t)(SX + 1) * (SY + 1));
    if (lds > 150 * 1024) return EC_E_UNSUPPORTED;
    dim3 grid(n / TX, n / TY, n / ZC);
    ec_prof_scope pr(ctx, "filter_vel", sizeof(T) * (p ? 4.0 : 3.0) * ec_cube(n));
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_filter_vel<T, C0, K0, C1, K1, C2, K2, C3, K3>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_filter_vel<T, C0, K0, C1, K1, C2, K2, C3, K3><<<grid, threads, lds, ctx->stream>>>(
        (const T *)v[0], (const T *)v[1], (const T *)v[2], (const T *)p, n, TX, TY, ZC, n / l, A);
    EC_CHECK_LAUNCH(ctx, "k_filter_vel");
    return EC_OK;
}

// one set with compression c (instantiated for the common compressions)
template <typename T, int K>
int launch_single(ec_ctx *ctx, int n, double l, void *const v[3], const void *p, const FArgs &A, int c, int tx,
                  int tz) {
    switch (c) {
        case 2: return launch_fv<T, 2, K, 0, 0, 0, 0, 0, 0>(ctx, n, l, v, p, A, tx, tz);
        case 3: return launch_fv<T, 3, K, 0, 0, 0, 0, 0, 0>(ctx, n, l, v, p, A, tx, tz);
        case 4: return launch_fv<T, 4, K, 0, 0, 0, 0, 0, 0>(ctx, n, l, v, p, A, tx, tz);
        case 5: return launch_fv<T, 5, K, 0, 0, 0, 0, 0, 0>(ctx, n, l, v, p, A, tx, tz);
        case 8: return launch_fv<T, 8, K, 0, 0, 0, 0, 0, 0>(ctx, n, l, v, p, A, tx, tz);
        case 16: return launch_fv<T, 16, K, 0, 0, 0, 0, 0, 0>(ctx, n, l, v, p, A, tx, tz);
        default: return EC_E_UNSUPPORTED;
    }
}

template <typename T>
int filter_vel_t(ec_ctx *ctx, int n, double l, void *const v[3], const void *p, int nsets, const ec_vel_set *sets,
                 int tx, int tz) {
    // the four-set volume/face configuration of the 1024^3 workload in one launch
    if (nsets == 4) {
        FArgs A;
        int cs[4], ks[4];
        for (int s = 0; s < 4; s++) {
            cs[s] = n / sets[s].nbar;
            ks[s] = sets[s].kind;
            for (int a = 0; a < 3; a++) A.s[s].out[a] = (double *)sets[s].out[a];
            A.s[s].nb = sets[s].nbar;
        }
        if (cs[0] == 16 && ks[0] == 0 && cs[1] == 8 && ks[1] == 0 && cs[2] == 4 && ks[2] == 0 && cs[3] == 4 && ks[3] == 1)
            return launch_fv<T, 16, 0, 8, 0, 4, 0, 4, 1>(ctx, n, l, v, p, A, tx, tz);
    }
    // otherwise one launch per set
    for (int s = 0; s < nsets; s++) {
        FArgs A;
        for (int a = 0; a < 3; a++) A.s[0].out[a] = (double *)sets[s].out[a];
        A.s[0].nb = sets[s].nbar;
        const int c = n / sets[s].nbar;
        int rc = sets[s].kind == 0 ? launch_single<T, 0>(ctx, n, l, v, p, A, c, tx, tz)
                                   : launch_single<T, 1>(ctx, n, l, v, p, A, c, tx, tz);
        if (rc) return rc;
    }
    return EC_OK;
}

}  // namespace

static int g_fv_tx = 16, g_fv_tz = 64;

extern "C" int ec_filter_vel_set_tiles(ec_ctx *ctx, int tx, int tz) {
    EC_REQUIRE(ctx, ctx && tx >= 1 && tz >= 1, "ec_filter_vel_set_tiles: bad argument");
    g_fv_tx = tx;
    g_fv_tz = tz;
    return EC_OK;
}

extern "C" int ec_ns_filter_vel(ec_ctx *ctx, int dtype, int n, double l, void *const v[3], const void *p, int nsets,
                                const ec_vel_set *sets) {
    EC_REQUIRE(ctx, ctx && v && (nsets == 0 || sets), "ec_ns_filter_vel: null argument");
    for (int s = 0; s < nsets; s++) {
        EC_REQUIRE(ctx, sets[s].nbar >= 1 && n % sets[s].nbar == 0, "set %d: n=%d not a multiple of nbar=%d", s, n,
                   sets[s].nbar);
        EC_REQUIRE(ctx, sets[s].kind == 0 || sets[s].kind == 1, "set %d: kind %d", s, sets[s].kind);
        for (int a = 0; a < 3; a++) EC_REQUIRE(ctx, sets[s].out[a], "set %d: null output", s);
    }
    int rc = dtype == EC_F64 ? filter_vel_t<double>(ctx, n, l, v, p, nsets, sets, g_fv_tx, g_fv_tz)
                             : filter_vel_t<float>(ctx, n, l, v, p, nsets, sets, g_fv_tx, g_fv_tz);
    if (rc != EC_E_UNSUPPORTED) return rc;
    // any other compression: the per-quantity separable filter
    static const int src_k[3] = {EC_SRC_K1, EC_SRC_K2, EC_SRC_K3};
    for (int s = 0; s < nsets; s++) {
        const int c = n / sets[s].nbar;
        std::vector<std::vector<double>> keep;
        std::vector<ec_filter_quantity> q(3);
        for (int a = 0; a < 3; a++) {
            q[a].source = p ? src_k[a] : EC_SRC_U1 + a;
            for (int d = 0; d < 3; d++) {
                std::vector<double> w;
                int off;
                if (d != a) {
                    w.assign(c, 1.0 / c);
                    off = -(c - 1);
                } else if (sets[s].kind == 1) {
                    w.assign(1, 1.0);
                    off = 0;
                } else if (c & 1) {
                    w.assign(c, 1.0 / c);
                    off = -(c / 2);
                } else {
                    w.assign(c + 1, 1.0 / c);
                    w[0] = w[c] = 0.5 / c;
                    off = -(c / 2);
                }
                keep.push_back(w);
                q[a].len[d] = (int)w.size();
                q[a].off[d] = off;
            }
            q[a].out = sets[s].out[a];
        }
        for (int a = 0; a < 3; a++)
            for (int d = 0; d < 3; d++) q[a].w[d] = keep[a * 3 + d].data();
        // velocity sources come from v (the filter kernel reads them as u)
        rc = ec_filter_coarse(ctx, dtype, n, sets[s].nbar, l, 0.0, v, p, 3, q.data());
        if (rc) return rc;
    }
    return EC_OK;
}
