// Exact-closure filters of several LES grids in one launch (get_rDHi,
// apply_DHpi, src/N.jl:1038-1073, 1228-1237, generalised to any compression
// with the volume-average and face-average pieces).
//
// Every quantity is a separable product of four 1D pieces (box over cells,
// face top-hat/trapezoid, centre point/pair, face point), each written as
// taps t in [lo, hi] with weight we at both ends and wi inside.  A workgroup
// owns a TX x TY x ZC block of fine cells (a multiple of every compression)
// and marches through its planes:
//   1. velocity planes k-1, k, k+1 (+ halo) and p(k) are resident in LDS;
//      plane k+2 and p(k+1) are fetched into registers meanwhile
//   2. the nine sources u1..u3, r11..r33 = sigma_aa + p, s12, s23, s31 are
//      formed at plane k (stresses in FP64 from the stored values)
//   3. x-partials: every (source, x-piece, coarse x) sum for every row
//   4. y-partials per (set, coarse x, coarse y, quantity), weighted by the
//      z-piece and accumulated in registers; a coarse cell is written once
//      its last plane has been seen.
// The per-partial and per-slot descriptors live in a small device table, so
// nothing in the kernel indexes its arguments dynamically.  The fine fields
// are read once per block (plus an L2-served halo of c_max/2 + 1), independent
// of the number of sets; all sums are FP64.
#include "ec_internal.h"

#include <algorithm>
#include <map>
#include <numeric>

namespace {

constexpr int MAXSETS = 4;
enum { PB = 0, PF = 1, PC = 2, PP = 3 };  // box, face, centre, point

struct Piece {
    int lo, hi;
    double wi, we;
};

Piece make_piece(int pc, int c) {
    switch (pc) {
        case PB: return {-(c - 1), 0, 1.0 / c, 1.0 / c};
        case PF:
            if (c & 1) return {-(c / 2), c / 2, 1.0 / c, 1.0 / c};
            return {-(c / 2), c / 2, 1.0 / c, 0.5 / c};
        case PC:
            if (c & 1) return {-(c / 2), -(c / 2), 1.0, 1.0};
            return {-(c / 2), -(c / 2) + 1, 0.5, 0.5};
        default: return {0, 0, 1.0, 1.0};
    }
}

// x-partial descriptor: S[src][row][f+lo..f+hi] weighted
struct XD {
    int soff;   // source * source-plane stride
    int lo, hi; // absolute columns f+lo, f+hi in the source tile
    int pad;
    double wi, we;
};

// slot descriptor: (set, coarse x, coarse y, quantity)
struct SD {
    double *out;
    double wiy, wey, wiz, wez;
    int xcol;        // column in the x-partial table
    int ylo, yhi;    // absolute rows f_y+lo .. f_y+hi
    int zlo, zhi;    // z taps relative to the coarse face
    int c, Ix, Iy, nb;
    int pad;
};

__device__ __forceinline__ int wrapn(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}


// Sum of L values p[0], p[stride], ..., unrolled with four independent
// accumulators so all LDS loads issue back to back.
template <int L, typename V>
__device__ __forceinline__ double usum(const V *p, int stride) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
    for (int t = 0; t < L; t++) {
        const double v = (double)p[t * stride];
        if ((t & 3) == 0) a0 += v;
        else if ((t & 3) == 1) a1 += v;
        else if ((t & 3) == 2) a2 += v;
        else a3 += v;
    }
    return (a0 + a1) + (a2 + a3);
}

// we * (p[lo] + p[hi]) + wi * (p[lo+1] + ... + p[hi-1]); lo == hi: wi * p[lo].
// Interior lengths of the box / face / centre pieces at power-of-two
// compressions are unrolled; any other length takes the loop.
template <typename V>
__device__ __forceinline__ double piece_sum(const V *p, int stride, int lo, int hi, double wi, double we) {
    if (lo == hi) return wi * (double)p[lo * stride];
    const V *q = p + (lo + 1) * stride;
    double in;
    switch (hi - lo - 1) {
        case 0: in = 0.0; break;
        case 1: in = usum<1>(q, stride); break;
        case 2: in = usum<2>(q, stride); break;
        case 3: in = usum<3>(q, stride); break;
        case 6: in = usum<6>(q, stride); break;
        case 7: in = usum<7>(q, stride); break;
        case 14: in = usum<14>(q, stride); break;
        case 15: in = usum<15>(q, stride); break;
        case 30: in = usum<30>(q, stride); break;
        case 31: in = usum<31>(q, stride); break;
        default: {
            in = 0.0;
            for (int t = 0; t < hi - lo - 1; t++) in += (double)q[t * stride];
        }
    }
    return we * ((double)p[lo * stride] + (double)p[hi * stride]) + wi * in;
}

template <typename T, int MAXS, int MAXX, int LW, int NP>
__global__ void __launch_bounds__(512) k_closure_tiled(const T *__restrict__ u1, const T *__restrict__ u2,
                                                      const T *__restrict__ u3, const T *__restrict__ p,
                                                      const XD *__restrict__ xd, const SD *__restrict__ sd, int n,
                                                      int TX, int TY, int ZC, int SX, int SY, int UX, int UY,
                                                      int NXP, int nslots, int hmax, int SST, int SXP, int SYP,
                                                      int4 ibase, int4 rows, int4 cbase, double ih, double visc) {
    extern __shared__ double smem_d[];
    double *XP = smem_d;
    XD *XDs = reinterpret_cast<XD *>(XP + (size_t)SYP * NXP);
    T *U = reinterpret_cast<T *>(XDs + NXP);
    T *PS = U + 9 * UX * UY;
    T *S = PS + SX * SY;  // nine source planes, stride SST (odd: planes start on different banks)
    const size_t nn = (size_t)n, n2 = nn * nn;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, z0 = blockIdx.z * ZC;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int plane_u = UX * UY, plane_s = SX * SY;
    const int nfetch = 3 * plane_u + plane_s;

    for (int j = tid; j < NXP; j += nthr) XDs[j] = xd[j];
    // a / b for the small non-negative integers here (a < 2^16): the float
    // product is never within 2^-10 of an integer boundary, so truncation is exact
    const float r_pu = 1.0f / plane_u, r_ux = 1.0f / UX, r_sx = 1.0f / SX;
    auto q = [](int a, float rb) { return (int)(((float)a + 0.5f) * rb); };
    // x-partial items: the columns of set s cover rows [0, rows.s) only
    // (TY + c/2: the rows its y pieces reach); set s starts at item ibase.s
    // and column cbase.s
    const int nxitems = ibase.w + rows.w * (NXP - cbase.w);

    // fetch layout: LW lanes along x, rows ly0, ly0 + RP, ... (NP passes); the
    // three velocity components share one row offset per pass
    constexpr int RP = 512 / LW;
    const int lx = tid % LW, ly0 = tid / LW;
    const size_t Xu = (size_t)wrapn(x0 - 1 + lx, n), Xs = (size_t)wrapn(x0 + lx, n);
    const bool in_ux = lx < UX, in_sx = lx < SX;

    // planes z0-1, z0, z0+1 into ring slots 0, 1, 2 and p(z0)
#pragma unroll
    for (int i = 0; i < NP; i++) {
        const int ly = ly0 + i * RP;
        if (in_ux && ly < UY) {
            const size_t rowoff = Xu + nn * (size_t)wrapn(y0 - 1 + ly, n);
#pragma unroll
            for (int sl = 0; sl < 3; sl++) {
                const size_t off = rowoff + n2 * (size_t)wrapn(z0 - 1 + sl, n);
                U[(sl * 3 + 0) * plane_u + ly * UX + lx] = u1[off];
                U[(sl * 3 + 1) * plane_u + ly * UX + lx] = u2[off];
                U[(sl * 3 + 2) * plane_u + ly * UX + lx] = u3[off];
            }
        }
        if (in_sx && ly < SY) PS[ly * SX + lx] = p[Xs + nn * (size_t)wrapn(y0 + ly, n) + n2 * (size_t)wrapn(z0, n)];
    }
    __syncthreads();

    double acc0[MAXS], acc1[MAXS];
#pragma unroll
    for (int m = 0; m < MAXS; m++) acc0[m] = acc1[m] = 0.0;

    const int kend = z0 + ZC - 1 + hmax;
    for (int k = z0; k <= kend; ++k) {
        const int lk = k - (z0 - 1);
        const int sm = (lk - 1) % 3, s0 = lk % 3, sp = (lk + 1) % 3;
        // prefetch plane k+2 and p(k+1) into registers; they land in slot sm
        // (plane k-1) and the p tile after the slot phase, so the whole plane's
        // work covers the memory latency
        T pre[4 * NP];
        {
            const size_t ku = (size_t)wrapn(k + 2, n) * n2, kp = (size_t)wrapn(k + 1, n) * n2;
#pragma unroll
            for (int i = 0; i < NP; i++) {
                const int ly = ly0 + i * RP;
                if (in_ux && ly < UY) {
                    const size_t off = Xu + nn * (size_t)wrapn(y0 - 1 + ly, n) + ku;
                    pre[0 * NP + i] = u1[off];
                    pre[1 * NP + i] = u2[off];
                    pre[2 * NP + i] = u3[off];
                }
                if (in_sx && ly < SY) pre[3 * NP + i] = p[Xs + nn * (size_t)wrapn(y0 + ly, n) + kp];
            }
        }
        // ---- sources at plane k
        for (int idx = tid; idx < plane_s; idx += nthr) {
            const int ly = idx / SX, lx = idx - ly * SX;
            const int a = lx + 1, b = ly + 1;
            const T *Us0 = U + s0 * 3 * plane_u, *Usp = U + sp * 3 * plane_u, *Usm = U + sm * 3 * plane_u;
            const int c0 = b * UX + a;
            const double u1c = Us0[c0], u1xm = Us0[c0 - 1], u1yp = Us0[c0 + UX], u1zp = Usp[c0];
            const double u2c = Us0[plane_u + c0], u2ym = Us0[plane_u + c0 - UX], u2xp = Us0[plane_u + c0 + 1],
                         u2zp = Usp[plane_u + c0];
            const double u3c = Us0[2 * plane_u + c0], u3zm = Usm[2 * plane_u + c0],
                         u3xp = Us0[2 * plane_u + c0 + 1], u3yp = Us0[2 * plane_u + c0 + UX];
            const double e1 = (u1c + u1xm) / 2, d1 = (u1c - u1xm) * ih;
            const double e2 = (u2c + u2ym) / 2, d2 = (u2c - u2ym) * ih;
            const double e3 = (u3c + u3zm) / 2, d3 = (u3c - u3zm) * ih;
            const double pp = (double)PS[idx];
            const int so = ly * SXP + lx;  // padded position in the source planes
            S[0 * SST + so] = (T)u1c;
            S[1 * SST + so] = (T)u2c;
            S[2 * SST + so] = (T)u3c;
            S[3 * SST + so] = (T)(e1 * e1 - visc * (d1 + d1) + pp);
            S[4 * SST + so] = (T)(e2 * e2 - visc * (d2 + d2) + pp);
            S[5 * SST + so] = (T)(e3 * e3 - visc * (d3 + d3) + pp);
            S[6 * SST + so] =
                (T)(((u1yp + u1c) / 2) * ((u2xp + u2c) / 2) - visc * ((u1yp - u1c) * ih + (u2xp - u2c) * ih));
            S[7 * SST + so] =
                (T)(((u2zp + u2c) / 2) * ((u3yp + u3c) / 2) - visc * ((u2zp - u2c) * ih + (u3yp - u3c) * ih));
            S[8 * SST + so] =
                (T)(((u3xp + u3c) / 2) * ((u1zp + u1c) / 2) - visc * ((u3xp - u3c) * ih + (u1zp - u1c) * ih));
        }
        __syncthreads();
        // ---- x-partials for every row of the source tile
#pragma unroll 1
        for (int m = 0; m < MAXX; m++) {
            const int it = tid + m * nthr;
            if (it < nxitems) {
                // rows fastest: the lanes of a wave share one (source, piece) column
                int b = 0, R = rows.x, cb = 0;
                if (it >= ibase.y) { b = ibase.y; R = rows.y; cb = cbase.y; }
                if (it >= ibase.z) { b = ibase.z; R = rows.z; cb = cbase.z; }
                if (it >= ibase.w) { b = ibase.w; R = rows.w; cb = cbase.w; }
                const int jl = q(it - b, 1.0f / R), row = it - b - jl * R;
                const int j = cb + jl;
                const XD d = XDs[j];
                const T *srow = S + d.soff + row * SXP;  // d.soff = source * SST
                XP[j * SYP + row] = piece_sum(srow, 1, d.lo, d.hi, d.wi, d.we);
            }
        }
        __syncthreads();
        // ---- y-partials, z accumulation, write-out
#pragma unroll
        for (int m = 0; m < MAXS; m++) {
            const int sl = tid + m * nthr;
            if (sl < nslots) {
                const SD d = sd[sl];
                const double *xcol = XP + (size_t)d.xcol * SYP;
                const double P = piece_sum(xcol, 1, d.ylo, d.yhi, d.wiy, d.wey);
                const int c = d.c;
                const int Kmin = z0 / c, Kmax = (z0 + ZC) / c;
                const int K1 = k / c;
#pragma unroll
                for (int dd = 0; dd < 2; dd++) {
                    const int K = K1 - dd;
                    if (K < Kmin || K >= Kmax) continue;
                    const int t = k - (c * (K + 1) - 1);
                    if (t >= d.zlo && t <= d.zhi) {
                        const double w = (t == d.zlo || t == d.zhi) ? d.wez : d.wiz;
                        if (K & 1) acc1[m] += w * P;
                        else acc0[m] += w * P;
                    }
                    if (t == c / 2) {
                        const size_t nb = (size_t)d.nb;
                        const size_t o = (size_t)(x0 / c + d.Ix) + nb * ((size_t)(y0 / c + d.Iy) + nb * (size_t)K);
                        if (K & 1) {
                            d.out[o] = acc1[m];
                            acc1[m] = 0.0;
                        } else {
                            d.out[o] = acc0[m];
                            acc0[m] = 0.0;
                        }
                    }
                }
            }
        }
        // land the prefetch (slot sm and the p tile were last read by the sources of k)
#pragma unroll
        for (int i = 0; i < NP; i++) {
            const int ly = ly0 + i * RP;
            if (in_ux && ly < UY) {
                U[(sm * 3 + 0) * plane_u + ly * UX + lx] = pre[0 * NP + i];
                U[(sm * 3 + 1) * plane_u + ly * UX + lx] = pre[1 * NP + i];
                U[(sm * 3 + 2) * plane_u + ly * UX + lx] = pre[2 * NP + i];
            }
            if (in_sx && ly < SY) PS[ly * SX + lx] = pre[3 * NP + i];
        }
        __syncthreads();
    }
}

int src_index(int a, int b) {  // source of r_ab in the LDS source array
    if (a == b) return 3 + a;
    const int lo = std::min(a, b), hi = std::max(a, b);
    if (lo == 0 && hi == 1) return 6;
    if (lo == 1 && hi == 2) return 7;
    return 8;  // (0, 2)
}

int largest_tile(int n, int L, int target) {
    int best = 0;
    for (int t = L; t <= n; t += L)
        if (n % t == 0 && t <= target) best = t;
    if (best == 0) {
        for (int t = L; t <= n; t += L)
            if (n % t == 0) return t;
    }
    return best;
}

// Device copies of descriptor tables, keyed by their bytes (the same sets
// repeat every stage), kept for the life of the process.
const void *table_on_device(ec_ctx *ctx, const std::vector<char> &bytes, int *rc) {
    static std::map<std::pair<int, std::vector<char>>, void *> cache;
    auto key = std::make_pair(ctx->device, bytes);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    void *d = nullptr;
    if (hipMalloc(&d, bytes.size()) != hipSuccess ||
        hipMemcpy(d, bytes.data(), bytes.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        *rc = ec_fail(ctx, EC_E_HIP, "closure descriptor upload failed");
        return nullptr;
    }
    if (cache.size() > 64) cache.clear();  // bounded: stale tables are simply re-uploaded
    cache[key] = d;
    return d;
}

template <typename T, int MAXS, int MAXX, int LW, int NP>
int launch_one(ec_ctx *ctx, dim3 grid, int threads, size_t lds, const void *const u[3], const void *p, const XD *xd,
               const SD *sd, int n, int TX, int TY, int ZC, int SX, int SY, int UX, int UY, int NXP, int nslots,
               int hmax, int4 ibase, int4 rows, int4 cbase, double ih, double visc) {
    const int SXP = SX | 1, SYP = SY | 1, SST = (SXP * SY) | 1;
    EC_HIP(ctx, hipFuncSetAttribute((const void *)k_closure_tiled<T, MAXS, MAXX, LW, NP>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_closure_tiled<T, MAXS, MAXX, LW, NP><<<grid, threads, lds, ctx->stream>>>(
        (const T *)u[0], (const T *)u[1], (const T *)u[2], (const T *)p, xd, sd, n, TX, TY, ZC, SX, SY, UX, UY, NXP,
        nslots, hmax, SST, SXP, SYP, ibase, rows, cbase, ih, visc);
    EC_CHECK_LAUNCH(ctx, "k_closure_tiled");
    return EC_OK;
}

template <typename T, int MAXS>
int launch_tiles(ec_ctx *ctx, dim3 grid, int threads, size_t lds, const void *const u[3], const void *p, const XD *xd,
                 const SD *sd, int n, int TX, int TY, int ZC, int SX, int SY, int UX, int UY, int NXP, int nslots,
                 int hmax, int4 ibase, int4 rows, int4 cbase, int nitems, double ih, double visc) {
    const int mx = (nitems + threads - 1) / threads;
#define EC_ARGS \
    ctx, grid, threads, lds, u, p, xd, sd, n, TX, TY, ZC, SX, SY, UX, UY, NXP, nslots, hmax, ibase, rows, cbase, ih, visc
    if (threads != 512 || mx > 16) return EC_E_UNSUPPORTED;
    const int LW = UX <= 32 ? 32 : 64;
    if (UX > 64) return EC_E_UNSUPPORTED;
    const int np = (UY + 512 / LW - 1) / (512 / LW);
    if (LW == 32 && np <= 2) return launch_one<T, MAXS, 16, 32, 2>(EC_ARGS);
    if (LW == 32 && np <= 3) return launch_one<T, MAXS, 16, 32, 3>(EC_ARGS);
    if (LW == 64 && np <= 4) return launch_one<T, MAXS, 16, 64, 4>(EC_ARGS);
    if (LW == 64 && np <= 6) return launch_one<T, MAXS, 16, 64, 6>(EC_ARGS);
#undef EC_ARGS
    return EC_E_UNSUPPORTED;
}

template <typename T>
int tiled_launch(ec_ctx *ctx, int n, double l, double visc, void *const u[3], const void *p, int nsets,
                 const ec_closure_set *sets, int tx_target, int ty_target, int zc_target) {
    int L = 1, cmax = 1;
    for (int s = 0; s < nsets; s++) {
        const int c = n / sets[s].nbar;
        L = std::lcm(L, c);
        cmax = std::max(cmax, c);
    }
    if (L > n || n % L) return EC_E_UNSUPPORTED;
    const int hmax = cmax / 2;
    const size_t bt = sizeof(T);
    for (int attempt = 0; attempt < 4; attempt++) {
        const int TX = largest_tile(n, L, tx_target), TY = largest_tile(n, L, ty_target);
        const int ZC = largest_tile(n, L, zc_target);
        if (!TX || !TY || !ZC) return EC_E_UNSUPPORTED;
        const int SX = TX + hmax, SY = TY + hmax, UX = SX + 2, UY = SY + 2;
        const int sxp = SX | 1, syp = SY | 1, sst = (sxp * SY) | 1;  // the kernel's padded strides
        std::vector<XD> xds;
        std::vector<SD> sds;
        int ib[4] = {0, 0, 0, 0}, rw[4] = {1, 1, 1, 1}, cbs[4] = {0, 0, 0, 0};
        int nitems = 0;
        for (int s = 0; s < nsets; s++) {
            const ec_closure_set &S = sets[s];
            const int c = n / S.nbar;
            int qsrc[12], qp[12][3];
            double *qout[12];
            for (int a = 0; a < 3; a++) {
                qsrc[a] = a;
                for (int d = 0; d < 3; d++) qp[a][d] = d == a ? (S.kind == 0 ? PF : PP) : PB;
                qout[a] = (double *)S.ubar[a];
            }
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++) {
                    const int q = 3 + a + 3 * b;
                    qsrc[q] = src_index(a, b);
                    for (int d = 0; d < 3; d++) {
                        if (a == b) qp[q][d] = d == a ? PC : PB;
                        else if (S.kind == 0) qp[q][d] = d == b ? PP : (d == a ? PF : PB);
                        else qp[q][d] = (d == a || d == b) ? PP : PB;
                    }
                    qout[q] = (double *)S.flux[a + 3 * b];
                }
            // distinct (source, x piece) types
            int ntypes = 0, tsrc[12], tpc[12], qt[12];
            for (int q = 0; q < 12; q++) {
                int found = -1;
                for (int t = 0; t < ntypes; t++)
                    if (tsrc[t] == qsrc[q] && tpc[t] == qp[q][0]) found = t;
                if (found < 0) {
                    found = ntypes++;
                    tsrc[found] = qsrc[q];
                    tpc[found] = qp[q][0];
                }
                qt[q] = found;
            }
            const int xp_off = (int)xds.size();
            ib[s] = nitems;
            rw[s] = std::min(SY, TY + c / 2);
            cbs[s] = xp_off;
            nitems += rw[s] * ntypes * (TX / c);
            for (int t = 0; t < ntypes; t++)
                for (int Ix = 0; Ix < TX / c; Ix++) {
                    const Piece pc = make_piece(tpc[t], c);
                    const int f = c * (Ix + 1) - 1;
                    xds.push_back(XD{tsrc[t] * sst, f + pc.lo, f + pc.hi, 0, pc.wi, pc.we});
                }
            for (int q = 0; q < 12; q++)
                for (int Iy = 0; Iy < TY / c; Iy++)
                    for (int Ix = 0; Ix < TX / c; Ix++) {
                        const Piece py = make_piece(qp[q][1], c), pz = make_piece(qp[q][2], c);
                        const int fy = c * (Iy + 1) - 1;
                        SD d;
                        memset(&d, 0, sizeof(d));
                        d.out = qout[q];
                        d.wiy = py.wi;
                        d.wey = py.we;
                        d.wiz = pz.wi;
                        d.wez = pz.we;
                        d.xcol = xp_off + qt[q] * (TX / c) + Ix;
                        d.ylo = fy + py.lo;
                        d.yhi = fy + py.hi;
                        d.zlo = pz.lo;
                        d.zhi = pz.hi;
                        d.c = c;
                        d.Ix = Ix;
                        d.Iy = Iy;
                        d.nb = S.nbar;
                        sds.push_back(d);
                    }
        }
        const int NXP = (int)xds.size(), nslots = (int)sds.size();
        const size_t lds = (size_t)syp * NXP * 8 + (size_t)NXP * sizeof(XD) +
                           bt * (9 * (size_t)UX * UY + (size_t)SX * SY + 9 * (size_t)sst);
        const int threads = 512;
        const int maxs = (nslots + threads - 1) / threads;
        const int pre = (3 * UX * UY + SX * SY + threads - 1) / threads;
        for (int s = nsets; s < 4; s++) {  // unused sets: empty ranges past the end
            ib[s] = nitems;
            rw[s] = 1;
            cbs[s] = (int)xds.size();
        }
        const int mx = (nitems + threads - 1) / threads;
        if (lds <= 152 * 1024 && maxs <= 2 && pre <= 16 && mx <= 16) {
            const int4 ibase = make_int4(ib[0], ib[1], ib[2], ib[3]);
            const int4 rows = make_int4(rw[0], rw[1], rw[2], rw[3]);
            const int4 cbase = make_int4(cbs[0], cbs[1], cbs[2], cbs[3]);
            int rc = EC_OK;
            std::vector<char> xb((char *)xds.data(), (char *)(xds.data() + xds.size()));
            std::vector<char> sb((char *)sds.data(), (char *)(sds.data() + sds.size()));
            const XD *dxd = (const XD *)table_on_device(ctx, xb, &rc);
            if (rc) return rc;
            const SD *dsd = (const SD *)table_on_device(ctx, sb, &rc);
            if (rc) return rc;
            dim3 grid(n / TX, n / TY, n / ZC);
            ec_prof_scope pr(ctx, "closure_tiled", bt * 4.0 * ec_cube(n));
            const double ih = n / l;
            if (maxs <= 1)
                return launch_tiles<T, 1>(ctx, grid, threads, lds, u, p, dxd, dsd, n, TX, TY, ZC, SX, SY, UX, UY, NXP,
                                          nslots, hmax, ibase, rows, cbase, nitems, ih, visc);
            return launch_tiles<T, 2>(ctx, grid, threads, lds, u, p, dxd, dsd, n, TX, TY, ZC, SX, SY, UX, UY, NXP,
                                      nslots, hmax, ibase, rows, cbase, nitems, ih, visc);
        }
        tx_target = std::max(L, tx_target / 2);
        ty_target = std::max(L, ty_target / 2);
    }
    return EC_E_UNSUPPORTED;
}

}  // namespace

extern "C" int ec_closure_set_tiles(ec_ctx *ctx, int tx, int ty, int tz) {
    EC_REQUIRE(ctx, ctx, "null context");
    EC_REQUIRE(ctx, tx >= 1 && ty >= 1 && tz >= 1, "tile targets must be positive");
    ctx->cl_tx = tx;
    ctx->cl_ty = ty;
    ctx->cl_tz = tz;
    return EC_OK;
}

// Returns EC_E_UNSUPPORTED (without touching any output) when the sets do not
// fit the tiled scheme; the caller then uses the per-quantity path.
int ec_closure_tiled(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3], const void *p,
                     int nsets, const ec_closure_set *sets) {
    if (nsets < 1 || nsets > MAXSETS) return EC_E_UNSUPPORTED;
    for (int s = 0; s < nsets; s++) {
        if (sets[s].nbar < 1 || n % sets[s].nbar) return EC_E_UNSUPPORTED;
        if (sets[s].kind != 0 && sets[s].kind != 1) return EC_E_UNSUPPORTED;
        for (int a = 0; a < 3; a++)
            if (!sets[s].ubar[a]) return EC_E_UNSUPPORTED;
        for (int m = 0; m < 9; m++)
            if (!sets[s].flux[m]) return EC_E_UNSUPPORTED;
    }
    return dtype == EC_F64 ? tiled_launch<double>(ctx, n, l, visc, u, p, nsets, sets, ctx->cl_tx, ctx->cl_ty, ctx->cl_tz)
                           : tiled_launch<float>(ctx, n, l, visc, u, p, nsets, sets, ctx->cl_tx, ctx->cl_ty, ctx->cl_tz);
}
