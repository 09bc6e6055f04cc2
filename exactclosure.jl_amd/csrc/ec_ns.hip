// Staggered-grid Navier-Stokes stencils: stress, tensor divergence, fused
// right-hand side, divergence, pressure gradient and the RK4 stage update.
// Bandwidth-bound: one thread per cell, x fastest so every wave reads 64
// consecutive values of each face array; the neighbour loads hit L1/L2.
#include "ec_internal.h"

#include <type_traits>

namespace {

struct Grid3 {
    dim3 block, grid;
};

Grid3 grid3(int n) {
    Grid3 g;
    int bx = n >= 64 ? 64 : n;
    int by = 256 / bx;
    if (by > n) by = n;
    g.block = dim3(bx, by, 1);
    g.grid = dim3((n + bx - 1) / bx, (n + by - 1) / by, n);
    return g;
}

struct Cell {
    int i, j, k, im, ip, jm, jp, km, kp;
    size_t n1, n2;
    __device__ __forceinline__ size_t at(int a, int b, int c) const { return (size_t)a + n1 * b + n2 * c; }
};

__device__ __forceinline__ bool cell_of(int n, Cell &c) {
    c.i = blockIdx.x * blockDim.x + threadIdx.x;
    c.j = blockIdx.y * blockDim.y + threadIdx.y;
    c.k = blockIdx.z;
    if (c.i >= n || c.j >= n) return false;
    c.im = ec_dn(c.i, n);
    c.ip = ec_up(c.i, n);
    c.jm = ec_dn(c.j, n);
    c.jp = ec_up(c.j, n);
    c.km = ec_dn(c.k, n);
    c.kp = ec_up(c.k, n);
    c.n1 = (size_t)n;
    c.n2 = (size_t)n * n;
    return true;
}

// stress(u, visc, i, i) and stress(u, visc, i, j) of src/N.jl:106-118 from
// the right/left neighbour values.
template <typename T>
__device__ __forceinline__ T s_diag(T r, T l, T visc, T ih) {
    const T e = (r + l) / 2;
    const T d = (r - l) * ih;
    return e * e - visc * (d + d);
}
template <typename T>
__device__ __forceinline__ T s_off(T ar, T al, T br, T bl, T visc, T ih) {
    return ((ar + al) / 2) * ((br + bl) / 2) - visc * ((ar - al) * ih + (br - bl) * ih);
}

template <typename T>
__global__ void k_stresstensor(int n, T ih, T visc, const T *__restrict__ u1, const T *__restrict__ u2,
                               const T *__restrict__ u3, T *__restrict__ s11, T *__restrict__ s22,
                               T *__restrict__ s33, T *__restrict__ s12, T *__restrict__ s23,
                               T *__restrict__ s31) {
    Cell c;
    if (!cell_of(n, c)) return;
    const size_t I = c.at(c.i, c.j, c.k);
    s11[I] = s_diag(u1[I], u1[c.at(c.im, c.j, c.k)], visc, ih);
    s22[I] = s_diag(u2[I], u2[c.at(c.i, c.jm, c.k)], visc, ih);
    s33[I] = s_diag(u3[I], u3[c.at(c.i, c.j, c.km)], visc, ih);
    s12[I] = s_off(u1[c.at(c.i, c.jp, c.k)], u1[I], u2[c.at(c.ip, c.j, c.k)], u2[I], visc, ih);
    s23[I] = s_off(u2[c.at(c.i, c.j, c.kp)], u2[I], u3[c.at(c.i, c.jp, c.k)], u3[I], visc, ih);
    s31[I] = s_off(u3[c.at(c.ip, c.j, c.k)], u3[I], u1[c.at(c.i, c.j, c.kp)], u1[I], visc, ih);
}

template <typename T>
__global__ void k_tensordivergence(int n, T ih, const T *__restrict__ s11, const T *__restrict__ s22,
                                   const T *__restrict__ s33, const T *__restrict__ s12,
                                   const T *__restrict__ s23, const T *__restrict__ s31,
                                   T *__restrict__ du1, T *__restrict__ du2, T *__restrict__ du3, int doadd) {
    Cell c;
    if (!cell_of(n, c)) return;
    const size_t I = c.at(c.i, c.j, c.k);
    T a1 = doadd ? du1[I] : T(0);
    a1 = a1 - (s11[c.at(c.ip, c.j, c.k)] - s11[I]) * ih - (s12[I] - s12[c.at(c.i, c.jm, c.k)]) * ih -
         (s31[I] - s31[c.at(c.i, c.j, c.km)]) * ih;
    T a2 = doadd ? du2[I] : T(0);
    a2 = a2 - (s12[I] - s12[c.at(c.im, c.j, c.k)]) * ih - (s22[c.at(c.i, c.jp, c.k)] - s22[I]) * ih -
         (s23[I] - s23[c.at(c.i, c.j, c.km)]) * ih;
    T a3 = doadd ? du3[I] : T(0);
    a3 = a3 - (s31[I] - s31[c.at(c.im, c.j, c.k)]) * ih - (s23[I] - s23[c.at(c.i, c.jm, c.k)]) * ih -
         (s33[c.at(c.i, c.j, c.kp)] - s33[I]) * ih;
    du1[I] = a1;
    du2[I] = a2;
    du3[I] = a3;
}

struct Ptr9 {
    const void *p[9];
};

template <typename T, typename S>
__global__ void k_tensordivergence_nonsym(int n, T ih, Ptr9 sp, T *__restrict__ du1, T *__restrict__ du2,
                                          T *__restrict__ du3, int doadd) {
    Cell c;
    if (!cell_of(n, c)) return;
    const size_t I = c.at(c.i, c.j, c.k);
    T *du[3] = {du1, du2, du3};
    const size_t nb_up[3] = {c.at(c.ip, c.j, c.k), c.at(c.i, c.jp, c.k), c.at(c.i, c.j, c.kp)};
    const size_t nb_dn[3] = {c.at(c.im, c.j, c.k), c.at(c.i, c.jm, c.k), c.at(c.i, c.j, c.km)};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        T acc = doadd ? du[a][I] : T(0);
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const S *s = (const S *)sp.p[a + 3 * b];
            const S diff = a == b ? s[nb_up[b]] - s[I] : s[I] - s[nb_dn[b]];
            acc = acc - (T)(diff * (S)ih);
        }
        du[a][I] = acc;
    }
}

// Fused stresstensor + tensordivergence!: 27 neighbour loads, 13 stresses,
// nothing materialised.
// Output modes of the RHS kernels (doadd): 0 writes du = F(u); 1 adds, du +=
// F(u); 2 writes du = s F(u) + add with an FP64 addend, rounded once to T (the
// arithmetic of an FP64 axpy applied to the written F, in the same pass).
struct RhsAdd {
    double s;
    const double *a1, *a2, *a3;
};

template <typename T>
__global__ void k_rhs(int n, T ih, T visc, const T *__restrict__ u1, const T *__restrict__ u2,
                      const T *__restrict__ u3, T *__restrict__ du1, T *__restrict__ du2,
                      T *__restrict__ du3, int doadd, RhsAdd ad) {
    Cell c;
    if (!cell_of(n, c)) return;
    const int i = c.i, j = c.j, k = c.k;
    const size_t I = c.at(i, j, k);
    // u1 neighbourhood
    const T u1_0 = u1[I], u1_xp = u1[c.at(c.ip, j, k)], u1_xm = u1[c.at(c.im, j, k)];
    const T u1_yp = u1[c.at(i, c.jp, k)], u1_ym = u1[c.at(i, c.jm, k)];
    const T u1_zp = u1[c.at(i, j, c.kp)], u1_zm = u1[c.at(i, j, c.km)];
    const T u1_xm_yp = u1[c.at(c.im, c.jp, k)], u1_xm_zp = u1[c.at(c.im, j, c.kp)];
    // u2 neighbourhood
    const T u2_0 = u2[I], u2_xp = u2[c.at(c.ip, j, k)], u2_xm = u2[c.at(c.im, j, k)];
    const T u2_yp = u2[c.at(i, c.jp, k)], u2_ym = u2[c.at(i, c.jm, k)];
    const T u2_zp = u2[c.at(i, j, c.kp)], u2_zm = u2[c.at(i, j, c.km)];
    const T u2_xp_ym = u2[c.at(c.ip, c.jm, k)], u2_ym_zp = u2[c.at(i, c.jm, c.kp)];
    // u3 neighbourhood
    const T u3_0 = u3[I], u3_xp = u3[c.at(c.ip, j, k)], u3_xm = u3[c.at(c.im, j, k)];
    const T u3_yp = u3[c.at(i, c.jp, k)], u3_ym = u3[c.at(i, c.jm, k)];
    const T u3_zp = u3[c.at(i, j, c.kp)], u3_zm = u3[c.at(i, j, c.km)];
    const T u3_xp_zm = u3[c.at(c.ip, j, c.km)], u3_yp_zm = u3[c.at(i, c.jp, c.km)];

    const T s11_p = s_diag(u1_xp, u1_0, visc, ih), s11_0 = s_diag(u1_0, u1_xm, visc, ih);
    const T s22_p = s_diag(u2_yp, u2_0, visc, ih), s22_0 = s_diag(u2_0, u2_ym, visc, ih);
    const T s33_p = s_diag(u3_zp, u3_0, visc, ih), s33_0 = s_diag(u3_0, u3_zm, visc, ih);
    const T s12_0 = s_off(u1_yp, u1_0, u2_xp, u2_0, visc, ih);
    const T s12_ym = s_off(u1_0, u1_ym, u2_xp_ym, u2_ym, visc, ih);
    const T s12_xm = s_off(u1_xm_yp, u1_xm, u2_0, u2_xm, visc, ih);
    const T s31_0 = s_off(u3_xp, u3_0, u1_zp, u1_0, visc, ih);
    const T s31_zm = s_off(u3_xp_zm, u3_zm, u1_0, u1_zm, visc, ih);
    const T s31_xm = s_off(u3_0, u3_xm, u1_xm_zp, u1_xm, visc, ih);
    const T s23_0 = s_off(u2_zp, u2_0, u3_yp, u3_0, visc, ih);
    const T s23_zm = s_off(u2_0, u2_zm, u3_yp_zm, u3_zm, visc, ih);
    const T s23_ym = s_off(u2_ym_zp, u2_ym, u3_0, u3_ym, visc, ih);

    T a1 = doadd == 1 ? du1[I] : T(0);
    a1 = a1 - (s11_p - s11_0) * ih - (s12_0 - s12_ym) * ih - (s31_0 - s31_zm) * ih;
    T a2 = doadd == 1 ? du2[I] : T(0);
    a2 = a2 - (s12_0 - s12_xm) * ih - (s22_p - s22_0) * ih - (s23_0 - s23_zm) * ih;
    T a3 = doadd == 1 ? du3[I] : T(0);
    a3 = a3 - (s31_0 - s31_xm) * ih - (s23_0 - s23_ym) * ih - (s33_p - s33_0) * ih;
    if (doadd == 2) {
        a1 = (T)(ad.s * (double)a1 + ad.a1[I]);
        a2 = (T)(ad.s * (double)a2 + ad.a2[I]);
        a3 = (T)(ad.s * (double)a3 + ad.a3[I]);
    }
    du1[I] = a1;
    du2[I] = a2;
    du3[I] = a3;
}

// -div sigma(u) at one cell of a haloed plane triple (pm, p0, pp: planes k-1,
// k, k+1 of the three components, each component HP values apart, rows HX
// apart); o is the cell's offset within a plane.  The stress differences of
// k_rhs are regrouped so that no stress is formed on its own: with averages
// taken as sums and the 1/4 applied once,
//   convective  ((r + l) / 2)^2 differences -> (u+ - u-) (u+ + 2 u0 + u-) / 4,
//               off-diagonal products (a + a')(b + b') / 4, each formed once;
//   viscous     the nu (d_a u_b + d_b u_a) differences of one component sum to
//               ih^2 (2 S_a + S_b + S_c - 8 u0 + two mixed differences),
//               S_d = sum of the two neighbours along d;
// so du_a = ih (nu ih V_a - C_a / 4).  Same values as the stress form up to
// rounding.
template <typename T>
__device__ __forceinline__ void rhs_cell(const T *pm, const T *p0, const T *pp, int o, int HX, int HP, T ih, T visc,
                                         T &a1, T &a2, T &a3, T &divu) {
    const T *b0 = p0 + HP, *c0 = p0 + 2 * HP;
    const T u1_0 = p0[o], u1_xp = p0[o + 1], u1_xm = p0[o - 1], u1_yp = p0[o + HX], u1_ym = p0[o - HX];
    const T u1_xm_yp = p0[o - 1 + HX], u1_zp = pp[o], u1_zm = pm[o], u1_xm_zp = pp[o - 1];
    const T u2_0 = b0[o], u2_xp = b0[o + 1], u2_xm = b0[o - 1], u2_yp = b0[o + HX], u2_ym = b0[o - HX];
    const T u2_xp_ym = b0[o + 1 - HX], u2_zp = pp[HP + o], u2_zm = pm[HP + o], u2_ym_zp = pp[HP + o - HX];
    const T u3_0 = c0[o], u3_xp = c0[o + 1], u3_xm = c0[o - 1], u3_yp = c0[o + HX], u3_ym = c0[o - HX];
    const T u3_zp = pp[2 * HP + o], u3_zm = pm[2 * HP + o], u3_xp_zm = pm[2 * HP + o + 1];
    const T u3_yp_zm = pm[2 * HP + o + HX];
    // convective, 4x: the nine off-diagonal products and the three diagonal differences
    const T p12 = (u1_yp + u1_0) * (u2_xp + u2_0), q12y = (u1_0 + u1_ym) * (u2_xp_ym + u2_ym);
    const T q12x = (u1_xm_yp + u1_xm) * (u2_0 + u2_xm);
    const T p31 = (u3_xp + u3_0) * (u1_zp + u1_0), q31z = (u3_xp_zm + u3_zm) * (u1_0 + u1_zm);
    const T q31x = (u3_0 + u3_xm) * (u1_xm_zp + u1_xm);
    const T p23 = (u2_zp + u2_0) * (u3_yp + u3_0), q23z = (u2_0 + u2_zm) * (u3_yp_zm + u3_zm);
    const T q23y = (u2_ym_zp + u2_ym) * (u3_0 + u3_ym);
    const T s1x = u1_xp + u1_xm, s2y = u2_yp + u2_ym, s3z = u3_zp + u3_zm;
    const T c1 = (u1_xp - u1_xm) * (s1x + 2 * u1_0) + (p12 - q12y) + (p31 - q31z);
    const T c2 = (p12 - q12x) + (u2_yp - u2_ym) * (s2y + 2 * u2_0) + (p23 - q23z);
    const T c3 = (p31 - q31x) + (p23 - q23y) + (u3_zp - u3_zm) * (s3z + 2 * u3_0);
    // viscous, divided by ih^2
    const T v1 = 2 * s1x + (u1_yp + u1_ym) + (u1_zp + u1_zm) - 8 * u1_0 + ((u2_xp - u2_0) - (u2_xp_ym - u2_ym)) +
                 ((u3_xp - u3_0) - (u3_xp_zm - u3_zm));
    const T v2 = (u2_xp + u2_xm) + 2 * s2y + (u2_zp + u2_zm) - 8 * u2_0 + ((u1_yp - u1_0) - (u1_xm_yp - u1_xm)) +
                 ((u3_yp - u3_0) - (u3_yp_zm - u3_zm));
    const T v3 = (u3_xp + u3_xm) + (u3_yp + u3_ym) + 2 * s3z - 8 * u3_0 + ((u1_zp - u1_0) - (u1_xm_zp - u1_xm)) +
                 ((u2_zp - u2_0) - (u2_ym_zp - u2_ym));
    const T vi = visc * ih;
    a1 = ih * (vi * v1 - T(0.25) * c1);
    a2 = ih * (vi * v2 - T(0.25) * c2);
    a3 = ih * (vi * v3 - T(0.25) * c3);
    // h div_h u at the cell centre (divergence!, src/N.jl:222-238), from values already loaded
    divu = (u1_0 - u1_xm) + (u2_0 - u2_ym) + (u3_0 - u3_zm);
}

template <typename T>
__device__ __forceinline__ void rhs_cell(const T *pm, const T *p0, const T *pp, int o, int HX, int HP, T ih, T visc,
                                         T &a1, T &a2, T &a3) {
    T divu;
    rhs_cell(pm, p0, pp, o, HX, HP, ih, visc, a1, a2, a3, divu);
}

// Fused RHS for large grids, z-marching: a block owns a TX x TY column tile
// and marches ZC planes.  Planes k-1 .. k+2 of the three velocity components
// (with a one-cell x/y halo) rotate through a four-slot LDS ring; the loads of
// plane k+2 are issued before the stencil of plane k and land after it, so
// every u value is read from HBM once per tile (the plain kernel re-reads the
// neighbouring planes, about 4x the compulsory traffic).
template <typename T, int TX, int TY, int DOADD>
__global__ void __launch_bounds__(256) k_rhs_march(int n, int ZC, T ih, T visc, const T *__restrict__ u1,
                                                   const T *__restrict__ u2, const T *__restrict__ u3,
                                                   T *__restrict__ du1, T *__restrict__ du2, T *__restrict__ du3,
                                                   int doadd, RhsAdd ad) {
    constexpr int HX = TX + 2, HY = TY + 2, HP = HX * HY;
    constexpr int NLC = (HP + 255) / 256, NLD = 3 * NLC;  // loads per component and plane, all components
    constexpr int CPT = TX * TY / 256;  // cells per thread and plane
    static_assert(TX * TY % 256 == 0 && 256 % TX == 0, "tile shape");
    __shared__ T ring[4][3 * HP];
    const int gx = n / TX, gy = n / TY;
    const int L = (int)ec_xcd_tile(blockIdx.x, gridDim.x);
    const int x0 = (L % gx) * TX, y0 = ((L / gx) % gy) * TY, z0 = (L / (gx * gy)) * ZC;
    const int tid = threadIdx.x;
    const unsigned nn = (unsigned)n;
    const size_t n2 = (size_t)nn * nn;
    // loader: load slot jj of every component takes element q = tid + 256 jj of
    // the haloed plane (threads past the plane load its last element and do not
    // store it); a load is the uniform plane base plus a loop-invariant byte offset
    unsigned loff[NLC];
#pragma unroll
    for (int jj = 0; jj < NLC; jj++) {
        const int q = min(tid + 256 * jj, HP - 1), hx = q % HX, hy = q / HX;
        loff[jj] = ((unsigned)ec_wrap_any(x0 + hx - 1, n) + nn * (unsigned)ec_wrap_any(y0 + hy - 1, n)) *
                   (unsigned)sizeof(T);
    }
    auto load = [&](int k, T *r) {
        const size_t kk = (size_t)ec_wrap_any(k, n) * n2;
        const char *b1 = (const char *)(u1 + kk), *b2 = (const char *)(u2 + kk), *b3 = (const char *)(u3 + kk);
#pragma unroll
        for (int jj = 0; jj < NLC; jj++) {
            r[jj] = *(const T *)(b1 + loff[jj]);
            r[NLC + jj] = *(const T *)(b2 + loff[jj]);
            r[2 * NLC + jj] = *(const T *)(b3 + loff[jj]);
        }
    };
    auto store = [&](int slot, const T *r) {
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int jj = 0; jj < NLC; jj++)
                if (HP % 256 == 0 || tid + 256 * jj < HP) ring[slot][c * HP + tid + 256 * jj] = r[c * NLC + jj];
    };
    {
        T r[NLD];
        load(z0 - 1, r);
        store(0, r);
        load(z0, r);
        store(1, r);
        load(z0 + 1, r);
        store(2, r);
    }
    __syncthreads();
    const int cx = tid % TX, cy0 = tid / TX;
    // the addend of the plane's cells (DOADD 1: du itself, 2: the FP64 ad.a)
    typedef typename std::conditional<DOADD == 2, double, T>::type AT;
    for (int rr = 0; rr < ZC; ++rr) {
        const int k = z0 + rr;
        // the addend loads go out before the next plane's: the in-order memory
        // counter then lets the plane wait for them without waiting for the
        // next plane (and DOADD is a template parameter: a load under a
        // runtime branch would make the waits drain everything in flight)
        AT add[CPT][3];
#pragma unroll
        for (int m = 0; m < CPT; m++) {
            const size_t I = (size_t)(x0 + cx) + (size_t)nn * (size_t)(y0 + cy0 + m * (256 / TX)) + n2 * (size_t)k;
            if constexpr (DOADD == 1) {
                add[m][0] = du1[I];
                add[m][1] = du2[I];
                add[m][2] = du3[I];
            } else if constexpr (DOADD == 2) {  // (read once: non-temporal)
                add[m][0] = ec_nt_load(ad.a1 + I);
                add[m][1] = ec_nt_load(ad.a2 + I);
                add[m][2] = ec_nt_load(ad.a3 + I);
            }
        }
        T nxt[NLD];
        load(k + 2, nxt);  // unconditional: see k_rhs_div_march
        const T *pm = ring[rr & 3], *p0 = ring[(rr + 1) & 3], *pp = ring[(rr + 2) & 3];
#pragma unroll
        for (int m = 0; m < CPT; m++) {
            const int cy = cy0 + m * (256 / TX);
            const int o = (cy + 1) * HX + (cx + 1);  // the cell in a haloed plane
            T a1, a2, a3;
            rhs_cell(pm, p0, pp, o, HX, HP, ih, visc, a1, a2, a3);
            const size_t I = (size_t)(x0 + cx) + (size_t)nn * (size_t)(y0 + cy) + n2 * (size_t)k;
            if constexpr (DOADD == 1) {
                a1 += add[m][0];
                a2 += add[m][1];
                a3 += add[m][2];
            } else if constexpr (DOADD == 2) {
                a1 = (T)(ad.s * (double)a1 + add[m][0]);
                a2 = (T)(ad.s * (double)a2 + add[m][1]);
                a3 = (T)(ad.s * (double)a3 + add[m][2]);
            }
            du1[I] = a1;
            du2[I] = a2;
            du3[I] = a3;
        }
        if (rr + 1 < ZC) store((rr + 3) & 3, nxt);
        __syncthreads();
    }
}

// Fused RHS and its divergence, z-marching like k_rhs_march: du = -div sigma(u)
// and d = div_h du (the right-hand side of the pressure Poisson equation).  The
// planes carry a two-cell halo on the low x/y sides so that the block also
// evaluates du1 on the column x0 - 1 and du2 on the row y0 - 1; du3 of the
// previous plane stays in registers (one extra plane z0 - 1 starts the march).
// du1 / du2 go through a double-buffered LDS exchange, one barrier per plane.
// With rdh = r / h, d = div_h du + r div_h u: the divergence of the stage
// state u is folded into the Poisson right-hand side (see ec_ns_rhs_div).
template <typename T, int TX, int TY, int NT = 256>
__global__ void __launch_bounds__(NT) k_rhs_div_march(int n, int ZC, T ih, T visc, const T *__restrict__ u1,
                                                       const T *__restrict__ u2, const T *__restrict__ u3,
                                                       T *__restrict__ du1, T *__restrict__ du2, T *__restrict__ du3,
                                                       T *__restrict__ dout, T rdh) {
    constexpr int HX = TX + 3, HY = TY + 3, HP = HX * HY;
    constexpr int NLC = (HP + NT - 1) / NT, NLD = 3 * NLC;  // loads per component and plane, all components
    constexpr int CPT = TX * TY / NT;
    constexpr int D1N = TY * (TX + 1), D2N = (TY + 1) * TX;
    static_assert(TX * TY % NT == 0 && NT % TX == 0 && NT % 64 == 0 && NT >= 256, "tile shape");
    __shared__ T ring[4][3 * HP];
    __shared__ T dx1[2][D1N], dx2[2][D2N];  // du1 at x in [x0-1, x0+TX), du2 at y in [y0-1, y0+TY)
    const int gx = n / TX, gy = n / TY;
    const int L = (int)ec_xcd_tile(blockIdx.x, gridDim.x);
    const int x0 = (L % gx) * TX, y0 = ((L / gx) % gy) * TY, z0 = (L / (gx * gy)) * ZC;
    const int tid = threadIdx.x;
    const unsigned nn = (unsigned)n;
    const size_t n2 = (size_t)nn * nn;

    // loader: load slot jj of every component takes element q = tid + 256 jj of
    // the haloed plane (threads past the plane load its last element and do not
    // store it).  The byte offset of q within a z-plane is loop invariant and the
    // plane base is uniform, so a load is a scalar base plus one vector offset:
    // no per-load address arithmetic or component select.  (Byte offsets stay
    // below 2^32: n^3 < 2^32 on the marching path.)
    unsigned loff[NLC];
#pragma unroll
    for (int jj = 0; jj < NLC; jj++) {
        const int q = min(tid + NT * jj, HP - 1), hx = q % HX, hy = q / HX;
        loff[jj] = ((unsigned)ec_wrap_any(x0 + hx - 2, n) + nn * (unsigned)ec_wrap_any(y0 + hy - 2, n)) *
                   (unsigned)sizeof(T);
    }
    auto load = [&](int k, T *r) {
        const size_t kk = (size_t)ec_wrap_any(k, n) * n2;
        const char *b1 = (const char *)(u1 + kk), *b2 = (const char *)(u2 + kk), *b3 = (const char *)(u3 + kk);
#pragma unroll
        for (int jj = 0; jj < NLC; jj++) {
            r[jj] = *(const T *)(b1 + loff[jj]);
            r[NLC + jj] = *(const T *)(b2 + loff[jj]);
            r[2 * NLC + jj] = *(const T *)(b3 + loff[jj]);
        }
    };
    auto store = [&](int slot, const T *r) {
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int jj = 0; jj < NLC; jj++)
                if (HP % NT == 0 || tid + NT * jj < HP) ring[slot][c * HP + tid + NT * jj] = r[c * NLC + jj];
    };
    {
        T r[NLD];
        load(z0 - 2, r);
        store(0, r);
        load(z0 - 1, r);
        store(1, r);
        load(z0, r);
        store(2, r);
    }
    __syncthreads();
    const int cx = tid % TX, cy0 = tid / TX;
    // halo cells: wave 0 evaluates du1 on the column x0 - 1 (TY cells), waves
    // 1..3 du2 on the row y0 - 1 (TX cells); one component per wave, so each
    // wave forms only the terms of that component
    static_assert(TY <= 64 && 3 * ((TX + 2) / 3) <= 192, "halo split");
    constexpr int RPW = (TX + 2) / 3;  // row cells per wave
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), ln = tid & 63;
    const bool hcol = wv == 0 && ln < TY;
    const int ridx = (wv - 1) * RPW + ln;
    const bool hrow = wv > 0 && ln < RPW && ridx < TX;
    const int hcx = wv == 0 ? -1 : ridx, hcy = wv == 0 ? ln : -1;
    T a3prev[CPT];
    // outputs: byte offset of cell m within the tile's part of a z-plane
    // (loop invariant); the plane base is uniform
    unsigned soff[CPT];
#pragma unroll
    for (int m = 0; m < CPT; m++) soff[m] = (unsigned)(cx + (int)nn * (cy0 + m * (NT / TX))) * (unsigned)sizeof(T);
    // plane k (ring index q = rr + 2): plane k + 2 arrives in `cur` (loaded one
    // plane earlier) and goes to the ring; plane k + 3 is loaded into `ahead`,
    // so two planes of loads are in flight while a plane is evaluated
    auto plane = [&](int rr, T *cur, T *ahead) {
        const int k = z0 + rr, q = rr + 2;
        const size_t pk = n2 * (size_t)k + (size_t)x0 + (size_t)nn * (size_t)y0;  // tile corner in plane k
        // unconditional (the plane index wraps; past the chunk the values go
        // unused): a load under a branch leaves the wait counts of the join
        // to the path without it, which drains the other plane in flight
        load(k + 3, ahead);
        const auto r1 = ec_rsrc(du1 + pk), r2 = ec_rsrc(du2 + pk), r3 = ec_rsrc(du3 + pk), rd = ec_rsrc(dout + pk);
        const T *pm = ring[(q - 1) & 3], *p0 = ring[q & 3], *pp = ring[(q + 1) & 3];
        T a1[CPT], a2[CPT], a3[CPT], dv[CPT];
#pragma unroll
        for (int m = 0; m < CPT; m++) {
            const int cy = cy0 + m * (NT / TX);
            rhs_cell(pm, p0, pp, (cy + 2) * HX + (cx + 2), HX, HP, ih, visc, a1[m], a2[m], a3[m], dv[m]);
        }
        const int buf = (rr + 2) & 1;
        if (rr >= 0) {
#pragma unroll
            for (int m = 0; m < CPT; m++) {
                const int cy = cy0 + m * (NT / TX);
                // streaming stores: the outputs are not re-read here, and keeping
                // them out of L2 leaves it to the halo lines the neighbouring tiles share
                ec_nt_store_at(r1, soff[m], a1[m]);
                ec_nt_store_at(r2, soff[m], a2[m]);
                ec_nt_store_at(r3, soff[m], a3[m]);
                dx1[buf][cy * (TX + 1) + cx + 1] = a1[m];
                dx2[buf][(cy + 1) * TX + cx] = a2[m];
            }
            if (wv == 0) {
                if (hcol) {
                    T h1, h2, h3;
                    rhs_cell(pm, p0, pp, (hcy + 2) * HX + 1, HX, HP, ih, visc, h1, h2, h3);
                    dx1[buf][hcy * (TX + 1)] = h1;
                }
            } else if (hrow) {
                T h1, h2, h3;
                rhs_cell(pm, p0, pp, HX + (hcx + 2), HX, HP, ih, visc, h1, h2, h3);
                dx2[buf][hcx] = h2;
            }
        }
        if (rr + 1 < ZC) store((q + 2) & 3, cur);
        __syncthreads();
        if (rr >= 0) {
#pragma unroll
            for (int m = 0; m < CPT; m++) {
                const int cy = cy0 + m * (NT / TX);
                ec_nt_store_at(rd, soff[m],
                               (a1[m] - dx1[buf][cy * (TX + 1) + cx]) * ih + (a2[m] - dx2[buf][cy * TX + cx]) * ih +
                                   (a3[m] - a3prev[m]) * ih + rdh * dv[m]);
            }
        }
#pragma unroll
        for (int m = 0; m < CPT; m++) a3prev[m] = a3[m];
    };
    T ba[NLD], bb[NLD];
    if (ZC > 0) load(z0 + 1, ba);  // plane k + 2 of the first step (rr = -1)
    for (int rr = -1; rr < ZC; rr += 2) {
        plane(rr, ba, bb);
        if (rr + 1 < ZC) plane(rr + 1, bb, ba);
    }
}

template <typename T>
__global__ void k_divergence(int n, T ih, const T *__restrict__ u1, const T *__restrict__ u2,
                             const T *__restrict__ u3, T *__restrict__ d) {
    Cell c;
    if (!cell_of(n, c)) return;
    const size_t I = c.at(c.i, c.j, c.k);
    d[I] = (u1[I] - u1[c.at(c.im, c.j, c.k)]) * ih + (u2[I] - u2[c.at(c.i, c.jm, c.k)]) * ih +
           (u3[I] - u3[c.at(c.i, c.j, c.km)]) * ih;
}

// d += s div_h u (the divergence term of ec_ns_rhs_div on the plain path)
template <typename T>
__global__ void k_divergence_add(int n, T sih, const T *__restrict__ u1, const T *__restrict__ u2,
                                 const T *__restrict__ u3, T *__restrict__ d) {
    Cell c;
    if (!cell_of(n, c)) return;
    const size_t I = c.at(c.i, c.j, c.k);
    d[I] += sih * ((u1[I] - u1[c.at(c.im, c.j, c.k)]) + (u2[I] - u2[c.at(c.i, c.jm, c.k)]) +
                   (u3[I] - u3[c.at(c.i, c.j, c.km)]));
}

template <typename T>
__global__ void k_pressuregradient(int n, T ih, T *__restrict__ u1, T *__restrict__ u2, T *__restrict__ u3,
                                   const T *__restrict__ p) {
    Cell c;
    if (!cell_of(n, c)) return;
    const size_t I = c.at(c.i, c.j, c.k);
    const T p0 = p[I];
    u1[I] -= (p[c.at(c.ip, c.j, c.k)] - p0) * ih;
    u2[I] -= (p[c.at(c.i, c.jp, c.k)] - p0) * ih;
    u3[I] -= (p[c.at(c.i, c.j, c.kp)] - p0) * ih;
}

// out-of-place: o = a - grad p (a unchanged)
template <typename T>
__global__ void k_pressuregradient_into(int n, T ih, const T *__restrict__ a1, const T *__restrict__ a2,
                                        const T *__restrict__ a3, T *__restrict__ o1, T *__restrict__ o2,
                                        T *__restrict__ o3, const T *__restrict__ p) {
    Cell c;
    if (!cell_of(n, c)) return;
    const size_t I = c.at(c.i, c.j, c.k);
    const T p0 = p[I];
    o1[I] = a1[I] - (p[c.at(c.ip, c.j, c.k)] - p0) * ih;
    o2[I] = a2[I] - (p[c.at(c.i, c.jp, c.k)] - p0) * ih;
    o3[I] = a3[I] - (p[c.at(c.i, c.j, c.kp)] - p0) * ih;
}

struct Ptr3 {
    void *p[3];
};

// RK4 stage: k = du - grad p (the projected right-hand side), then the
// tableau of rk4_timestep! (src/Burgers.jl:120-142).
template <typename T, int STAGE>
__global__ void k_rk4_update(int n, T ih, T dt, Ptr3 u0p, Ptr3 usp, Ptr3 dup, const T *__restrict__ p,
                             Ptr3 accp) {
    Cell c;
    if (!cell_of(n, c)) return;
    const size_t I = c.at(c.i, c.j, c.k);
    const T p0 = p[I];
    const T gp[3] = {(p[c.at(c.ip, c.j, c.k)] - p0) * ih, (p[c.at(c.i, c.jp, c.k)] - p0) * ih,
                     (p[c.at(c.i, c.j, c.kp)] - p0) * ih};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const T kk = ((const T *)dup.p[a])[I] - gp[a];
        T *u0 = (T *)u0p.p[a];
        T *acc = (T *)accp.p[a];
        if (STAGE == 0) {
            acc[I] = kk;
            ((T *)usp.p[a])[I] = u0[I] + (dt / 2) * kk;
        } else if (STAGE == 1) {
            acc[I] = acc[I] + 2 * kk;
            ((T *)usp.p[a])[I] = u0[I] + (dt / 2) * kk;
        } else if (STAGE == 2) {
            acc[I] = acc[I] + 2 * kk;
            ((T *)usp.p[a])[I] = u0[I] + dt * kk;
        } else {
            u0[I] = u0[I] + (dt / 6) * (acc[I] + kk);
        }
    }
}

}  // namespace

#define EC_CHECK_FIELD(ctx, n) EC_REQUIRE(ctx, (n) >= 2, "grid size n=%d must be >= 2", (n))

// z-marching kernels: grids of at least 256 cells per side (n % 64 == 0);
static bool march_ok(int n) { return n >= 256 && n % 64 == 0 && (double)n * n * n < 4294967296.0; }
// planes per block of the marching RHS with 64 x ty column tiles: at most 64
// (neighbouring tiles then start and finish nearly together, so their shared
// halo lines still hit in L2; 128-plane chunks measured 7% slower at 1024^3),
// fewer where needed for at least 2048 tiles (chunks of 16..64)
static int march_zc(int n, int ty = 8) {
    int zc = 64;
    while (zc > 16 && (long)(n / 64) * (n / ty) * (n / zc) < 2048) zc /= 2;
    while (n % zc) zc /= 2;
    return zc;
}

extern "C" int ec_ns_stresstensor(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3],
                                  void *const s[6]) {
    EC_REQUIRE(ctx, ctx && u && s, "ec_ns_stresstensor: null argument");
    EC_CHECK_FIELD(ctx, n);
    Grid3 g = grid3(n);
    const double ih = n / l;
    ec_prof_scope ps(ctx, "stresstensor", 9.0 * ec_cube(n) * (dtype == EC_F64 ? 8 : 4));
    if (dtype == EC_F64)
        k_stresstensor<double><<<g.grid, g.block, 0, ctx->stream>>>(
            n, ih, visc, (const double *)u[0], (const double *)u[1], (const double *)u[2], (double *)s[0],
            (double *)s[1], (double *)s[2], (double *)s[3], (double *)s[4], (double *)s[5]);
    else
        k_stresstensor<float><<<g.grid, g.block, 0, ctx->stream>>>(
            n, (float)ih, (float)visc, (const float *)u[0], (const float *)u[1], (const float *)u[2],
            (float *)s[0], (float *)s[1], (float *)s[2], (float *)s[3], (float *)s[4], (float *)s[5]);
    EC_CHECK_LAUNCH(ctx, "k_stresstensor");
    return EC_OK;
}

extern "C" int ec_ns_tensordivergence(ec_ctx *ctx, int dtype, int n, double l, void *const s[6],
                                      void *const du[3], int doadd) {
    EC_REQUIRE(ctx, ctx && s && du, "ec_ns_tensordivergence: null argument");
    EC_CHECK_FIELD(ctx, n);
    Grid3 g = grid3(n);
    const double ih = n / l;
    if (dtype == EC_F64)
        k_tensordivergence<double><<<g.grid, g.block, 0, ctx->stream>>>(
            n, ih, (const double *)s[0], (const double *)s[1], (const double *)s[2], (const double *)s[3],
            (const double *)s[4], (const double *)s[5], (double *)du[0], (double *)du[1], (double *)du[2], doadd);
    else
        k_tensordivergence<float><<<g.grid, g.block, 0, ctx->stream>>>(
            n, (float)ih, (const float *)s[0], (const float *)s[1], (const float *)s[2], (const float *)s[3],
            (const float *)s[4], (const float *)s[5], (float *)du[0], (float *)du[1], (float *)du[2], doadd);
    EC_CHECK_LAUNCH(ctx, "k_tensordivergence");
    return EC_OK;
}

extern "C" int ec_ns_tensordivergence_nonsym(ec_ctx *ctx, int dtype, int n, double l, int sigma_dtype,
                                             void *const s9[9], void *const du[3], int doadd) {
    EC_REQUIRE(ctx, ctx && s9 && du, "ec_ns_tensordivergence_nonsym: null argument");
    EC_CHECK_FIELD(ctx, n);
    Grid3 g = grid3(n);
    const double ih = n / l;
    Ptr9 sp;
    for (int q = 0; q < 9; q++) sp.p[q] = s9[q];
    if (dtype == EC_F64 && sigma_dtype == EC_F64)
        k_tensordivergence_nonsym<double, double><<<g.grid, g.block, 0, ctx->stream>>>(
            n, ih, sp, (double *)du[0], (double *)du[1], (double *)du[2], doadd);
    else if (dtype == EC_F32 && sigma_dtype == EC_F64)
        k_tensordivergence_nonsym<float, double><<<g.grid, g.block, 0, ctx->stream>>>(
            n, (float)ih, sp, (float *)du[0], (float *)du[1], (float *)du[2], doadd);
    else if (dtype == EC_F32 && sigma_dtype == EC_F32)
        k_tensordivergence_nonsym<float, float><<<g.grid, g.block, 0, ctx->stream>>>(
            n, (float)ih, sp, (float *)du[0], (float *)du[1], (float *)du[2], doadd);
    else
        return ec_fail(ctx, EC_E_UNSUPPORTED, "nonsym divergence: FP32 sigma into FP64 du");
    EC_CHECK_LAUNCH(ctx, "k_tensordivergence_nonsym");
    return EC_OK;
}

static int rhs_launch(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3], void *const du[3],
                      int doadd, const RhsAdd &ad) {
    Grid3 g = grid3(n);
    const double ih = n / l;
    const double b = dtype == EC_F64 ? 8 : 4;
    ec_prof_scope ps(ctx, "ns_rhs", (doadd == 1 ? 9.0 * b : (doadd == 2 ? 6.0 * b + 24.0 : 6.0 * b)) * ec_cube(n));
    if (march_ok(n)) {
        // z-marching tiles of 64 x 8 columns
        const int ZC = march_zc(n);
        const unsigned tiles = (unsigned)((n / 64) * (n / 8) * (n / ZC));
#define EC_RHS_MARCH(T, D)                                                                                       \
    k_rhs_march<T, 64, 8, D><<<tiles, 256, 0, ctx->stream>>>(n, ZC, (T)ih, (T)visc, (const T *)u[0], (const T *)u[1], \
                                                             (const T *)u[2], (T *)du[0], (T *)du[1], (T *)du[2],    \
                                                             doadd, ad)
        if (dtype == EC_F64) {
            if (doadd == 1) EC_RHS_MARCH(double, 1);
            else if (doadd == 2) EC_RHS_MARCH(double, 2);
            else EC_RHS_MARCH(double, 0);
        } else {
            if (doadd == 1) EC_RHS_MARCH(float, 1);
            else if (doadd == 2) EC_RHS_MARCH(float, 2);
            else EC_RHS_MARCH(float, 0);
        }
#undef EC_RHS_MARCH
        EC_CHECK_LAUNCH(ctx, "k_rhs_march");
        return EC_OK;
    }
    if (dtype == EC_F64)
        k_rhs<double><<<g.grid, g.block, 0, ctx->stream>>>(n, ih, visc, (const double *)u[0], (const double *)u[1],
                                                           (const double *)u[2], (double *)du[0], (double *)du[1],
                                                           (double *)du[2], doadd, ad);
    else
        k_rhs<float><<<g.grid, g.block, 0, ctx->stream>>>(n, (float)ih, (float)visc, (const float *)u[0],
                                                          (const float *)u[1], (const float *)u[2], (float *)du[0],
                                                          (float *)du[1], (float *)du[2], doadd, ad);
    EC_CHECK_LAUNCH(ctx, "k_rhs");
    return EC_OK;
}

extern "C" int ec_ns_rhs(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3],
                         void *const du[3], int doadd) {
    EC_REQUIRE(ctx, ctx && u && du, "ec_ns_rhs: null argument");
    EC_CHECK_FIELD(ctx, n);
    return rhs_launch(ctx, dtype, n, l, visc, u, du, doadd ? 1 : 0, RhsAdd{1.0, nullptr, nullptr, nullptr});
}

extern "C" int ec_ns_rhs_add(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3],
                             void *const du[3], double s, const void *const add[3]) {
    EC_REQUIRE(ctx, ctx && u && du && add && add[0] && add[1] && add[2], "ec_ns_rhs_add: null argument");
    EC_CHECK_FIELD(ctx, n);
    return rhs_launch(ctx, dtype, n, l, visc, u, du, 2,
                      RhsAdd{s, (const double *)add[0], (const double *)add[1], (const double *)add[2]});
}

extern "C" int ec_ns_rhs_div(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3],
                             void *const du[3], void *d, double rdiv) {
    EC_REQUIRE(ctx, ctx && u && du && d, "ec_ns_rhs_div: null argument");
    EC_CHECK_FIELD(ctx, n);
    if (!march_ok(n)) {
        // small grids: the plain kernels
        int rc = ec_ns_rhs(ctx, dtype, n, l, visc, u, du, 0);
        if (!rc) rc = ec_ns_divergence(ctx, dtype, n, l, du, d);
        if (rc || rdiv == 0.0) return rc;
        Grid3 g = grid3(n);
        const double sih = rdiv * n / l;
        ec_prof_scope ps(ctx, "divergence", 5.0 * ec_cube(n) * (dtype == EC_F64 ? 8 : 4));
        if (dtype == EC_F64)
            k_divergence_add<double><<<g.grid, g.block, 0, ctx->stream>>>(
                n, sih, (const double *)u[0], (const double *)u[1], (const double *)u[2], (double *)d);
        else
            k_divergence_add<float><<<g.grid, g.block, 0, ctx->stream>>>(
                n, (float)sih, (const float *)u[0], (const float *)u[1], (const float *)u[2], (float *)d);
        EC_CHECK_LAUNCH(ctx, "k_divergence_add");
        return EC_OK;
    }
    const double ih = n / l;
    const double b = dtype == EC_F64 ? 8 : 4;
    ec_prof_scope ps(ctx, "ns_rhs_div", 7.0 * ec_cube(n) * b);
    // FP32 takes 128 x 16 tiles when n allows (the 3-cell x halo touches two
    // partial 128-byte lines per row of 4 instead of per row of 2; 150 KB of
    // LDS, one block of 1024 threads per CU) and 64 x 16 otherwise (halo rows
    // 19/16 of the tile instead of 11/8; 78 KB, two blocks per CU); FP64 keeps
    // 64 x 8, whose 88 KB already hold it to one block per CU
    const int TY = dtype == EC_F64 ? 8 : 16;
    const int ZC = march_zc(n, TY);
    const bool wide = dtype == EC_F32 && n % 128 == 0 && n >= 512;
    const unsigned tiles = (unsigned)((n / (wide ? 128 : 64)) * (n / TY) * (n / ZC));
    if (dtype == EC_F64)
        k_rhs_div_march<double, 64, 8><<<tiles, 256, 0, ctx->stream>>>(
            n, ZC, ih, visc, (const double *)u[0], (const double *)u[1], (const double *)u[2], (double *)du[0],
            (double *)du[1], (double *)du[2], (double *)d, rdiv * ih);
    else if (wide)
        k_rhs_div_march<float, 128, 16, 1024><<<tiles, 1024, 0, ctx->stream>>>(
            n, ZC, (float)ih, (float)visc, (const float *)u[0], (const float *)u[1], (const float *)u[2],
            (float *)du[0], (float *)du[1], (float *)du[2], (float *)d, (float)(rdiv * ih));
    else
        k_rhs_div_march<float, 64, 16, 512><<<tiles, 512, 0, ctx->stream>>>(
            n, ZC, (float)ih, (float)visc, (const float *)u[0], (const float *)u[1], (const float *)u[2],
            (float *)du[0], (float *)du[1], (float *)du[2], (float *)d, (float)(rdiv * ih));
    EC_CHECK_LAUNCH(ctx, "k_rhs_div_march");
    return EC_OK;
}

extern "C" int ec_ns_divergence(ec_ctx *ctx, int dtype, int n, double l, void *const u[3], void *d) {
    EC_REQUIRE(ctx, ctx && u && d, "ec_ns_divergence: null argument");
    EC_CHECK_FIELD(ctx, n);
    Grid3 g = grid3(n);
    const double ih = n / l;
    ec_prof_scope ps(ctx, "divergence", 4.0 * ec_cube(n) * (dtype == EC_F64 ? 8 : 4));
    if (dtype == EC_F64)
        k_divergence<double><<<g.grid, g.block, 0, ctx->stream>>>(n, ih, (const double *)u[0], (const double *)u[1],
                                                                  (const double *)u[2], (double *)d);
    else
        k_divergence<float><<<g.grid, g.block, 0, ctx->stream>>>(n, (float)ih, (const float *)u[0],
                                                                 (const float *)u[1], (const float *)u[2], (float *)d);
    EC_CHECK_LAUNCH(ctx, "k_divergence");
    return EC_OK;
}

extern "C" int ec_ns_pressuregradient(ec_ctx *ctx, int dtype, int n, double l, void *const u[3], const void *p) {
    EC_REQUIRE(ctx, ctx && u && p, "ec_ns_pressuregradient: null argument");
    EC_CHECK_FIELD(ctx, n);
    Grid3 g = grid3(n);
    const double ih = n / l;
    ec_prof_scope ps(ctx, "pressuregradient", 7.0 * ec_cube(n) * (dtype == EC_F64 ? 8 : 4));
    if (dtype == EC_F64)
        k_pressuregradient<double><<<g.grid, g.block, 0, ctx->stream>>>(n, ih, (double *)u[0], (double *)u[1],
                                                                        (double *)u[2], (const double *)p);
    else
        k_pressuregradient<float><<<g.grid, g.block, 0, ctx->stream>>>(n, (float)ih, (float *)u[0], (float *)u[1],
                                                                       (float *)u[2], (const float *)p);
    EC_CHECK_LAUNCH(ctx, "k_pressuregradient");
    return EC_OK;
}

extern "C" int ec_ns_pressuregradient_into(ec_ctx *ctx, int dtype, int n, double l, const void *const u[3],
                                           const void *p, void *const out[3]) {
    EC_REQUIRE(ctx, ctx && u && p && out, "ec_ns_pressuregradient_into: null argument");
    EC_CHECK_FIELD(ctx, n);
    Grid3 g = grid3(n);
    const double ih = n / l;
    ec_prof_scope ps(ctx, "pressuregradient", 7.0 * ec_cube(n) * (dtype == EC_F64 ? 8 : 4));
    if (dtype == EC_F64)
        k_pressuregradient_into<double><<<g.grid, g.block, 0, ctx->stream>>>(
            n, ih, (const double *)u[0], (const double *)u[1], (const double *)u[2], (double *)out[0],
            (double *)out[1], (double *)out[2], (const double *)p);
    else
        k_pressuregradient_into<float><<<g.grid, g.block, 0, ctx->stream>>>(
            n, (float)ih, (const float *)u[0], (const float *)u[1], (const float *)u[2], (float *)out[0],
            (float *)out[1], (float *)out[2], (const float *)p);
    EC_CHECK_LAUNCH(ctx, "k_pressuregradient_into");
    return EC_OK;
}

// The same stage update with 16-byte accesses: a thread owns V = 16 / sizeof(T)
// consecutive x cells (n % V == 0), so every field is read and written as one
// vector per thread; the x + 1 pressure neighbour of the last cell is one
// extra scalar load.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float4 type;
    static constexpr int V = 4;
};
template <>
struct Vec16<double> {
    typedef double2 type;
    static constexpr int V = 2;
};

template <typename T, int STAGE>
__global__ void __launch_bounds__(256) k_rk4_update_vec(int n, T ih, T dt, Ptr3 u0p, Ptr3 usp, Ptr3 dup,
                                                        const T *__restrict__ p, Ptr3 accp, size_t nvec) {
    typedef typename Vec16<T>::type VT;
    constexpr int V = Vec16<T>::V;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= nvec) return;
    const unsigned nx = (unsigned)n / V;
    const unsigned iv = (unsigned)(t % nx);
    const size_t row = t / nx;  // j + n k
    const unsigned j = (unsigned)(row % (unsigned)n), k = (unsigned)(row / (unsigned)n);
    const size_t n1 = (size_t)n, n2 = n1 * n1;
    const size_t I = (size_t)iv * V + n1 * row;
    const size_t Iy = (size_t)iv * V + n1 * ((j + 1 == (unsigned)n ? 0 : j + 1) + n1 * k);
    const size_t Iz = (size_t)iv * V + n1 * (j + n1 * (k + 1 == (unsigned)n ? 0 : k + 1));
    const unsigned xl = iv * V + V;  // x + 1 of the last cell
    const T pxl = p[(xl == (unsigned)n ? 0 : xl) + n1 * row];
    const VT p0 = *(const VT *)(p + I), py = *(const VT *)(p + Iy), pz = *(const VT *)(p + Iz);
    const T *p0s = (const T *)&p0, *pys = (const T *)&py, *pzs = (const T *)&pz;
    T gp[3][V];
#pragma unroll
    for (int e = 0; e < V; e++) {
        const T pxp = e + 1 < V ? p0s[e + 1] : pxl;
        gp[0][e] = (pxp - p0s[e]) * ih;
        gp[1][e] = (pys[e] - p0s[e]) * ih;
        gp[2][e] = (pzs[e] - p0s[e]) * ih;
    }
    // all loads first: the three components' fields may alias as far as the
    // compiler knows, so loads placed after a store would wait for it (read
    // once: non-temporal; p's neighbour rows are shared and stay cached)
    VT d[3], uv[3], av[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        d[a] = ec_nt_ld((const VT *)((const T *)dup.p[a] + I));
        uv[a] = ec_nt_ld((const VT *)((const T *)u0p.p[a] + I));
        if (STAGE != 0) av[a] = ec_nt_ld((const VT *)((const T *)accp.p[a] + I));
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const T *ds = (const T *)&d[a], *us_ = (const T *)&uv[a], *as = (const T *)&av[a];
        T kk[V];
#pragma unroll
        for (int e = 0; e < V; e++) kk[e] = ds[e] - gp[a][e];
        VT o1, o2;
        T *r1 = (T *)&o1, *r2 = (T *)&o2;
        if (STAGE == 0) {
#pragma unroll
            for (int e = 0; e < V; e++) {
                r1[e] = kk[e];
                r2[e] = us_[e] + (dt / 2) * kk[e];
            }
        } else if (STAGE == 1 || STAGE == 2) {
#pragma unroll
            for (int e = 0; e < V; e++) {
                r1[e] = as[e] + 2 * kk[e];
                r2[e] = us_[e] + (STAGE == 1 ? dt / 2 : dt) * kk[e];
            }
        } else {
#pragma unroll
            for (int e = 0; e < V; e++) r1[e] = us_[e] + (dt / 6) * (as[e] + kk[e]);
        }
        // streaming stores: nothing here reads them back
        if (STAGE < 3) {
            ec_nt_store16((VT *)((T *)accp.p[a] + I), o1);
            ec_nt_store16((VT *)((T *)usp.p[a] + I), o2);
        } else {
            ec_nt_store16((VT *)((T *)u0p.p[a] + I), o1);
        }
    }
}

// The stage update of RK4 in stored stage states (ec_rk4_combine): k =
// du - grad p; stage s < 3 writes u_{s+1} = u0 + c_s dt k into its own
// buffer, the last stage reads u1, u2, u3 and writes the new u0.  16-byte
// accesses as k_rk4_update_vec, or one cell per thread (VEC false) where the
// fields are not 16-byte aligned.
template <typename T, int STAGE, bool VEC, bool HASP = true>
__global__ void __launch_bounds__(256) k_rk4_states(int n, T ih, T dt, Ptr3 u0p, Ptr3 u1p, Ptr3 u2p, Ptr3 u3p,
                                                    Ptr3 dup, const T *__restrict__ p, size_t nvec) {
    typedef typename std::conditional<VEC, typename Vec16<T>::type, T>::type VT;
    constexpr int V = VEC ? Vec16<T>::V : 1;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= nvec) return;
    const unsigned nx = (unsigned)n / V;
    const unsigned iv = (unsigned)(t % nx);
    const size_t row = t / nx;  // j + n k
    const unsigned j = (unsigned)(row % (unsigned)n), k = (unsigned)(row / (unsigned)n);
    const size_t n1 = (size_t)n;
    const size_t I = (size_t)iv * V + n1 * row;
    const size_t Iy = (size_t)iv * V + n1 * ((j + 1 == (unsigned)n ? 0 : j + 1) + n1 * k);
    const size_t Iz = (size_t)iv * V + n1 * (j + n1 * (k + 1 == (unsigned)n ? 0 : k + 1));
    T gp[3][V];
    if constexpr (HASP) {
        const unsigned xl = iv * V + V;  // x + 1 of the last cell
        const T pxl = p[(xl == (unsigned)n ? 0 : xl) + n1 * row];
        const VT p0 = *(const VT *)(p + I), py = *(const VT *)(p + Iy), pz = *(const VT *)(p + Iz);
        const T *p0s = (const T *)&p0, *pys = (const T *)&py, *pzs = (const T *)&pz;
#pragma unroll
        for (int e = 0; e < V; e++) {
            const T pxp = e + 1 < V ? p0s[e + 1] : pxl;
            gp[0][e] = (pxp - p0s[e]) * ih;
            gp[1][e] = (pys[e] - p0s[e]) * ih;
            gp[2][e] = (pzs[e] - p0s[e]) * ih;
        }
    }
    // all loads first (the fields may alias as far as the compiler knows);
    // read once: non-temporal
    VT d[3], uv[3], s1[3], s2[3], s3[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        d[a] = ec_nt_ld((const VT *)((const T *)dup.p[a] + I));
        uv[a] = ec_nt_ld((const VT *)((const T *)u0p.p[a] + I));
        if constexpr (STAGE == 3) {
            s1[a] = ec_nt_ld((const VT *)((const T *)u1p.p[a] + I));
            s2[a] = ec_nt_ld((const VT *)((const T *)u2p.p[a] + I));
            s3[a] = ec_nt_ld((const VT *)((const T *)u3p.p[a] + I));
        }
    }
    const Ptr3 &outp = STAGE == 0 ? u1p : (STAGE == 1 ? u2p : (STAGE == 2 ? u3p : u0p));
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const T *ds = (const T *)&d[a], *us_ = (const T *)&uv[a];
        const T *a1 = (const T *)&s1[a], *a2 = (const T *)&s2[a], *a3 = (const T *)&s3[a];
        VT o;
        T *r = (T *)&o;
#pragma unroll
        for (int e = 0; e < V; e++) {
            T kk = ds[e];
            if constexpr (HASP) kk = ds[e] - gp[a][e];
            if constexpr (STAGE < 3) r[e] = us_[e] + ec_rk4_stage_coef<T>(STAGE, dt) * kk;
            else r[e] = ec_rk4_combine<T>(us_[e], a1[e], a2[e], a3[e], kk, dt);
        }
        if constexpr (VEC) ec_nt_store16((VT *)((T *)outp.p[a] + I), o);
        else ec_nt_store((T *)outp.p[a] + I, r[0]);
    }
}

template <typename T>
static int rk4_states_launch(ec_ctx *ctx, int n, double l, int stage, double dt, void *const u0[3],
                             void *const ust[9], void *const du[3], const void *p) {
    const T ih = (T)(n / l);
    Ptr3 U0{{u0[0], u0[1], u0[2]}}, DU{{du[0], du[1], du[2]}};
    Ptr3 S[3];
    for (int s = 0; s < 3; s++)
        for (int a = 0; a < 3; a++) S[s].p[a] = ust[3 * s + a];
    constexpr int V = Vec16<T>::V;
    bool aligned = n % V == 0 && ((uintptr_t)p % 16) == 0;
    for (int a = 0; a < 3; a++) {
        aligned = aligned && ((uintptr_t)u0[a] % 16) == 0 && ((uintptr_t)du[a] % 16) == 0;
        for (int s = 0; s < 3; s++) aligned = aligned && ((uintptr_t)ust[3 * s + a] % 16) == 0;
    }
    const size_t nvec = ec_cube(n) / (aligned ? V : 1);
    const unsigned blocks = (unsigned)((nvec + 255) / 256);
    // without p, k = du (a tendency that is already projected)
#define EC_RK4S(ST, VEC)                                                                                          \
    do {                                                                                                          \
        if (p)                                                                                                    \
            k_rk4_states<T, ST, VEC, true><<<blocks, 256, 0, ctx->stream>>>(n, ih, (T)dt, U0, S[0], S[1], S[2], DU, \
                                                                             (const T *)p, nvec);                 \
        else                                                                                                      \
            k_rk4_states<T, ST, VEC, false><<<blocks, 256, 0, ctx->stream>>>(n, ih, (T)dt, U0, S[0], S[1], S[2],  \
                                                                              DU, nullptr, nvec);                 \
    } while (0)
    if (aligned) {
        switch (stage) {
            case 0: EC_RK4S(0, true); break;
            case 1: EC_RK4S(1, true); break;
            case 2: EC_RK4S(2, true); break;
            default: EC_RK4S(3, true); break;
        }
    } else {
        switch (stage) {
            case 0: EC_RK4S(0, false); break;
            case 1: EC_RK4S(1, false); break;
            case 2: EC_RK4S(2, false); break;
            default: EC_RK4S(3, false); break;
        }
    }
#undef EC_RK4S
    EC_CHECK_LAUNCH(ctx, "k_rk4_states");
    return EC_OK;
}

template <typename T>
static int rk4_launch(ec_ctx *ctx, int n, double l, int stage, double dt, void *const u0[3], void *const us[3],
                      void *const du[3], const void *p, void *const acc[3]) {
    Grid3 g = grid3(n);
    const T ih = (T)(n / l);
    Ptr3 U0{{u0[0], u0[1], u0[2]}}, US{{us ? us[0] : nullptr, us ? us[1] : nullptr, us ? us[2] : nullptr}},
        DU{{du[0], du[1], du[2]}}, AC{{acc[0], acc[1], acc[2]}};
    constexpr int V = Vec16<T>::V;
    bool aligned = n % V == 0 && ((uintptr_t)p % 16) == 0;
    for (int a = 0; a < 3; a++)
        aligned = aligned && ((uintptr_t)u0[a] % 16) == 0 && ((uintptr_t)du[a] % 16) == 0 &&
                  ((uintptr_t)acc[a] % 16) == 0 && (!us || ((uintptr_t)us[a] % 16) == 0);
    if (aligned) {
        const size_t nvec = ec_cube(n) / V;
        const unsigned blocks = (unsigned)((nvec + 255) / 256);
        switch (stage) {
            case 0: k_rk4_update_vec<T, 0><<<blocks, 256, 0, ctx->stream>>>(n, ih, (T)dt, U0, US, DU, (const T *)p, AC, nvec); break;
            case 1: k_rk4_update_vec<T, 1><<<blocks, 256, 0, ctx->stream>>>(n, ih, (T)dt, U0, US, DU, (const T *)p, AC, nvec); break;
            case 2: k_rk4_update_vec<T, 2><<<blocks, 256, 0, ctx->stream>>>(n, ih, (T)dt, U0, US, DU, (const T *)p, AC, nvec); break;
            default: k_rk4_update_vec<T, 3><<<blocks, 256, 0, ctx->stream>>>(n, ih, (T)dt, U0, US, DU, (const T *)p, AC, nvec); break;
        }
        EC_CHECK_LAUNCH(ctx, "k_rk4_update_vec");
        return EC_OK;
    }
    switch (stage) {
        case 0: k_rk4_update<T, 0><<<g.grid, g.block, 0, ctx->stream>>>(n, ih, (T)dt, U0, US, DU, (const T *)p, AC); break;
        case 1: k_rk4_update<T, 1><<<g.grid, g.block, 0, ctx->stream>>>(n, ih, (T)dt, U0, US, DU, (const T *)p, AC); break;
        case 2: k_rk4_update<T, 2><<<g.grid, g.block, 0, ctx->stream>>>(n, ih, (T)dt, U0, US, DU, (const T *)p, AC); break;
        default: k_rk4_update<T, 3><<<g.grid, g.block, 0, ctx->stream>>>(n, ih, (T)dt, U0, US, DU, (const T *)p, AC); break;
    }
    EC_CHECK_LAUNCH(ctx, "k_rk4_update");
    return EC_OK;
}

extern "C" int ec_ns_rk4_update(ec_ctx *ctx, int dtype, int n, double l, int stage, double dt, void *const u0[3],
                                void *const us[3], void *const du[3], const void *p, void *const acc[3]) {
    EC_REQUIRE(ctx, ctx && u0 && du && p && acc, "ec_ns_rk4_update: null argument");
    EC_REQUIRE(ctx, stage >= 0 && stage <= 3, "ec_ns_rk4_update: stage %d out of range", stage);
    EC_REQUIRE(ctx, stage == 3 || us, "ec_ns_rk4_update: stage state required before the last stage");
    EC_CHECK_FIELD(ctx, n);
    const double b = dtype == EC_F64 ? 8 : 4;
    // reads p, du(3), u0(3), acc(3 unless stage 0); writes us(3) + acc(3), or u0(3) at stage 3
    const double units = stage == 0 ? 13 : (stage == 3 ? 13 : 16);
    ec_prof_scope ps(ctx, "rk4_update", units * ec_cube(n) * b);
    return dtype == EC_F64 ? rk4_launch<double>(ctx, n, l, stage, dt, u0, us, du, p, acc)
                           : rk4_launch<float>(ctx, n, l, stage, dt, u0, us, du, p, acc);
}

extern "C" int ec_ns_rk4_update_states(ec_ctx *ctx, int dtype, int n, double l, int stage, double dt,
                                       void *const u0[3], void *const ust[9], void *const du[3], const void *p) {
    EC_REQUIRE(ctx, ctx && u0 && ust && du, "ec_ns_rk4_update_states: null argument");
    EC_REQUIRE(ctx, stage >= 0 && stage <= 3, "ec_ns_rk4_update_states: stage %d out of range", stage);
    for (int q = 0; q < 9; q++) {
        // the buffers a stage touches: its own state (stage < 3), all three at stage 3
        if (stage == 3 || q / 3 == stage) EC_REQUIRE(ctx, ust[q], "ec_ns_rk4_update_states: null stage state %d", q);
    }
    EC_CHECK_FIELD(ctx, n);
    EC_REQUIRE(ctx, ec_cube(n) / 256 < 2147483647.0, "ec_ns_rk4_update_states: n=%d too large", n);
    const double b = dtype == EC_F64 ? 8 : 4;
    // reads p, du(3), u0(3) (+ u1, u2, u3 at stage 3); writes one state (3)
    const double units = (stage == 3 ? 19 : 10) - (p ? 0 : 1);
    ec_prof_scope ps(ctx, "rk4_update", units * ec_cube(n) * b);
    return dtype == EC_F64 ? rk4_states_launch<double>(ctx, n, l, stage, dt, u0, ust, du, p)
                           : rk4_states_launch<float>(ctx, n, l, stage, dt, u0, ust, du, p);
}
