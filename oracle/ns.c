/*
 * Oracle: 3D staggered finite-volume Navier-Stokes operators, FP64.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates src/N.jl; each function
 * cites the lines it follows.  0-based indices: the reference's I (1-based)
 * is our i+1, so every shift pattern carries over unchanged.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define IDX(i, j, k) ((size_t)(i) + (size_t)n * ((size_t)(j) + (size_t)n * (size_t)(k)))

/* periodic wrap for |shift| < n (src/N.jl:64-77, mod1) */
static inline int pw(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

/* stress(u, visc, i, j) (src/N.jl:106-118), evaluated at cell I */
static inline double stress_at(int n, double h, double visc, const double *const u[3], int a,
                               int b, int i, int j, int k) {
    const int I[3] = {i, j, k};
    const int s = (a != b);
    int R[3] = {i, j, k}, L[3] = {i, j, k};
    R[b] = pw(I[b] + s, n);
    L[b] = pw(I[b] + s - 1, n);
    const double ua_r = u[a][IDX(R[0], R[1], R[2])];
    const double ua_l = u[a][IDX(L[0], L[1], L[2])];
    int R2[3] = {i, j, k}, L2[3] = {i, j, k};
    R2[a] = pw(I[a] + s, n);
    L2[a] = pw(I[a] + s - 1, n);
    const double ub_r = u[b][IDX(R2[0], R2[1], R2[2])];
    const double ub_l = u[b][IDX(L2[0], L2[1], L2[2])];
    const double eta_b_ua = (ua_r + ua_l) / 2;
    const double eta_a_ub = (ub_r + ub_l) / 2;
    const double d_b_ua = (ua_r - ua_l) / h;
    const double d_a_ub = (ub_r - ub_l) / h;
    return eta_b_ua * eta_a_ub - visc * (d_b_ua + d_a_ub);
}

/* stresstensor (src/N.jl:94-104): order 11, 22, 33, 12, 23, 31 */
void or_stresstensor(int n, double l, double visc, const double *u1, const double *u2,
                     const double *u3, double *s11, double *s22, double *s33, double *s12,
                     double *s23, double *s31) {
    const double h = l / n;
    const double *const u[3] = {u1, u2, u3};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const size_t I = IDX(i, j, k);
                s11[I] = stress_at(n, h, visc, u, 0, 0, i, j, k);
                s22[I] = stress_at(n, h, visc, u, 1, 1, i, j, k);
                s33[I] = stress_at(n, h, visc, u, 2, 2, i, j, k);
                s12[I] = stress_at(n, h, visc, u, 0, 1, i, j, k);
                s23[I] = stress_at(n, h, visc, u, 1, 2, i, j, k);
                s31[I] = stress_at(n, h, visc, u, 2, 0, i, j, k);
            }
}

/* tensordivergence!(du, sigma, doadd) (src/N.jl:120-160).  doadd = false is a
 * strong zero in Julia (false * NaN == 0), so it is a branch here. */
void or_tensordivergence(int n, double l, const double *s11, const double *s22,
                         const double *s33, const double *s12, const double *s23,
                         const double *s31, double *du1, double *du2, double *du3, int doadd) {
    const double h = l / n;
    const double *s21 = s12, *s32 = s23, *s13 = s31;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const size_t I = IDX(i, j, k);
                const int ip = pw(i + 1, n), im = pw(i - 1, n);
                const int jp = pw(j + 1, n), jm = pw(j - 1, n);
                const int kp = pw(k + 1, n), km = pw(k - 1, n);
                double a1 = doadd ? du1[I] : 0.0;
                a1 = a1 - (s11[IDX(ip, j, k)] - s11[I]) / h - (s12[I] - s12[IDX(i, jm, k)]) / h -
                     (s13[I] - s13[IDX(i, j, km)]) / h;
                double a2 = doadd ? du2[I] : 0.0;
                a2 = a2 - (s21[I] - s21[IDX(im, j, k)]) / h - (s22[IDX(i, jp, k)] - s22[I]) / h -
                     (s23[I] - s23[IDX(i, j, km)]) / h;
                double a3 = doadd ? du3[I] : 0.0;
                a3 = a3 - (s31[I] - s31[IDX(im, j, k)]) / h - (s32[I] - s32[IDX(i, jm, k)]) / h -
                     (s33[IDX(i, j, kp)] - s33[I]) / h;
                du1[I] = a1;
                du2[I] = a2;
                du3[I] = a3;
            }
}

/* tensordivergence_nonsym! (src/N.jl:162-200).  s9 in column-major order
 * 11, 21, 31, 12, 22, 32, 13, 23, 33: sigma_ab = s9[a + 3 b]. */
void or_tensordivergence_nonsym(int n, double l, const double *const *s9, double *du1,
                                double *du2, double *du3, int doadd) {
    const double h = l / n;
    double *du[3] = {du1, du2, du3};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const size_t I = IDX(i, j, k);
                const int c[3] = {i, j, k};
                double out[3];
                for (int a = 0; a < 3; a++) {
                    double acc = doadd ? du[a][I] : 0.0;
                    for (int b = 0; b < 3; b++) {
                        const double *s = s9[a + 3 * b];
                        int q[3] = {c[0], c[1], c[2]};
                        double diff;
                        if (a == b) {
                            q[b] = pw(c[b] + 1, n);
                            diff = s[IDX(q[0], q[1], q[2])] - s[I];
                        } else {
                            q[b] = pw(c[b] - 1, n);
                            diff = s[I] - s[IDX(q[0], q[1], q[2])];
                        }
                        acc = acc - diff / h;
                    }
                    out[a] = acc;
                }
                du1[I] = out[0];
                du2[I] = out[1];
                du3[I] = out[2];
            }
}

/* divergence! (src/N.jl:222-238) */
void or_divergence(int n, double l, const double *u1, const double *u2, const double *u3,
                   double *d) {
    const double h = l / n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const size_t I = IDX(i, j, k);
                d[I] = (u1[I] - u1[IDX(pw(i - 1, n), j, k)]) / h +
                       (u2[I] - u2[IDX(i, pw(j - 1, n), k)]) / h +
                       (u3[I] - u3[IDX(i, j, pw(k - 1, n))]) / h;
            }
}

/* pressuregradient! (src/N.jl:240-254) */
void or_pressuregradient(int n, double l, double *u1, double *u2, double *u3, const double *p) {
    const double h = l / n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const size_t I = IDX(i, j, k);
                u1[I] -= (p[IDX(pw(i + 1, n), j, k)] - p[I]) / h;
                u2[I] -= (p[IDX(i, pw(j + 1, n), k)] - p[I]) / h;
                u3[I] -= (p[IDX(i, j, pw(k + 1, n))] - p[I]) / h;
            }
}

static inline int wavenumber_int(int n, int i0) {
    /* src/N.jl:32 with 1-based i = i0 + 1 */
    const int i = i0 + 1;
    return i - 1 - ((i <= ((n + 1) >> 1)) ? 0 : n);
}

/* project!(u, poisson) (src/N.jl:256-307).  p receives the pressure. */
int or_project(int n, double l, double *u1, double *u2, double *u3, double *p) {
    const double h = l / n;
    const int nh = n / 2 + 1;
    const size_t N = (size_t)n * n * n;
    (void)N;
    or_divergence(n, l, u1, u2, u3, p);
    double *phat = (double *)malloc(sizeof(double) * 2 * (size_t)nh * n * n);
    if (!phat) return -1;
    or_rfft3(n, p, phat);
    double *sx = (double *)malloc(sizeof(double) * n);
    for (int q = 0; q < n; q++) {
        const double s = sin(M_PI * (double)wavenumber_int(n, q) / (double)n);
        sx[q] = s * s;
    }
#pragma omp parallel for schedule(static)
    for (int kz = 0; kz < n; kz++)
        for (int ky = 0; ky < n; ky++)
            for (int kx = 0; kx < nh; kx++) {
                const size_t I = (size_t)kx + (size_t)nh * ((size_t)ky + (size_t)n * kz);
                const double s1 = sin(M_PI * (double)kx / (double)n);
                const double denom = -4 / (h * h) * (s1 * s1 + sx[ky] + sx[kz]);
                phat[2 * I] /= denom;
                phat[2 * I + 1] /= denom;
            }
    phat[0] = 0.0;
    phat[1] = 0.0;
    free(sx);
    or_irfft3(n, phat, p);
    free(phat);
    or_pressuregradient(n, l, u1, u2, u3, p);
    return 0;
}

double or_maxabs(size_t count, const double *x) {
    double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (size_t i = 0; i < count; i++) {
        const double a = fabs(x[i]);
        if (a > m) m = a;
    }
    return m;
}

double or_sumsq(size_t count, const double *x) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t i = 0; i < count; i++) s += x[i] * x[i];
    return s;
}

/* ---- counter-based normal generator (replaces Julia's Xoshiro/randn!,
 *      src/N.jl:406; the GPU path uses the identical integer recipe) ---- */
static inline unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

static inline double normal_at(unsigned long long seed, unsigned long long ctr) {
    const unsigned long long h1 = splitmix64(seed ^ splitmix64(2 * ctr));
    const unsigned long long h2 = splitmix64(seed ^ splitmix64(2 * ctr + 1));
    const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

void or_randn(size_t count, unsigned long long seed, unsigned long long offset, double *out) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < count; i++) out[i] = normal_at(seed, offset + i);
}

static double profile_eval(int profile, double param, int k) {
    /* peak_profile (src/N.jl:366) or the power law of src/N.jl:1245 */
    if (profile == 0) return pow((double)k, 4) * exp(-2 * pow((double)k / param, 2));
    return k > 0 ? pow((double)k, param) : 0.0;
}

/* randomfield (src/N.jl:394-476) */
int or_randomfield(int n, double l, int profile, double param, double totalenergy,
                   unsigned long long seed, double *u1, double *u2, double *u3) {
    const size_t N = (size_t)n * n * n;
    const int nh = n / 2 + 1, kmax = n / 2;
    const size_t ns = (size_t)nh * n * n;
    double *u[3] = {u1, u2, u3};
    for (int c = 0; c < 3; c++) or_randn(N, seed, (unsigned long long)c * N, u[c]);
    double *p = (double *)malloc(sizeof(double) * N);
    or_project(n, l, u1, u2, u3, p);
    double *uh[3];
    const double fft_factor = (double)n * n * n;
    for (int c = 0; c < 3; c++) {
        uh[c] = (double *)malloc(sizeof(double) * 2 * ns);
        or_rfft3(n, u[c], uh[c]);
        for (size_t q = 0; q < 2 * ns; q++) uh[c][q] /= fft_factor;
    }
    const int kdiag = (int)floor(sqrt(3.0) * kmax);
    double totalprofile = 0.0;
    for (int k = 0; k <= kdiag; k++) totalprofile += profile_eval(profile, param, k);
    double *eshell = (double *)calloc((size_t)kdiag + 2, sizeof(double));
    int *shell = (int *)malloc(sizeof(int) * ns);
    for (int kz = 0; kz < n; kz++)
        for (int ky = 0; ky < n; ky++)
            for (int kx = 0; kx < nh; kx++) {
                const size_t I = (size_t)kx + (size_t)nh * ((size_t)ky + (size_t)n * kz);
                const long k2 = wavenumber_int(n, ky), k3 = wavenumber_int(n, kz);
                const long kk = (long)kx * kx + k2 * k2 + k3 * k3;
                int s = (int)floor(sqrt((double)kk));
                while ((long)s * s > kk) s--;
                while ((long)(s + 1) * (s + 1) <= kk) s++;
                shell[I] = s;
                double E = 0.0;
                for (int c = 0; c < 3; c++)
                    E += uh[c][2 * I] * uh[c][2 * I] + uh[c][2 * I + 1] * uh[c][2 * I + 1];
                E /= 2;
                /* sum(Emask) + sum(selectdim(Emask, 1, 2:kmax)) */
                const double wgt = (kx >= 1 && kx <= kmax - 1) ? 2.0 : 1.0;
                eshell[s] += wgt * E;
            }
    double *factor = (double *)malloc(sizeof(double) * ((size_t)kdiag + 1));
    for (int k = 0; k <= kdiag; k++) {
        const double E0 = totalenergy * profile_eval(profile, param, k) / totalprofile;
        factor[k] = sqrt(E0 / eshell[k]);
    }
    for (int c = 0; c < 3; c++) {
        for (size_t I = 0; I < ns; I++) {
            uh[c][2 * I] *= factor[shell[I]];
            uh[c][2 * I + 1] *= factor[shell[I]];
        }
        or_irfft3(n, uh[c], u[c]);
        for (size_t q = 0; q < N; q++) u[c][q] *= fft_factor;
        free(uh[c]);
    }
    or_project(n, l, u1, u2, u3, p);
    free(p);
    free(factor);
    free(shell);
    free(eshell);
    return 0;
}

/* ---- filters ---- */

/* applyfilter! with a kernelproduct kernel (src/N.jl:627-653): the dense
 * (2R+1)^3 array is the outer product of the three 1D factors, summed in
 * Julia's column-major order over the offsets. */
void or_applyfilter(int n, const double *w1, int r1, const double *w2, int r2,
                    const double *w3, int r3, const double *in, double *out) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                double s = 0.0;
                for (int c = -r3; c <= r3; c++)
                    for (int b = -r2; b <= r2; b++)
                        for (int a = -r1; a <= r1; a++) {
                            const double w = w1[a + r1] * w2[b + r2] * w3[c + r3];
                            const int ii = (int)(((long)i + a) % n + n) % n;
                            const int jj = (int)(((long)j + b) % n + n) % n;
                            const int kk = (int)(((long)k + c) % n + n) % n;
                            s += w * in[IDX(ii, jj, kk)];
                        }
                out[IDX(i, j, k)] = s;
            }
}

/* coarsegrain! (src/N.jl:670-684) */
void or_coarsegrain(int n, int nbar, const int *stag, const double *in, double *out) {
    const int c = n / nbar, a = c / 2;
#pragma omp parallel for schedule(static)
    for (int K = 0; K < nbar; K++)
        for (int J = 0; J < nbar; J++)
            for (int I = 0; I < nbar; I++) {
                const int i = c * (I + 1) - 1 - (!stag[0]) * a;
                const int j = c * (J + 1) - 1 - (!stag[1]) * a;
                const int k = c * (K + 1) - 1 - (!stag[2]) * a;
                out[(size_t)I + (size_t)nbar * ((size_t)J + (size_t)nbar * K)] = in[IDX(i, j, k)];
            }
}

/* lazycoarsegrain(lazyfilter(.)) with separable weights (src/N.jl:655-698):
 * each axis has its own weights and offset relative to the coarse face
 * index c*(J+1)-1, which also expresses the centre point (offset -c/2) and
 * the even-compression trapezoid / two-point forms. */
void or_filter_coarse(int n, int nbar, const double *w1, int len1, int off1,
                      const double *w2, int len2, int off2, const double *w3, int len3,
                      int off3, const double *in, double *out) {
    const int c = n / nbar;
#pragma omp parallel for schedule(static)
    for (int K = 0; K < nbar; K++)
        for (int J = 0; J < nbar; J++)
            for (int I = 0; I < nbar; I++) {
                double s = 0.0;
                for (int t3 = 0; t3 < len3; t3++) {
                    const int kk = (int)(((long)c * (K + 1) - 1 + off3 + t3) % n + n) % n;
                    for (int t2 = 0; t2 < len2; t2++) {
                        const int jj = (int)(((long)c * (J + 1) - 1 + off2 + t2) % n + n) % n;
                        double sx = 0.0;
                        for (int t1 = 0; t1 < len1; t1++) {
                            const int ii = (int)(((long)c * (I + 1) - 1 + off1 + t1) % n + n) % n;
                            sx += w1[t1] * in[IDX(ii, jj, kk)];
                        }
                        s += w3[t3] * w2[t2] * sx;
                    }
                }
                out[(size_t)I + (size_t)nbar * ((size_t)J + (size_t)nbar * K)] = s;
            }
}
