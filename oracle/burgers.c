/*
 * Oracle: 1D viscous Burgers, staggered finite volume, FP64.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Restates src/Burgers.jl.
 * u[i] sits on the right face of cell i (x = (i+1) h, 0-based), the stress
 * s[i] at the centre of cell i (src/Burgers.jl:82-94).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static inline int pw1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

/* stress(u, visc) (src/Burgers.jl:89-94) */
void or_burgers_stress(int n, double L, double visc, const double *u, double *s) {
    const double h = L / n;
    for (int i = 0; i < n; i++) {
        const double ua = u[pw1(i - 1, n)], ub = u[i];
        s[i] = -((ua + ub) * (ua + ub)) / 8 + visc * (ub - ua) / h;
    }
}

/* rhs! (src/Burgers.jl:103-109): du[i] = delta_coll(s, i) */
void or_burgers_rhs(int n, double L, double visc, const double *u, double *du) {
    const double h = L / n;
    double *s = (double *)malloc(sizeof(double) * n);
    or_burgers_stress(n, L, visc, u, s);
    for (int i = 0; i < n; i++) du[i] = (s[pw1(i + 1, n)] - s[i]) / h;
    free(s);
}

/* cfl (src/Burgers.jl:97-101) */
double or_burgers_cfl(int n, double L, double visc, const double *u) {
    const double h = L / n;
    double m = 0.0;
    for (int i = 0; i < n; i++) m = fabs(u[i]) > m ? fabs(u[i]) : m;
    const double a = h / m, b = h * h / visc;
    return a < b ? a : b;
}

/* rk4_timestep! (src/Burgers.jl:120-142), applied nsteps times */
void or_burgers_rk4(int n, double L, double visc, double dt, int nsteps, double *u) {
    double *uold = (double *)malloc(sizeof(double) * n);
    double *k1 = (double *)malloc(sizeof(double) * n), *k2 = (double *)malloc(sizeof(double) * n);
    double *k3 = (double *)malloc(sizeof(double) * n), *k4 = (double *)malloc(sizeof(double) * n);
    for (int step = 0; step < nsteps; step++) {
        memcpy(uold, u, sizeof(double) * n);
        or_burgers_rhs(n, L, visc, u, k1);
        for (int i = 0; i < n; i++) u[i] = uold[i] + dt / 2 * k1[i];
        or_burgers_rhs(n, L, visc, u, k2);
        for (int i = 0; i < n; i++) u[i] = uold[i] + dt / 2 * k2[i];
        or_burgers_rhs(n, L, visc, u, k3);
        for (int i = 0; i < n; i++) u[i] = uold[i] + dt * k3[i];
        or_burgers_rhs(n, L, visc, u, k4);
        for (int i = 0; i < n; i++)
            u[i] = uold[i] + dt / 6 * (k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i]);
    }
    free(uold); free(k1); free(k2); free(k3); free(k4);
}

/* lazycoarsegrain(lazyfilter(u, (R, w))) (src/Burgers.jl:164-180) with the
 * offset written relative to the coarse face index c*(J+1)-1. */
void or_burgers_filter_coarse(int n, int nbar, const double *w, int len, int off,
                              const double *in, double *out) {
    const int c = n / nbar;
    for (int J = 0; J < nbar; J++) {
        double s = 0.0;
        for (int t = 0; t < len; t++) {
            const int i = (int)((((long)c * (J + 1) - 1 + off + t) % n + n) % n);
            s += w[t] * in[i];
        }
        out[J] = s;
    }
}

/* Volume-average pair for any compression c (even c generalises the
 * odd-only tophat_weights of src/Burgers.jl:158-162):
 *   V  (velocity, coarse face):   odd  c: c faces, 1/c;  even c: c+1 faces,
 *                                  trapezoid 1/(2c), 1/c, ..., 1/(2c)
 *   Pc (stress, coarse centre):    odd  c: the centre cell;  even c: mean of
 *                                  the two cells straddling the centre.
 * With these, V(delta_coll s) = delta_coll_H(Pc s) holds exactly. */
static void va_weights(int c, double *wv, int *lenv, int *offv, double *wp, int *lenp,
                       int *offp) {
    if (c % 2) {
        *lenv = c; *offv = -(c / 2);
        for (int t = 0; t < c; t++) wv[t] = 1.0 / c;
        *lenp = 1; *offp = -(c / 2); wp[0] = 1.0;
    } else {
        *lenv = c + 1; *offv = -(c / 2);
        for (int t = 0; t <= c; t++) wv[t] = 1.0 / c;
        wv[0] = wv[c] = 0.5 / c;
        *lenp = 2; *offp = -(c / 2); wp[0] = wp[1] = 0.5;
    }
}

static void les_stress(int nH, double H, double visc, const double *u, double *s) {
    for (int J = 0; J < nH; J++) {
        const double ua = u[pw1(J - 1, nH)], ub = u[J];
        s[J] = -((ua + ub) * (ua + ub)) / 8 + visc * (ub - ua) / H;
    }
}

/* DNS-aided LES with classical RK4 (src/Burgers.jl:350-445 restated with the
 * rk4_timestep! tableau of src/Burgers.jl:120-142 and the closure evaluated
 * at every stage).  Outputs: u (DNS, n), u_swap (exact closure, n/c),
 * u_nomo (no model, n/c).  On entry only u is read; the LES states start
 * from V(u). */
void or_burgers_dnsaid_rk4(int n, int c, double L, double visc, double dt, int nsteps,
                           double *u, double *u_swap, double *u_nomo) {
    const int nH = n / c;
    const double h = L / n, H = L / nH;
    double wv[4096], wp[2];
    int lenv, offv, lenp, offp;
    va_weights(c, wv, &lenv, &offv, wp, &lenp, &offp);
    or_burgers_filter_coarse(n, nH, wv, lenv, offv, u, u_swap);
    memcpy(u_nomo, u_swap, sizeof(double) * nH);

    double *u0 = malloc(sizeof(double) * n), *acc = malloc(sizeof(double) * n);
    double *k = malloc(sizeof(double) * n), *s = malloc(sizeof(double) * n);
    double *w0 = malloc(sizeof(double) * nH), *wacc = malloc(sizeof(double) * nH);
    double *m0 = malloc(sizeof(double) * nH), *macc = malloc(sizeof(double) * nH);
    double *kw = malloc(sizeof(double) * nH), *km = malloc(sizeof(double) * nH);
    double *fsu = malloc(sizeof(double) * nH), *vu = malloc(sizeof(double) * nH);
    double *svu = malloc(sizeof(double) * nH), *sl = malloc(sizeof(double) * nH);
    const double a_next[4] = {0.5, 0.5, 1.0, 0.0};
    const double b[4] = {1.0, 2.0, 2.0, 1.0};
    for (int step = 0; step < nsteps; step++) {
        memcpy(u0, u, sizeof(double) * n);
        memcpy(w0, u_swap, sizeof(double) * nH);
        memcpy(m0, u_nomo, sizeof(double) * nH);
        for (int st = 0; st < 4; st++) {
            /* DNS right-hand side */
            or_burgers_stress(n, L, visc, u, s);
            for (int i = 0; i < n; i++) k[i] = (s[pw1(i + 1, n)] - s[i]) / h;
            /* closure commutator tau = Pc(s) - s_H(V u) */
            or_burgers_filter_coarse(n, nH, wp, lenp, offp, s, fsu);
            or_burgers_filter_coarse(n, nH, wv, lenv, offv, u, vu);
            les_stress(nH, H, visc, vu, svu);
            /* exact-closure LES */
            les_stress(nH, H, visc, u_swap, sl);
            for (int J = 0; J < nH; J++) sl[J] += fsu[J] - svu[J];
            for (int J = 0; J < nH; J++) kw[J] = (sl[pw1(J + 1, nH)] - sl[J]) / H;
            /* no-model LES */
            les_stress(nH, H, visc, u_nomo, sl);
            for (int J = 0; J < nH; J++) km[J] = (sl[pw1(J + 1, nH)] - sl[J]) / H;
            /* RK4 bookkeeping */
            if (st == 0) {
                for (int i = 0; i < n; i++) acc[i] = b[0] * k[i];
                for (int J = 0; J < nH; J++) { wacc[J] = b[0] * kw[J]; macc[J] = b[0] * km[J]; }
            } else {
                for (int i = 0; i < n; i++) acc[i] += b[st] * k[i];
                for (int J = 0; J < nH; J++) { wacc[J] += b[st] * kw[J]; macc[J] += b[st] * km[J]; }
            }
            if (st < 3) {
                for (int i = 0; i < n; i++) u[i] = u0[i] + a_next[st] * dt * k[i];
                for (int J = 0; J < nH; J++) {
                    u_swap[J] = w0[J] + a_next[st] * dt * kw[J];
                    u_nomo[J] = m0[J] + a_next[st] * dt * km[J];
                }
            } else {
                for (int i = 0; i < n; i++) u[i] = u0[i] + dt / 6 * acc[i];
                for (int J = 0; J < nH; J++) {
                    u_swap[J] = w0[J] + dt / 6 * wacc[J];
                    u_nomo[J] = m0[J] + dt / 6 * macc[J];
                }
            }
        }
    }
    free(u0); free(acc); free(k); free(s); free(w0); free(wacc); free(m0); free(macc);
    free(kw); free(km); free(fsu); free(vu); free(svu); free(sl);
}

/* Initial condition of the batched run: modes 1..kmax with amplitude 1/k and
 * a uniform phase per (instance, k), u = irfft(c * n) as in randomfield
 * (src/Burgers.jl:195-202), written as the explicit cosine sum. */
static inline unsigned long long sm64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

void or_burgers_ic(int n, int kmax, unsigned long long seed, double *u) {
    double ph[256];
    for (int k = 1; k <= kmax && k < 256; k++)
        ph[k] = (double)(sm64(seed ^ sm64((unsigned long long)k)) >> 11) * 0x1.0p-53;
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int k = 1; k <= kmax; k++) {
            /* angle reduced exactly in integers before the FP64 cosine */
            const long m = ((long)k * i) % n;
            s += 2.0 / k * cos(2.0 * M_PI * ((double)m / n + ph[k]));
        }
        u[i] = s;
    }
}
