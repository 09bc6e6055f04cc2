/*
 * Oracle FFT: mixed-radix Stockham autosort, FP64.  TEST INFRASTRUCTURE ONLY.
 *
 * The reference calls FFTW through FFTW.jl (plan_rfft / mul! / ldiv!,
 * src/N.jl:217,267,301,409,458).  FFTW is not available here, so this file
 * restates the transform it computes: X[k] = sum_x x[x] exp(sign*2*pi*i*k*x/n)
 * with sign = -1 forward (unnormalised) and the inverse normalised by 1/n^3
 * (AbstractFFTs ldiv! semantics).  The half spectrum has shape (n/2+1, n, n)
 * with kx fastest, like rfft of a Julia Array{Float64,3}.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef struct {
    int n;
    int nfac;
    int fac[40];
    double *tw; /* exp(+2 pi i t / n), t in [0, n), interleaved; sign applied at use */
} plan1d;

static int factorise(int n, int *fac) {
    int nf = 0;
    while (n % 4 == 0) { fac[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { fac[nf++] = 2; n /= 2; }
    for (int p = 3; n > 1; p += 2) {
        while (n % p == 0) { fac[nf++] = p; n /= p; }
    }
    return nf;
}

static void plan_init(plan1d *P, int n) {
    P->n = n;
    P->nfac = factorise(n, P->fac);
    P->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int t = 0; t < n; t++) {
        /* angle reduced to [0, pi/4] symmetry is unnecessary at FP64 for n <= 4096 */
        double a = 2.0 * M_PI * (double)t / (double)n;
        P->tw[2 * t] = cos(a);
        P->tw[2 * t + 1] = sin(a);
    }
}

static void plan_free(plan1d *P) { free(P->tw); }

/* One line, in place in x (length n complex), work buffer w (length n). */
static void fft_line(const plan1d *P, int sign, double *x, double *w) {
    const int n = P->n;
    double *src = x, *dst = w;
    int Ns = 1;
    for (int f = 0; f < P->nfac; f++) {
        const int R = P->fac[f];
        const int nR = n / R;
        double vr[64], vi[64];
        for (int j = 0; j < nR; j++) {
            const int jm = j % Ns;
            /* twiddle exponent: r * jm * n/(Ns*R) */
            const int step = jm * (n / (Ns * R));
            for (int r = 0; r < R; r++) {
                double a = src[2 * (j + r * nR)], b = src[2 * (j + r * nR) + 1];
                int t = (r * step) % n;
                double c = P->tw[2 * t], s = sign * P->tw[2 * t + 1];
                vr[r] = a * c - b * s;
                vi[r] = a * s + b * c;
            }
            const int idxD = (j / Ns) * Ns * R + jm;
            for (int q = 0; q < R; q++) {
                double sr = 0.0, si = 0.0;
                for (int r = 0; r < R; r++) {
                    int t = ((r * q) % R) * nR;
                    double c = P->tw[2 * t], s = sign * P->tw[2 * t + 1];
                    sr += vr[r] * c - vi[r] * s;
                    si += vr[r] * s + vi[r] * c;
                }
                dst[2 * (idxD + q * Ns)] = sr;
                dst[2 * (idxD + q * Ns) + 1] = si;
            }
        }
        double *tmp = src; src = dst; dst = tmp;
        Ns *= R;
    }
    if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)n);
}

int or_fft1d(int n, int sign, double *re_im) {
    if (n < 1) return -1;
    plan1d P;
    plan_init(&P, n);
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    fft_line(&P, sign, re_im, w);
    free(w);
    plan_free(&P);
    return 0;
}

/* Transform along y or z of a half spectrum (nh, n, n) in place. */
static void fft_strided(const plan1d *P, int sign, double *a, int nh, int n, int axis) {
    const size_t sx = 1, sy = (size_t)nh, sz = (size_t)nh * n;
    const size_t stride = axis == 1 ? sy : sz;
    const long nlines = (long)nh * n; /* lines per axis */
#pragma omp parallel
    {
        double *line = (double *)malloc(sizeof(double) * 2 * (size_t)n);
        double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for schedule(static)
        for (long L = 0; L < nlines; L++) {
            size_t base;
            if (axis == 1) { long kx = L % nh, kz = L / nh; base = kx * sx + (size_t)kz * sz; }
            else { long kx = L % nh, ky = L / nh; base = kx * sx + (size_t)ky * sy; }
            for (int t = 0; t < n; t++) {
                line[2 * t] = a[2 * (base + t * stride)];
                line[2 * t + 1] = a[2 * (base + t * stride) + 1];
            }
            fft_line(P, sign, line, w);
            for (int t = 0; t < n; t++) {
                a[2 * (base + t * stride)] = line[2 * t];
                a[2 * (base + t * stride) + 1] = line[2 * t + 1];
            }
        }
        free(line);
        free(w);
    }
}

int or_rfft3(int n, const double *in, double *out) {
    if (n < 1) return -1;
    plan1d P;
    plan_init(&P, n);
    const int nh = n / 2 + 1;
#pragma omp parallel
    {
        double *line = (double *)malloc(sizeof(double) * 2 * (size_t)n);
        double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for schedule(static)
        for (long L = 0; L < (long)n * n; L++) {
            const double *src = in + (size_t)L * n;
            for (int t = 0; t < n; t++) { line[2 * t] = src[t]; line[2 * t + 1] = 0.0; }
            fft_line(&P, -1, line, w);
            double *dst = out + 2 * (size_t)L * nh;
            memcpy(dst, line, sizeof(double) * 2 * (size_t)nh);
        }
        free(line);
        free(w);
    }
    fft_strided(&P, -1, out, nh, n, 1);
    fft_strided(&P, -1, out, nh, n, 2);
    plan_free(&P);
    return 0;
}

int or_irfft3(int n, const double *in, double *out) {
    if (n < 1) return -1;
    plan1d P;
    plan_init(&P, n);
    const int nh = n / 2 + 1;
    const size_t nspec = (size_t)nh * n * n;
    double *tmp = (double *)malloc(sizeof(double) * 2 * nspec);
    memcpy(tmp, in, sizeof(double) * 2 * nspec);
    fft_strided(&P, +1, tmp, nh, n, 2);
    fft_strided(&P, +1, tmp, nh, n, 1);
    const double scale = 1.0 / ((double)n * (double)n * (double)n);
#pragma omp parallel
    {
        double *line = (double *)malloc(sizeof(double) * 2 * (size_t)n);
        double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for schedule(static)
        for (long L = 0; L < (long)n * n; L++) {
            const double *src = tmp + 2 * (size_t)L * nh;
            /* Hermitian extension along x; the real part of the inverse then
             * equals what a c2r transform returns (imaginary parts of the
             * self-conjugate bins drop out). */
            for (int k = 0; k < n; k++) {
                if (k < nh) { line[2 * k] = src[2 * k]; line[2 * k + 1] = src[2 * k + 1]; }
                else { line[2 * k] = src[2 * (n - k)]; line[2 * k + 1] = -src[2 * (n - k) + 1]; }
            }
            fft_line(&P, +1, line, w);
            double *dst = out + (size_t)L * n;
            for (int t = 0; t < n; t++) dst[t] = line[2 * t] * scale;
        }
        free(line);
        free(w);
    }
    free(tmp);
    plan_free(&P);
    return 0;
}
