/*
 * FP64 CPU oracle for the DNS-aided LES hot path of ExactClosure.jl.
 *
 * TEST INFRASTRUCTURE ONLY.  The parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark call it.  The
 * product path (libexactclosure, the HIP kernels) never links or calls it.
 *
 * Every function restates one reference function (file:line given at each
 * definition) in plain C99 + OpenMP, FP64 throughout, with the reference's
 * evaluation order where it matters.  Indexing is 0-based; the reference is
 * 1-based.  Layout is x-fastest, idx = i + n*(j + n*k), one array per vector
 * or tensor component, periodic with no ghost cells.
 */
#ifndef EC_ORACLE_H
#define EC_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FFT (FFTW/AbstractFFTs semantics: forward unnormalised, inverse
 *      normalised by 1/n^3; half spectrum (n/2+1, n, n), kx fastest) ---- */
int or_fft1d(int n, int sign, double *re_im);            /* in place, unnormalised */
int or_rfft3(int n, const double *in, double *out);      /* out: 2*(n/2+1)*n*n */
int or_irfft3(int n, const double *in, double *out);     /* normalised */

/* ---- Navier-Stokes operators (src/N.jl) ---- */
void or_stresstensor(int n, double l, double visc, const double *u1, const double *u2,
                     const double *u3, double *s11, double *s22, double *s33, double *s12,
                     double *s23, double *s31);
void or_tensordivergence(int n, double l, const double *s11, const double *s22,
                         const double *s33, const double *s12, const double *s23,
                         const double *s31, double *du1, double *du2, double *du3, int doadd);
void or_tensordivergence_nonsym(int n, double l, const double *const *s9, double *du1,
                                double *du2, double *du3, int doadd);
void or_divergence(int n, double l, const double *u1, const double *u2, const double *u3,
                   double *d);
void or_pressuregradient(int n, double l, double *u1, double *u2, double *u3, const double *p);
int or_project(int n, double l, double *u1, double *u2, double *u3, double *p);
double or_maxabs(size_t count, const double *x);
double or_sumsq(size_t count, const double *x);

/* Normal variates of the counter-based generator shared with the GPU path. */
void or_randn(size_t count, unsigned long long seed, unsigned long long offset, double *out);
/* randomfield (src/N.jl:394-476); profile 0: k^4 exp(-2 (k/kp)^2), 1: (k>0) k^kp */
int or_randomfield(int n, double l, int profile, double param, double totalenergy,
                   unsigned long long seed, double *u1, double *u2, double *u3);

/* ---- Filters (src/N.jl:592-698) ---- */
/* Dense separable kernel product applied at every fine point (applyfilter!). */
void or_applyfilter(int n, const double *w1, int r1, const double *w2, int r2,
                    const double *w3, int r3, const double *in, double *out);
/* coarsegrain!: out[J] = in[c*(J+1)-1 - (!stag_d)*(c/2)] */
void or_coarsegrain(int n, int nbar, const int *stag, const double *in, double *out);
/* Separable filter evaluated only at coarse points: out[J] = sum_t w1[t1] w2[t2] w3[t3]
 * in[c*(J+1)-1 + off_d + t_d].  Covers lazycoarsegrain(lazyfilter(.)). */
void or_filter_coarse(int n, int nbar, const double *w1, int len1, int off1,
                      const double *w2, int len2, int off2, const double *w3, int len3,
                      int off3, const double *in, double *out);

/* ---- Burgers (src/Burgers.jl) ---- */
void or_burgers_stress(int n, double L, double visc, const double *u, double *s);
void or_burgers_rhs(int n, double L, double visc, const double *u, double *du);
void or_burgers_rk4(int n, double L, double visc, double dt, int nsteps, double *u);
double or_burgers_cfl(int n, double L, double visc, const double *u);
void or_burgers_filter_coarse(int n, int nbar, const double *w, int len, int off,
                              const double *in, double *out);
void or_burgers_dnsaid_rk4(int n, int c, double L, double visc, double dt, int nsteps,
                           double *u, double *u_swap, double *u_nomo);
void or_burgers_ic(int n, int kmax, unsigned long long seed, double *u);

#ifdef __cplusplus
}
#endif
#endif
