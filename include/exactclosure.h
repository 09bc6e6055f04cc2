/*
 * libexactclosure -- MI355X (gfx950) implementation of the DNS-aided LES hot
 * path of ExactClosure.jl behind a plain C ABI.
 *
 * The reference has no FFI: its seam is Julia dispatch on the array type
 * inside Field/LazyField (src/N.jl:35-59) with every loop behind
 * myforeachindex (src/N.jl:16, src/Burgers.jl:21) and the FFT behind
 * plan_rfft/mul!/ldiv! (src/N.jl:217,267,301).  Each entry point below names
 * the reference function it replaces; a Julia `ccall` (or Python ctypes)
 * binding passes device pointers from ec_alloc.
 *
 * Conventions
 *  - Every function returns 0 (EC_OK) or a negative EC_E_* code; the message
 *    is available from ec_last_error(ctx).  No C++ exception crosses the ABI.
 *  - One context = one device + one HIP stream; calls are stream-ordered and
 *    asynchronous, except reductions and ec_d2h which synchronise.  A context
 *    must not be used from two host threads at once.
 *  - Fields: n^3 values, x fastest (idx = i + n*(j + n*k)), periodic with no
 *    ghost cells, one buffer per vector/tensor component (the layout of a
 *    Julia Array{T,3}).  dtype is EC_F32 or EC_F64.
 *  - Vector arguments (u[3], du[3], sigma[6|9]) are arrays of device pointers
 *    on the host.  Closure outputs are always FP64.
 *  - The caller owns field buffers; the library owns contexts, Poisson
 *    solvers (FFT plans, spectra) and Burgers batches.
 */
#ifndef EXACTCLOSURE_H
#define EXACTCLOSURE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EC_OK 0
#define EC_E_INVALID (-1)
#define EC_E_HIP (-2)
#define EC_E_OOM (-3)
#define EC_E_UNSUPPORTED (-4)

#define EC_F32 0
#define EC_F64 1

typedef struct ec_ctx ec_ctx;
typedef struct ec_poisson ec_poisson;
typedef struct ec_burgers_batch ec_burgers_batch;

/* ---- context (replaces defaultbackend(), src/N.jl:19, src/Burgers.jl:19) ---- */
/* stream: the hipStream_t every call of this context is ordered on (NULL: the
 * device's default stream). */
int ec_create(ec_ctx **ctx, int device, void *stream);
int ec_destroy(ec_ctx *ctx);
const char *ec_last_error(ec_ctx *ctx);
int ec_sync(ec_ctx *ctx);
int ec_version(void);

/* ---- graph capture (no reference counterpart; the reference launches every
 * kernel of a step from Julia) ----
 * Between ec_capture_begin and ec_capture_end the context's asynchronous calls
 * are recorded, not run, into one HIP graph (e.g. a whole DNS-aided RK4 step,
 * the ec_ns_dnsaid_step of the boundary); ec_graph_launch replays it on the
 * context's stream.  Kernel arguments (sizes, dt, device pointers) are fixed
 * at capture.  Synchronising calls (reductions, ec_d2h, spectrum) must not be
 * captured; per-launch profiling must be off. */
int ec_capture_begin(ec_ctx *ctx);
int ec_capture_end(ec_ctx *ctx, void **graph);
int ec_graph_launch(ec_ctx *ctx, void *graph);
int ec_graph_destroy(ec_ctx *ctx, void *graph);
/* Stream order between two contexts of one device (no reference counterpart):
 * the calls made on ctx after this one run after everything queued on other
 * so far (an event recorded on other's stream, waited on by ctx's).  Inside a
 * capture of either context the other stream joins the same graph as a
 * parallel branch; every branch must be joined back into the capturing
 * context's stream this way before ec_capture_end. */
int ec_stream_after(ec_ctx *ctx, ec_ctx *other);

/* ---- memory (replaces KernelAbstractions.zeros/allocate and Adapt moves) ---- */
int ec_alloc(ec_ctx *ctx, size_t bytes, void **ptr);
int ec_free(ec_ctx *ctx, void *ptr);
int ec_h2d(ec_ctx *ctx, void *dst, const void *src, size_t bytes);
int ec_d2h(ec_ctx *ctx, void *dst, const void *src, size_t bytes);
int ec_d2d(ec_ctx *ctx, void *dst, const void *src, size_t bytes);
int ec_zero(ec_ctx *ctx, void *dst, size_t bytes);
/* element-wise y = a*x + y (LinearAlgebra.axpy!, src/N.jl:337,1161,1322) */
int ec_axpy(ec_ctx *ctx, int dtype, size_t count, double a, const void *x, void *y);
/* y = a*x + y with x and y of different dtypes (FP64 closure terms into FP32 fields) */
int ec_axpy2(ec_ctx *ctx, int ydtype, size_t count, double a, int xdtype, const void *x, void *y);
/* element-wise y = a*x + b*y in place (the closure forcing m = V k - F_H(P_H v)
 * of src/NavierStokes.jl's DNS-aided step as one pass over x and y) */
int ec_axpby(ec_ctx *ctx, int dtype, size_t count, double a, const void *x, double b, void *y);
/* dtype conversion copy (FP64 <-> FP32) */
int ec_convert(ec_ctx *ctx, int dst_dtype, void *dst, int src_dtype, const void *src, size_t count);

/* ---- per-launch timing with HIP events on the context stream ---- */
int ec_profile_begin(ec_ctx *ctx);
/* Writes up to cap records: name (static string), milliseconds, algorithmic bytes. */
int ec_profile_end(ec_ctx *ctx, const char **names, double *ms, double *bytes, int cap, int *count);

/* ---- Navier-Stokes operators (src/N.jl) ---- */
/* stresstensor (src/N.jl:94-118) materialised: sigma[6] = 11,22,33,12,23,31 */
int ec_ns_stresstensor(ec_ctx *ctx, int dtype, int n, double l, double visc,
                       void *const u[3], void *const sigma[6]);
/* tensordivergence! (src/N.jl:120-160) */
int ec_ns_tensordivergence(ec_ctx *ctx, int dtype, int n, double l, void *const sigma[6],
                           void *const du[3], int doadd);
/* tensordivergence_nonsym! (src/N.jl:162-200); sigma9 order 11,21,31,12,22,32,13,23,33.
 * sigma_dtype may differ from dtype (FP64 closure fluxes into FP32 fields). */
int ec_ns_tensordivergence_nonsym(ec_ctx *ctx, int dtype, int n, double l, int sigma_dtype,
                                  void *const sigma9[9], void *const du[3], int doadd);
/* stresstensor + tensordivergence!(du, sigma, doadd) fused: du = doadd*du - div sigma(u)
 * (src/N.jl:310-311, 1306-1307) */
int ec_ns_rhs(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3],
              void *const du[3], int doadd);
/* No counterpart in the reference (it forms the RHS and then adds a forcing in
 * a second broadcast): du = s * (-div sigma(u)) + add in one pass, add three
 * FP64 fields, the sum formed in FP64 and rounded once to dtype.  Used for the
 * exact forcing m = V k - F_H(P_H v) (s = -1, add = V k) and for the LES
 * right-hand side F_H(u) + m (s = 1, add = m). */
int ec_ns_rhs_add(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3],
                  void *const du[3], double s, const void *const add[3]);
/* ec_ns_rhs (doadd 0) and divergence!(d, du) fused: du = -div sigma(u) and
 * d = div_h du + rdiv * div_h u, the right-hand side of the pressure Poisson
 * equation of project!(du) (src/N.jl:256-262).  d is typically
 * ec_poisson_pressure(ps), followed by ec_poisson_solve.  With rdiv = 0 this
 * is the reference's projection of du; rdiv = 1/tau makes k = du - grad p
 * also remove the divergence of the stage state u at rate 1/tau (the
 * divergence that FP32 rounding of the projected k leaves behind and that
 * would otherwise accumulate step after step). */
int ec_ns_rhs_div(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3],
                  void *const du[3], void *d, double rdiv);
/* divergence! (src/N.jl:222-238) */
int ec_ns_divergence(ec_ctx *ctx, int dtype, int n, double l, void *const u[3], void *d);
/* pressuregradient! (src/N.jl:240-254) */
int ec_ns_pressuregradient(ec_ctx *ctx, int dtype, int n, double l, void *const u[3],
                           const void *p);
/* out = u - grad p with u unchanged (the gradient step of ec_ns_project_into) */
int ec_ns_pressuregradient_into(ec_ctx *ctx, int dtype, int n, double l, const void *const u[3], const void *p,
                                void *const out[3]);

/* ---- FFT Poisson projection (poissonsolver / project!, src/N.jl:202-220,256-307) ---- */
int ec_poisson_create(ec_ctx *ctx, int dtype, int n, double l, ec_poisson **ps);
int ec_poisson_destroy(ec_ctx *ctx, ec_poisson *ps);
/* The solver's pressure buffer (poisson.p), valid until destroy. */
void *ec_poisson_pressure(ec_poisson *ps);
/* project!(u, poisson): leaves the pressure in ec_poisson_pressure(ps) */
int ec_ns_project(ec_ctx *ctx, ec_poisson *ps, void *const u[3]);
/* dst = P src with src unchanged (project! of a copy, without the copy); no
 * counterpart entry in the reference, which projects in place */
int ec_ns_project_into(ec_ctx *ctx, ec_poisson *ps, const void *const src[3], void *const dst[3]);
/* The first half of project!: divergence! and the FFT Poisson solve, leaving p in
 * ec_poisson_pressure(ps) and u unchanged (the gradient is then applied by
 * ec_ns_pressuregradient or fused into ec_ns_rk4_update). */
int ec_ns_pressure_solve(ec_ctx *ctx, ec_poisson *ps, void *const u[3]);
/* The FFT half of the pressure solve of project! (src/N.jl:263-303): the
 * pressure buffer ec_poisson_pressure(ps) holds div u on entry and p on exit. */
int ec_poisson_solve(ec_ctx *ctx, ec_poisson *ps);
/* plan_rfft * x (unnormalised) and ldiv! (normalised by 1/n^3); spectrum (n/2+1, n, n) */
int ec_fft_r2c(ec_ctx *ctx, ec_poisson *ps, const void *in, void *out);
int ec_fft_c2r(ec_ctx *ctx, ec_poisson *ps, const void *in, void *out);

/* ---- reductions (maximum(abs, .), sum(abs2, .), dot; src/N.jl:352,1346).  FP64 result,
 *      synchronising. ---- */
int ec_reduce_maxabs(ec_ctx *ctx, int dtype, size_t count, const void *x, double *out);
int ec_reduce_sumsq(ec_ctx *ctx, int dtype, size_t count, const void *x, double *out);
int ec_reduce_diffsq(ec_ctx *ctx, int dtype, size_t count, const void *x, const void *y,
                     double *out);

/* ---- random initial field (randomfield, src/N.jl:394-476) ----
 * profile 0: k^4 exp(-2 (k/param)^2) (peak_profile, src/N.jl:366);
 * profile 1: (k > 0) k^param (src/N.jl:1245).  Counter-based normal variates. */
int ec_ns_randomfield(ec_ctx *ctx, ec_poisson *ps, int profile, double param,
                      double totalenergy, uint64_t seed, void *const u[3]);
/* spectrum(u, stuff, poisson) (src/N.jl:576-590) over the shells of getshells
 * (src/N.jl:478-545): out[s] = sum over s^2 <= |k|^2 < (s+1)^2 of
 * |u^(k) / n^3|^2 / 2 summed over the components (the conjugate half of the
 * rfft counted), s in [0, nshell), 1 <= nshell <= 8192.  The host mirror picks
 * the shells 1..n/3 of spectral_stuff.  Uses (overwrites) the plan's spectrum;
 * synchronises the context's stream. */
int ec_ns_spectrum(ec_ctx *ctx, ec_poisson *ps, void *const u[3], int nshell, double *out);
int ec_randn(ec_ctx *ctx, int dtype, size_t count, uint64_t seed, uint64_t offset, void *out);

/* ---- separable filters and closure (src/N.jl:592-698, 999-1135) ----
 * A quantity is one source evaluated on the fine grid, filtered with three 1D
 * weight vectors and sampled at the coarse points:
 *   out[J] = sum_t w1[t1] w2[t2] w3[t3] src[c*(J_d+1)-1 + off_d + t_d]
 * Sources: EC_SRC_U1..U3 (velocity), EC_SRC_S11..S31 (stress sigma(u, visc)),
 * EC_SRC_R11..R33 (sigma_aa + p), EC_SRC_P (pressure). */
#define EC_SRC_U1 0
#define EC_SRC_U2 1
#define EC_SRC_U3 2
#define EC_SRC_S11 3
#define EC_SRC_S22 4
#define EC_SRC_S33 5
#define EC_SRC_S12 6
#define EC_SRC_S23 7
#define EC_SRC_S31 8
#define EC_SRC_R11 9
#define EC_SRC_R22 10
#define EC_SRC_R33 11
#define EC_SRC_P 12
#define EC_SRC_K1 13 /* u_a - delta_a p: the projected tendency when u = du */
#define EC_SRC_K2 14
#define EC_SRC_K3 15

typedef struct {
    int source;
    int len[3];
    int off[3];
    const double *w[3]; /* host pointers, copied at call time */
    void *out;          /* FP64 device array, nbar^3 */
} ec_filter_quantity;

/* lazycoarsegrain(lazyfilter(.)) of many quantities (get_rDHi :1038,
 * get_sigDH :1088, apply_DHpi :1228).  Narrow kernels (at most 64 taps in
 * all) in one dense launch; wider ones (Gaussian-composed) as three
 * separable passes per quantity in a context-owned FP64 workspace of
 * nbar n^2 + nbar^2 n doubles. */
int ec_filter_coarse(ec_ctx *ctx, int dtype, int n, int nbar, double l, double visc,
                     void *const u[3], const void *p, int nq, const ec_filter_quantity *q);
/* No counterpart in the reference: ec_filter_coarse for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid hipMalloc).  It calls no hipMalloc/hipFree and no allocating
 * library routine: every quantity takes the separable passes in a fixed
 * per-GPU scratch array of 4 Mi doubles (32 MiB, held until the process
 * ends), with the weights passed to the device as kernel arguments.  Limits:
 * at most 64 taps per axis and nbar n^2 + nbar^2 n + (taps) <= 4 Mi doubles
 * per quantity (e.g. n = 256 with nbar <= 32); larger requests are refused
 * with EC_E_INVALID before any launch.  Results equal ec_filter_coarse's
 * separable path bit for bit.  All calls on one GPU share the scratch: calls
 * that overlap on one GPU (several streams or threads) race. */
int ec_filter_coarse_noalloc(ec_ctx *ctx, int dtype, int n, int nbar, double l, double visc,
                             void *const u[3], const void *p, int nq, const ec_filter_quantity *q);
/* applyfilter!(out, in, kernelproduct(w1, w2, w3)) (src/N.jl:627-653) with
 * centred odd-length 1D kernels (host arrays), as three periodic 1D passes,
 * FP64 sums; in and out are n^3 fields of dtype (not in place). */
int ec_filter_apply(ec_ctx *ctx, int dtype, int n, const int len[3], const double *const w[3], const void *in,
                    void *out);

/* Volume-average (kind 0) or face-average (kind 1) closure of one or more LES
 * grids from one DNS state, in FP64:
 *   ubar[s][3]  filtered velocity V u (not yet projected)
 *   flux[s][9]  S r with r = sigma(u) + p I, nonsym order 11,21,31,12,22,32,13,23,33
 * Any compression c = n / nbar (odd: top-hat; even: trapezoid / two-point). */
typedef struct {
    int nbar;
    int kind;
    void *ubar[3];
    void *flux[9];
} ec_closure_set;
int ec_ns_closure(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3],
                  const void *p, int nsets, const ec_closure_set *sets);
/* Velocity-type filter of one or more LES grids in one launch:
 *   out[s][a] = V_s(v_a - delta_a p)   (p may be NULL: V_s v_a)
 * V_s is the LES velocity filter of set s (kind 0 volume average, 1 face
 * average; apply_DHpi, src/N.jl:1228-1237, for any compression).  With v the
 * DNS right-hand side and p its pressure, V(k) equals -div_H^ns (S r) of
 * get_rDHi (src/N.jl:1038-1073), the exact closure forcing in divergence form.
 * FP64 outputs. */
typedef struct {
    int nbar;
    int kind;
    void *out[3];
} ec_vel_set;
int ec_ns_filter_vel(ec_ctx *ctx, int dtype, int n, double l, void *const v[3], const void *p,
                     int nsets, const ec_vel_set *sets);
/* ec_ns_filter_vel of du - grad p followed by ec_ns_rk4_update(stage, dt, u0, us,
 * du, p, acc) of the same DNS fields, in one pass over du and p where the set
 * list allows (the volume-16/8/4 + face-4 list); otherwise the two calls.  No
 * counterpart entry in the reference (its step filters and updates separately,
 * src/NavierStokes.jl's DNS-aided stepper). */
int ec_ns_filter_vel_rk4(ec_ctx *ctx, int dtype, int n, double l, void *const du[3], const void *p,
                         int nsets, const ec_vel_set *sets, int stage, double dt, void *const u0[3],
                         void *const us[3], void *const acc[3]);
/* ec_ns_filter_vel of du - grad p followed by ec_ns_rk4_update_states(stage,
 * dt, u0, ust, du, p), fused like ec_ns_filter_vel_rk4.  No counterpart entry
 * in the reference. */
int ec_ns_filter_vel_rk4_states(ec_ctx *ctx, int dtype, int n, double l, void *const du[3],
                                const void *p, int nsets, const ec_vel_set *sets, int stage, double dt,
                                void *const u0[3], void *const ust[9]);
/* Plain column passes (y) of FP32 spectra with n >= 512: persistent blocks,
 * as many as fit on the device, each prefetching its next column tile while
 * it transforms the current one (on = 1, the default), or one block per tile
 * (on = 0).  Both give the same bits.  A launch-shape setting of this context
 * with no counterpart in the reference (the internals of plan_rfft,
 * src/N.jl:217). */
int ec_fft_set_cols_persistent(ec_ctx *ctx, int on);
/* The z pass of FP32 Poisson solves (poissonsolver, src/N.jl:202-220; n a
 * power of two >= 512): on = 1 (the default) solves each (kx, ky) column's
 * cyclic tridiagonal system (the discrete d^2/dz^2 plus the column's x/y
 * eigenvalues) directly, by two first-order recursions in FP64; on = 0 runs the
 * forward z transform, the division by the eigenvalues and the inverse
 * transform.  The same solution up to rounding (the direct solve is the more
 * accurate: FP64 arithmetic); a setting of this context. */
int ec_poisson_set_z_tridiagonal(ec_ctx *ctx, int on);
/* Target block size (x = y, z) of ec_ns_filter_vel and ec_ns_filter_vel_rk4 on
 * this context (default z 256); set by the tile autotuner. */
int ec_filter_vel_set_tiles(ec_ctx *ctx, int tx, int tz);
/* Target block size (fine cells x, y, z) of the one-launch closure kernel on
 * this context (default 16, 16, 64); the library picks the largest multiples
 * of every compression not above them.  Set by the tile autotuner; no
 * counterpart in the reference. */
int ec_closure_set_tiles(ec_ctx *ctx, int tx, int ty, int tz);
/* tau = flux - sigma_H(ubar) in FP64 on the LES grid (the closure commutator) */
int ec_ns_commutator(ec_ctx *ctx, int nbar, double l, double visc, void *const ubar[3],
                     void *const flux9[9], void *const tau9[9]);

/* ---- sub-filter stresses (src/NavierStokes.jl:470-593, process.jl:107-193) ----
 * The DNS-side filters of sfs_tensors_volume (kind 0) and sfs_tensors_surface
 * (kind 1) in one launch, r = sigma(u) + p I with the DNS pressure p, FP64
 * outputs on the LES grid, tensors in nonsym order 11,21,31,12,22,32,13,23,33:
 *   kind 0: fu = volumefilter!(u), fr = volumefilter!(r),
 *           sr = surfacefilter!(r, false) (faces normal to the derivative
 *           direction: the exact closure flux of the volume average);
 *   kind 1: fu = surfacefilter!(u), fr = surfacefilter!(r, true) (faces normal
 *           to the row direction), sr = the diagonal of fr and linefilter!(r)
 *           off the diagonal (the exact closure flux of the face average).
 * The filters are Turbulox's (not in the reference tree), re-derived for any
 * compression: odd c top-hat windows, even c trapezoid windows at faces and
 * two-point means at centres.  A null output component is skipped. */
int ec_ns_sfs_filter(ec_ctx *ctx, int dtype, int n, int nbar, double l, double visc, void *const u[3],
                     const void *p, int kind, void *const fu[3], void *const fr9[9], void *const sr9[9]);
/* Turbulox.tensordissipation_staggered!(d, sigma, u) (process.jl:160-180):
 * d = sigma : grad u at the LES cell centres, FP64; sigma in nonsym order,
 * off-diagonal products averaged from the four edges around each centre so
 * that the sum of d equals the sum of u . tensordivergence_nonsym(sigma). */
int ec_ns_tensordissipation(ec_ctx *ctx, int nbar, double l, void *const sigma9[9], void *const u[3], void *d);

/* ---- RK4 stage update fused with the pressure gradient ----
 * k = du - grad p.  stage 0..2: acc (= or +=) b_s k; us = u0 + a_{s+1} dt k.
 * stage 3: u0 += dt/6 (acc + k).  (rk4_timestep! tableau, src/Burgers.jl:120-142,
 * with the projection of step_wray3!, src/N.jl:320-346, at every stage.) */
int ec_ns_rk4_update(ec_ctx *ctx, int dtype, int n, double l, int stage, double dt,
                     void *const u0[3], void *const us[3], void *const du[3], const void *p,
                     void *const acc[3]);
/* The same RK4 step in stored stage states (no accumulator): k = du - grad p;
 * stage s = 0..2 writes its state u_{s+1} = u0 + c_s dt k (c = 1/2, 1/2, 1)
 * to ust[3s..3s+2]; stage 3 reads the three states and writes
 * u0 = u0 + ((u1 - u0) + 2 (u2 - u0) + (u3 - u0)) / 3 + dt/6 k, which is
 * u0 + dt/6 (k1 + 2 k2 + 2 k3 + k4) of rk4_timestep! (src/Burgers.jl:120-142;
 * projection at every stage as in step_wray3!, src/N.jl:320-346).  The states
 * u1..u3 equal the stage states us of ec_ns_rk4_update bit for bit.  Per step
 * the update moves 49 fields instead of 58 (no accumulator read and written
 * at every stage).  p may be NULL: k = du (a tendency already projected). */
int ec_ns_rk4_update_states(ec_ctx *ctx, int dtype, int n, double l, int stage, double dt,
                            void *const u0[3], void *const ust[9], void *const du[3], const void *p);

/* ---- batched 1D Burgers DNS-aided LES (src/Burgers.jl:89-142,350-445) ----
 * count independent instances, instance i: n[i] points, compression c[i],
 * state FP32 or FP64, exact-closure commutator in FP64.  One persistent
 * launch advances every instance nsteps RK4 steps (optionally replayed from a
 * captured hipGraph). */
int ec_burgers_batch_create(ec_ctx *ctx, int dtype, int count, const int *n, const int *c,
                            double L, double visc, double cfl, int kmax, uint64_t seed,
                            ec_burgers_batch **batch);
int ec_burgers_batch_destroy(ec_ctx *ctx, ec_burgers_batch *batch);
int ec_burgers_batch_set_threads(ec_ctx *ctx, ec_burgers_batch *batch, int threads);
int ec_burgers_batch_run(ec_ctx *ctx, ec_burgers_batch *batch, int nsteps, int use_graph);
/* which: 0 DNS, 1 exact-closure LES, 2 no-model LES; FP64 host output */
int ec_burgers_batch_get(ec_ctx *ctx, ec_burgers_batch *batch, int instance, int which,
                         double *out, int cap);
/* The time step of an instance (cfl, src/Burgers.jl:103-109, at its initial
 * state); read on the context stream after the work queued before it. */
int ec_burgers_batch_dt(ec_ctx *ctx, ec_burgers_batch *batch, int instance, double *dt);
/* randomfield with bump_profile (src/Burgers.jl:192-202) for nsamples samples
 * (sample q seeded by seed0 + q, counter-based phases): u (device, nsamples x
 * n FP64) = sum_k 2 (k/kp)^4 exp(-2 (k/kp)^2) cos(2 pi (k m / n + phi_k)),
 * the irfft of the reference's coefficients, scaled to energy totalenergy
 * (sum u^2 / 2 * L/n); modes past 6 kpeak (below 1e-27 of the peak) are
 * omitted. */
int ec_burgers_randomfield(ec_ctx *ctx, int n, double L, double kpeak, double totalenergy, int nsamples,
                           uint64_t seed0, void *u);
/* dns_aided_les (src/Burgers.jl:350-445) for nsamples DNS initial fields
 * (ustart, device, nsamples x nh FP64), one persistent block per sample:
 * forward Euler with dt = cfl_factor min(h / max|u|, h^2 / visc) cut at tstop,
 * the DNS and the nomodel, class_m, class_p and swapfil LES with the LES
 * kernel (Rl, wl) and the double kernel (K, wd) (host arrays).  out (device,
 * per sample 2 nh + 5 nH doubles): dns_ref (nh), dns_ref_fil = D * u (nh),
 * dns_fil (nH), nomodel, class_m, class_p, swapfil (nH each).  steps
 * (host, nullable): the number of time steps of each sample.  nh <= 16384.
 * Returns when the results are written (allocates a work area for the call). */
int ec_burgers_dns_aided_les(ec_ctx *ctx, int nh, int nH, double L, double visc, double tstop, double cfl_factor,
                             int nsamples, const void *ustart, int Rl, const double *wl, int K, const double *wd,
                             void *out, int *steps);
/* Single-instance operators for parity tests (src/Burgers.jl:89-142). */
int ec_burgers_rhs(ec_ctx *ctx, int dtype, int n, double L, double visc, const void *u, void *du);
/* compute_fractions (src/Burgers.jl:1022-1099) over nsamples FP64 DNS
 * snapshots u (n points each, stored one after another, device memory): the
 * mean over coarse points and samples of |x| / (|classic| + |flux| + |div| +
 * eps) for x = classic (vsu - sVU), flux (sVU - svu), div (fsu - vsu), with the
 * LES kernel (Rl, wl[2 Rl + 1]) and the double kernel (K, wd[2 K + 1]) given
 * as host arrays (tophat/gaussian_weights and build_doublekernel); nH divides
 * n.  out = {classic, flux, div}.  FP64; synchronises the context's stream. */
int ec_burgers_fractions(ec_ctx *ctx, int n, int nH, double L, double visc, int nsamples,
                         const void *u, int Rl, const double *wl, int K, const double *wd,
                         double out[3]);
/* SFS dissipation coefficients (compute_dissipation, src/Burgers.jl:447-516):
 * per coarse point and sample, sigma * (vu_J - vu_{J-1}) / H for the three
 * sub-filter stresses class_m = vsu - s_h(D*u), class_p = vsu - s_H(vu) and
 * swapfil = fsu - s_H(vu), FP64.  out (device, 3 nH nsamples doubles) holds
 * out[(m nsamples + q) nH + J], m = 0 class_m, 1 class_p, 2 swapfil: each
 * model's block is the reference's vec(d[:, m, :]).  Kernels as for
 * ec_burgers_fractions; returns when out is written. */
int ec_burgers_dissipation(ec_ctx *ctx, int n, int nH, double L, double visc, int nsamples,
                           const void *u, int Rl, const double *wl, int K, const double *wd,
                           void *out);
/* Smagorinsky coefficient fit (smagorinsky_fields, src/Burgers.jl:547-680, and
 * smagorinsky.jl:24-36) over nsamples FP64 DNS snapshots u (device): out =
 * {<S_h,T_h>, <S_h,S_h>, <S_H,T_H>, <S_H,S_H>} summed over points and samples
 * for the classic (fine-grid) and discretisation-consistent (coarse-stencil)
 * sub-filter stress T and Smagorinsky stress S = -(delta^2 + H^2)|dU| dU;
 * theta^2 = -<S,T>/<S,S>.  Kernels as for ec_burgers_fractions. */
int ec_burgers_smag_fit(ec_ctx *ctx, int n, int nH, double L, double visc, double delta, int nsamples,
                        const void *u, int Rl, const double *wl, int K, const double *wd, double out[4]);
/* solve_smagorinsky (src/Burgers.jl:682-772): per sample, ubar = C_f(D * u)
 * on the coarse grid advanced to tstop by the reference's RK4 (stage states
 * accumulate on the running state) with dt = min(cfl min(H/max|ubar|,
 * H^2/visc), tstop - t), closure coefficient C = theta^2 (delta^2 + H^2);
 * ubar_out (device, nsamples x nH, FP64).  Asynchronous; nH doubles of LDS
 * per sample (nH up to 4096). */
int ec_burgers_smag_solve(ec_ctx *ctx, int n, int nH, double L, double visc, double C, double tstop,
                          double cfl, int nsamples, const void *ustart, int K, const double *wd,
                          void *ubar_out);

/* ---- learned closure: 3D CNN 3 -> 32 -> 32 -> 32 -> 3, 3x3x3 kernels, periodic
 * padding, GELU; BF16 weights and activations, FP32 accumulation (MFMA), Adam.
 * No reference counterpart (the reference has no learned closure); the
 * network maps the filtered LES velocity to the exact closure forcing.
 * Parameters (FP32 master copy) are one flat array, layer by layer:
 * W[27][cout][cin] (tap t = 9 (dz+1) + 3 (dy+1) + (dx+1), cross-correlation
 * y[p] = sum_t W[t] x[p + (dx, dy, dz)]) then b[cout]. ---- */
typedef struct ec_cnn ec_cnn;
int ec_cnn_create(ec_ctx *ctx, uint64_t seed, ec_cnn **net);
int ec_cnn_destroy(ec_ctx *ctx, ec_cnn *net);
int ec_cnn_num_params(ec_cnn *net);
/* which 0: parameters, 1: gradients of the last backward; get 1 reads, 0 writes */
int ec_cnn_params(ec_ctx *ctx, ec_cnn *net, int which, int get, float *host, int count);
/* u: 3*batch FP64 n^3 device fields (sample s, component a at u[3s+a]), scaled by
 * in_scale; out (nullable): 3*batch FP64 fields receiving prediction / out_scale.
 * Keeps the activations for ec_cnn_backward.  n a multiple of 16. */
int ec_cnn_forward(ec_ctx *ctx, ec_cnn *net, int n, int batch, const void *const *u, double in_scale,
                   void *const *out, double out_scale);
/* Mean squared error against target * tgt_scale (3*batch FP64 fields) of the
 * last forward; overwrites the gradients; loss (nullable, synchronising). */
int ec_cnn_backward(ec_ctx *ctx, ec_cnn *net, const void *const *target, double tgt_scale, double *loss);
int ec_cnn_adam(ec_ctx *ctx, ec_cnn *net, double lr, double beta1, double beta2, double eps);

#ifdef __cplusplus
}
#endif
#endif
