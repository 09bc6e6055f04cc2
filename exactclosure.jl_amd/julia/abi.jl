# The ccall layer of libexactclosure: one Julia function per C entry point of
# include/exactclosure.h, same name, C parameter names, argument types as in
# the header.  Written by gen_abi.py from the header and _abi.py; edit those,
# not this file.  Included by ExactClosureAMD.jl, which defines
# `libexactclosure` (the path of the shared library).

# ec_filter_quantity
struct FilterQuantity
    source::Cint
    len::NTuple{3,Cint}
    off::NTuple{3,Cint}
    w::NTuple{3,Ptr{Cdouble}}
    out::Ptr{Cvoid}
end

# ec_vel_set
struct VelSet
    nbar::Cint
    kind::Cint
    out::NTuple{3,Ptr{Cvoid}}
end

# ec_closure_set
struct ClosureSet
    nbar::Cint
    kind::Cint
    ubar::NTuple{3,Ptr{Cvoid}}
    flux::NTuple{9,Ptr{Cvoid}}
end

ec_create(ctx, device, stream) =
    ccall((:ec_create, libexactclosure), Cint, (Ptr{Ptr{Cvoid}}, Cint, Ptr{Cvoid}), ctx, device, stream)
ec_destroy(ctx) =
    ccall((:ec_destroy, libexactclosure), Cint, (Ptr{Cvoid},), ctx)
ec_last_error(ctx) =
    ccall((:ec_last_error, libexactclosure), Cstring, (Ptr{Cvoid},), ctx)
ec_sync(ctx) =
    ccall((:ec_sync, libexactclosure), Cint, (Ptr{Cvoid},), ctx)
ec_version() =
    ccall((:ec_version, libexactclosure), Cint, ())
ec_capture_begin(ctx) =
    ccall((:ec_capture_begin, libexactclosure), Cint, (Ptr{Cvoid},), ctx)
ec_capture_end(ctx, graph) =
    ccall((:ec_capture_end, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}), ctx, graph)
ec_graph_launch(ctx, graph) =
    ccall((:ec_graph_launch, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx, graph)
ec_graph_destroy(ctx, graph) =
    ccall((:ec_graph_destroy, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx, graph)
ec_stream_after(ctx, other) =
    ccall((:ec_stream_after, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx, other)
ec_alloc(ctx, bytes, ptr) =
    ccall((:ec_alloc, libexactclosure), Cint, (Ptr{Cvoid}, Csize_t, Ptr{Ptr{Cvoid}}), ctx, bytes, ptr)
ec_free(ctx, ptr) =
    ccall((:ec_free, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx, ptr)
ec_h2d(ctx, dst, src, bytes) =
    ccall((:ec_h2d, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), ctx, dst, src, bytes)
ec_d2h(ctx, dst, src, bytes) =
    ccall((:ec_d2h, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), ctx, dst, src, bytes)
ec_d2d(ctx, dst, src, bytes) =
    ccall((:ec_d2d, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), ctx, dst, src, bytes)
ec_zero(ctx, dst, bytes) =
    ccall((:ec_zero, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), ctx, dst, bytes)
ec_axpy(ctx, dtype, count, a, x, y) =
    ccall((:ec_axpy, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Csize_t, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}), ctx, dtype, count, a, x, y)
ec_axpy2(ctx, ydtype, count, a, xdtype, x, y) =
    ccall((:ec_axpy2, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Csize_t, Cdouble, Cint, Ptr{Cvoid}, Ptr{Cvoid}), ctx, ydtype, count, a, xdtype, x, y)
ec_axpby(ctx, dtype, count, a, x, b, y) =
    ccall((:ec_axpby, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Csize_t, Cdouble, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}), ctx, dtype, count, a, x, b, y)
ec_convert(ctx, dst_dtype, dst, src_dtype, src, count) =
    ccall((:ec_convert, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Csize_t), ctx, dst_dtype, dst, src_dtype, src, count)
ec_profile_begin(ctx) =
    ccall((:ec_profile_begin, libexactclosure), Cint, (Ptr{Cvoid},), ctx)
ec_profile_end(ctx, names, ms, bytes, cap, count) =
    ccall((:ec_profile_end, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cstring}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Ptr{Cint}), ctx, names, ms, bytes, cap, count)
ec_ns_stresstensor(ctx, dtype, n, l, visc, u, sigma) =
    ccall((:ec_ns_stresstensor, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{6,Ptr{Cvoid}}}), ctx, dtype, n, l, visc, u, sigma)
ec_ns_tensordivergence(ctx, dtype, n, l, sigma, du, doadd) =
    ccall((:ec_ns_tensordivergence, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Ref{NTuple{6,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}, Cint), ctx, dtype, n, l, sigma, du, doadd)
ec_ns_tensordivergence_nonsym(ctx, dtype, n, l, sigma_dtype, sigma9, du, doadd) =
    ccall((:ec_ns_tensordivergence_nonsym, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cint, Ref{NTuple{9,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}, Cint), ctx, dtype, n, l, sigma_dtype, sigma9, du, doadd)
ec_ns_rhs(ctx, dtype, n, l, visc, u, du, doadd) =
    ccall((:ec_ns_rhs, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}, Cint), ctx, dtype, n, l, visc, u, du, doadd)
ec_ns_rhs_add(ctx, dtype, n, l, visc, u, du, s, add) =
    ccall((:ec_ns_rhs_add, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}), ctx, dtype, n, l, visc, u, du, s, add)
ec_ns_rhs_div(ctx, dtype, n, l, visc, u, du, d, rdiv) =
    ccall((:ec_ns_rhs_div, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}, Cdouble), ctx, dtype, n, l, visc, u, du, d, rdiv)
ec_ns_divergence(ctx, dtype, n, l, u, d) =
    ccall((:ec_ns_divergence, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}), ctx, dtype, n, l, u, d)
ec_ns_pressuregradient(ctx, dtype, n, l, u, p) =
    ccall((:ec_ns_pressuregradient, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}), ctx, dtype, n, l, u, p)
ec_ns_pressuregradient_into(ctx, dtype, n, l, u, p, out) =
    ccall((:ec_ns_pressuregradient_into, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}, Ref{NTuple{3,Ptr{Cvoid}}}), ctx, dtype, n, l, u, p, out)
ec_poisson_create(ctx, dtype, n, l, ps) =
    ccall((:ec_poisson_create, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Ptr{Ptr{Cvoid}}), ctx, dtype, n, l, ps)
ec_poisson_destroy(ctx, ps) =
    ccall((:ec_poisson_destroy, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx, ps)
ec_poisson_pressure(ps) =
    ccall((:ec_poisson_pressure, libexactclosure), Ptr{Cvoid}, (Ptr{Cvoid},), ps)
ec_ns_project(ctx, ps, u) =
    ccall((:ec_ns_project, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{NTuple{3,Ptr{Cvoid}}}), ctx, ps, u)
ec_ns_project_into(ctx, ps, src, dst) =
    ccall((:ec_ns_project_into, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}), ctx, ps, src, dst)
ec_ns_pressure_solve(ctx, ps, u) =
    ccall((:ec_ns_pressure_solve, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{NTuple{3,Ptr{Cvoid}}}), ctx, ps, u)
ec_poisson_solve(ctx, ps) =
    ccall((:ec_poisson_solve, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx, ps)
ec_fft_r2c(ctx, ps, in_, out) =
    ccall((:ec_fft_r2c, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), ctx, ps, in_, out)
ec_fft_c2r(ctx, ps, in_, out) =
    ccall((:ec_fft_c2r, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), ctx, ps, in_, out)
ec_reduce_maxabs(ctx, dtype, count, x, out) =
    ccall((:ec_reduce_maxabs, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Csize_t, Ptr{Cvoid}, Ptr{Cdouble}), ctx, dtype, count, x, out)
ec_reduce_sumsq(ctx, dtype, count, x, out) =
    ccall((:ec_reduce_sumsq, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Csize_t, Ptr{Cvoid}, Ptr{Cdouble}), ctx, dtype, count, x, out)
ec_reduce_diffsq(ctx, dtype, count, x, y, out) =
    ccall((:ec_reduce_diffsq, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Csize_t, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}), ctx, dtype, count, x, y, out)
ec_ns_randomfield(ctx, ps, profile, param, totalenergy, seed, u) =
    ccall((:ec_ns_randomfield, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Cdouble, UInt64, Ref{NTuple{3,Ptr{Cvoid}}}), ctx, ps, profile, param, totalenergy, seed, u)
ec_ns_spectrum(ctx, ps, u, nshell, out) =
    ccall((:ec_ns_spectrum, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{NTuple{3,Ptr{Cvoid}}}, Cint, Ptr{Cdouble}), ctx, ps, u, nshell, out)
ec_randn(ctx, dtype, count, seed, offset, out) =
    ccall((:ec_randn, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Csize_t, UInt64, UInt64, Ptr{Cvoid}), ctx, dtype, count, seed, offset, out)
ec_filter_coarse(ctx, dtype, n, nbar, l, visc, u, p, nq, q) =
    ccall((:ec_filter_coarse, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Cdouble, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}, Cint, Ptr{FilterQuantity}), ctx, dtype, n, nbar, l, visc, u, p, nq, q)
ec_filter_coarse_noalloc(ctx, dtype, n, nbar, l, visc, u, p, nq, q) =
    ccall((:ec_filter_coarse_noalloc, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Cdouble, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}, Cint, Ptr{FilterQuantity}), ctx, dtype, n, nbar, l, visc, u, p, nq, q)
ec_filter_apply(ctx, dtype, n, len, w, in_, out) =
    ccall((:ec_filter_apply, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Cint}, Ptr{Ptr{Cdouble}}, Ptr{Cvoid}, Ptr{Cvoid}), ctx, dtype, n, len, w, in_, out)
ec_ns_closure(ctx, dtype, n, l, visc, u, p, nsets, sets) =
    ccall((:ec_ns_closure, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}, Cint, Ptr{ClosureSet}), ctx, dtype, n, l, visc, u, p, nsets, sets)
ec_ns_filter_vel(ctx, dtype, n, l, v, p, nsets, sets) =
    ccall((:ec_ns_filter_vel, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}, Cint, Ptr{VelSet}), ctx, dtype, n, l, v, p, nsets, sets)
ec_ns_filter_vel_rk4(ctx, dtype, n, l, du, p, nsets, sets, stage, dt, u0, us, acc) =
    ccall((:ec_ns_filter_vel_rk4, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}, Cint, Ptr{VelSet}, Cint, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}), ctx, dtype, n, l, du, p, nsets, sets, stage, dt, u0, us, acc)
ec_ns_filter_vel_rk4_states(ctx, dtype, n, l, du, p, nsets, sets, stage, dt, u0, ust) =
    ccall((:ec_ns_filter_vel_rk4_states, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}, Cint, Ptr{VelSet}, Cint, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{9,Ptr{Cvoid}}}), ctx, dtype, n, l, du, p, nsets, sets, stage, dt, u0, ust)
ec_fft_set_cols_persistent(ctx, on) =
    ccall((:ec_fft_set_cols_persistent, libexactclosure), Cint, (Ptr{Cvoid}, Cint), ctx, on)
ec_poisson_set_z_tridiagonal(ctx, on) =
    ccall((:ec_poisson_set_z_tridiagonal, libexactclosure), Cint, (Ptr{Cvoid}, Cint), ctx, on)
ec_filter_vel_set_tiles(ctx, tx, tz) =
    ccall((:ec_filter_vel_set_tiles, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint), ctx, tx, tz)
ec_closure_set_tiles(ctx, tx, ty, tz) =
    ccall((:ec_closure_set_tiles, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cint), ctx, tx, ty, tz)
ec_ns_commutator(ctx, nbar, l, visc, ubar, flux9, tau9) =
    ccall((:ec_ns_commutator, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{9,Ptr{Cvoid}}}, Ref{NTuple{9,Ptr{Cvoid}}}), ctx, nbar, l, visc, ubar, flux9, tau9)
ec_ns_sfs_filter(ctx, dtype, n, nbar, l, visc, u, p, kind, fu, fr9, sr9) =
    ccall((:ec_ns_sfs_filter, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Cdouble, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}, Cint, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{9,Ptr{Cvoid}}}, Ref{NTuple{9,Ptr{Cvoid}}}), ctx, dtype, n, nbar, l, visc, u, p, kind, fu, fr9, sr9)
ec_ns_tensordissipation(ctx, nbar, l, sigma9, u, d) =
    ccall((:ec_ns_tensordissipation, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cdouble, Ref{NTuple{9,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}), ctx, nbar, l, sigma9, u, d)
ec_ns_rk4_update(ctx, dtype, n, l, stage, dt, u0, us, du, p, acc) =
    ccall((:ec_ns_rk4_update, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cint, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}, Ref{NTuple{3,Ptr{Cvoid}}}), ctx, dtype, n, l, stage, dt, u0, us, du, p, acc)
ec_ns_rk4_update_states(ctx, dtype, n, l, stage, dt, u0, ust, du, p) =
    ccall((:ec_ns_rk4_update_states, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cint, Cdouble, Ref{NTuple{3,Ptr{Cvoid}}}, Ref{NTuple{9,Ptr{Cvoid}}}, Ref{NTuple{3,Ptr{Cvoid}}}, Ptr{Cvoid}), ctx, dtype, n, l, stage, dt, u0, ust, du, p)
ec_burgers_batch_create(ctx, dtype, count, n, c, L, visc, cfl, kmax, seed, batch) =
    ccall((:ec_burgers_batch_create, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Cint}, Ptr{Cint}, Cdouble, Cdouble, Cdouble, Cint, UInt64, Ptr{Ptr{Cvoid}}), ctx, dtype, count, n, c, L, visc, cfl, kmax, seed, batch)
ec_burgers_batch_destroy(ctx, batch) =
    ccall((:ec_burgers_batch_destroy, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx, batch)
ec_burgers_batch_set_threads(ctx, batch, threads) =
    ccall((:ec_burgers_batch_set_threads, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint), ctx, batch, threads)
ec_burgers_batch_run(ctx, batch, nsteps, use_graph) =
    ccall((:ec_burgers_batch_run, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint), ctx, batch, nsteps, use_graph)
ec_burgers_batch_get(ctx, batch, instance, which, out, cap) =
    ccall((:ec_burgers_batch_get, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}, Cint), ctx, batch, instance, which, out, cap)
ec_burgers_batch_dt(ctx, batch, instance, dt) =
    ccall((:ec_burgers_batch_dt, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cdouble}), ctx, batch, instance, dt)
ec_burgers_randomfield(ctx, n, L, kpeak, totalenergy, nsamples, seed0, u) =
    ccall((:ec_burgers_randomfield, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cint, UInt64, Ptr{Cvoid}), ctx, n, L, kpeak, totalenergy, nsamples, seed0, u)
ec_burgers_dns_aided_les(ctx, nh, nH, L, visc, tstop, cfl_factor, nsamples, ustart, Rl, wl, K, wd, out, steps) =
    ccall((:ec_burgers_dns_aided_les, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Ptr{Cvoid}, Ptr{Cint}), ctx, nh, nH, L, visc, tstop, cfl_factor, nsamples, ustart, Rl, wl, K, wd, out, steps)
ec_burgers_rhs(ctx, dtype, n, L, visc, u, du) =
    ccall((:ec_burgers_rhs, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}), ctx, dtype, n, L, visc, u, du)
ec_burgers_fractions(ctx, n, nH, L, visc, nsamples, u, Rl, wl, K, wd, out) =
    ccall((:ec_burgers_fractions, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cint, Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Ptr{Cdouble}), ctx, n, nH, L, visc, nsamples, u, Rl, wl, K, wd, out)
ec_burgers_dissipation(ctx, n, nH, L, visc, nsamples, u, Rl, wl, K, wd, out) =
    ccall((:ec_burgers_dissipation, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cint, Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Ptr{Cvoid}), ctx, n, nH, L, visc, nsamples, u, Rl, wl, K, wd, out)
ec_burgers_smag_fit(ctx, n, nH, L, visc, delta, nsamples, u, Rl, wl, K, wd, out) =
    ccall((:ec_burgers_smag_fit, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cdouble, Cint, Ptr{Cvoid}, Cint, Ptr{Cdouble}, Cint, Ptr{Cdouble}, Ptr{Cdouble}), ctx, n, nH, L, visc, delta, nsamples, u, Rl, wl, K, wd, out)
ec_burgers_smag_solve(ctx, n, nH, L, visc, C, tstop, cfl, nsamples, ustart, K, wd, ubar_out) =
    ccall((:ec_burgers_smag_solve, libexactclosure), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{Cvoid}), ctx, n, nH, L, visc, C, tstop, cfl, nsamples, ustart, K, wd, ubar_out)
ec_cnn_create(ctx, seed, net) =
    ccall((:ec_cnn_create, libexactclosure), Cint, (Ptr{Cvoid}, UInt64, Ptr{Ptr{Cvoid}}), ctx, seed, net)
ec_cnn_destroy(ctx, net) =
    ccall((:ec_cnn_destroy, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx, net)
ec_cnn_num_params(net) =
    ccall((:ec_cnn_num_params, libexactclosure), Cint, (Ptr{Cvoid},), net)
ec_cnn_params(ctx, net, which, get, host, count) =
    ccall((:ec_cnn_params, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cfloat}, Cint), ctx, net, which, get, host, count)
ec_cnn_forward(ctx, net, n, batch, u, in_scale, out, out_scale) =
    ccall((:ec_cnn_forward, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Ptr{Cvoid}}, Cdouble, Ptr{Ptr{Cvoid}}, Cdouble), ctx, net, n, batch, u, in_scale, out, out_scale)
ec_cnn_backward(ctx, net, target, tgt_scale, loss) =
    ccall((:ec_cnn_backward, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cdouble, Ptr{Cdouble}), ctx, net, target, tgt_scale, loss)
ec_cnn_adam(ctx, net, lr, beta1, beta2, eps) =
    ccall((:ec_cnn_adam, libexactclosure), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble), ctx, net, lr, beta1, beta2, eps)
