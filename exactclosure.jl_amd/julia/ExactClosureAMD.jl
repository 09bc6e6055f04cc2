"""
    ExactClosureAMD

MI355X backend of ExactClosure.jl's DNS-aided LES.  The submodules
`NavierStokes` and `Burgers` keep the public names, arguments and results of
the reference's `NavierStokes` (src/N.jl) and `Burgers` (src/Burgers.jl)
for the DNS-aided path; every loop over a field runs in a HIP kernel of
libexactclosure (gfx950), reached through the C ABI of
include/exactclosure.h (bound one function per entry point in abi.jl).
Fields live in device memory (`HipArray`, column-major like `Array{T,3}`);
closure and filter outputs are FP64.

    include("exactclosure.jl_amd/julia/ExactClosureAMD.jl")
    using .ExactClosureAMD: NavierStokes as NS, Burgers as B

    setup = NS.getsetup()                      # navierstokes.jl:10
    uaid = NS.dnsaid_project(setup)            # navierstokes.jl:13
    NS.compute_errors(uaid) |> pairs           # navierstokes.jl:28

    bsetup = B.getsetup()                      # burgers.jl:146
    series = B.run_dns_aided_les(bsetup)       # burgers.jl:149-151
    B.compute_errors(series) |> pairs          # burgers.jl:154

There is no CPU path: without the library or a GPU every call throws.
"""
module ExactClosureAMD

const libexactclosure = normpath(joinpath(@__DIR__, "..", "libexactclosure.so"))
const EC_F32 = Cint(0)
const EC_F64 = Cint(1)

include("abi.jl")

dtcode(::Type{Float32}) = EC_F32
dtcode(::Type{Float64}) = EC_F64

struct ECError <: Exception
    code::Cint
    msg::String
end
Base.showerror(io::IO, e::ECError) = print(io, "libexactclosure error ", e.code, ": ", e.msg)

# One context per process: device 0 and its default stream.  It is never
# destroyed: device arrays free themselves through it in their finalizers.
mutable struct Context
    h::Ptr{Cvoid}
end
const CONTEXT = Ref{Union{Nothing,Context}}(nothing)

function context()
    c = CONTEXT[]
    c === nothing || return c
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ec_create(r, Cint(0), C_NULL)
    rc == 0 || throw(ECError(rc, "ec_create(device 0) failed: no usable GPU"))
    c = Context(r[])
    CONTEXT[] = c
    return c
end

"`call(f, args...)`: entry point `f` with the context handle first; throws on a non-zero code."
function call(f, args...)
    h = context().h
    rc = f(h, args...)
    rc == 0 || throw(ECError(rc, string(nameof(f), ": ", unsafe_string(ec_last_error(h)))))
    return nothing
end

synchronize() = call(ec_sync)

# ---------------------------------------------------------------------------
# Device arrays
# ---------------------------------------------------------------------------

"""
    HipArray{T}(undef, dims)

Device memory from ec_alloc holding `prod(dims)` values of `T`, laid out as
`Array{T,N}`; freed by its finalizer.  Copy with `copyto!` and `Array`.
"""
mutable struct HipArray{T,N}
    ptr::Ptr{T}
    dims::NTuple{N,Int}
    function HipArray{T}(::UndefInitializer, dims::NTuple{N,Int}) where {T,N}
        r = Ref{Ptr{Cvoid}}(C_NULL)
        call(ec_alloc, max(sizeof(T) * prod(dims), 1), r)
        a = new{T,N}(Ptr{T}(r[]), dims)
        finalizer(x -> ec_free(context().h, x.ptr), a)
        return a
    end
end
HipArray{T}(::UndefInitializer, dims::Int...) where {T} = HipArray{T}(undef, dims)
HipArray(x::Array{T,N}) where {T,N} = copyto!(HipArray{T}(undef, size(x)), x)

Base.size(a::HipArray) = a.dims
Base.size(a::HipArray, d::Integer) = a.dims[d]
Base.length(a::HipArray) = prod(a.dims)
Base.eltype(::HipArray{T}) where {T} = T
Base.pointer(a::HipArray) = a.ptr
Base.sizeof(a::HipArray{T}) where {T} = sizeof(T) * length(a)
Base.unsafe_convert(::Type{Ptr{Cvoid}}, a::HipArray) = Ptr{Cvoid}(a.ptr)
Base.show(io::IO, a::HipArray{T}) where {T} = print(io, "HipArray{", T, "}", a.dims)

function hipzeros(::Type{T}, dims::Int...) where {T}
    a = HipArray{T}(undef, dims)
    call(ec_zero, a, sizeof(a))
    return a
end
function Base.copyto!(dst::HipArray{T}, src::Array{T}) where {T}
    length(dst) == length(src) || throw(DimensionMismatch("copyto!: $(dst.dims) <- $(size(src))"))
    call(ec_h2d, dst, src, sizeof(src))
    return dst
end
function Base.copyto!(dst::Array{T}, src::HipArray{T}) where {T}
    length(dst) == length(src) || throw(DimensionMismatch("copyto!: $(size(dst)) <- $(src.dims)"))
    call(ec_d2h, dst, src, sizeof(dst))
    return dst
end
function Base.copyto!(dst::HipArray{T}, src::HipArray{T}) where {T}
    length(dst) == length(src) || throw(DimensionMismatch("copyto!: $(dst.dims) <- $(src.dims)"))
    call(ec_d2d, dst, src, sizeof(dst))
    return dst
end
Base.Array(a::HipArray{T,N}) where {T,N} = copyto!(Array{T,N}(undef, a.dims), a)

"y .+= a .* x on the device (LinearAlgebra.axpy!)."
function axpy!(a, x::HipArray{T}, y::HipArray{T}) where {T}
    call(ec_axpy, dtcode(T), length(y), Float64(a), x, y)
    return y
end
"y .= a .* x .+ b .* y on the device (LinearAlgebra.axpby!)."
function axpby!(a, x::HipArray{T}, b, y::HipArray{T}) where {T}
    call(ec_axpby, dtcode(T), length(y), Float64(a), x, Float64(b), y)
    return y
end
function sumsq(x::HipArray{T}) where {T}
    r = Ref{Cdouble}(0)
    call(ec_reduce_sumsq, dtcode(T), length(x), x, r)
    return r[]
end
function diffsq(x::HipArray{T}, y::HipArray{T}) where {T}
    r = Ref{Cdouble}(0)
    call(ec_reduce_diffsq, dtcode(T), length(x), x, y, r)
    return r[]
end
function maxabs(x::HipArray{T}) where {T}
    r = Ref{Cdouble}(0)
    call(ec_reduce_maxabs, dtcode(T), length(x), x, r)
    return r[]
end

# ---------------------------------------------------------------------------
# NavierStokes (src/N.jl)
# ---------------------------------------------------------------------------

module NavierStokes

using Random: AbstractRNG, Xoshiro
using ..ExactClosureAMD: ExactClosureAMD as EC, HipArray, call, dtcode

export Grid, Field, vectorfield, tensorfield, poissonsolver, project!, dnsaid_project, compute_errors

"The backend of this module: fields in libexactclosure device memory."
struct AMDBackend end
defaultbackend() = AMDBackend()

"Grid{D}(l, n) (src/N.jl:22-32); the kernels implement D = 3."
struct Grid{D}
    l::Float64
    n::Int
end
Grid(::Val{D}, l, n) where {D} = Grid{D}(l, n)
dim(::Grid{D}) where {D} = D
spacing(g::Grid) = g.l / g.n
wavenumber_int(g::Grid, i) = i - 1 - ifelse(i <= (g.n + 1) >> 1, 0, g.n)

"Scalar field on a grid (src/N.jl:35-43) with device data."
struct Field{D,T}
    grid::Grid{D}
    data::HipArray{T,D}
end
Field(g::Grid{3}, ::AMDBackend, ::Type{T} = Float64) where {T} = Field(g, EC.hipzeros(T, g.n, g.n, g.n))
Field(g::Grid, ::AMDBackend, ::Type = Float64) =
    throw(ArgumentError("libexactclosure implements three-dimensional grids (got D = $(dim(g)))"))
Field(g::Grid{3}, a::Array{T,3}) where {T} = Field(g, HipArray(a))
Base.eltype(::Field{D,T}) where {D,T} = T
Base.Array(f::Field) = Array(f.data)

vectorfield(g::Grid, b::AMDBackend, T = Float64) = ntuple(_ -> Field(g, b, T), dim(g))
tensorfield(g::Grid{3}, b::AMDBackend, T = Float64) = ntuple(_ -> Field(g, b, T), 6)

"Host copies of every field of a named tuple (the `adapt(CPU(), x)` of the scripts)."
to_host(x::NamedTuple) = map(to_host, x)
to_host(x::Tuple) = map(to_host, x)
to_host(f::Field) = Array(f)

dcode(f::Field) = dtcode(eltype(f))
"The host array of device pointers for a `void *const x[k]` argument."
ptrs(fs) = Ref(ntuple(i -> Ptr{Cvoid}(pointer(fs[i] isa Field ? fs[i].data : fs[i])), length(fs)))

# --- operators (src/N.jl:94-307) ---

"stresstensor(u, visc) (src/N.jl:94-118), materialised in the order 11, 22, 33, 12, 23, 31."
function stresstensor(u, visc)
    g = u[1].grid
    s = tensorfield(g, AMDBackend(), eltype(u[1]))
    GC.@preserve u s call(EC.ec_ns_stresstensor, dcode(u[1]), g.n, g.l, visc, ptrs(u), ptrs(s))
    return s
end

"tensordivergence!(du, σ, doadd) (src/N.jl:120-160)."
function tensordivergence!(du, σ, doadd)
    g = du[1].grid
    GC.@preserve du σ call(EC.ec_ns_tensordivergence, dcode(du[1]), g.n, g.l, ptrs(σ), ptrs(du), Cint(doadd))
    return du
end

"tensordivergence_nonsym!(du, σ, doadd) (src/N.jl:162-200), σ = (11, 21, 31, 12, 22, 32, 13, 23, 33)."
function tensordivergence_nonsym!(du, σ, doadd)
    g = du[1].grid
    sd = dtcode(eltype(σ[1] isa Field ? σ[1].data : σ[1]))
    GC.@preserve du σ call(EC.ec_ns_tensordivergence_nonsym, dcode(du[1]), g.n, g.l, sd, ptrs(σ), ptrs(du),
                           Cint(doadd))
    return du
end

"The fused tensordivergence!(du, stresstensor(u, visc), doadd)."
function rhs!(du, u, visc, doadd = false)
    g = u[1].grid
    GC.@preserve du u call(EC.ec_ns_rhs, dcode(u[1]), g.n, g.l, visc, ptrs(u), ptrs(du), Cint(doadd))
    return du
end

"divergence!(d, u) (src/N.jl:222-238)."
function divergence!(d::Field, u)
    g = d.grid
    GC.@preserve d u call(EC.ec_ns_divergence, dcode(d), g.n, g.l, ptrs(u), d.data)
    return d
end

"pressuregradient!(u, p) (src/N.jl:240-254): u .-= grad p."
function pressuregradient!(u, p::Field)
    g = p.grid
    GC.@preserve u p call(EC.ec_ns_pressuregradient, dcode(p), g.n, g.l, ptrs(u), p.data)
    return u
end

"poissonsolver(grid, backend) (src/N.jl:202-220): FFT plan, half spectrum and pressure."
mutable struct PoissonSolver{T}
    h::Ptr{Cvoid}
    grid::Grid{3}
end
function poissonsolver(g::Grid{3}, ::AMDBackend, ::Type{T} = Float64) where {T}
    r = Ref{Ptr{Cvoid}}(C_NULL)
    call(EC.ec_poisson_create, dtcode(T), g.n, g.l, r)
    ps = PoissonSolver{T}(r[], g)
    finalizer(x -> EC.ec_poisson_destroy(EC.context().h, x.h), ps)
    return ps
end
"The device pointer of the solver's pressure (poisson.p), valid while the solver lives."
pressure(ps::PoissonSolver) = EC.ec_poisson_pressure(ps.h)

"project!(u, poisson) (src/N.jl:256-307); the pressure stays in the solver."
function project!(u, poisson::PoissonSolver)
    GC.@preserve u call(EC.ec_ns_project, poisson.h, ptrs(u))
    return u
end

"(k > 0) k^p, the power-law spectrum of the reference's setups."
struct PowerProfile
    p::Float64
end
(f::PowerProfile)(k) = (k > 0) * k^f.p
"peak_profile (src/N.jl:366)."
peak_profile(k; kpeak) = k^4 * exp(-2 * (k / kpeak)^2)

"""
    randomfield(profile, g, backend, poisson; rng = Xoshiro(0), totalenergy = 1.0, kpeak = 5)

randomfield (src/N.jl:394-476) for `PowerProfile(p)` or `peak_profile`: a
divergence-free field with that shell spectrum and energy `totalenergy`.
The normal variates come from the library's counter-based generator seeded
by one draw of `rng`, so the field differs from Julia's Xoshiro stream.
"""
function randomfield(profile, g::Grid{3}, b::AMDBackend, poisson::PoissonSolver{T};
                     rng::AbstractRNG = Xoshiro(0), totalenergy = 1.0, kpeak = 5) where {T}
    pid, param = if profile isa PowerProfile
        Cint(1), profile.p
    elseif profile === peak_profile
        Cint(0), Float64(kpeak)
    else
        throw(ArgumentError("randomfield: the device generator takes PowerProfile(p) or peak_profile"))
    end
    u = vectorfield(g, b, T)
    GC.@preserve u call(EC.ec_ns_randomfield, poisson.h, pid, param, Float64(totalenergy), rand(rng, UInt64),
                        ptrs(u))
    return u
end

"propose_timestep (src/N.jl:348-356); the maxima are FP64 device reductions."
function propose_timestep(u, visc, cfl)
    g = u[1].grid
    h = spacing(g)
    umax = maximum(x -> EC.maxabs(x.data), u)
    return min(cfl * h / umax, cfl * h^2 / visc / 2dim(g))
end

"step_forwardeuler! (src/N.jl:309-318)."
function step_forwardeuler!(u, du, poisson, visc, dt)
    rhs!(du, u, visc, false)
    for (x, d) in zip(u, du)
        EC.axpy!(dt, d.data, x.data)
    end
    return project!(u, poisson)
end

"step_wray3! (src/N.jl:320-346)."
function step_wray3!(u, du, u0, poisson, visc, dt)
    a = (8 / 15, 5 / 12, 3 / 4)
    b = (1 / 4, 0 / 1)
    foreach((x0, x) -> copyto!(x0.data, x.data), u0, u)
    for i in 1:3
        rhs!(du, u, visc, false)
        for (d, x, x0) in zip(du, u, u0)
            i == 1 || copyto!(x.data, x0.data)
            EC.axpy!(a[i] * dt, d.data, x.data)
        end
        project!(u, poisson)
        if i < 3
            foreach((d, x0) -> EC.axpy!(b[i] * dt, d.data, x0.data), du, u0)
        end
    end
    return u
end

"spectrum(u, stuff, poisson) (src/N.jl:576-590) over the shells `stuff.shells`."
function spectrum(u, stuff, poisson::PoissonSolver)
    nshell = maximum(stuff.shells) + 1
    out = Vector{Cdouble}(undef, nshell)
    GC.@preserve u call(EC.ec_ns_spectrum, poisson.h, ptrs(u), nshell, out)
    return (; k = stuff.k, s = out[stuff.shells .+ 1])
end

# --- filters (src/N.jl:592-698, 999-1020) ---

"tophat (src/N.jl:592-597): odd compression, weights 1/c."
function tophat(grid, compression)
    isodd(compression) || throw(ArgumentError("tophat: odd compression (got $compression)"))
    return fill(1 / compression, compression)
end

"gaussian (src/N.jl:599-607); Δ = 0 gives the identity kernel [1]."
function gaussian(grid, Δ, nσ)
    Δ < 1e-15 && return [1.0]
    h = spacing(grid)
    σ = Δ / sqrt(12)
    r = round(Int, nσ * σ / h)
    w = [exp(-(h * i)^2 / 2σ^2) for i in (-r):r]
    return w ./ sum(w)
end

"composekernel (src/N.jl:609-625): the full discrete convolution."
function composekernel(f, g)
    out = zeros(length(f) + length(g) - 1)
    for (i, a) in enumerate(f), (j, b) in enumerate(g)
        out[i + j - 1] += a * b
    end
    return out
end

"kernelproduct(grid, kernels) (src/N.jl:627-637), kept as its three factors."
struct KernelProduct
    kernels::NTuple{3,Vector{Float64}}
end
kernelproduct(::Grid{3}, kernels) = KernelProduct(ntuple(i -> Vector{Float64}(kernels[i]), 3))

"get_kernels (src/N.jl:1008-1020): GΔH = FΔH ⊗ FΔH ⊗ FΔH and GΔHi (FΔ along i)."
function get_kernels(g_dns, g_les, Δ_ratio, nσ)
    FH = tophat(g_dns, div(g_dns.n, g_les.n))
    FΔ = gaussian(g_dns, Δ_ratio * spacing(g_les), nσ)
    FΔH = composekernel(FH, FΔ)
    GΔH = kernelproduct(g_dns, (FΔH, FΔH, FΔH))
    GΔHi = ntuple(i -> kernelproduct(g_dns, ntuple(j -> i == j ? FΔ : FΔH, 3)), 3)
    return (; FΔ, FΔH, GΔH, GΔHi)
end

"applyfilter!(ubar, u, kernel) (src/N.jl:639-653) on the full grid (three periodic 1D passes)."
function applyfilter!(ubar::Field, u::Field, kernel::KernelProduct)
    ks = kernel.kernels
    lens = Cint[length(k) for k in ks]
    ws = Ptr{Cdouble}[pointer(k) for k in ks]
    GC.@preserve ks lens ws ubar u call(EC.ec_filter_apply, dcode(u), u.grid.n, lens, ws, u.data, ubar.data)
    return ubar
end

# get_positions (src/N.jl:999-1006): faces, centre and the edge positions corners3D
const CENTER = (false, false, false)
const CORNER23, CORNER31, CORNER12 = (false, true, true), (true, false, true), (true, true, false)
face(i) = ntuple(d -> d == i, 3)

"Offset of a (2R+1)-tap kernel at the coarse-grain point of coarsegrain! (src/N.jl:670-684)."
coarse_offset(c, stag, R) = -R - (stag ? 0 : div(c, 2))

# the source fields of ec_filter_quantity
const SOURCES = (u1 = 0, u2 = 1, u3 = 2, s11 = 3, s22 = 4, s33 = 5, s12 = 6, s23 = 7, s31 = 8,
                 r11 = 9, r22 = 10, r33 = 11, p = 12)

"""
    filter_coarse(u, p, g_les, visc, quantities)

lazycoarsegrain(lazyfilter(q, kernel)) of several quantities of the DNS in
one launch.  `quantities`: `(source, ((w1, off1), (w2, off2), (w3, off3)))`
with `source` a key of `SOURCES` (u, the stresses σ(u), r = σ + p I) and the
offsets relative to the coarse face index.  `p` is a device pointer to the
pressure or `C_NULL`.  Returns FP64 n_les^3 device arrays.
"""
function filter_coarse(u, p::Ptr{Cvoid}, gbar::Grid{3}, visc, quantities)
    g = u[1].grid
    outs = [HipArray{Float64}(undef, (gbar.n, gbar.n, gbar.n)) for _ in quantities]
    ws = [ntuple(d -> Vector{Float64}(pcs[d][1]), 3) for (_, pcs) in quantities]
    qs = map(enumerate(quantities)) do (i, (src, pcs))
        EC.FilterQuantity(Cint(SOURCES[src]), ntuple(d -> Cint(length(ws[i][d])), 3), ntuple(d -> Cint(pcs[d][2]), 3),
                          ntuple(d -> pointer(ws[i][d]), 3), Ptr{Cvoid}(pointer(outs[i])))
    end
    GC.@preserve u outs ws qs call(EC.ec_filter_coarse, dcode(u[1]), g.n, gbar.n, g.l, visc, ptrs(u), p,
                                   length(qs), qs)
    return outs
end

"""
The coarse quantities of one DNS-aided LES step (name => (source, pieces)):
ubar_a (apply_ΔHπ, src/N.jl:1228-1237), σΔH_ab at the stress positions
(get_σΔH, src/N.jl:1088-1115) in the order 11, 22, 33, 12, 23, 31, and rΔHi_ab
(get_rΔHi, src/N.jl:1038-1073) in the order 11, 21, 31, 12, 22, 32, 13, 23, 33.
"""
function dnsaid_quantities(c, FΔ, FΔH)
    GΔH = (FΔH, FΔH, FΔH)
    GΔHi = ntuple(i -> ntuple(d -> d == i ? FΔ : FΔH, 3), 3)
    pieces(ks, stag) = ntuple(d -> (ks[d], coarse_offset(c, stag[d], div(length(ks[d]), 2))), 3)
    q = Pair{Symbol,Any}[]
    for a in 1:3
        push!(q, Symbol(:ubar, a) => (Symbol(:u, a), pieces(GΔH, face(a))))
    end
    for (name, src, pos) in (("11", :s11, CENTER), ("22", :s22, CENTER), ("33", :s33, CENTER),
                             ("12", :s12, CORNER12), ("23", :s23, CORNER23), ("31", :s31, CORNER31))
        push!(q, Symbol(:σΔH, name) => (src, pieces(GΔH, pos)))
    end
    for (name, src, b, pos) in (("11", :r11, 1, CENTER), ("21", :s12, 1, CORNER12), ("31", :s31, 1, CORNER31),
                                ("12", :s12, 2, CORNER12), ("22", :r22, 2, CENTER), ("32", :s23, 2, CORNER23),
                                ("13", :s31, 3, CORNER31), ("23", :s23, 3, CORNER23), ("33", :r33, 3, CENTER))
        push!(q, Symbol(:rΔHi, name) => (src, pieces(GΔHi[b], pos)))
    end
    return q
end

"lazytensorcoarsegrain (src/N.jl:1117-1135): the six stresses of a fine field at their LES positions."
function stress_positions(c)
    one = [1.0]
    return [(src, ntuple(d -> (one, coarse_offset(c, pos[d], 0)), 3))
            for (src, pos) in ((:s11, CENTER), (:s22, CENTER), (:s33, CENTER), (:s12, CORNER12),
                               (:s23, CORNER23), (:s31, CORNER31))]
end

"apply_ΔHπ(ubar, u, poisson, GΔH) (src/N.jl:1228-1237): filter, sample at the faces, project."
function apply_ΔHπ(ubar, u, poisson::PoissonSolver, GΔH::KernelProduct)
    gl = ubar[1].grid
    c = div(u[1].grid.n, gl.n)
    ks = GΔH.kernels
    qs = [(Symbol(:u, a), ntuple(d -> (ks[d], coarse_offset(c, face(a)[d], div(length(ks[d]), 2))), 3)) for a in 1:3]
    outs = filter_coarse(u, C_NULL, gl, 0.0, qs)
    foreach((x, o) -> copyto!(x.data, o), ubar, outs)
    return project!(ubar, poisson)
end

# --- the DNS-aided LES comparison (src/N.jl:1239-1349) ---

"getsetup (src/N.jl:712-723)."
getsetup(; n_dns = 500, n_les = 100, visc = 2.0e-4) = (;
    D = Val(3),
    l = 1.0,
    n_dns,
    n_les,
    visc,
    Δ_ratio = 2,
    nσ = 2,
    outdir = joinpath(pwd(), "output", "navierstokes", "n_dns=$(n_dns)") |> mkpath,
)

function _les_forward_euler!(u_les, du, poisson, dt)
    project!(du, poisson)
    foreach((x, d) -> EC.axpy!(dt, d.data, x.data), u_les, du)
    return u_les
end

"""
    dnsaid_project(setup; twarm = 0.1, tstop = 0.1, cfl = 0.4, rng = Xoshiro(0))

dnsaid_project (src/N.jl:1239-1332): a (k > 0) k^(-5/3) DNS warmed up to
`twarm` with forward Euler, then the DNS and five LES from apply_ΔHπ of it
advanced to `tstop` with the CFL time step: no model, classic, classic+flux,
classic+flux+div (the exact closure) and its symmetrised form.  Returns
`(; u_les, u_nomo, u_c, u_cf, u_cfd, u_cfd_symm)`, u_les = apply_ΔHπ of the
final DNS.  FP64 throughout; odd compressions n_dns / n_les.
"""
function dnsaid_project(setup; twarm = 0.1, tstop = 0.1, cfl = 0.4, rng::AbstractRNG = Xoshiro(0))
    (; l, n_dns, n_les, visc, Δ_ratio, nσ) = setup
    g_dns = Grid(setup.D, l, n_dns)
    g_les = Grid(setup.D, l, n_les)
    c = div(n_dns, n_les)
    c * n_les == n_dns || throw(ArgumentError("n_les must divide n_dns"))
    backend = defaultbackend()
    poisson_dns = poissonsolver(g_dns, backend)
    poisson_les = poissonsolver(g_les, backend)
    u_dns = randomfield(PowerProfile(-5 / 3), g_dns, backend, poisson_dns; rng, totalenergy = 1.0)
    du_dns = vectorfield(g_dns, backend)
    du_les = vectorfield(g_les, backend)
    ubar = vectorfield(g_les, backend)
    uf = vectorfield(g_dns, backend)
    (; FΔ, FΔH, GΔH) = get_kernels(g_dns, g_les, Δ_ratio, nσ)
    quantities = dnsaid_quantities(c, FΔ, FΔH)

    # warm-up of the DNS
    t = 0.0
    while t < twarm
        dt = min(propose_timestep(u_dns, visc, cfl), twarm - t)
        t += dt
        step_forwardeuler!(u_dns, du_dns, poisson_dns, visc, dt)
    end

    # the five LES start from the filtered DNS
    apply_ΔHπ(ubar, u_dns, poisson_les, GΔH)
    les = map((:u_nomo, :u_c, :u_cf, :u_cfd, :u_cfd_symm)) do _
        v = vectorfield(g_les, backend)
        foreach((x, y) -> copyto!(x.data, y.data), v, ubar)
        v
    end
    u_nomo, u_c, u_cf, u_cfd, u_cfd_symm = les

    t = 0.0
    while t < tstop
        dt = min(propose_timestep(u_dns, visc, cfl), tstop - t)
        t += dt
        # DNS right-hand side, projected; its pressure stays in the solver
        rhs!(du_dns, u_dns, visc, false)
        project!(du_dns, poisson_dns)
        outs = filter_coarse(u_dns, pressure(poisson_dns), g_les, visc, map(last, quantities))
        q = Dict(first(p) => o for (p, o) in zip(quantities, outs))
        foreach((a, x) -> copyto!(x.data, q[Symbol(:ubar, a)]), 1:3, ubar)
        project!(ubar, poisson_les)
        σΔH = [q[Symbol(:σΔH, ab)] for ab in ("11", "22", "33", "12", "23", "31")]
        rΔHi = [q[Symbol(:rΔHi, ab)] for ab in ("11", "21", "31", "12", "22", "32", "13", "23", "33")]
        σH = stresstensor(ubar, visc)

        # no model
        rhs!(du_les, u_nomo, visc, false)
        _les_forward_euler!(u_nomo, du_les, poisson_les, dt)
        # classic: σ_les + σΔH - coarse(σ(GΔH u)) (step_ΔHπ_classic!, src/N.jl:1137-1162)
        foreach((x, y) -> applyfilter!(x, y, GΔH), uf, u_dns)
        s2 = filter_coarse(uf, C_NULL, g_les, visc, stress_positions(c))
        σ = stresstensor(u_c, visc)
        for m in 1:6
            EC.axpy!(1.0, σΔH[m], σ[m].data)
            EC.axpy!(-1.0, s2[m], σ[m].data)
        end
        tensordivergence!(du_les, σ, false)
        _les_forward_euler!(u_c, du_les, poisson_les, dt)
        # classic + flux: σ_les + σΔH - σ_H(ubar) (step_ΔHπ_classic_flux!, src/N.jl:1164-1187)
        σ = stresstensor(u_cf, visc)
        for m in 1:6
            EC.axpy!(1.0, σΔH[m], σ[m].data)
            EC.axpy!(-1.0, σH[m].data, σ[m].data)
        end
        tensordivergence!(du_les, σ, false)
        _les_forward_euler!(u_cf, du_les, poisson_les, dt)
        # classic + flux + div (step_ΔHπ_classic_flux_div!, src/N.jl:1189-1206)
        σ = stresstensor(u_cfd, visc)
        foreach(m -> EC.axpy!(-1.0, σH[m].data, σ[m].data), 1:6)
        tensordivergence!(du_les, σ, false)
        tensordivergence_nonsym!(du_les, rΔHi, true)
        _les_forward_euler!(u_cfd, du_les, poisson_les, dt)
        # symmetrised rΔHi (step_ΔHπ_classic_flux_div_symm!, src/N.jl:1208-1226)
        σ = stresstensor(u_cfd_symm, visc)
        foreach(m -> EC.axpy!(-1.0, σH[m].data, σ[m].data), 1:6)
        tensordivergence!(du_les, σ, false)
        R = Dict(zip(("11", "21", "31", "12", "22", "32", "13", "23", "33"), rΔHi))
        symm = map(("11", "22", "33", "21", "32", "31")) do ab
            x = HipArray{Float64}(undef, R[ab].dims)
            copyto!(x, R[ab])
        end
        for (m, other) in ((4, "12"), (5, "23"), (6, "13"))
            EC.axpby!(0.5, R[other], 0.5, symm[m])
        end
        tensordivergence!(du_les, map(x -> Field(g_les, x), symm), true)
        _les_forward_euler!(u_cfd_symm, du_les, poisson_les, dt)
        # DNS
        foreach((x, d) -> EC.axpy!(dt, d.data, x.data), u_dns, du_dns)
    end
    u_les = vectorfield(g_les, backend)
    apply_ΔHπ(u_les, u_dns, poisson_les, GΔH)
    return (; u_les, u_nomo, u_c, u_cf, u_cfd, u_cfd_symm)
end

"compute_errors (src/N.jl:1334-1349): relative L2 error of each LES against u_les (FP64 reductions)."
function compute_errors(uaid)
    err(ubar) = sqrt(sum(zip(ubar, uaid.u_les)) do (a, b)
        EC.diffsq(a.data, b.data) / EC.sumsq(b.data)
    end / length(ubar))
    return map(err, (; uaid.u_nomo, uaid.u_c, uaid.u_cf, uaid.u_cfd, uaid.u_cfd_symm))
end

# --- sub-filter stresses and their dissipation (src/NavierStokes.jl:470-593, process.jl:107-193) ---

const NONSYM = ("11", "21", "31", "12", "22", "32", "13", "23", "33")

function _sfs_tensors(u, g_les::Grid{3}, poisson_dns, poisson_les, visc, experiment::Symbol)
    g = u[1].grid
    rem(g.n, g_les.n) == 0 || throw(ArgumentError("n_les = $(g_les.n) does not divide n_dns = $(g.n)"))
    du = vectorfield(g, AMDBackend(), eltype(u[1]))
    rhs!(du, u, visc, false)
    project!(du, poisson_dns)
    nb = g_les.n
    new() = HipArray{Float64}(undef, (nb, nb, nb))
    fu, fr, sr = ntuple(_ -> new(), 3), ntuple(_ -> new(), 9), ntuple(_ -> new(), 9)
    kind = experiment === :surfavg ? Cint(1) : Cint(0)
    GC.@preserve u fu fr sr call(EC.ec_ns_sfs_filter, dcode(u[1]), g.n, nb, g.l, visc, ptrs(u), pressure(poisson_dns),
                                 kind, ptrs(fu), ptrs(fr), ptrs(sr))
    fuF = map(x -> Field(g_les, x), fu)
    experiment === :project_volavg && project!(fuF, poisson_les)
    dul = vectorfield(g_les, AMDBackend())
    rhs!(dul, fuF, visc, false)
    project!(dul, poisson_les)
    pl = pressure(poisson_les)
    # σ = filtered r - r_H(fu), r_H = σ_H(fu) + p_les I
    σ_classic, σ_swapfil = map((fr, sr)) do q
        τ = ntuple(_ -> new(), 9)
        GC.@preserve fu q τ call(EC.ec_ns_commutator, nb, g_les.l, visc, ptrs(fu), ptrs(q), ptrs(τ))
        foreach(a -> call(EC.ec_axpy, EC.EC_F64, length(τ[4a + 1]), -1.0, pl, τ[4a + 1]), 0:2)
        τ
    end
    σ_swapfil_symm = ntuple(9) do m
        x = copyto!(new(), σ_swapfil[m])
        ab = NONSYM[m]
        ab[1] == ab[2] || EC.axpby!(0.5, σ_swapfil[findfirst(==(reverse(ab)), NONSYM)], 0.5, x)
        x
    end
    return (; σ_classic, σ_swapfil, σ_swapfil_symm, fu = fuF)
end

"""
    sfs_tensors_volume(; ustart, g_dns, g_les, poisson_dns, poisson_les, viscosity, compression, doproject)

sfs_tensors_volume (src/NavierStokes.jl:470-528): the classic, swap-filter
and symmetrised swap-filter sub-filter stresses of the volume average (fu
projected when `doproject`), 9-component FP64 device tensors in the order
11, 21, 31, 12, 22, 32, 13, 23, 33, and the filtered velocity `fu`.
"""
sfs_tensors_volume(; ustart, g_dns = nothing, g_les, poisson_dns, poisson_les, viscosity, compression = nothing,
                   doproject) =
    _sfs_tensors(ustart, g_les, poisson_dns, poisson_les, viscosity, doproject ? :project_volavg : :volavg)

"sfs_tensors_surface (src/NavierStokes.jl:530-593): as sfs_tensors_volume for the face average."
sfs_tensors_surface(; ustart, g_dns = nothing, g_les, poisson_dns, poisson_les, viscosity, compression = nothing,
                    doproject = false) =
    _sfs_tensors(ustart, g_les, poisson_dns, poisson_les, viscosity, :surfavg)

"Turbulox.tensordissipation_staggered!(d, σ, u) as process.jl:160-180 applies it: d = σ : ∇u at the centres (FP64)."
function tensordissipation_staggered!(d::Field, σ, u)
    g = d.grid
    GC.@preserve d σ u call(EC.ec_ns_tensordissipation, g.n, g.l, ptrs(σ), ptrs(u), d.data)
    return d
end

"""
The dissipation fields of one LES grid and experiment (:volavg,
:project_volavg or :surfavg) of process.jl:107-193: `(; classic, swapfil,
swapfil_symm)` FP64 fields.  The kernel density estimate is left to the caller.
"""
function sfs_dissipation(u, g_les::Grid{3}, viscosity, experiment::Symbol; poisson_dns = nothing,
                         poisson_les = nothing)
    ps = something(poisson_dns, poissonsolver(u[1].grid, AMDBackend(), eltype(u[1])))
    psl = something(poisson_les, poissonsolver(g_les, AMDBackend()))
    t = _sfs_tensors(u, g_les, ps, psl, viscosity, experiment)
    d(σ) = tensordissipation_staggered!(Field(g_les, AMDBackend()), σ, t.fu)
    return (; classic = d(t.σ_classic), swapfil = d(t.σ_swapfil), swapfil_symm = d(t.σ_swapfil_symm))
end

end # module NavierStokes

# ---------------------------------------------------------------------------
# Burgers (src/Burgers.jl)
# ---------------------------------------------------------------------------

module Burgers

using LinearAlgebra: norm
using ..ExactClosureAMD: ExactClosureAMD as EC, HipArray, call

export getsetup, randomfield, dns_aided_les, run_dns_aided_les, compute_errors

"Grid(L, n) of src/Burgers.jl."
struct Grid
    l::Float64
    n::Int
end
spacing(g::Grid) = g.l / g.n

"tophat_weights (src/Burgers.jl:158-162) for odd widths; even widths span c + 1 faces with half end weights."
function tophat_weights(comp)
    isodd(comp) && return div(comp, 2), fill(1 / comp, comp)
    w = fill(1 / comp, comp + 1)
    w[1] = w[end] = 0.5 / comp
    return div(comp, 2), w
end

"gaussian_weights (src/Burgers.jl:146-156)."
function gaussian_weights(h, Δ; nσ = 3)
    s2 = (Δ / h)^2 / 12
    R = round(Int, nσ * sqrt(s2))
    Δ < 1e-15 && return R, [1.0]
    w = [exp(-i^2 / 2s2) for i in (-R):R]
    return R, w ./ sum(w)
end

"build_doublekernel (src/Burgers.jl:217-233)."
function build_doublekernel(F, G)
    (I, f), (J, g) = F, G
    w = zeros(length(f) + length(g) - 1)
    for (i, a) in enumerate(f), (j, b) in enumerate(g)
        w[i + j - 1] += a * b
    end
    return I + J, w
end

"The LES kernel and the double kernel with the finite-volume top-hat (src/Burgers.jl:350-370, 1030-1048)."
function les_kernels(nh, nH, L, Δ_ratio, lesfiltertype)
    h, H = L / nh, L / nH
    Δ = H * Δ_ratio
    fvm = tophat_weights(div(nh, nH))
    les = if lesfiltertype == :tophat
        tophat_weights(2 * div(round(Int, Δ / h), 2) + 1)
    elseif lesfiltertype == :gaussian
        gaussian_weights(h, Δ)
    else
        throw(ArgumentError("lesfiltertype $lesfiltertype"))
    end
    return les, build_doublekernel(fvm, les)
end

"getsetup (src/Burgers.jl:235-265)."
getsetup(; nsample = 1000) = (;
    L = 2π,
    nh = 100 * 3^3 * 5,
    nH = 100 .* 3 .^ (1:3),
    Δ_ratios = [2],
    visc = 5.0e-4,
    kpeak = 10,
    initialenergy = 2.0,
    tstop = 0.1,
    nsample,
    lesfiltertype = :gaussian,
)

"""
    randomfield(g, kpeak, totalenergy, nsample; seed = 1)

randomfield with bump_profile (src/Burgers.jl:192-202) for `nsample`
samples: an (n, nsample) FP64 device array, column i seeded by seed + i - 1
(the reference seeds Xoshiro(i) for sample i).  The phases come from the
library's counter-based hash, not Julia's stream.
"""
function randomfield(g::Grid, kpeak, totalenergy, nsample; seed = 1)
    u = HipArray{Float64}(undef, (g.n, nsample))
    call(EC.ec_burgers_randomfield, g.n, g.l, Float64(kpeak), Float64(totalenergy), nsample, UInt64(seed), u)
    return u
end

const LABELS = (dns_ref = "DNS", dns_ref_fil = "Filtered DNS", dns_fil = "Filtered DNS", nomodel = "No model",
                class_m = "Classic", class_p = "Classic+Flux", swapfil = "Classic+Flux+Div (ours)")

"""
    dns_aided_les(ustart, gh, gH, visc; Δ, tstop, cfl_factor, lesfiltertype)

dns_aided_les (src/Burgers.jl:350-445) for every column of `ustart` (an
(nh, nsample) FP64 device array) in one launch: forward Euler with the CFL
step, the DNS and the no-model, classic, classic+flux and swap LES.  Returns
the fields as the reference's named tuple `(; g, u, label)` with host
(n, nsample) matrices, and the step count of each sample.
"""
function dns_aided_les(ustart::HipArray{Float64,2}, gh::Grid, gH::Grid, visc; Δ, tstop, cfl_factor = 0.3,
                       lesfiltertype = :gaussian)
    nh, ns = size(ustart)
    nh == gh.n || throw(DimensionMismatch("ustart has $nh rows, the DNS grid $(gh.n) points"))
    nH = gH.n
    (Rl, wl), (K, wd) = les_kernels(nh, nH, gh.l, Δ / spacing(gH), lesfiltertype)
    m = 2nh + 5nH
    out = HipArray{Float64}(undef, (m, ns))
    steps = zeros(Cint, ns)
    GC.@preserve wl wd call(EC.ec_burgers_dns_aided_les, nh, nH, gh.l, visc, tstop, cfl_factor, ns, ustart,
                            Rl, wl, K, wd, out, steps)
    o = Array(out)
    # per sample: dns_ref, dns_ref_fil (nh each), then five nH blocks
    rows = (1:nh, (nh + 1):(2nh), ntuple(i -> (2nh + (i - 1) * nH + 1):(2nh + i * nH), 5)...)
    names = keys(LABELS)
    fields = NamedTuple{names}(ntuple(j -> (; g = j <= 2 ? gh : gH, u = o[rows[j], :], label = LABELS[j]), 7))
    return fields, Int.(steps)
end

"""
    run_dns_aided_les(setup)

run_dns_aided_les (src/Burgers.jl:267-328): every (nH, Δ_ratio) of the setup
over all samples (each (nH, Δ_ratio) in one launch).  Returns the series
`[(; Δ_ratio, nH, fields)]` that the reference saves and
load_dns_aided_les reads back.
"""
function run_dns_aided_les(setup)
    (; L, nh, visc, kpeak, initialenergy, tstop, nsample, lesfiltertype) = setup
    gh = Grid(L, nh)
    ustart = randomfield(gh, kpeak, initialenergy, nsample)
    series = []
    for nH in setup.nH, Δ_ratio in setup.Δ_ratios
        gH = Grid(L, nH)
        fields, _ = dns_aided_les(ustart, gh, gH, visc; Δ = spacing(gH) * Δ_ratio, tstop, cfl_factor = 0.3,
                                  lesfiltertype)
        push!(series, (; Δ_ratio, nH, fields))
    end
    return series
end

"compute_errors (src/Burgers.jl:909-917)."
compute_errors(series) = NamedTuple(map((:nomodel, :class_m, :class_p, :swapfil)) do key
    e = map(series) do (; fields)
        norm(fields[key].u - fields.dns_fil.u) / norm(fields.dns_fil.u)
    end
    key => (; e, series[1].fields[key].label)
end)

"""
    compute_fractions(u, nh, nH, L, visc, Δ_ratio, lesfiltertype)

compute_fractions (src/Burgers.jl:1022-1099) over the columns of the
(nh, nsample) device array `u`: the classic, flux and div fractions.
"""
function compute_fractions(u::HipArray{Float64,2}, nh, nH, L, visc, Δ_ratio, lesfiltertype)
    (Rl, wl), (K, wd) = les_kernels(nh, nH, L, Δ_ratio, lesfiltertype)
    out = zeros(Cdouble, 3)
    GC.@preserve wl wd call(EC.ec_burgers_fractions, nh, nH, L, visc, size(u, 2), u, Rl, wl, K, wd, out)
    return (; classic = out[1], flux = out[2], div = out[3])
end

"""
compute_dissipation (src/Burgers.jl:447-516): the per-point dissipation of
class_m, class_p and swapfil, each an (nH, nsample) host matrix.
"""
function compute_dissipation(u::HipArray{Float64,2}, nh, nH, L, visc, Δ_ratio, lesfiltertype)
    (Rl, wl), (K, wd) = les_kernels(nh, nH, L, Δ_ratio, lesfiltertype)
    ns = size(u, 2)
    d = HipArray{Float64}(undef, (nH, ns, 3))
    GC.@preserve wl wd call(EC.ec_burgers_dissipation, nh, nH, L, visc, ns, u, Rl, wl, K, wd, d)
    dh = Array(d)
    return (; class_m = dh[:, :, 1], class_p = dh[:, :, 2], swapfil = dh[:, :, 3])
end

"""
The Smagorinsky coefficients of smagorinsky.jl:24-36 (classic and discrete
forms) from all samples of `U`.
"""
function smagorinsky_coefficients(U::HipArray{Float64,2}, nh, nH, L, visc, Δ_ratio, lesfiltertype)
    (Rl, wl), (K, wd) = les_kernels(nh, nH, L, Δ_ratio, lesfiltertype)
    out = zeros(Cdouble, 4)
    GC.@preserve wl wd call(EC.ec_burgers_smag_fit, nh, nH, L, visc, Δ_ratio * L / nH, size(U, 2), U, Rl, wl, K,
                            wd, out)
    return (; classic = sqrt(-out[1] / out[2]), discret = sqrt(-out[3] / out[4]))
end

"""
solve_smagorinsky (src/Burgers.jl:682-772): the Smagorinsky LES with
coefficient θ (C = θ² (Δ² + H²)) from the filtered initial fields, to
`tstop`; an (nH, nsample) device array.
"""
function solve_smagorinsky(ustart::HipArray{Float64,2}, nh, nH, θ, L, visc, Δ_ratio, lesfiltertype, tstop;
                           cfl = 0.3)
    _, (K, wd) = les_kernels(nh, nH, L, Δ_ratio, lesfiltertype)
    H = L / nH
    C = θ^2 * ((H * Δ_ratio)^2 + H^2)
    ns = size(ustart, 2)
    ubar = HipArray{Float64}(undef, (nH, ns))
    GC.@preserve wd call(EC.ec_burgers_smag_solve, nh, nH, L, visc, C, tstop, cfl, ns, ustart, K, wd, ubar)
    return ubar
end

end # module Burgers

end # module ExactClosureAMD
