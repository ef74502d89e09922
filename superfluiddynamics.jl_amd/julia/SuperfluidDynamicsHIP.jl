# SuperfluidDynamicsHIP.jl -- Julia side of the drop-in: HIP-array methods of
# the split-step path of SuperfluidDynamics.jl, calling libsfdhip.so through
# `ccall`.  Load after SuperfluidDynamics and AMDGPU:
#
#   using SuperfluidDynamics, AMDGPU
#   include("superfluiddynamics.jl_amd/julia/SuperfluidDynamicsHIP.jl")
#   grid  = Grid((256, 256, 256), ((0, 2π), (0, 2π), (0, 2π)); FT=Float32, array_type=ROCArray)
#   field = Field(grid, ComplexField())
#   ...   NumModelSplit2(field, param, Δt, niter, freqbckp) / timeStep! / solve! as usual
#
# Julia is not installed in the build container, so this file is not exercised
# by the test suite; the same entry points are exercised through ctypes
# (superfluiddynamics.jl_amd/sfdhip, tests/).  Dependencies of the drop-in:
# SuperfluidDynamics, AMDGPU, AbstractFFTs, PencilArrays, ChainRulesCore.
module SuperfluidDynamicsHIP

using SuperfluidDynamics
using SuperfluidDynamics: AbstractField, Field, Grid, GrossPitaevskiiParameters,
                          NumModelSplit1, NumModelSplit2, PotentialZero
using AMDGPU: ROCArray, stream
using AbstractFFTs: fftfreq
using PencilArrays: PencilArray
import ChainRulesCore

const libsfd = normpath(joinpath(@__DIR__, "..", "sfdhip", "libsfdhip.so"))

const SFD_F32, SFD_F64, SFD_F32X = Cint(1), Cint(2), Cint(3)
const SFD_REAL_TIME, SFD_IMAG_TIME = Cint(0), Cint(1)
const SFD_FORWARD, SFD_BACKWARD = Cint(-1), Cint(1)
const SFD_BATCH_GRAPH, SFD_BATCH_ESCALATE = Cint(1), Cint(2)

# Arithmetic of Float32 fields: SFD_F32 (FP32, lifting rotations, FP64
# reductions) or SFD_F32X (FP32 storage, FP64 arithmetic).  Set before the
# first step of a field.
const FLOAT32_MODE = Ref(SFD_F32)

last_error() = unsafe_string(ccall((:sfd_last_error, libsfd), Cstring, ()))

# Nonzero codes become the exceptions the reference raises for the same
# conditions (ArgumentError: Fields.jl:65-70; error(): Plans.jl:874-884).
function check(rc::Cint)
    rc == 0 && return nothing
    rc in (Cint(-1), Cint(-2), Cint(-3)) && throw(ArgumentError("libsfdhip: " * last_error()))
    rc == Cint(-4) && throw(OutOfMemoryError())
    error("libsfdhip: " * last_error())
end

mutable struct SfdContext
    handle::Ptr{Cvoid}
end

# One context per (field, stream): Grid + Plan state of Grids.jl:64-129 and
# Plans.jl:958-999 (wavenumbers, twiddles, phase tables, reduction workspace).
function SfdContext(g::Grid{N,FT}) where {N,FT}
    h = Ref{Ptr{Cvoid}}(C_NULL)
    n = Int64[g.n..., ntuple(_ -> 1, 3 - N)...]
    lo = Cdouble[g.min..., ntuple(_ -> 0.0, 3 - N)...]
    hi = Cdouble[g.max..., ntuple(_ -> 1.0, 3 - N)...]
    prec = FT === Float32 ? FLOAT32_MODE[] : SFD_F64
    check(ccall((:sfd_ctx_create, libsfd), Cint,
                (Ptr{Ptr{Cvoid}}, Cint, Ptr{Int64}, Ptr{Cdouble}, Ptr{Cdouble}, Cint, Ptr{Cvoid}),
                h, N, n, lo, hi, prec, stream().stream))
    ctx = SfdContext(h[])
    finalizer(c -> ccall((:sfd_ctx_destroy, libsfd), Cint, (Ptr{Cvoid},), c.handle), ctx)
    return ctx
end

const CONTEXTS = IdDict{Any,SfdContext}()
context(f::AbstractField) = get!(() -> SfdContext(f.g), CONTEXTS, f)

const HipField = Field{N,1,FT,Complex{FT},<:ROCArray} where {N,FT}

devptr(a) = reinterpret(Ptr{Cvoid}, pointer(parent(a)))
vptr(p::GrossPitaevskiiParameters) = p.pot isa PotentialZero ? C_NULL : devptr(p.pot.V)

function split!(n, order::Integer, mode::Cint, nsteps::Integer=1)
    c = context(n.f)
    p = n.param
    check(ccall((:sfd_split_step, libsfd), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Int64, Cint, Cint),
                c.handle, devptr(n.f.ϕ), vptr(p), p.coeffΔ, p.β, p.Ω, n.Δt, nsteps, order, mode))
    return order
end

# timeStep! of splitting.jl:78-82 and 149-154 for fields in HIP memory.
SuperfluidDynamics.timeStep!(n::NumModelSplit2{<:HipField}) = split!(n, 2, SFD_REAL_TIME)
SuperfluidDynamics.timeStep!(n::NumModelSplit1{<:HipField}) = split!(n, 1, SFD_REAL_TIME)

# k timeStep! calls in one call of the library: with Ω = 0 it merges the
# half-step flows that meet between two nonlinear phases (three passes over
# the field per 3D Strang step instead of five); returns Σ of the returns.
timeSteps!(n::NumModelSplit2{<:HipField}, k::Integer) = (split!(n, 2, SFD_REAL_TIME, k); 2k)
timeSteps!(n::NumModelSplit1{<:HipField}, k::Integer) = (split!(n, 1, SFD_REAL_TIME, k); k)

# solveLapRot! / solveNL! of splitting.jl:156-209.
function SuperfluidDynamics.solveLapRot!(n::Union{NumModelSplit1{<:HipField},NumModelSplit2{<:HipField}}, Δt)
    p = n.param
    check(ccall((:sfd_lap_rot, libsfd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cint),
                context(n.f).handle, devptr(n.f.ϕ), p.coeffΔ, p.Ω, Δt, SFD_REAL_TIME))
end
function SuperfluidDynamics.solveNL!(n::Union{NumModelSplit1{<:HipField},NumModelSplit2{<:HipField}}, Δt)
    p = n.param
    check(ccall((:sfd_solve_nl, libsfd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cint),
                context(n.f).handle, devptr(n.f.ϕ), vptr(p), p.β, Δt, SFD_REAL_TIME))
end

# Directional transforms of Plans.jl:1132-1250 (mul_d! forward, ldiv_d!
# normalised inverse), in the pen_x layout: the AbstractFFTs plug for other
# FFT-based models.
for (d, ax) in ((:x, 0), (:y, 1), (:z, 2))
    mul = Symbol(:hip_mul_, d, :!)
    ldv = Symbol(:hip_ldiv_, d, :!)
    @eval $mul(out::HipField, f::HipField) =
        check(ccall((:sfd_fft_axis, libsfd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint),
                    context(f).handle, devptr(f.ϕ), devptr(out.ϕ), $ax, SFD_FORWARD))
    @eval $ldv(out::HipField, f::HipField) =
        check(ccall((:sfd_fft_axis, libsfd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint),
                    context(f).handle, devptr(f.ϕ), devptr(out.ϕ), $ax, SFD_BACKWARD))
end

# Plan(f; t=FFTPlan()) for fields in HIP memory (Plans.jl:821-854,
# 958-999): the AbstractFFTs plug of the other FFT-based models
# (NumModelGPRK, NumModelExternalVelocity, computeDerivatives!).  Every
# buffer stays in the pen_x layout: mul_d!/ldiv_d! call sfd_fft_axis along
# axis d with the spectrum left where the samples were (no transposes), and
# grid_d returns the coordinates and wavenumbers broadcast in that layout,
# so the reference's multiplier broadcasts line up with it.  The buffer
# accessors (ϕx_hat, a_tmpx, a_tmp2x, ...; Plans.jl:1078-1124) come from
# AbstractFFTPlan through the datax/datay/dataz fields.
struct HipFFTPlan{N,F,V,B} <: SuperfluidDynamics.AbstractFFTPlan{N}
    f::F
    x::V
    ξ::V
    datax::Vector{B}
    datay::Vector{B}
    dataz::Vector{B}
end

function SuperfluidDynamics.Plan(f::HipField; t=SuperfluidDynamics.FFTPlan())
    t isa SuperfluidDynamics.FFTPlan || error("HIP fields take FFTPlan only")
    g = f.g
    nd = ndims(f.ϕ)
    ns = size(f.ϕ)
    Δs = (g.Δx, g.Δy, g.Δz)[1:nd]
    xs = (g.x, g.y, g.z)[1:nd]
    bc(v, a) = reshape(ROCArray(collect(v)), ntuple(i -> i == a ? ns[a] : 1, nd))
    x = Tuple(bc(xs[a], a) for a in 1:nd)
    ξ = Tuple(bc(fftfreq(ns[a], 2π / Δs[a]), a) for a in 1:nd)
    buf() = [similar(f.ϕ) for _ in 1:3]
    return HipFFTPlan{1,typeof(f),typeof(x),typeof(f.ϕ)}(f, x, ξ, buf(), buf(), buf())
end

function hip_fft!(p::HipFFTPlan, out, src, ax::Integer, dir::Cint)
    check(ccall((:sfd_fft_axis, libsfd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint),
                context(p.f).handle, devptr(src), devptr(out), ax, dir))
    return nothing
end

for (d, ax) in ((:x, 0), (:y, 1), (:z, 2))
    mul = Symbol(:mul_, d, :!)
    ldv = Symbol(:ldiv_, d, :!)
    grd = Symbol(:grid_, d)
    # PencilArray arguments, as the reference's methods take (Plans.jl:1132-1250)
    @eval SuperfluidDynamics.$mul(out::PencilArray, p::HipFFTPlan, src::PencilArray) =
        hip_fft!(p, out, src, $ax, SFD_FORWARD)
    @eval SuperfluidDynamics.$ldv(out::PencilArray, p::HipFFTPlan, src::PencilArray) =
        hip_fft!(p, out, src, $ax, SFD_BACKWARD)
    @eval SuperfluidDynamics.$grd(p::HipFFTPlan) = (p.x..., p.ξ...)
end

# Imaginary-time Strang steps with FP64 renormalisation after every step.
imag_steps!(n::NumModelSplit2{<:HipField}, nsteps::Integer) = split!(n, 2, SFD_IMAG_TIME, nsteps)

# Autotuner (like FFTW planning): fastest launch variant per axis, timed on
# scratch data; the field is not touched.
function autotune!(f::HipField)
    times = fill(-1f0, 24)
    check(ccall((:sfd_autotune, libsfd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cfloat}),
                context(f).handle, C_NULL, times))
    return reshape(times, 8, 3)
end

# Batched imaginary-time ground states of uneven sizes, 2D and 3D, with
# rotation (sfd_batch_*): advances every field until ||ψ_{n+1} - ψ_n|| < tol
# or max_steps.  With escalate (Float32 fields) each instance runs in FP32
# down to its rounding floor and finishes in FP64, so tol may be 1e-9.
# Returns (steps, ||Δψ||, finished_in_fp64) per field.
function batch_ground_states!(fields::Vector{<:HipField}, params::Vector{<:GrossPitaevskiiParameters}, Δt;
                              max_steps=20000, tol=1e-9, check_every=100, use_graph=true, tune=true,
                              escalate=true)
    ni = length(fields)
    nd = Cint[length(f.g.n) for f in fields]
    n = Int64[a <= length(f.g.n) ? f.g.n[a] : 1 for f in fields for a in 1:3]
    box = Cdouble[v for f in fields for a in 1:3
                  for v in (a <= length(f.g.n) ? (f.g.min[a], f.g.max[a]) : (0.0, 0.0))]
    FT = eltype(real(first(fields).ϕ))
    prec = FT === Float32 ? FLOAT32_MODE[] : SFD_F64
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sfd_batch_create_nd, libsfd), Cint,
                (Ptr{Ptr{Cvoid}}, Cint, Ptr{Cint}, Ptr{Int64}, Ptr{Cdouble}, Cint, Ptr{Cvoid}),
                h, ni, nd, n, box, prec, stream().stream))
    try
        tune && check(ccall((:sfd_batch_autotune, libsfd), Cint, (Ptr{Cvoid}, Ptr{Cfloat}), h[], C_NULL))
        psi = Ptr{Cvoid}[devptr(f.ϕ) for f in fields]
        V = Ptr{Cvoid}[vptr(p) for p in params]
        β = Cdouble[p.β for p in params]
        Ω = Cdouble[p.Ω for p in params]
        flags = Cint((use_graph ? SFD_BATCH_GRAPH : 0) | (escalate ? SFD_BATCH_ESCALATE : 0))
        steps = zeros(Int64, ni)
        dpsi = zeros(Cdouble, ni)
        level = zeros(Cint, ni)
        check(ccall((:sfd_batch_imag_solve, libsfd), Cint,
                    (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble, Int64,
                     Cdouble, Cint, Cint, Ptr{Int64}, Ptr{Cdouble}, Ptr{Ptr{Cvoid}}, Ptr{Cint}),
                    h[], psi, V, β, Ω, first(params).coeffΔ, Δt, max_steps, tol, check_every, flags, steps, dpsi,
                    C_NULL, level))
        return steps, dpsi, level .== 1
    finally
        ccall((:sfd_batch_destroy, libsfd), Cint, (Ptr{Cvoid},), h[])
    end
end

# Final-state fidelity L(V, β) = 1 - |<target|ψ(T)>|² after nsteps Strang
# steps of n from its current field, with the reverse-mode rule below: the
# adjoint split step (sfd_adjoint, √N checkpoint/recompute) forms dL/dV and
# dL/dβ together with L.
function adjoint_call(n::NumModelSplit2{<:HipField}, V, β, target::HipField, nsteps::Integer, every::Integer=0)
    p = n.param
    gV = ROCArray{Float64}(undef, size(n.f.ϕ))
    gβ, L = Ref{Cdouble}(0), Ref{Cdouble}(0)
    Vp = V === nothing ? C_NULL : reinterpret(Ptr{Cvoid}, pointer(V))
    check(ccall((:sfd_adjoint, libsfd), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Int64, Int64,
                 Ptr{Cvoid}, Ref{Cdouble}, Ref{Cdouble}, Ptr{Cvoid}),
                context(n.f).handle, devptr(n.f.ϕ), Vp, devptr(target.ϕ), p.coeffΔ, β, p.Ω, n.Δt, nsteps, every,
                reinterpret(Ptr{Cvoid}, pointer(gV)), gβ, L, C_NULL))
    return L[], gV, gβ[]
end

fidelity_loss(n::NumModelSplit2{<:HipField}, V, β, target::HipField, nsteps::Integer) =
    first(adjoint_call(n, V, β, target, nsteps))

function ChainRulesCore.rrule(::typeof(fidelity_loss), n::NumModelSplit2{<:HipField}, V, β, target::HipField,
                              nsteps::Integer)
    L, gV, gβ = adjoint_call(n, V, β, target, nsteps)
    pullback(ȳ) = (ChainRulesCore.NoTangent(), ChainRulesCore.NoTangent(), ȳ .* gV, ȳ * gβ,
                   ChainRulesCore.NoTangent(), ChainRulesCore.NoTangent())
    return L, pullback
end

# norm / normalize! of Fields.jl:172-200 with an FP64 reduction.
function SuperfluidDynamics.norm(f::HipField)
    s = Ref{Cdouble}(0)
    check(ccall((:sfd_norm2, libsfd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Cdouble}),
                context(f).handle, devptr(f.ϕ), s))
    return sqrt(s[]) * sqrt(prod(f.g.Δ))
end

function SuperfluidDynamics.normalize!(f::HipField)
    check(ccall((:sfd_normalize, libsfd), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}),
                context(f).handle, devptr(f.ϕ), C_NULL))
    return nothing
end

# energy of GrossPitaevskii.jl:61-136 (returns EΩ, EΔ, Eβ, E; prints μ in 2D).
function SuperfluidDynamics.energy(n::Union{NumModelSplit1{<:HipField},NumModelSplit2{<:HipField}},
                                   showEnergy=false)
    out = zeros(Cdouble, 5)
    p = n.param
    check(ccall((:sfd_energy, libsfd), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Ptr{Cdouble}),
                context(n.f).handle, devptr(n.f.ϕ), vptr(p), p.coeffΔ, p.β, p.Ω, out))
    EΩ, EΔ, Eβ, E, μ = out
    if showEnergy
        ndims(n.f.ϕ) == 2 && println("Chemical potential : $(μ)")
        println("Angular Momentum Energy : $(EΩ)")
        println("Kinetic + Potential Energy : $(EΔ)")
        println("Interaction Energy : $(Eβ)")
        println("Total Energy: $(E)")
    end
    return EΩ, EΔ, Eβ, E
end

# solve! of NumModels.jl:23-46 for the split models on HIP fields.  The
# iterations up to each backup (it % freqbckp == 0) run in one
# sfd_split_step_energy call: with Ω = 0 every step's sweeps also give the
# energy of the state before it, so no separate energy pass is made (with
# Ω ≠ 0 the library evaluates the energy before each step).
function SuperfluidDynamics.solve!(n::Union{NumModelSplit1{<:HipField},NumModelSplit2{<:HipField}};
                                   istart=1, plot=false)
    plot && error("plot=true: use the reference's solve! on host arrays")
    order = n isa NumModelSplit2 ? 2 : 1
    c = context(n.f)
    p = n.param
    res = Tuple{Int64,Int64,Float64,Float64,Float64,Float64}[]
    total = 0
    SuperfluidDynamics.write!(n.writers; prefix="res", istep=istart, Δt=n.Δt)
    it, last = istart, istart + n.niter
    while it <= last
        stop = min(last, it + mod(-it, n.freqbckp))
        k = stop - it + 1
        E = zeros(Cdouble, 5, k)
        check(ccall((:sfd_split_step_energy, libsfd), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Int64, Cint, Cint,
                     Ptr{Cdouble}),
                    c.handle, devptr(n.f.ϕ), vptr(p), p.coeffΔ, p.β, p.Ω, n.Δt, k, order, SFD_REAL_TIME, E))
        for j in 1:k
            EΩ, EΔ, Eβ, Et, μ = E[:, j]
            println("iteration $(it + j - 1)")
            ndims(n.f.ϕ) == 2 && println("Chemical potential : $(μ)")
            println("Angular Momentum Energy : $(EΩ)")
            println("Kinetic + Potential Energy : $(EΔ)")
            println("Interaction Energy : $(Eβ)")
            println("Total Energy: $(Et)")
            total += order
            push!(res, (order, total, EΩ, EΔ, Eβ, Et))
        end
        stop % n.freqbckp == 0 && SuperfluidDynamics.write!(n.writers; prefix="res", istep=stop, Δt=n.Δt)
        it = stop + 1
    end
    return res
end

end # module
