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
# by the test suite; the same entry points are exercised through ctypes.
module SuperfluidDynamicsHIP

using SuperfluidDynamics
using SuperfluidDynamics: AbstractField, Field, Grid, GrossPitaevskiiParameters,
                          NumModelSplit1, NumModelSplit2, PotentialZero
using AMDGPU: ROCArray, stream

const libsfd = normpath(joinpath(@__DIR__, "..", "sfdhip", "libsfdhip.so"))

const SFD_F32, SFD_F64 = Cint(1), Cint(2)
const SFD_REAL_TIME, SFD_IMAG_TIME = Cint(0), Cint(1)

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
    prec = FT === Float32 ? SFD_F32 : SFD_F64
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

end # module
