# HIP (MI355X / gfx950) backend for MaxwellEigenmodes: a new AbstractEigensolver whose
# solve_ω², solve_k and solve_k rrule call libomx through its C ABI (include/omx.h).
#
# Drop-in for the CPU path: same signatures, keyword arguments, layouts and return types as
#   solve_ω²   lib/MaxwellEigenmodes/src/solve.jl:38-63, solvers/gpu.jl:201-211
#   solve_k    lib/MaxwellEigenmodes/src/solve.jl:182-196, solvers/gpu.jl:213-242
#   rrule      lib/MaxwellEigenmodes/src/solve.jl:213-289, solvers/gpu.jl:292-365
#   3D cells   solve_k on Grid{3} and its rrule, rrule(solve_k_periodic) (src/grads/period.jl:144-249)
#   learned    HIPWarmNet + hip_solve_k_learned: the learned warm start (no reference counterpart)
# Host data stays Julia-owned Float64 / ComplexF64 column-major arrays; the library copies
# them to the device.  Non-zero statuses become `error(...)` with the library's message.
module MaxwellEigenmodesHIPExt

using LinearAlgebra
using Logging: AbstractLogger, NullLogger
using ChainRulesCore
using ChainRulesCore: NoTangent, AbstractZero
using DielectricSmoothing: Grid, N
import MaxwellEigenmodes: solve_ω², solve_k, solve_k_periodic, AbstractEigensolver, ModeSolver

export HIPSolver, hip_group_index, hip_ng_gvd, hip_ng_gvd_E, hip_build_modes, hip_overlap_matrix,
       hip_period_derivatives, hip_sliceinv_3x3, hip_sliceinv_3x3_pullback, HIPWarmNet, hip_warmnet_params,
       hip_warmnet_params!, hip_warmnet_forward, hip_warmnet_train!, hip_solve_k_learned

const libomx = get(ENV, "OMX_LIBRARY", "libomx.so")

struct OmxInfo
    converged::Int32
    n_newton::Int32
    lobpcg_iters_fp32::Int32
    lobpcg_iters_fp64::Int32
    n_matvec::Int64
    n_precon::Int64
    max_residual::Float64
    device_ms::Float64
end
OmxInfo() = OmxInfo(0, 0, 0, 0, 0, 0, 0.0, 0.0)

"""
    HIPSolver(; device=0, precision=:mixed, logger=NullLogger())

Eigensolver backend running every step of `solve_ω²` / `solve_k` and the `solve_k`
adjoint on an AMD MI355X through libomx.  `precision=:mixed` runs the operator in FP32 and
refines in FP64; `:fp64` runs FP64 throughout.
"""
mutable struct HIPSolver{L<:AbstractLogger} <: AbstractEigensolver{L}
    logger::L
    device::Int
    precision::Symbol
    handle::Ptr{Cvoid}
    key::Tuple{Int,Int,Int,Float64,Float64,Float64}
end

function HIPSolver(; device::Integer=0, precision::Symbol=:mixed, logger::L=NullLogger()) where {L}
    precision in (:mixed, :fp64) || error("HIPSolver precision must be :mixed or :fp64")
    s = HIPSolver{L}(logger, Int(device), precision, C_NULL, (0, 0, 0, 0.0, 0.0, 0.0))
    finalizer(s) do x
        x.handle == C_NULL || ccall((:omx_destroy, libomx), Cint, (Ptr{Cvoid},), x.handle)
        x.handle = C_NULL
    end
    return s
end

_lasterr(h) = unsafe_string(ccall((:omx_last_error, libomx), Cstring, (Ptr{Cvoid},), h))
_check(st, h) = st == 0 || error("HIPSolver (omx status $st): " * _lasterr(h))

# one handle per (grid, device); ε⁻¹ is re-uploaded when it changes (update_ε⁻¹)
function _handle!(s::HIPSolver, grid::Grid{2}, ε⁻¹::AbstractArray{<:Real,4})
    ei = Array{Float64,4}(ε⁻¹)
    key = (grid.Nx, grid.Ny, 1, Float64(grid.Δx), Float64(grid.Δy), 0.0)
    if s.handle == C_NULL || s.key != key
        s.handle == C_NULL || ccall((:omx_destroy, libomx), Cint, (Ptr{Cvoid},), s.handle)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:omx_create, libomx), Cint,
                   (Ref{Ptr{Cvoid}}, Cint, Cint, Cint, Cdouble, Cdouble, Ptr{Cdouble}),
                   h, s.device, grid.Nx, grid.Ny, grid.Δx, grid.Δy, ei)
        _check(st, C_NULL)
        s.handle, s.key = h[], key
    else
        _check(ccall((:omx_set_epsinv, libomx), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), s.handle, ei), s.handle)
    end
    _check(ccall((:omx_set_option, libomx), Cint, (Ptr{Cvoid}, Cstring, Cdouble),
                 s.handle, "fp64_only", s.precision === :fp64 ? 1.0 : 0.0), s.handle)
    return s.handle
end

# a 3D cell periodic along z (Grid{3}; Δz is the period Λ of grads/period.jl)
function _handle!(s::HIPSolver, grid::Grid{3}, ε⁻¹::AbstractArray{<:Real,5})
    ei = Array{Float64,5}(ε⁻¹)
    key = (grid.Nx, grid.Ny, grid.Nz, Float64(grid.Δx), Float64(grid.Δy), Float64(grid.Δz))
    if s.handle == C_NULL || s.key != key
        s.handle == C_NULL || ccall((:omx_destroy, libomx), Cint, (Ptr{Cvoid},), s.handle)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:omx_create_3d, libomx), Cint,
                   (Ref{Ptr{Cvoid}}, Cint, Cint, Cint, Cint, Cdouble, Cdouble, Cdouble, Ptr{Cdouble}),
                   h, s.device, grid.Nx, grid.Ny, grid.Nz, grid.Δx, grid.Δy, grid.Δz, ei)
        _check(st, C_NULL)
        s.handle, s.key = h[], key
    else
        _check(ccall((:omx_set_epsinv, libomx), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), s.handle, ei), s.handle)
    end
    _check(ccall((:omx_set_option, libomx), Cint, (Ptr{Cvoid}, Cstring, Cdouble),
                 s.handle, "fp64_only", s.precision === :fp64 ? 1.0 : 0.0), s.handle)
    return s.handle
end

function solve_ω²(ms::ModeSolver{2,T}, solver::HIPSolver; nev=1, maxiter=100, tol=1e-8, log=false,
                  f_filter=nothing) where {T<:Real}
    h = _handle!(solver, ms.grid, ms.M̂.ε⁻¹)
    H = Matrix{ComplexF64}(ms.H⃗[:, 1:nev])
    ω² = Vector{Float64}(undef, nev)
    info = Ref(OmxInfo())
    st = ccall((:omx_solve_omega2, libomx), Cint,
               (Ptr{Cvoid}, Cdouble, Cint, Cdouble, Cint, Ptr{ComplexF64}, Ptr{Cdouble}, Ref{OmxInfo}),
               h, ms.M̂.k⃗[3], nev, tol, maxiter, H, ω², info)
    _check(st, h)
    ms.H⃗[:, 1:nev] .= H
    ms.ω²[1:nev] .= complex.(ω²)
    return ω², [H[:, j] for j in 1:nev]
end

function solve_k(ω::Real, ε⁻¹::Union{AbstractArray{Float64,4},AbstractArray{Float64,5}},
                 grid::Union{Grid{2,Float64},Grid{3,Float64}}, solver::HIPSolver; nev=1,
                 max_eigsolves=60, maxiter=100, k_tol=1e-8, eig_tol=1e-8, log=false, kguess=nothing,
                 Hguess=nothing, f_filter=nothing, overwrite=false)
    h = _handle!(solver, grid, ε⁻¹)
    n2 = 2 * N(grid)
    kmags = Vector{Float64}(undef, nev)
    evecs = Matrix{ComplexF64}(undef, n2, nev)
    cols = Hguess === nothing ? Vector{ComplexF64}[] :
           Hguess isa AbstractMatrix ? [Vector{ComplexF64}(c) for c in eachcol(Hguess)] :
           eltype(Hguess) <: Number ? [Vector{ComplexF64}(Hguess)] : [Vector{ComplexF64}(c) for c in Hguess]
    cols = [c for c in cols if length(c) == n2][1:min(end, nev)]
    Hg = isempty(cols) ? Matrix{ComplexF64}(undef, n2, 0) : reduce(hcat, cols)
    info = Ref(OmxInfo())
    kgv = kguess === nothing ? nothing : Float64[first(kguess)]
    GC.@preserve kgv Hg begin
        st = ccall((:omx_solve_k, libomx), Cint,
                   (Ptr{Cvoid}, Cdouble, Cint, Cdouble, Cdouble, Cint, Cint, Ptr{Cdouble}, Ptr{ComplexF64}, Cint,
                    Ptr{Cdouble}, Ptr{ComplexF64}, Ref{OmxInfo}),
                   h, ω, nev, k_tol, eig_tol, max_eigsolves, maxiter,
                   kgv === nothing ? Ptr{Cdouble}(C_NULL) : kgv,
                   isempty(cols) ? Ptr{ComplexF64}(C_NULL) : Hg, length(cols), kmags, evecs, info)
    end
    _check(st, h)
    return kmags, [evecs[:, j] for j in 1:nev]
end

# k̄ / ēvecs cotangents (AbstractZero entries are zero) packed for the C-ABI pull-backs
function _pack_cotangents(ΔΩ, nev, n2)
    k̄mags, ēvecs = ΔΩ
    kb = zeros(Float64, nev)
    if !(k̄mags isa AbstractZero)
        for (j, v) in enumerate(k̄mags)
            kb[j] = v isa AbstractZero ? 0.0 : Float64(v)
        end
    end
    eb = zeros(ComplexF64, n2, nev)
    mask = zeros(UInt8, nev)
    if !(ēvecs isa AbstractZero)
        for (j, v) in enumerate(ēvecs)
            v isa AbstractZero && continue
            eb[:, j] .= v
            mask[j] = 0x01
        end
    end
    return kb, eb, mask
end

function ChainRulesCore.rrule(::typeof(solve_k), ω::Real, ε⁻¹::Union{AbstractArray{Float64,4},AbstractArray{Float64,5}},
                              grid::Union{Grid{2,Float64},Grid{3,Float64}}, solver::HIPSolver; nev=1, kwargs...)
    kmags, evecs = solve_k(ω, ε⁻¹, grid, solver; nev, kwargs...)
    function hip_solve_k_pullback(ΔΩ)
        kb, eb, mask = _pack_cotangents(ΔΩ, nev, 2 * N(grid))
        E = reduce(hcat, evecs)
        ω̄ = Ref(0.0)
        ε̄⁻¹ = zeros(Float64, size(ε⁻¹))
        h = _handle!(solver, grid, ε⁻¹)
        st = ccall((:omx_solve_k_adjoint, libomx), Cint,
                   (Ptr{Cvoid}, Cdouble, Cint, Ptr{Cdouble}, Ptr{ComplexF64}, Ptr{Cdouble}, Ptr{ComplexF64},
                    Ptr{UInt8}, Ref{Cdouble}, Ptr{Cdouble}),
                   h, ω, nev, kmags, E, kb, eb, mask, ω̄, ε̄⁻¹)
        _check(st, h)
        return (NoTangent(), ω̄[], ε̄⁻¹, NoTangent(), NoTangent())
    end
    return (kmags, evecs), hip_solve_k_pullback
end

"""
    hip_period_derivatives(solver, kmags, evecs, ε⁻¹, grid::Grid{3}) -> (∂ω²∂k, ∂ω²∂Λ)

Per mode, `2 HMₖH(ev)` and `∂ω²_∂Λ(ev)` (src/grads/period.jl:69-90) on the GPU.
"""
function hip_period_derivatives(solver::HIPSolver, kmags::AbstractVector{<:Real}, evecs,
                                ε⁻¹::AbstractArray{Float64,5}, grid::Grid{3,Float64})
    h = _handle!(solver, grid, ε⁻¹)
    km = Vector{Float64}(kmags)
    E = reduce(hcat, [Vector{ComplexF64}(v) for v in evecs])
    dk = Vector{Float64}(undef, length(km))
    dL = Vector{Float64}(undef, length(km))
    st = ccall((:omx_period_derivatives, libomx), Cint,
               (Ptr{Cvoid}, Cint, Ptr{Cdouble}, Ptr{ComplexF64}, Ptr{Cdouble}, Ptr{Cdouble}),
               h, length(km), km, E, dk, dL)
    _check(st, h)
    return dk, dL
end

# rrule(solve_k_periodic, ...) (src/grads/period.jl:236-249): both channels of
# _solve_k_period_grads (:163-234) on the GPU -- ω̄, ε̄⁻¹ (3D convention: off-diagonal sums
# above the diagonal) and Λ̄.
function ChainRulesCore.rrule(::typeof(solve_k_periodic), ω::Real, ε⁻¹::AbstractArray{Float64,5}, Λ::Real,
                              grid::Grid{3,Float64}, solver::HIPSolver; nev=1, kwargs...)
    g = Grid(grid.Δx, grid.Δy, Float64(Λ), grid.Nx, grid.Ny, grid.Nz)
    kmags, evecs = solve_k(ω, ε⁻¹, g, solver; nev, kwargs...)
    function hip_solve_k_periodic_pullback(ΔΩ)
        kb, eb, mask = _pack_cotangents(ΔΩ, nev, 2 * N(g))
        E = reduce(hcat, evecs)
        ω̄ = Ref(0.0)
        Λ̄ = Ref(0.0)
        ε̄⁻¹ = zeros(Float64, size(ε⁻¹))
        h = _handle!(solver, g, ε⁻¹)
        st = ccall((:omx_solve_k_periodic_adjoint, libomx), Cint,
                   (Ptr{Cvoid}, Cdouble, Cint, Ptr{Cdouble}, Ptr{ComplexF64}, Ptr{Cdouble}, Ptr{ComplexF64},
                    Ptr{UInt8}, Ref{Cdouble}, Ptr{Cdouble}, Ref{Cdouble}),
                   h, ω, nev, kmags, E, kb, eb, mask, ω̄, ε̄⁻¹, Λ̄)
        _check(st, h)
        return (NoTangent(), ω̄[], ε̄⁻¹, Λ̄[], NoTangent(), NoTangent())
    end
    return (kmags, evecs), hip_solve_k_periodic_pullback
end

"""
    hip_sliceinv_3x3(A; device=0)

`sliceinv_3x3(A)` (src/linalg.jl:126-149) on the GPU: the per-pixel inverse of a
(3,3,Nx,Ny[,Nz]) tensor field.
"""
function hip_sliceinv_3x3(A::AbstractArray{Float64}; device::Integer=0)
    a = Array{Float64}(A)
    y = similar(a)
    st = ccall((:omx_sliceinv_3x3, libomx), Cint, (Cint, Int64, Ptr{Cdouble}, Ptr{Cdouble}),
               device, length(a) ÷ 9, a, y)
    _check(st, C_NULL)
    return y
end

"""
    hip_sliceinv_3x3_pullback(Y, Ȳ; device=0)

Pull-back of `sliceinv_3x3` (src/grads/linalg.jl:40-47): Ā = -Yᵀ Ȳ Yᵀ per pixel, Y = A⁻¹.
"""
function hip_sliceinv_3x3_pullback(Y::AbstractArray{Float64}, Ȳ::AbstractArray{Float64}; device::Integer=0)
    size(Y) == size(Ȳ) || error("hip_sliceinv_3x3_pullback: Y and Ȳ differ in shape")
    y = Array{Float64}(Y)
    yb = Array{Float64}(Ȳ)
    ab = similar(y)
    st = ccall((:omx_sliceinv_3x3_pullback, libomx), Cint, (Cint, Int64, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
               device, length(y) ÷ 9, y, yb, ab)
    _check(st, C_NULL)
    return ab
end

"""
    hip_group_index(solver, kmags, evecs, ω, ε⁻¹, ∂ε_∂ω, grid)

`group_index` (lib/ModeAnalysis/src/group_index.jl:21-29) for every mode on the GPU:
n_g = (ω + HMH(ev, ε⁻¹ ∂ωε ε⁻¹)/2) / HMₖH(ev); `∂ε_∂ω = nothing` for a non-dispersive medium.
"""
function hip_group_index(solver::HIPSolver, kmags::AbstractVector{<:Real}, evecs, ω::Real,
                         ε⁻¹::AbstractArray{Float64,4}, ∂ε_∂ω, grid::Grid{2,Float64})
    h = _handle!(solver, grid, ε⁻¹)
    km = Vector{Float64}(kmags)
    E = reduce(hcat, [Vector{ComplexF64}(v) for v in evecs])
    D = ∂ε_∂ω === nothing ? nothing : Array{Float64,4}(∂ε_∂ω)
    ng = Vector{Float64}(undef, length(km))
    GC.@preserve D begin
        st = ccall((:omx_group_index, libomx), Cint,
                   (Ptr{Cvoid}, Cdouble, Cint, Ptr{Cdouble}, Ptr{ComplexF64}, Ptr{Cdouble}, Ptr{Cdouble}),
                   h, ω, length(km), km, E, D === nothing ? Ptr{Cdouble}(C_NULL) : D, ng)
    end
    _check(st, h)
    return ng
end

function _ng_gvd(solver, kmags, evecs, ω, ε⁻¹, ∂ε_∂ω, ∂²ε_∂ω², grid, want_E::Bool)
    h = _handle!(solver, grid, ε⁻¹)
    km = Vector{Float64}(kmags)
    E = reduce(hcat, [Vector{ComplexF64}(v) for v in evecs])
    D1 = Array{Float64,4}(∂ε_∂ω)
    D2 = Array{Float64,4}(∂²ε_∂ω²)
    ng = Vector{Float64}(undef, length(km))
    gvd = Vector{Float64}(undef, length(km))
    Ef = want_E ? Array{ComplexF64,4}(undef, 3, size(grid)..., length(km)) : nothing
    GC.@preserve Ef begin
        st = ccall((:omx_ng_gvd, libomx), Cint,
                   (Ptr{Cvoid}, Cdouble, Cint, Ptr{Cdouble}, Ptr{ComplexF64}, Ptr{Cdouble}, Ptr{Cdouble},
                    Ptr{Cdouble}, Ptr{Cdouble}, Ptr{ComplexF64}),
                   h, ω, length(km), km, E, D1, D2, ng, gvd, Ef === nothing ? Ptr{ComplexF64}(C_NULL) : Ef)
    end
    _check(st, h)
    return ng, gvd, Ef
end

"""
    hip_ng_gvd(solver, kmags, evecs, ω, ε⁻¹, ∂ε_∂ω, ∂²ε_∂ω², grid)

`ng_gvd` (lib/ModeAnalysis/src/group_index.jl:126-162) for every mode on the GPU: returns
the vectors `(ng, gvd)`, `gvd = ∂n_g/∂ω` from one linear solve per mode.
"""
function hip_ng_gvd(solver::HIPSolver, kmags::AbstractVector{<:Real}, evecs, ω::Real,
                    ε⁻¹::AbstractArray{Float64,4}, ∂ε_∂ω, ∂²ε_∂ω², grid::Grid{2,Float64})
    ng, gvd, _ = _ng_gvd(solver, kmags, evecs, ω, ε⁻¹, ∂ε_∂ω, ∂²ε_∂ω², grid, false)
    return ng, gvd
end

"""
    hip_ng_gvd_E(solver, kmags, evecs, ω, ε⁻¹, ∂ε_∂ω, ∂²ε_∂ω², grid)

`ng_gvd_E` (group_index.jl:55-95): `(ng, gvd, E)` with `E[:, :, :, j]` the real-space
field ε⁻¹ fft(kx_tc(ev_j)) of mode j, size (3, Nx, Ny).
"""
function hip_ng_gvd_E(solver::HIPSolver, kmags::AbstractVector{<:Real}, evecs, ω::Real,
                      ε⁻¹::AbstractArray{Float64,4}, ∂ε_∂ω, ∂²ε_∂ω², grid::Grid{2,Float64})
    return _ng_gvd(solver, kmags, evecs, ω, ε⁻¹, ∂ε_∂ω, ∂²ε_∂ω², grid, true)
end

"""
    hip_build_modes(solver, slot, kmags, ε⁻¹, grid; evecs=nothing, conjugate=false, normalize=true)

`build_mode` (lib/EigenmodeExpansion/src/modes.jl:116-122) for every mode into device mode
set `slot` (0..3): E = i ε⁻¹ fft(kx_tc(ev)), H = fft(tc(ev)), both divided by
sqrt(inner_product(m, m)).  `evecs = nothing` uses the eigenvectors of the handle's last
`solve_k`.  Returns the complex norms.
"""
function hip_build_modes(solver::HIPSolver, slot::Integer, kmags::AbstractVector{<:Real},
                         ε⁻¹::AbstractArray{Float64,4}, grid::Grid{2,Float64}; evecs=nothing,
                         conjugate::Bool=false, normalize::Bool=true)
    h = _handle!(solver, grid, ε⁻¹)
    km = Vector{Float64}(kmags)
    E = evecs === nothing ? nothing : reduce(hcat, [Vector{ComplexF64}(v) for v in evecs])
    norms = Vector{ComplexF64}(undef, length(km))
    GC.@preserve E begin
        st = ccall((:omx_build_modes, libomx), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}, Ptr{ComplexF64}, Cint, Cint, Ptr{ComplexF64}),
                   h, slot, length(km), km, E === nothing ? Ptr{ComplexF64}(C_NULL) : E, conjugate, normalize, norms)
    end
    _check(st, h)
    return norms
end

"""
    hip_overlap_matrix(solver, slot_a, na, slot_b, nb; conjugate=false)

`overlap_matrix(modes_a, modes_b; conjugate)` (modes.jl:156-170) of two device mode sets.
"""
function hip_overlap_matrix(solver::HIPSolver, slot_a::Integer, na::Integer, slot_b::Integer, nb::Integer;
                            conjugate::Bool=false)
    h = solver.handle
    O = Matrix{ComplexF64}(undef, na, nb)
    st = ccall((:omx_overlap_matrix, libomx), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Ptr{ComplexF64}),
               h, slot_a, slot_b, conjugate, O)
    _check(st, h)
    return O
end

# ---------------------------------------------------------------------------------------
# Learned warm start (omx_warmnet_*): a small CNN on the GPU predicting per-band effective
# indices and low-plane-wave start vectors from the ε⁻¹ map and ω.  Arrays follow the C
# layout [batch][channel][a][b] (b fastest), i.e. Julia arrays of size (b, a, channel, batch);
# the maps here use a = block index along x, b = block index along y.  The normalisation
# constants are those of omx/warmstart.py, with which the shipped network was trained.
# ---------------------------------------------------------------------------------------

const WS_N = (1.8, 0.2)          # effective index = N[1] + N[2] * output
const WS_OMEGA = (0.65, 0.05)
const WS_DIAG = (0.55, 0.25)
const WS_OFF = 0.05

mutable struct HIPWarmNet
    handle::Ptr{Cvoid}
    hw::Int
    nev::Int
end

function _wcheck(st)
    st == 0 || error("omx warm-start network: ", unsafe_string(ccall((:omx_last_error, libomx), Cstring, (Ptr{Cvoid},), C_NULL)))
end

"""
    HIPWarmNet(hw=64, nev=8; seed=0, device=0)

The warm-start network on the GPU; load trained parameters with `hip_warmnet_params!`
(the shipped `omx/data/warmnet_tfln256.npz` holds them under `params`).
"""
function HIPWarmNet(hw::Integer=64, nev::Integer=8; seed::Integer=0, device::Integer=0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    _wcheck(ccall((:omx_warmnet_create, libomx), Cint, (Cint, Cint, Cint, UInt64, Ref{Ptr{Cvoid}}),
                  device, hw, nev, seed, h))
    n = HIPWarmNet(h[], Int(hw), Int(nev))
    finalizer(n) do x
        x.handle == C_NULL || ccall((:omx_warmnet_destroy, libomx), Cint, (Ptr{Cvoid},), x.handle)
        x.handle = C_NULL
    end
    return n
end

function hip_warmnet_params(net::HIPWarmNet)
    n = Ref{Int64}(0)
    _wcheck(ccall((:omx_warmnet_num_params, libomx), Cint, (Ptr{Cvoid}, Ref{Int64}), net.handle, n))
    p = Vector{Float32}(undef, n[])
    _wcheck(ccall((:omx_warmnet_get_params, libomx), Cint, (Ptr{Cvoid}, Ptr{Float32}), net.handle, p))
    return p
end

function hip_warmnet_params!(net::HIPWarmNet, p::AbstractVector{<:Real})
    q = Vector{Float32}(p)
    length(q) == length(hip_warmnet_params(net)) || error("wrong parameter count")
    _wcheck(ccall((:omx_warmnet_set_params, libomx), Cint, (Ptr{Cvoid}, Ptr{Float32}), net.handle, q))
    return net
end

"""
    hip_warmnet_forward(net, x) -> (fields, k)

x: (hw, hw, 7, B) Float32 -> fields (hw, hw, 4nev, B), k (nev, B).
"""
function hip_warmnet_forward(net::HIPWarmNet, x::AbstractArray{<:Real,4})
    X = Array{Float32}(x)
    size(X)[1:3] == (net.hw, net.hw, 7) || error("expected (hw, hw, 7, B) inputs")
    B = size(X, 4)
    Y = Array{Float32}(undef, net.hw, net.hw, 4 * net.nev, B)
    K = Array{Float32}(undef, net.nev, B)
    _wcheck(ccall((:omx_warmnet_forward, libomx), Cint, (Ptr{Cvoid}, Cint, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}),
                  net.handle, B, X, Y, K))
    return Y, K
end

"""
    hip_warmnet_train!(net, x, t_fields, t_k; wk=1, epochs=100, batch=16, lr=1e-3, seed=0) -> loss per epoch

Adam on the GPU over the whole set (x: (hw, hw, 7, n), t_fields: (hw, hw, 4nev, n), t_k: (nev, n)).
"""
function hip_warmnet_train!(net::HIPWarmNet, x, tf, tk; wk::Real=1.0, epochs::Integer=100, batch::Integer=16,
                            lr::Real=1e-3, seed::Integer=0)
    X, TF, TK = Array{Float32}(x), Array{Float32}(tf), Array{Float32}(tk)
    n = size(X, 4)
    hist = Vector{Float64}(undef, epochs)
    _wcheck(ccall((:omx_warmnet_train, libomx), Cint,
                  (Ptr{Cvoid}, Cint, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Cfloat, Cint, Cint, Cfloat, UInt64,
                   Ptr{Float64}), net.handle, n, X, TF, TK, wk, epochs, batch, lr, seed, hist))
    return hist
end

"""
    hip_solve_k_learned(ω, ε⁻¹, grid, solver, net; nev=net.nev, k0_scale=1.0, kwargs...)

`solve_k` from the network's guess, built on the GPU from the solver's ε⁻¹ (features, one
forward, inverse DFT into the start block's lowest plane waves; `omx_solve_k_learned`): the
shared first eigensolve starts at `k0_scale` times the predicted band-1 k; the stopping rules
and the converged result are those of `solve_k`.
"""
function hip_solve_k_learned(ω::Real, ε⁻¹::AbstractArray{Float64,4}, grid::Grid{2,Float64}, solver::HIPSolver,
                             net::HIPWarmNet; nev::Integer=net.nev, k0_scale::Real=1.0, max_eigsolves=60,
                             maxiter=100, k_tol=1e-8, eig_tol=1e-8)
    h = _handle!(solver, grid, ε⁻¹)
    kmags = Vector{Float64}(undef, nev)
    evecs = Matrix{ComplexF64}(undef, 2 * N(grid), nev)
    info = Ref(OmxInfo())
    st = ccall((:omx_solve_k_learned, libomx), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint, Cdouble, Cdouble, Cdouble, Cint, Cint, Ptr{Cdouble},
                Ptr{ComplexF64}, Ref{OmxInfo}),
               h, net.handle, ω, nev, k0_scale, k_tol, eig_tol, max_eigsolves, maxiter, kmags, evecs, info)
    _check(st, h)
    return kmags, [evecs[:, j] for j in 1:nev]
end

end # module
