# AOKHip.jl -- drop-in `ccall` shim that routes AtomicOrbitalKernels.jl's kernel-facing
# layer to libaok.so (include/aok.h).  Loaded as a package extension when AMDGPU.jl is
# present; the public API (`evaluate`, `evaluate_ed`, `compile_basis`, `batch_overlap[!]`,
# `batch_overlap_3c[!]`, `pullback_ps`, both `rrule`s, the Lux ps/st layout) is unchanged:
# only the L3 methods that launch KernelAbstractions kernels get ROCArray methods here.
#
# This file cannot be executed in the build container (no Julia); its contract is
# exercised through the same C ABI by the Python tests.
module AOKHip

using AMDGPU: ROCArray, ROCVector, ROCMatrix, stream
using StaticArrays: SVector
using DecoratedParticles: PState
import Unitful
import ChainRulesCore: rrule, NoTangent, unthunk
import AtomicOrbitalKernels as AOK
import AtomicOrbitalKernels: evaluate, evaluate_ed, batch_overlap!, batch_overlap_3c!, pullback_ps

const libaok = joinpath(@__DIR__, "..", "lib", "libaok.so")

const AOK_OK = Cint(0)
const AOK_F32 = Cint(0); const AOK_F64 = Cint(1)
const AOK_I64 = Cint(0); const AOK_I32 = Cint(1)
const AOK_FAST = Cuint(0)

_dt(::Type{Float32}) = AOK_F32
_dt(::Type{Float64}) = AOK_F64

mutable struct Ctx
    h::Ptr{Cvoid}
end
const _CTX = Ref{Union{Nothing, Ctx}}(nothing)

function ctx()
    if _CTX[] === nothing
        h = Ref{Ptr{Cvoid}}(C_NULL)
        # borrow AMDGPU's task-local stream so ordering with Julia-side work is preserved
        st = Ptr{Cvoid}(stream().stream)
        rc = ccall((:aok_create, libaok), Cint, (Ref{Ptr{Cvoid}}, Cint, Ptr{Cvoid}), h, 0, st)
        rc == AOK_OK || error("aok_create failed ($rc)")
        _CTX[] = Ctx(h[])
    end
    return _CTX[]::Ctx
end

# status -> the reference's exception types (errors are raised before any launch there)
function check(rc, what)
    rc == AOK_OK && return nothing
    msg = unsafe_string(ccall((:aok_last_error, libaok), Cstring, (Ptr{Cvoid},), ctx().h))
    rc == 2 && error("$what: $msg")                  # ErrorException (unknown species)
    rc == 1 && throw(ArgumentError("$what: $msg"))
    error("$what: $msg (status $rc)")
end

_ip(v::AbstractVector{<:Integer}) = Vector{Int64}(Array(v))   # host int64 (structural arrays)

# ---- H1: evaluate / evaluate_ed (atomicorbitals.jl:148-162, 192-210) -------------------
function _ao_eval(basis, Rs::ROCVector{SVector{3, T}}, sidx::ROCVector{<:Integer}, ps, st, grad) where {T}
    nX = length(Rs); nOrb = length(basis)
    ζ = ps.Rnl.ζ; D = ps.Rnl.D
    nRad, K, NZ = size(ζ)
    P = similar(Rs, T, nX, nOrb)
    dP = grad ? similar(Rs, SVector{3, T}, nX, nOrb) : nothing
    it = eltype(sidx) == Int32 ? AOK_I32 : AOK_I64
    poly = _ip(st.Rnl.poly); iR = _ip(st.iR); iY = _ip(st.iY)
    L = length(st.Ylm.Flm) > 0 ? AOK._ylm_degree(basis.Ylm) : 0
    decay = basis.Rnl isa AOK.GaussianTypeRadials ? Cint(0) : Cint(1)
    rc = ccall((:aok_ao_eval, libaok), Cint,
               (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64,
                Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64, Cint, Cdouble, Cint, Ptr{Cvoid}, Ptr{Cvoid}),
               ctx().h, _dt(T), pointer(Rs), pointer(sidx), it, nX, pointer(ζ), pointer(D), nRad, K, NZ,
               poly, iR, iY, nOrb, L, Float64(basis.lengthscale), decay, pointer(P),
               grad ? pointer(dP) : C_NULL)
    check(rc, grad ? "evaluate_ed" : "evaluate")
    return grad ? (P, dP) : P
end

evaluate(basis::AOK.AtomicOrbitals, Rs::ROCVector{<:SVector{3}}, sidx::ROCVector{<:Integer}, ps, st) =
    _ao_eval(basis, Rs, sidx, ps, st, false)
evaluate_ed(basis::AOK.AtomicOrbitals, Rs::ROCVector{<:SVector{3}}, sidx::ROCVector{<:Integer}, ps, st) =
    _ao_eval(basis, Rs, sidx, ps, st, true)

# ---- H1 pullback: pullback_ps + _aorb_pullback_x (atomicorbitals.jl:254-299) ----------
# One aok_ao_pullback call gives the parameter cotangents and, with want_x, the position
# cotangent sum_o dP[i,o] grad P[i,o] that _aorb_pullback_x forms from a stored dP.
function _ao_pullback(∂P::ROCMatrix{T}, basis::AOK.AtomicOrbitals, Rs::ROCVector{SVector{3, T}},
                      sidx::ROCVector{<:Integer}, ps, st, want_x::Bool) where {T}
    ζ = ps.Rnl.ζ; D = ps.Rnl.D
    nRad, K, NZ = size(ζ)
    ∂ζ = similar(ζ); ∂D = similar(D)
    ∂Rs = want_x ? similar(Rs) : nothing
    rc = ccall((:aok_ao_pullback, libaok), Cint,
               (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Int64,
                Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64, Cint, Cdouble, Cint, Ptr{Cvoid}, Ptr{Cvoid},
                Ptr{Cvoid}, Ptr{Cvoid}),
               ctx().h, _dt(T), pointer(Rs), pointer(sidx), eltype(sidx) == Int32 ? AOK_I32 : AOK_I64,
               length(Rs), pointer(ζ), pointer(D), nRad, K, NZ, _ip(st.Rnl.poly), _ip(st.iR), _ip(st.iY),
               length(basis), AOK._ylm_degree(basis.Ylm), Float64(basis.lengthscale),
               basis.Rnl isa AOK.GaussianTypeRadials ? Cint(0) : Cint(1), pointer(∂P),
               want_x ? pointer(∂Rs) : C_NULL, pointer(∂ζ), pointer(∂D))
    check(rc, want_x ? "rrule(evaluate) pullback" : "pullback_ps")
    return ∂Rs, (Rnl = (ζ = ∂ζ, D = ∂D), Ylm = NamedTuple())
end

pullback_ps(∂P::ROCMatrix{T}, basis::AOK.AtomicOrbitals, Rs::ROCVector{SVector{3, T}},
            sidx::ROCVector{<:Integer}, ps::NamedTuple, st) where {T} =
    _ao_pullback(∂P, basis, Rs, sidx, ps, st, false)[2]

# rrule(evaluate, basis, X, ps, st) (atomicorbitals.jl:307-328) dispatches to _aorb_rrule
# with the extracted device positions; the forward keeps no dP (the pullback kernel forms
# the spatial Jacobian in registers), the cotangent mirrors the input kind (_xtangent).
function AOK._aorb_rrule(basis::AOK.AtomicOrbitals, X, Rs::ROCVector{SVector{3, T}},
                         sidx::ROCVector{<:Integer}, ps, st) where {T}
    P = _ao_eval(basis, Rs, sidx, ps, st, false)
    function _aorb_pullback(_∂)
        ∂P = ROCArray{T}(unthunk(_∂))
        ∂Rs, ∂ps = _ao_pullback(∂P, basis, Rs, sidx, ps, st, true)
        return (NoTangent(), NoTangent(), AOK._xtangent(X, ∂Rs), ∂ps, NoTangent())
    end
    return P, _aorb_pullback
end

# ---- H2: batch_overlap!(out, ::CartesianGTOBasis, XA, XB) (overlap_2c.jl:171-197) ----------
function batch_overlap!(out::ROCArray{FT, 3}, basis::AOK.CartesianGTOBasis, XA::AbstractVector{<:PState},
                        XB::AbstractVector{<:PState}; kwargs...) where {FT}
    B = length(XA)
    length(XB) == B || error("XA and XB must have equal length; got $(length(XA)) and $(length(XB))")
    posA = ROCArray(AOK._cgto_posmat(XA, FT)); posB = ROCArray(AOK._cgto_posmat(XB, FT))
    sA = ROCArray(AOK._cgto_sidx(basis, XA)); sB = ROCArray(AOK._cgto_sidx(basis, XB))
    coef = ROCArray(FT.(basis.coef)); α = ROCArray(FT.(basis.ζ))
    ns, K, NZ = size(basis.coef)
    rc = ccall((:aok_ovl2c, libaok), Cint,
               (Ptr{Cvoid}, Cint, Cuint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Int64, Int64,
                Int64, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}),
               ctx().h, _dt(FT), AOK_FAST, pointer(posA), pointer(posB), pointer(sA), pointer(sB), AOK_I64, B, ns,
               K, NZ, _ip(basis.ls), pointer(coef), pointer(α), Float64(basis.lengthscale), pointer(out))
    check(rc, "batch_overlap!")
    return out
end

# ---- H2 kernel pullback: _overlap_pb_coef_alpha(∂S, basis, XA, XB) (overlap_2c_grad.jl:218-242)
function AOK._overlap_pb_coef_alpha(∂S::ROCArray{FT, 3}, basis::AOK.CartesianGTOBasis,
                                    XA::AbstractVector{<:PState}, XB::AbstractVector{<:PState};
                                    kwargs...) where {FT}
    B = length(XA)
    posA = ROCArray(AOK._cgto_posmat(XA, FT)); posB = ROCArray(AOK._cgto_posmat(XB, FT))
    sA = ROCArray(AOK._cgto_sidx(basis, XA)); sB = ROCArray(AOK._cgto_sidx(basis, XB))
    coef = ROCArray(FT.(basis.coef)); α = ROCArray(FT.(basis.ζ))
    ns, K, NZ = size(basis.coef)
    ∂c = similar(coef); ∂α = similar(α)
    rc = ccall((:aok_ovl2c_pullback, libaok), Cint,
               (Ptr{Cvoid}, Cint, Cuint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64,
                Int64, Int64, Int64, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}),
               ctx().h, _dt(FT), AOK_FAST, pointer(∂S), pointer(posA), pointer(posB), pointer(sA), pointer(sB),
               AOK_I64, B, ns, K, NZ, _ip(basis.ls), pointer(coef), pointer(α), Float64(basis.lengthscale),
               pointer(∂c), pointer(∂α))
    check(rc, "_overlap_pb_coef_alpha")
    return ∂c, ∂α
end

# ---- H2 pullback: pullback_ps(∂S, orb, XA, XB, ps, st) (overlap_2c_grad.jl:296-304) ------
# kernel pullback + normalisation VJP in one call
function pullback_ps(∂S::ROCArray{FT, 3}, orb::AOK.GaussianTypeOrbitals, XA::AbstractVector{<:PState},
                     XB::AbstractVector{<:PState}, ps, st) where {FT}
    str = AOK._compile_struct(orb)
    B = length(XA)
    posA = ROCArray(AOK._cgto_posmat(XA, FT)); posB = ROCArray(AOK._cgto_posmat(XB, FT))
    zl = orb.Rnl.zlist
    sA = ROCArray([AOK._z2i(orb.Rnl, x.S) for x in XA]); sB = ROCArray([AOK._z2i(orb.Rnl, x.S) for x in XB])
    ζ = ps.Rnl.ζ; D = ps.Rnl.D
    ns, K, NZ = size(ζ)
    ∂ζ = similar(ζ); ∂D = similar(D)
    rc = ccall((:aok_ovl2c_pullback_ps, libaok), Cint,
               (Ptr{Cvoid}, Cint, Cuint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64,
                Int64, Int64, Int64, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}),
               ctx().h, _dt(FT), AOK_FAST, pointer(∂S), pointer(posA), pointer(posB), pointer(sA), pointer(sB),
               AOK_I64, B, ns, K, NZ, _ip(str.ls), pointer(ζ), pointer(D), Float64(orb.lengthscale),
               pointer(∂ζ), pointer(∂D))
    check(rc, "pullback_ps")
    return (Rnl = (ζ = ∂ζ, D = ∂D), Ylm = NamedTuple())
end

# ---- H2 rrule: forward + pullback on one plan (overlap_2c_grad.jl:309-318) -------------
# Taken when the parameters live on the device (ps.Rnl.ζ, ps.Rnl.D ROCArrays): strictly
# more specific than the reference's method, which stays in charge of host parameters.
const AOK_PARAMS_PS = Cint(1)
const RocPS = NamedTuple{(:Rnl, :Ylm), <:Tuple{NamedTuple{(:ζ, :D), <:Tuple{ROCArray, ROCArray}}, Any}}

mutable struct Plan2c
    h::Ptr{Cvoid}
    keep::Any          # sidx / parameter arrays the plan references until destroyed
end
function _plan_destroy(p::Plan2c)
    p.h == C_NULL || ccall((:aok_ovl2c_plan_destroy, libaok), Cint, (Ptr{Cvoid},), p.h)
    p.h = C_NULL
end

# the plan of one (orb, XA, XB, ps): device copies of positions, species and (ζ, D)
function _plan2c(orb::AOK.GaussianTypeOrbitals, XA::AbstractVector, XB::AbstractVector, ps)
    str = AOK._compile_struct(orb)
    FT = promote_type(eltype(ps.Rnl.ζ), eltype(ps.Rnl.D))
    B = length(XA)
    length(XB) == B || error("XA and XB must have equal length; got $(length(XA)) and $(length(XB))")
    posA = ROCArray(AOK._cgto_posmat(XA, FT)); posB = ROCArray(AOK._cgto_posmat(XB, FT))
    sA = ROCArray([AOK._z2i(orb.Rnl, x.S) for x in XA]); sB = ROCArray([AOK._z2i(orb.Rnl, x.S) for x in XB])
    ζ = ROCArray{FT}(ps.Rnl.ζ); D = ROCArray{FT}(ps.Rnl.D)
    ns, K, NZ = size(ζ)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:aok_ovl2c_plan_create, libaok), Cint,
               (Ptr{Cvoid}, Cint, Cuint, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Int64, Int64, Int64, Ptr{Int64},
                Cint, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ref{Ptr{Cvoid}}),
               ctx().h, _dt(FT), AOK_FAST, pointer(sA), pointer(sB), AOK_I64, B, ns, K, NZ, _ip(str.ls),
               AOK_PARAMS_PS, pointer(ζ), pointer(D), Float64(orb.lengthscale), h)
    check(rc, "batch_overlap (plan)")
    plan = finalizer(_plan_destroy, Plan2c(h[], (sA, sB, ζ, D)))
    return plan, posA, posB, ζ, D, FT, str.nbf_total, B
end

_plan_forward(plan, posA, posB, S) =
    check(ccall((:aok_ovl2c_plan_forward, libaok), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                plan.h, pointer(posA), pointer(posB), pointer(S)), "batch_overlap")
_plan_pullback(plan, ∂S, posA, posB, ∂ζ, ∂D) =
    check(ccall((:aok_ovl2c_plan_pullback, libaok), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                plan.h, pointer(∂S), pointer(posA), pointer(posB), pointer(∂ζ), pointer(∂D)), "pullback_ps")

function rrule(::typeof(AOK.batch_overlap), orb::AOK.GaussianTypeOrbitals, XA::AbstractVector{<:PState},
               XB::AbstractVector{<:PState}, ps::RocPS, st)
    plan, posA, posB, ζ, D, FT, N, B = _plan2c(orb, XA, XB, ps)
    S = similar(posA, FT, N, N, B)
    _plan_forward(plan, posA, posB, S)
    function batch_overlap_pullback(_∂)
        ∂Sd = ROCArray{FT}(unthunk(_∂)); ∂ζ = similar(ζ); ∂D = similar(D)
        _plan_pullback(plan, ∂Sd, posA, posB, ∂ζ, ∂D)
        return NoTangent(), NoTangent(), NoTangent(), NoTangent(), (Rnl = (ζ = ∂ζ, D = ∂D), Ylm = NamedTuple()),
               NoTangent()
    end
    return S, batch_overlap_pullback
end

# ---- HIP-graph training step (no reference counterpart) --------------------------------
# The rrule's plan update + forward + (ζ, D) pullback recorded once (aok_graph_*) and
# replayed as one launch.  Between replays update `sg.ζ`, `sg.D` (the graph's device copies
# of ps.Rnl) and `sg.∂S` in place; `replay!` returns the graph's own S and cotangents.  The
# shim's context must run on a stream other than the NULL stream (AMDGPU's task-local
# streams are).
mutable struct StepGraph
    g::Ptr{Cvoid}
    plan::Plan2c
    posA::Any; posB::Any; ζ::Any; D::Any
    S::Any; ∂S::Any; ∂ζ::Any; ∂D::Any
end
function _graph_destroy(sg::StepGraph)
    sg.g == C_NULL || ccall((:aok_graph_destroy, libaok), Cint, (Ptr{Cvoid},), sg.g)
    sg.g = C_NULL
end

function StepGraph(orb::AOK.GaussianTypeOrbitals, XA::AbstractVector, XB::AbstractVector, ps)
    plan, posA, posB, ζ, D, FT, N, B = _plan2c(orb, XA, XB, ps)
    S = similar(posA, FT, N, N, B); ∂S = fill!(similar(S), zero(FT)); ∂ζ = similar(ζ); ∂D = similar(D)
    function record()
        check(ccall((:aok_ovl2c_plan_update, libaok), Cint, (Ptr{Cvoid},), plan.h), "plan update")
        _plan_forward(plan, posA, posB, S)
        _plan_pullback(plan, ∂S, posA, posB, ∂ζ, ∂D)
    end
    record()                                   # eager pass: sizes every workspace
    check(ccall((:aok_graph_begin, libaok), Cint, (Ptr{Cvoid},), ctx().h), "graph begin")
    g = Ref{Ptr{Cvoid}}(C_NULL)
    ok = false
    try
        record()
        ok = true
    finally                                    # always end the recording; report the first error
        rc = ccall((:aok_graph_end, libaok), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), ctx().h, g)
        ok && check(rc, "graph end")
    end
    return finalizer(_graph_destroy, StepGraph(g[], plan, posA, posB, ζ, D, S, ∂S, ∂ζ, ∂D))
end

function replay!(sg::StepGraph)
    check(ccall((:aok_graph_launch, libaok), Cint, (Ptr{Cvoid},), sg.g), "graph launch")
    return sg.S, (Rnl = (ζ = sg.∂ζ, D = sg.∂D), Ylm = NamedTuple())
end

# ---- H2 3-centre: batch_overlap_3c!(out, ::CompiledBasis, posA, posB, posC) (kernels_3c.jl:183-200) --
# Unitful (3, B) positions stripped by the reference's own to_bohr; the CompiledBasis
# (no species) becomes one species whose primitive slots are zero-padded to the longest
# contraction (_padded, as the 2-centre geometry-in method below).
function batch_overlap_3c!(out::ROCArray{FT, 4}, basis::AOK.CompiledBasis,
                           posA::AbstractMatrix{<:Unitful.Length}, posB::AbstractMatrix{<:Unitful.Length},
                           posC::AbstractMatrix{<:Unitful.Length}; kwargs...) where {FT}
    A = ROCArray(AOK.to_bohr(posA, FT)); Bm = ROCArray(AOK.to_bohr(posB, FT)); C = ROCArray(AOK.to_bohr(posC, FT))
    coef, α = _padded(basis, FT)
    ns, K, _ = size(coef)
    rc = ccall((:aok_ovl3c, libaok), Cint,
               (Ptr{Cvoid}, Cint, Cuint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint,
                Int64, Int64, Int64, Int64, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}),
               ctx().h, _dt(FT), AOK_FAST, pointer(A), pointer(Bm), pointer(C), C_NULL, C_NULL, C_NULL,
               AOK_I64, size(posA, 2), ns, K, 1, _ip(Vector{Int}(Array(basis.ls))), pointer(coef), pointer(α), 1.0,
               pointer(out))
    check(rc, "batch_overlap_3c!")
    return out
end

# ---- H2 geometry-in 2-centre: batch_overlap!(out, ::CompiledBasis, posA, posB) ------------
# (kernels_2c.jl:158-168): units stripped on the host by the reference's own to_bohr; the
# shells become one species whose primitive slots are zero-padded to the longest contraction.
function _padded(basis::AOK.CompiledBasis, FT)
    K = maximum(basis.nprim)
    coef = zeros(FT, basis.nshells, K, 1); α = ones(FT, basis.nshells, K, 1)
    for s in 1:basis.nshells, k in 1:basis.nprim[s]
        coef[s, k, 1] = basis.coef[basis.prim_offset[s] + k - 1]
        α[s, k, 1] = basis.α[basis.prim_offset[s] + k - 1]
    end
    return ROCArray(coef), ROCArray(α)
end

function batch_overlap!(out::ROCArray{FT, 3}, basis::AOK.CompiledBasis,
                        posA::AbstractMatrix{<:Unitful.Length}, posB::AbstractMatrix{<:Unitful.Length}) where {FT}
    A = ROCArray(AOK.to_bohr(posA, FT)); Bm = ROCArray(AOK.to_bohr(posB, FT))
    coef, α = _padded(basis, FT)
    ns, K, _ = size(coef)
    rc = ccall((:aok_ovl2c, libaok), Cint,
               (Ptr{Cvoid}, Cint, Cuint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Int64, Int64,
                Int64, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}),
               ctx().h, _dt(FT), AOK_FAST, pointer(A), pointer(Bm), C_NULL, C_NULL, AOK_I64, size(posA, 2), ns, K, 1,
               _ip(Vector{Int}(Array(basis.ls))), pointer(coef), pointer(α), 1.0, pointer(out))
    check(rc, "batch_overlap!")
    return out
end

# ---- no reference counterpart: position cotangents of the 2-centre overlap ---------------
# (dA, dB) [3 x B] per unit of the basis' length unit, dB = -dA.
function overlap_pullback_x(∂S::ROCArray{FT, 3}, basis::AOK.CartesianGTOBasis, XA::AbstractVector,
                            XB::AbstractVector) where {FT}
    B = length(XA)
    posA = ROCArray(AOK._cgto_posmat(XA, FT)); posB = ROCArray(AOK._cgto_posmat(XB, FT))
    sA = ROCArray(AOK._cgto_sidx(basis, XA)); sB = ROCArray(AOK._cgto_sidx(basis, XB))
    coef = ROCArray(FT.(basis.coef)); α = ROCArray(FT.(basis.ζ))
    ns, K, NZ = size(basis.coef)
    dA = similar(posA); dB = similar(posB)
    rc = ccall((:aok_ovl2c_pullback_x, libaok), Cint,
               (Ptr{Cvoid}, Cint, Cuint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64,
                Int64, Int64, Int64, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}),
               ctx().h, _dt(FT), AOK_FAST, pointer(∂S), pointer(posA), pointer(posB), pointer(sA), pointer(sB),
               AOK_I64, B, ns, K, NZ, _ip(basis.ls), pointer(coef), pointer(α), Float64(basis.lengthscale),
               pointer(dA), pointer(dB))
    check(rc, "overlap_pullback_x")
    return dA, dB
end

# ---- no reference counterpart: (coef, α) cotangents of the 3-centre overlap ---------------
function overlap_3c_pullback(∂T::ROCArray{FT, 4}, ls::Vector{Int}, coef::ROCArray{FT, 3}, α::ROCArray{FT, 3},
                             posA::ROCMatrix{FT}, posB::ROCMatrix{FT}, posC::ROCMatrix{FT}) where {FT}
    ns, K, NZ = size(coef)
    ∂c = similar(coef); ∂α = similar(α)
    rc = ccall((:aok_ovl3c_pullback, libaok), Cint,
               (Ptr{Cvoid}, Cint, Cuint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                Ptr{Cvoid}, Cint, Int64, Int64, Int64, Int64, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid},
                Ptr{Cvoid}),
               ctx().h, _dt(FT), AOK_FAST, pointer(∂T), pointer(posA), pointer(posB), pointer(posC), C_NULL, C_NULL,
               C_NULL, AOK_I64, size(posA, 2), ns, K, NZ, ls, pointer(coef), pointer(α), 1.0, pointer(∂c),
               pointer(∂α))
    check(rc, "overlap_3c_pullback")
    return ∂c, ∂α
end

end # module
