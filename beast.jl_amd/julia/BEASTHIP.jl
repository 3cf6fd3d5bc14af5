"""
    BEASTHIP

Julia side of the drop-in boundary: routes BEAST's `assemble(op, X, Y; quadstrat)`
for `Maxwell3D.singlelayer` / `Maxwell3D.doublelayer` on `raviartthomas` spaces
(BEAST src/operator.jl:100-116) to libbeast_hip through `ccall`, and offers the
resident-matrix solve (FP64 LU, or FP32 LU + FP64 GMRES-IR).

Selection mirrors the reference's threading traits (operator.jl:106-108):
`assemble(op, X, Y; threading=:gpu)` or `threading=BEAST.Threading{:gpu}` takes the
GPU path, also for the terms of linear combinations, transposes and direct-product
spaces (the `assemble!` methods below); any input the library does not cover (other
quadstrats or rule degrees, non-RT spaces, Float32 or refined meshes, dofs with more
than two shapes) falls back to BEAST's own multithreaded CPU assembly.

Not exercised in this repository (no Julia toolchain here); the C ABI it binds
is tested from Python with the same argument conventions.
"""
module BEASTHIP

using BEAST
using CompScienceMeshes
using LinearAlgebra

const libbeast_hip = get(ENV, "BEAST_HIP_LIB", joinpath(@__DIR__, "..", "csrc", "libbeast_hip.so"))

const BH_OK = Int32(0)
const BH_ERR_ASSERT = Int32(-4)
const BH_ERR_UNSUPPORTED = Int32(-5)

struct BHOperator
    kind::Int32
    gamma::NTuple{2,Float64}
    alpha::NTuple{2,Float64}
    beta::NTuple{2,Float64}
end

struct BHQuadStrat
    kind::Int32
    params::NTuple{8,Int32}
end

struct BHSolveStats
    iterations::Int32
    inner::Int32
    residual::Float64
    seconds_factor::Float64
    seconds_solve::Float64
end

struct Unsupported <: Exception
    msg::String
end

lasterror() = unsafe_string(ccall((:bh_last_error, libbeast_hip), Cstring, ()))

function check(rc)
    rc == BH_OK && return nothing
    rc == BH_ERR_ASSERT && throw(AssertionError(lasterror()))
    rc == BH_ERR_UNSUPPORTED && throw(Unsupported(lasterror()))
    error(lasterror())
end

mutable struct Context
    ptr::Ptr{Cvoid}
    function Context(device::Integer=0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:bh_ctx_create, libbeast_hip), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, r))
        c = new(r[])
        finalizer(c -> ccall((:bh_ctx_destroy, libbeast_hip), Cint, (Ptr{Cvoid},), c.ptr), c)
        return c
    end
end

const _ctx = Ref{Union{Nothing,Context}}(nothing)
context() = (_ctx[] === nothing && (_ctx[] = Context(0)); _ctx[])

c2(z) = (Float64(real(z)), Float64(imag(z)))

_gamma(op) = BEAST.isstatic(op.gamma) ? 0.0 : op.gamma

bhop(op::BEAST.MWSingleLayer3D) = BHOperator(0, c2(_gamma(op)), c2(op.α), c2(op.β))
bhop(op::BEAST.MWDoubleLayer3D) = BHOperator(1, c2(_gamma(op)), c2(op.alpha), (0.0, 0.0))
bhop(op::BEAST.NCross) = BHOperator(2, (0.0, 0.0), (0.0, 0.0), (0.0, 0.0))
bhop(op::BEAST.DoubleLayerRotatedMW3D) = BHOperator(3, c2(_gamma(op)), c2(op.alpha), (0.0, 0.0))
bhop(op::BEAST.Identity) = BHOperator(4, (0.0, 0.0), (0.0, 0.0), (0.0, 0.0))

pad8(t) = ntuple(i -> i <= length(t) ? Int32(t[i]) : Int32(0), 8)
bhqs(q::BEAST.DoubleNumWiltonSauterQStrat) = BHQuadStrat(0, pad8((q.outer_rule_far, q.inner_rule_far,
    q.outer_rule_near, q.inner_rule_near, q.sauter_schwab_common_tetr, q.sauter_schwab_common_face,
    q.sauter_schwab_common_edge, q.sauter_schwab_common_vert)))
bhqs(q::BEAST.DoubleNumSauterQstrat) = BHQuadStrat(1, pad8((q.outer_rule, q.inner_rule,
    q.sauter_schwab_common_tetr, q.sauter_schwab_common_face, q.sauter_schwab_common_edge,
    q.sauter_schwab_common_vert)))
bhqs(q::BEAST.DoubleNumQStrat) = BHQuadStrat(2, pad8((q.outer_rule, q.inner_rule)))
bhqs(q::BEAST.SingleNumQStrat) = BHQuadStrat(3, pad8((q.quad_rule,)))
bhqs(q) = throw(Unsupported("quadrature strategy $(typeof(q)) has no GPU path"))

# mesh and space handles, cached per (object, context)
const _handles = IdDict{Any,Ptr{Cvoid}}()

function meshhandle(Γ, ctx)
    get!(_handles, Γ) do
        V = Float64[v[k] for k in 1:3, v in CompScienceMeshes.vertices(Γ)]
        C = Int32[c[k] - 1 for k in 1:3, c in CompScienceMeshes.cells(Γ)]
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:bh_mesh_create, libbeast_hip), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int32}, Int64, Ptr{Ptr{Cvoid}}),
            ctx.ptr, V, size(V, 2), C, size(C, 2), r))
        r[]
    end
end

# RTBasis.fns (rtspace.jl:3-7) in CSR form: the dof map is BEAST's own, unchanged
function spacehandle(X::BEAST.RTBasis, ctx)
    get!(_handles, X) do
        ptr = Int32[0]; cell = Int32[]; ref = Int32[]; coeff = Float64[]
        for fn in X.fns
            for sh in fn
                push!(cell, sh.cellid - 1); push!(ref, sh.refid - 1); push!(coeff, sh.coeff)
            end
            push!(ptr, length(cell))
        end
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:bh_space_create, libbeast_hip), Cint,
            (Ptr{Cvoid}, Int64, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}, Ptr{Ptr{Cvoid}}),
            meshhandle(X.geo, ctx), length(X.fns), ptr, cell, ref, coeff, r))
        r[]
    end
end

# Operators with a GPU path on RT x RT: the integral ones (pair kernels) and the local ones.
const GPUIntegralOp = Union{BEAST.MWSingleLayer3D,BEAST.MWDoubleLayer3D,BEAST.DoubleLayerRotatedMW3D}
const GPULocalOp = Union{BEAST.NCross,BEAST.Identity}
const GPUOp = Union{GPUIntegralOp,GPULocalOp}

"""
    gpu_path(op, X, Y, qs) -> Bool

Whether the library covers `op` on `X x Y` with strategy `qs`; everything else takes
BEAST's CPU path: Float32 meshes (`test/test_assemblerow.jl:8-27`; the library
classifies and integrates in FP64), meshes refining one another (the reference's
`TestRefinesTrialQStrat`, `integralop.jl:93-99`), strategy types or rule degrees the
library lacks and operator/strategy pairs the reference does not define on this path
(`bh_check_strategy`: BH_ERR_UNSUPPORTED).
"""
function gpu_path(op::GPUOp, X::BEAST.RTBasis, Y::BEAST.RTBasis, qs)
    (X isa BEAST.RTBasis{Float64} && Y isa BEAST.RTBasis{Float64}) || return false
    tg, bg = BEAST.geometry(X), BEAST.geometry(Y)
    (CompScienceMeshes.refines(tg, bg) || CompScienceMeshes.refines(bg, tg)) && return false
    q = try bhqs(qs) catch e; e isa Unsupported || rethrow(); return false end
    rc = ccall((:bh_check_strategy, libbeast_hip), Cint, (Ref{BHOperator}, Ref{BHQuadStrat}), Ref(bhop(op)), Ref(q))
    rc == BH_ERR_UNSUPPORTED && return false
    check(rc)
    return true
end

"""
    gpu_assemble!(Z, op, X, Y, qs; scale=1, row_offset=0, col_offset=0)

`Z[row_offset+m, col_offset+n] += scale * A[m, n]` (the store of assemble!,
operator.jl:168-184, 319-356) computed on the GPU.
"""
function gpu_assemble!(Z::Matrix{ComplexF64}, op::GPUOp, X::BEAST.RTBasis, Y::BEAST.RTBasis, qs;
        scale=1.0, row_offset=0, col_offset=0, ctx=context())
    M, N = numfunctions(X), numfunctions(Y)
    # the block must lie inside Z, as the reference's store would check (BoundsError)
    (row_offset >= 0 && col_offset >= 0 && row_offset + M <= size(Z, 1) && col_offset + N <= size(Z, 2)) ||
        throw(BoundsError(Z, (row_offset + M, col_offset + N)))
    o = Ref(bhop(op)); q = Ref(bhqs(qs)); s = Float64[real(scale), imag(scale)]
    GC.@preserve Z begin
        check(ccall((:bh_assemble, libbeast_hip), Cint,
            (Ptr{Cvoid}, Ref{BHOperator}, Ref{BHQuadStrat}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Cint,
             Ptr{ComplexF64}, Int64, Int64, Int64),
            ctx.ptr, o, q, 0, spacehandle(X, ctx), spacehandle(Y, ctx), s, 1, Z, size(Z, 1), row_offset, col_offset))
    end
    return Z
end

# ---------------------------------------------------------------------------------------
# The seam: a threading trait, as :dofsplitting and :cellcoloring are mapped to
# Threading{:multi} and Threading{:cellcoloring} (operator.jl:106-108).  The reference's
# assemble passes any other threading value on unchanged, and its compositions pass it to
# their terms: LinearCombinationOfOperators (operator.jl:294-303), DirectProductSpace
# factors (:319-356), TransposedOperator (:284-291, handled below).  So
#     assemble(op, X, Y; threading=BEAST.Threading{:gpu})   (or threading=:gpu)
# reaches the methods here for every term; terms without a GPU path take Threading{:multi}.

# GPU operators on RT x RT: the dense block on the device, then the caller's store
function _gpu_store!(op::GPUOp, X::BEAST.RTBasis, Y::BEAST.RTBasis, store, quadstrat, kwargs)
    qs = quadstrat(op, X, Y)   # a strategy object returns itself (strategies/quadstrat.jl:3-5)
    if !gpu_path(op, X, Y, qs)
        return BEAST.assemble!(op, X, Y, store, BEAST.Threading{:multi}; quadstrat=qs, kwargs...)
    end
    (numfunctions(X) == 0 || numfunctions(Y) == 0) && return   # integralop.jl:72-73
    Z = zeros(ComplexF64, numfunctions(X), numfunctions(Y))
    gpu_assemble!(Z, op, X, Y, qs)
    @inbounds for n in axes(Z, 2), m in axes(Z, 1)
        store(Z[m, n], m, n)
    end
    return
end

BEAST.assemble!(op::GPUIntegralOp, X::BEAST.RTBasis, Y::BEAST.RTBasis, store, ::Type{BEAST.Threading{:gpu}};
    quadstrat=BEAST.defaultquadstrat, kwargs...) = _gpu_store!(op, X, Y, store, quadstrat, kwargs)

BEAST.assemble!(op::GPULocalOp, X::BEAST.RTBasis, Y::BEAST.RTBasis, store, ::Type{BEAST.Threading{:gpu}};
    quadstrat=BEAST.defaultquadstrat, kwargs...) = _gpu_store!(op, X, Y, store, quadstrat, kwargs)

# any other operator under the trait: the reference's own multithreaded CPU assembly
BEAST.assemble!(op::BEAST.Operator, X::BEAST.Space, Y::BEAST.Space, store, ::Type{BEAST.Threading{:gpu}};
    quadstrat=BEAST.defaultquadstrat, kwargs...) =
    BEAST.assemble!(op, X, Y, store, BEAST.Threading{:multi}; quadstrat, kwargs...)

# transposes: indices swapped, the inner operator assembled under the trait (operator.jl:284-291)
function BEAST.assemble!(op::BEAST.TransposedOperator, X::BEAST.Space, Y::BEAST.Space, store,
        ::Type{BEAST.Threading{:gpu}}; quadstrat=BEAST.defaultquadstrat(op, X, Y), kwargs...)
    store1(v, m, n) = store(v, n, m)
    BEAST.assemble!(op.op, Y, X, store1, BEAST.Threading{:gpu}; quadstrat, kwargs...)
end

# threading=:gpu as a Symbol, as assemble passes it on: converted to the trait
function _gpu_trait(threading::Symbol)
    threading == :gpu || throw(ArgumentError("unknown threading $(repr(threading))"))
    return BEAST.Threading{:gpu}
end
BEAST.assemble!(op::BEAST.Operator, X::BEAST.Space, Y::BEAST.Space, store, threading::Symbol; kwargs...) =
    BEAST.assemble!(op, X, Y, store, _gpu_trait(threading); kwargs...)
BEAST.assemble!(op::BEAST.LinearCombinationOfOperators, X::BEAST.Space, Y::BEAST.Space, store, threading::Symbol;
    kwargs...) = BEAST.assemble!(op, X, Y, store, _gpu_trait(threading); kwargs...)
BEAST.assemble!(op::BEAST.TransposedOperator, X::BEAST.Space, Y::BEAST.Space, store, threading::Symbol;
    kwargs...) = BEAST.assemble!(op, X, Y, store, _gpu_trait(threading); kwargs...)

# a single GPU operator on RT x RT: the dense matrix written by the library directly
function BEAST.assemble(op::GPUOp, X::BEAST.RTBasis, Y::BEAST.RTBasis;
        threading=:dofsplitting, quadstrat=BEAST.defaultquadstrat, kwargs...)
    if threading == :gpu || threading == BEAST.Threading{:gpu}
        qs = quadstrat(op, X, Y)
        if gpu_path(op, X, Y, qs)
            Z = zeros(ComplexF64, numfunctions(X), numfunctions(Y))
            (numfunctions(X) == 0 || numfunctions(Y) == 0) && return Z
            return gpu_assemble!(Z, op, X, Y, qs)
        end
        threading = :dofsplitting      # the reference's CPU path
    end
    return invoke(BEAST.assemble, Tuple{BEAST.AbstractOperator,Any,Any}, op, X, Y;
        threading=threading, quadstrat=quadstrat, kwargs...)
end

"""Resident system matrix (HBM) with its LU factors."""
mutable struct DeviceMatrix
    ptr::Ptr{Cvoid}
    n::Int
    function DeviceMatrix(ptr, n)
        M = new(ptr, n)
        finalizer(M -> ccall((:bh_matrix_destroy, libbeast_hip), Cint, (Ptr{Cvoid},), M.ptr), M)
        return M
    end
end

function DeviceMatrix(op::GPUOp, X::BEAST.RTBasis; quadstrat=BEAST.defaultquadstrat, ctx=context())
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bh_matrix_assemble, libbeast_hip), Cint,
        (Ptr{Cvoid}, Ref{BHOperator}, Ref{BHQuadStrat}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}),
        ctx.ptr, Ref(bhop(op)), Ref(bhqs(quadstrat(op, X, X))), spacehandle(X, ctx), r))
    return DeviceMatrix(r[], numfunctions(X))
end

# a linear combination of GPU operators (operator.jl:294-303), e.g. the MFIE n x K - Id/2 of
# examples/nxmfie.jl: the first term assembled in place, then the others accumulated, each
# with its own default strategy; a first coefficient other than 1 is applied by accumulating
# (c - 1) times the first term
function DeviceMatrix(op::BEAST.LinearCombinationOfOperators, X::BEAST.RTBasis; quadstrat=BEAST.defaultquadstrat,
        ctx=context())
    all(A -> A isa GPUOp, op.ops) || throw(Unsupported("linear combination with a term that has no GPU path"))
    M = DeviceMatrix(op.ops[1], X; quadstrat, ctx)
    c1 = op.coeffs[1]
    c1 == 1 || accumulate!(M, op.ops[1], X; scale=c1 - 1, quadstrat, ctx)
    for (a, A) in zip(op.coeffs[2:end], op.ops[2:end])
        accumulate!(M, A, X; scale=a, quadstrat, ctx)
    end
    return M
end

function DeviceMatrix(Z::Matrix{ComplexF64}; ctx=context())
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:bh_matrix_from_host, libbeast_hip), Cint,
        (Ptr{Cvoid}, Int64, Ptr{ComplexF64}, Int64, Ptr{Ptr{Cvoid}}), ctx.ptr, size(Z, 1), Z, size(Z, 1), r))
    return DeviceMatrix(r[], size(Z, 1))
end

const PRECISIONS = Dict(:fp64 => Cint(0), :fp32 => Cint(1), :fp32split => Cint(2))

"""`accumulate!(M, op, X; scale)`: M += scale * op(X, X) in HBM (a further term of a
linear combination, e.g. the N/2 of the MFIE K + N/2)."""
function accumulate!(M::DeviceMatrix, op::GPUOp, X::BEAST.RTBasis; scale=1.0, quadstrat=BEAST.defaultquadstrat,
        ctx=context())
    s = Float64[real(scale), imag(scale)]
    check(ccall((:bh_matrix_accumulate, libbeast_hip), Cint,
        (Ptr{Cvoid}, Ref{BHOperator}, Ref{BHQuadStrat}, Ptr{Cvoid}, Ptr{Float64}),
        M.ptr, Ref(bhop(op)), Ref(bhqs(quadstrat(op, X, X))), spacehandle(X, ctx), s))
    return M
end

"""`factor!(M, :fp64)` (getrf), `factor!(M, :fp32)` (FP32 LU for GMRES-IR) or
`factor!(M, :fp32split)` (FP32 LU whose trailing updates use split-bf16 products)."""
function factor!(M::DeviceMatrix, precision::Symbol=:fp64)
    haskey(PRECISIONS, precision) || throw(ArgumentError("precision must be one of $(keys(PRECISIONS))"))
    check(ccall((:bh_matrix_factor, libbeast_hip), Cint, (Ptr{Cvoid}, Cint), M.ptr, PRECISIONS[precision]))
    return M
end

function LinearAlgebra.ldiv!(x::AbstractVecOrMat{ComplexF64}, M::DeviceMatrix, b::AbstractVecOrMat{ComplexF64};
        tol=0.0, maxit=0)
    B = Matrix{ComplexF64}(reshape(b, M.n, :)); Xm = similar(B)
    st = Ref{BHSolveStats}()
    check(ccall((:bh_matrix_solve, libbeast_hip), Cint,
        (Ptr{Cvoid}, Int64, Ptr{ComplexF64}, Ptr{ComplexF64}, Float64, Cint, Ref{BHSolveStats}),
        M.ptr, size(B, 2), B, Xm, tol, maxit, st))
    copyto!(x, Xm)
    return x
end

Base.:\(M::DeviceMatrix, b::AbstractVecOrMat) = ldiv!(similar(b, ComplexF64), M, ComplexF64.(b))

# Plane-wave right-hand sides: assemble((n x E) x n, X) and assemble(n x E, X)
# (excitation.jl:17-115,145-180, mwexc.jl:1-54) for E a PlaneWaveMW (also curl E, a plane
# wave itself), on the device; rule = the SingleNumQStrat degree.
const GPUTrace = Union{BEAST.TangTraceMW,BEAST.CrossTraceMW}
_trace(::BEAST.TangTraceMW) = Cint(0)    # BH_TRACE_TANGENTIAL
_trace(::BEAST.CrossTraceMW) = Cint(1)   # BH_TRACE_CROSS

function gpu_assemble_rhs(e::GPUTrace, X::BEAST.RTBasis; quadstrat=BEAST.defaultquadstrat(e, X), ctx=context())
    E = e.field
    E isa BEAST.PlaneWaveMW || throw(Unsupported("right-hand side of $(typeof(E))"))
    X isa BEAST.RTBasis{Float64} || throw(Unsupported("Float32 meshes take the CPU path"))
    d = Float64[E.direction[k] for k in 1:3]
    pr = Float64[real(E.polarisation[k]) for k in 1:3]
    pi_ = Float64[imag(E.polarisation[k]) for k in 1:3]
    g = ComplexF64(E.gamma); a = ComplexF64(E.amplitude)
    b = zeros(ComplexF64, numfunctions(X))
    check(ccall((:bh_assemble_excitation, libbeast_hip), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{ComplexF64}, Ref{ComplexF64},
         Cint, Cint, Ptr{ComplexF64}),
        ctx.ptr, spacehandle(X, ctx), _trace(e), d, pr, pi_, g, a, quadstrat.quad_rule, 1, b))
    return b
end

function BEAST.assemble(e::GPUTrace, X::BEAST.RTBasis; threading=nothing, kwargs...)
    if threading == :gpu
        try
            return gpu_assemble_rhs(e, X; kwargs...)
        catch err
            err isa Unsupported || rethrow()
        end
    end
    return invoke(BEAST.assemble, Tuple{BEAST.Functional,Any}, e, X; kwargs...)
end

# Far fields: potential(op, points, coeffs, X) (postproc.jl:138-150,174-211) for the
# Maxwell far-field operators (maxwell/farfield.jl:1-103) on a raviartthomas space.
const FFKINDS = Dict{Any,Int32}(BEAST.MWFarField3D => 0, BEAST.MWDoubleLayerFarField3D => 1,
                                BEAST.MWDoubleLayerRotatedFarField3D => 2, BEAST.MWFarField3DDropConstant => 3)

"""
    gpu_potential(op, points, coeffs, X; quadstrat=defaultquadstrat(op, X))

Far field of the current `sum_n coeffs[n] f_n` at `points`, as `potential` returns it
(an array of `SVector{3,ComplexF64}` of the points' shape).
"""
function gpu_potential(op::BEAST.MWFarField, points, coeffs, X::BEAST.RTBasis;
        quadstrat=BEAST.defaultquadstrat(op, X), ctx=context(), type=BEAST.SVector{3,ComplexF64})
    type == BEAST.SVector{3,ComplexF64} || throw(Unsupported("far fields are computed as SVector{3,ComplexF64}"))
    kind = get(FFKINDS, typeof(op).name.wrapper, nothing)
    kind === nothing && throw(Unsupported("far-field operator $(typeof(op))"))
    amp = ComplexF64(op isa BEAST.MWFarField3DDropConstant ? op.coeff : op.amplitude)
    g = ComplexF64(op.gamma)
    @assert length(coeffs) == numfunctions(X)
    P = Float64[p[k] for k in 1:3, p in vec(points)]
    u = Vector{ComplexF64}(coeffs)
    ff = Matrix{ComplexF64}(undef, 3, length(points))
    check(ccall((:bh_farfield, libbeast_hip), Cint,
        (Ptr{Cvoid}, Cint, Ref{ComplexF64}, Ref{ComplexF64}, Ptr{Cvoid}, Ptr{ComplexF64}, Int64, Ptr{Float64}, Cint,
         Ptr{ComplexF64}),
        ctx.ptr, kind, g, amp, spacehandle(X, ctx), u, length(points), P, quadstrat.quad_rule, ff))
    return reshape([BEAST.SVector{3,ComplexF64}(ff[1, i], ff[2, i], ff[3, i]) for i in 1:length(points)], size(points))
end

function BEAST.potential(op::BEAST.MWFarField, points, coeffs, X::BEAST.RTBasis;
        threading=BEAST.Threading{:single}, kwargs...)
    threading == :gpu && return gpu_potential(op, points, coeffs, X; kwargs...)
    return invoke(BEAST.potential, Tuple{Any,Any,Any,Any}, op, points, coeffs, X; threading, kwargs...)
end

end # module
