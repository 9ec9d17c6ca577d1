# PureUMFPACKHIP.jl -- Julia binding of libmfgpu (include/mfgpu.h) behind PureUMFPACK's
# multifrontal entry points.  Julia keeps ordering, scaling and symbolic analysis
# (PureUMFPACK.symbolic_mf, row_scaling); every numeric step runs on the GPU.
#
# Drop-in: `hip_splu(A; scale, check)` returns a `HIPLU` with the fields of the `PureLU` of
# `splu(A; method=:multifrontal)` (src/interface.jl:93-120), L and U exported on first
# access; `solve(F, b; refine)` and `F \ b` keep their semantics (src/interface.jl:170-179).
# Only Float64 values with Int64 indices are routed here; other element types stay on the
# reference path.
module PureUMFPACKHIP

using PureUMFPACK, SparseArrays, LinearAlgebra
using PureUMFPACK: PureLU, symbolic_mf, row_scaling, SCALE_SUM, SCALE_NONE, ScaleKind

const libmfgpu = get(ENV, "MFGPU_LIB", "libmfgpu.so")

const MF_OK, MF_SINGULAR, MF_DIM, MF_BADARG = 0, 1, 2, 3

struct MFOptions
    precision::Int32
    small_nf::Int32
    use_graph::Int32
    nbatch::Int32
    pivoting::Int32
    pivot_tol::Float32
    solve_path::Int32
    reserved::Int32
end
MFOptions(; precision = 0, small_nf = 0, use_graph = 0, nbatch = 1, pivoting = 0, pivot_tol = 0.0f0, solve_path = 0) =
    MFOptions(precision, small_nf, use_graph, nbatch, pivoting, pivot_tol, solve_path, 0)

struct MFSolveReport
    iterations::Int32
    converged::Int32
    backward_error::Float64
    residual_2::Float64
    seconds::Float64
end

mutable struct MFPlan
    ptr::Ptr{Cvoid}
    n::Int
end

lasterr() = unsafe_string(ccall((:mf_last_error, libmfgpu), Cstring, ()))

function check(rc, info = 0)
    rc == MF_OK && return nothing
    rc == MF_SINGULAR && throw(SingularException(info))
    rc == MF_DIM && throw(DimensionMismatch(lasterr()))
    rc == MF_BADARG && throw(ArgumentError(lasterr()))
    error("libmfgpu: " * lasterr())
end

# flatten SymbolicMF (src/symbolic.jl:203-209) to the plan's CSR arrays (1-based)
function plan_arrays(S)
    ns = length(S.sstart) - 1
    upd_ptr = Vector{Int}(undef, ns + 1); upd_ptr[1] = 1
    child_ptr = Vector{Int}(undef, ns + 1); child_ptr[1] = 1
    for s in 1:ns
        upd_ptr[s + 1] = upd_ptr[s] + length(S.colstruct[S.sstart[s + 1] - 1])
        child_ptr[s + 1] = child_ptr[s] + length(S.childsupers[s])
    end
    upd_idx = reduce(vcat, (S.colstruct[S.sstart[s + 1] - 1] for s in 1:ns); init = Int[])
    child_idx = reduce(vcat, S.childsupers; init = Int[])
    return upd_ptr, upd_idx, child_ptr, child_idx
end

function MFPlan(A::SparseMatrixCSC{Float64, Int}, S; opts = MFOptions())
    n = size(A, 2)
    upd_ptr, upd_idx, child_ptr, child_idx = plan_arrays(S)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    o = Ref(opts)
    GC.@preserve S upd_ptr upd_idx child_ptr child_idx A begin
        rc = ccall((:mf_plan_create, libmfgpu), Cint,
            (Int64, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
             Ptr{Int64}, Ptr{Int64}, Ref{MFOptions}, Ref{Ptr{Cvoid}}),
            n, S.qf, length(S.sstart) - 1, S.sstart, upd_ptr, upd_idx, child_ptr, child_idx,
            SparseArrays.getcolptr(A), rowvals(A), o, out)
    end
    check(rc)
    P = MFPlan(out[], n)
    finalizer(p -> ccall((:mf_plan_destroy, libmfgpu), Cvoid, (Ptr{Cvoid},), p.ptr), P)
    return P
end

function factor!(P::MFPlan, A::SparseMatrixCSC{Float64, Int}, Rs; check_singular = true)
    info = Ref{Int64}(0)
    rc = GC.@preserve A Rs ccall((:mf_factor, libmfgpu), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int32, Ref{Int64}),
        P.ptr, nonzeros(A), Rs === nothing ? C_NULL : Rs, check_singular, info)
    check(rc, info[])
    return P
end

function export_factors(P::MFPlan)
    nL, nU = Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:mf_nnz, libmfgpu), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}), P.ptr, nL, nU))
    n = P.n
    Lp = Vector{Int}(undef, n + 1); Li = Vector{Int}(undef, nL[]); Lx = Vector{Float64}(undef, nL[])
    Up = Vector{Int}(undef, n + 1); Ui = Vector{Int}(undef, nU[]); Ux = Vector{Float64}(undef, nU[])
    p = Vector{Int}(undef, n); q = Vector{Int}(undef, n)
    check(ccall((:mf_export_factors, libmfgpu), Cint,
        (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64},
         Ptr{Int64}, Ptr{Int64}), P.ptr, Lp, Li, Lx, Up, Ui, Ux, p, q))
    return SparseMatrixCSC(n, n, Lp, Li, Lx), SparseMatrixCSC(n, n, Up, Ui, Ux), p, q
end

"""
    HIPLU

The factorization `hip_splu` returns.  The factors stay on the device; the fields of the
reference's `PureLU` (src/interface.jl:38-48) are there on demand: `F.L`, `F.U`, `F.p`, `F.q`
copy L, U, p and q to the host on first access (`mf_export_factors`, once), `F.Rs` and `F.A`
are held.  `F \ b` and `hip_solve(F.plan, b)` use the device factors without an export.
"""
mutable struct HIPLU
    plan::MFPlan
    Rs::Vector{Float64}
    A::SparseMatrixCSC{Float64, Int}
    lu::Union{Nothing, PureLU}
end

function materialize!(F::HIPLU)
    if getfield(F, :lu) === nothing
        L, U, p, q = export_factors(getfield(F, :plan))
        setfield!(F, :lu, PureLU(L, U, p, q, getfield(F, :Rs), getfield(F, :A)))
    end
    return getfield(F, :lu)
end

function Base.getproperty(F::HIPLU, s::Symbol)
    s in (:L, :U, :p, :q) && return getproperty(materialize!(F), s)
    return getfield(F, s)
end
Base.size(F::HIPLU, args...) = size(getfield(F, :A), args...)
Base.:\(F::HIPLU, b::AbstractVecOrMat{Float64}) = hip_solve(getfield(F, :plan), b)
PureUMFPACK.PureLU(F::HIPLU) = materialize!(F)

"""
    hip_splu(A; scale=SCALE_SUM, check=true, precision=0) -> (HIPLU, MFPlan)

GPU numeric phase behind `splu(A; method=:multifrontal)` (src/interface.jl:93-105).  L and U
are exported only when a caller reads them (`HIPLU`).
"""
function hip_splu(A::SparseMatrixCSC{Float64, Int}; scale::ScaleKind = SCALE_SUM, check = true,
                  precision = 0)
    n = size(A, 2)
    size(A, 1) == n || throw(DimensionMismatch("splu requires a square matrix"))
    Rs = row_scaling(A, scale)
    S = symbolic_mf(A)
    P = MFPlan(A, S; opts = MFOptions(precision = precision))
    factor!(P, A, scale == SCALE_NONE ? nothing : Rs; check_singular = check)
    return HIPLU(P, Rs, A, nothing), P
end

"""
    hip_solve(P, b; refine=0, tol=0.0) -> x

`PureUMFPACK.solve(F, b; refine)` on the device factors (src/interface.jl:170-177); b may be a
vector or an n x nrhs matrix.
"""
function hip_solve(P::MFPlan, b::AbstractVecOrMat{Float64}; refine::Integer = 0, tol::Real = 0.0)
    B = Matrix{Float64}(reshape(b, P.n, :))
    X = similar(B)
    rep = Ref{MFSolveReport}()
    check(ccall((:mf_solve, libmfgpu), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Int32, Float64, Ref{MFSolveReport}),
        P.ptr, size(B, 2), B, P.n, X, P.n, refine, tol, rep))
    return b isa AbstractVector ? vec(X) : X
end

"""
    hip_refactor_batch!(P, As; scale=SCALE_SUM) -> info

Refactor `length(As)` matrices sharing `P`'s pattern in one batch (`mf_refactor_batch`);
`P` must come from `MFPlan(A, S; opts=MFOptions(nbatch=...))`.  `info[m]` is the first singular
column of matrix m (0 if none).
"""
function hip_refactor_batch!(P::MFPlan, As::Vector{SparseMatrixCSC{Float64, Int}}; scale::ScaleKind = SCALE_SUM)
    nmat = length(As)
    Ax = reduce(hcat, (nonzeros(A) for A in As))             # nnz x nmat
    Rs = reduce(hcat, (row_scaling(A, scale) for A in As))   # n x nmat
    info = zeros(Int, nmat)
    rc = GC.@preserve Ax Rs ccall((:mf_refactor_batch, libmfgpu), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Int64}),
        P.ptr, nmat, Ax, scale == SCALE_NONE ? C_NULL : Rs, info)
    rc == MF_SINGULAR || check(rc)
    return info
end

"""
    hip_solve_batch(P, B; refine=0, tol=0.0) -> X

`B` is `n x nrhs x nmat`; matrix m of the last batch solves `B[:, :, m]` (`mf_solve_batch`).
"""
function hip_solve_batch(P::MFPlan, B::Array{Float64, 3}; refine::Integer = 0, tol::Real = 0.0)
    X = similar(B)
    rep = Ref{MFSolveReport}()
    check(ccall((:mf_solve_batch, libmfgpu), Cint,
        (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Int32, Float64, Ref{MFSolveReport}),
        P.ptr, size(B, 3), size(B, 2), B, X, refine, tol, rep))
    return X
end

end # module
