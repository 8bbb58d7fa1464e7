"""
    DiffMoMHIP

Julia binding of libdiffmom.so (MI355X / gfx950) for DiffMoM.jl's dense
surface-EFIE path.  Every function is a `ccall` into the C ABI declared in
include/diffmom.h; DiffMoM's own types (`TriMesh`, `RWGData`,
`PlaneWaveExcitation`, `ScatteringResult`) pass through unchanged because the
library takes Julia's in-memory layout (column-major 3xNv / 3xNt, 1-based
Int64 indices, interleaved ComplexF64).

The drop-in functions keep the reference's names, positional arguments and
keywords, and run the reference's own host-side validation (pair check, byte
gates, mesh quality, cache / adjacency limits, prefactors) before handing
the work to the GPU, so a call that fails in DiffMoM fails here with the same
exception.  Keyword values the GPU path cannot honour raise ArgumentError.

    using DiffMoM, DiffMoMHIP
    mesh = make_rect_plate(0.1, 0.1, 6, 6)
    rwg  = build_rwg(mesh)
    Z    = DiffMoMHIP.assemble_Z_efie(mesh, rwg, 2π/0.1)               # Matrix{ComplexF64}
    A    = DiffMoMHIP.assemble_Z_efie(mesh, rwg, 2π/0.1; device=true)  # GPU AbstractMatrix
    x, stats = solve_gmres(A, v)                  # Krylov.jl calls mul! on the GPU
    x, stats = DiffMoMHIP.solve_gmres(A, v)       # the whole solve on the GPU
    Amf  = DiffMoMHIP.matrixfree_efie_operator(mesh, rwg, 2π/0.1)   # no stored Z
"""
module DiffMoMHIP

using LinearAlgebra
using DiffMoM
using StaticArrays: SVector

const libdiffmom = joinpath(@__DIR__, "..", "libdiffmom.so")

const DM_Z64, DM_Z32, DM_Z64P, DM_Z32P = Cint(1), Cint(2), Cint(4), Cint(8)
const DM_METHOD_LU, DM_METHOD_GMRES, DM_METHOD_MIXED = Cint(0), Cint(1), Cint(2)
const DM_PRECOND_NONE, DM_PRECOND_DIAG, DM_PRECOND_NEARFIELD = Cint(0), Cint(1), Cint(2)

# Error codes map one-to-one onto the exception classes the reference throws.
function _check(rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:dm_last_error, libdiffmom), Cstring, ()))
    rc == 1 && throw(ArgumentError(msg))
    rc == 2 && throw(OverflowError(msg))
    rc == 3 && throw(DimensionMismatch(msg))
    rc == 7 && throw(DomainError(nothing, msg))
    error(msg)  # 4: convergence (ErrorException), 5/6: device / memory
end

# field for field the C struct dm_solve_opts (include/diffmom.h)
struct dm_solve_opts
    method::Cint; precond::Cint; precond_side::Cint; maxiter::Cint
    tol::Cdouble; nf_cutoff::Cdouble; restart::Cint; check_convergence::Cint
    inner_tol::Cdouble; max_outer::Cint; block_size::Cint
    batch::Cint; learned::Ptr{Cvoid}; adjoint::Cint; compress_tol::Cdouble
    storage::Cint; graph::Cint; relax::Cdouble
end
mutable struct dm_solve_stats
    iterations::Cint; outer_iterations::Cint; solved::Cint
    residual::Cdouble; precond_time_s::Cdouble; solve_time_s::Cdouble
    dm_solve_stats() = new(0, 0, 0, 0.0, 0.0, 0.0)
end

function _opts(; method=DM_METHOD_GMRES, precond=DM_PRECOND_NEARFIELD, precond_side=1, maxiter=300, tol=1e-6,
               nf_cutoff=0.0, restart=0, check_convergence=true, compress_tol=0.0, storage=0, block_size=128,
               graph=0, relax=0.0)
    return dm_solve_opts(Cint(method), Cint(precond), Cint(precond_side), Cint(maxiter), Float64(tol),
                         Float64(nf_cutoff), Cint(restart), Cint(check_convergence), 0.0, Cint(10), Cint(block_size),
                         Cint(0), C_NULL, Cint(0), Float64(compress_tol), Cint(storage), Cint(graph),
                         Float64(relax))
end

# ------------------------------------------------------------------ handles
# TriMesh and RWGData are immutable structs, which Julia cannot finalize.  A
# library handle is therefore tied to a mutable array the object owns
# (mesh.xyz, rwg.tplus): a finalizer on that array queues the handle for
# release, and the queue is drained (RWG handles before mesh handles) at the
# next call into the library, at `release!` and at exit.  Finalizers only
# push to the queue: they never block or call into the library.
# The key is the array's objectid, which is stable while the array lives (the
# GC does not move objects) and is dropped by the finalizer before the memory
# can be reused.  Handles assume the arrays are not mutated after first use;
# call `release!` after changing a mesh in place.
const _ctx = Ref{Ptr{Cvoid}}(C_NULL)
const _mesh_handles = Dict{UInt,Ptr{Cvoid}}()
const _rwg_handles = Dict{UInt,Ptr{Cvoid}}()
const _lock = ReentrantLock()
const _dead_lock = Threads.SpinLock()
const _dead_rwg = UInt[]
const _dead_mesh = UInt[]

function context()
    lock(_lock) do
        if _ctx[] == C_NULL
            out = Ref{Ptr{Cvoid}}(C_NULL)
            _check(ccall((:dm_create, libdiffmom), Cint, (Cint, Ptr{Ptr{Cvoid}}), 0, out))
            _ctx[] = out[]
        end
        return _ctx[]
    end
end

function _queue_dead!(list::Vector{UInt}, key::UInt)
    lock(_dead_lock)
    try
        push!(list, key)
    finally
        unlock(_dead_lock)
    end
    return nothing
end

# frees the handles whose arrays were collected (call with _lock held)
function _drain_dead!()
    lock(_dead_lock)
    rk = copy(_dead_rwg); mk = copy(_dead_mesh)
    empty!(_dead_rwg); empty!(_dead_mesh)
    unlock(_dead_lock)
    for key in rk
        h = pop!(_rwg_handles, key, C_NULL)
        h == C_NULL || ccall((:dm_rwg_free, libdiffmom), Cint, (Ptr{Cvoid},), h)
    end
    for key in mk
        h = pop!(_mesh_handles, key, C_NULL)
        h == C_NULL || ccall((:dm_mesh_free, libdiffmom), Cint, (Ptr{Cvoid},), h)
    end
    return nothing
end

function _mesh_handle(mesh::TriMesh)
    lock(_lock) do
        _drain_dead!()
        key = objectid(mesh.xyz)
        get!(_mesh_handles, key) do
            out = Ref{Ptr{Cvoid}}(C_NULL)
            xyz = mesh.xyz; tri = mesh.tri
            GC.@preserve xyz tri _check(ccall((:dm_mesh_create, libdiffmom), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int64}, Int64, Ptr{Ptr{Cvoid}}),
                context(), xyz, size(xyz, 2), tri, size(tri, 2), out))
            finalizer(_ -> _queue_dead!(_dead_mesh, key), mesh.xyz)
            out[]
        end
    end
end

function _rwg_handle(mesh::TriMesh, rwg::RWGData)
    DiffMoM._validate_mesh_rwg_pair(mesh, rwg)
    eltype(rwg.coeff_plus) <: Real ||
        throw(ArgumentError("Bloch-periodic (complex-coefficient) RWG data is not on the GPU path"))
    mh = _mesh_handle(mesh)
    lock(_lock) do
        key = objectid(rwg.tplus)
        get!(_rwg_handles, key) do
            out = Ref{Ptr{Cvoid}}(C_NULL)
            cp = Vector{Float64}(rwg.coeff_plus); cm = Vector{Float64}(rwg.coeff_minus)
            GC.@preserve rwg cp cm _check(ccall((:dm_rwg_create, libdiffmom), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Ptr{Cvoid}}),
                mh, rwg.nedges, rwg.tplus, rwg.tminus, rwg.evert, rwg.vplus_opp, rwg.vminus_opp,
                rwg.len, rwg.area_plus, rwg.area_minus, cp, cm, out))
            finalizer(_ -> _queue_dead!(_dead_rwg, key), rwg.tplus)
            out[]
        end
    end
end

"""
    release!(x)

Free the library handle held for a `TriMesh` / `RWGData` now (it is freed
anyway once the object is collected).  Device matrices and operators of that
RWG set must not be used afterwards.
"""
function release!(rwg::RWGData)
    lock(_lock) do
        h = pop!(_rwg_handles, objectid(rwg.tplus), C_NULL)
        h == C_NULL || ccall((:dm_rwg_free, libdiffmom), Cint, (Ptr{Cvoid},), h)
        _drain_dead!()
    end
    return nothing
end
function release!(mesh::TriMesh)
    lock(_lock) do
        h = pop!(_mesh_handles, objectid(mesh.xyz), C_NULL)
        h == C_NULL || ccall((:dm_mesh_free, libdiffmom), Cint, (Ptr{Cvoid},), h)
        _drain_dead!()
    end
    return nothing
end

function __init__()
    atexit() do
        lock(_lock) do
            for h in values(_rwg_handles)
                ccall((:dm_rwg_free, libdiffmom), Cint, (Ptr{Cvoid},), h)
            end
            for h in values(_mesh_handles)
                ccall((:dm_mesh_free, libdiffmom), Cint, (Ptr{Cvoid},), h)
            end
            empty!(_rwg_handles); empty!(_mesh_handles)
            if _ctx[] != C_NULL
                ccall((:dm_destroy, libdiffmom), Cint, (Ptr{Cvoid},), _ctx[])
                _ctx[] = C_NULL
            end
        end
    end
end

# The reference's own host-side checks of assemble_Z_efie /
# matrixfree_efie_operator (src/assembly/EFIE.jl:297-330,332-398,699-731,975-1000):
# same order, same exceptions.
function _reference_checks(mesh, rwg, k; quad_order, eta0, mesh_precheck, allow_boundary, require_closed,
                           area_tol_rel, max_output_bytes=nothing, max_cache_bytes, max_adjacency_pairs)
    DiffMoM._validate_mesh_rwg_pair(mesh, rwg)
    N = rwg.nedges
    if max_output_bytes !== nothing
        bytes = DiffMoM._checked_array_payload_bytes(ComplexF64, N, N; label="dense EFIE matrix")
        DiffMoM._enforce_payload_limit(bytes, max_output_bytes, "dense EFIE matrix", "max_output_bytes")
    end
    mesh_precheck && DiffMoM.assert_mesh_quality(mesh; allow_boundary=allow_boundary,
                                                 require_closed=require_closed, area_tol_rel=area_tol_rel)
    DiffMoM._validated_efie_prefactors(k, eta0)
    Tcoef = promote_type(eltype(rwg.coeff_plus), eltype(rwg.coeff_minus))
    Nq = length(DiffMoM.tri_quad_rule(quad_order)[2])
    Nq_hi = length(DiffMoM.tri_quad_rule(7)[2])
    fixed = DiffMoM._efie_cache_fixed_payload_bytes(N, DiffMoM.ntriangles(mesh), Nq, Nq_hi, Tcoef,
                                                    SVector{3,Tcoef})
    DiffMoM._build_triangle_adjacency(mesh; fixed_payload_bytes=fixed, max_cache_bytes=max_cache_bytes,
                                      max_adjacency_pairs=max_adjacency_pairs)
    return nothing
end

# --------------------------------------------------------- device matrix
"""GPU-resident EFIE matrix: an `AbstractMatrix{ComplexF64}` whose `mul!`
runs on the MI355X (so it drops into `solve_gmres` / `Krylov.gmres`)."""
mutable struct HIPEFIEMatrix <: AbstractMatrix{ComplexF64}
    h::Ptr{Cvoid}
    N::Int
    form::Cint
    rwg::RWGData   # keeps the RWG handle (and its mesh) alive while the matrix lives
end
Base.size(A::HIPEFIEMatrix) = (A.N, A.N)
function Base.getindex(A::HIPEFIEMatrix, i::Int, j::Int)
    checkbounds(A, i, j)
    out = Ref{ComplexF64}()
    _check(ccall((:dm_zmat_entries, libdiffmom), Cint,
        (Ptr{Cvoid}, Cint, Ref{Int64}, Ref{Int64}, Int64, Ref{ComplexF64}), A.h, A.form, i - 1, j - 1, 1, out))
    return out[]
end
function _apply!(y, A::HIPEFIEMatrix, x, adjoint::Integer)
    length(x) == A.N || throw(DimensionMismatch("x length $(length(x)) != $(A.N)"))
    length(y) == A.N || throw(DimensionMismatch("y length $(length(y)) != $(A.N)"))
    xs = (x === y || !(x isa Vector{ComplexF64})) ? Vector{ComplexF64}(x) : x
    yd = y isa Vector{ComplexF64} ? y : Vector{ComplexF64}(undef, A.N)
    GC.@preserve xs yd _check(ccall((:dm_apply, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Int64, Int64, Ptr{ComplexF64}, Int64),
        context(), A.h, A.form, adjoint, xs, A.N, 1, yd, A.N))
    yd === y || copyto!(y, yd)
    return y
end
LinearAlgebra.mul!(y::AbstractVector, A::HIPEFIEMatrix, x::AbstractVector) = _apply!(y, A, x, 0)
LinearAlgebra.mul!(y::AbstractVector, At::Adjoint{ComplexF64,HIPEFIEMatrix}, x::AbstractVector) =
    _apply!(y, parent(At), x, 1)
Base.:*(A::HIPEFIEMatrix, x::AbstractVector) = mul!(Vector{ComplexF64}(undef, A.N), A, x)

"""
    assemble_Z_efie(mesh, rwg, k; quad_order=3, eta0=376.730313668, mesh_precheck=true,
                    allow_boundary=true, require_closed=false, area_tol_rel=1e-12,
                    max_output_bytes=2_000_000_000, max_cache_bytes=2_000_000_000,
                    max_adjacency_pairs=20_000_000, device=false, precision=:fp64)

Drop-in for `DiffMoM.assemble_Z_efie` (src/assembly/EFIE.jl:699-753).
`device=true` keeps Z on the GPU and returns a `HIPEFIEMatrix`;
`precision=:fp32` stores the complex64 form (FP32 far quadrature, FP64
singular terms).
"""
function assemble_Z_efie(mesh::TriMesh, rwg::RWGData, k; quad_order::Int=3, eta0=376.730313668,
                         mesh_precheck::Bool=true, allow_boundary::Bool=true, require_closed::Bool=false,
                         area_tol_rel::Float64=1e-12,
                         max_output_bytes::Integer=DiffMoM._DEFAULT_MAX_DENSE_PAYLOAD_BYTES,
                         max_cache_bytes::Integer=DiffMoM._DEFAULT_MAX_EFIE_CACHE_BYTES,
                         max_adjacency_pairs::Integer=DiffMoM._DEFAULT_MAX_EFIE_ADJACENCY_PAIRS,
                         device::Bool=false, precision::Symbol=:fp64)
    precision in (:fp64, :fp32) || throw(ArgumentError("precision must be :fp64 or :fp32, got :$precision"))
    _reference_checks(mesh, rwg, k; quad_order, eta0, mesh_precheck, allow_boundary, require_closed,
                      area_tol_rel, max_output_bytes, max_cache_bytes, max_adjacency_pairs)
    rh = _rwg_handle(mesh, rwg)
    form = precision === :fp32 ? DM_Z32 : DM_Z64
    out = Ref{Ptr{Cvoid}}(C_NULL)
    kc = ComplexF64(k)
    _check(ccall((:dm_assemble_z, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cint, Cint, Int64, Ptr{Ptr{Cvoid}}),
        context(), rh, real(kc), imag(kc), Float64(eta0), quad_order, form, 0, out))
    A = HIPEFIEMatrix(out[], rwg.nedges, form, rwg)
    finalizer(a -> ccall((:dm_zmat_free, libdiffmom), Cint, (Ptr{Cvoid},), a.h), A)
    device && return A
    Z = Matrix{ComplexF64}(undef, A.N, A.N)
    _check(ccall((:dm_zmat_get, libdiffmom), Cint, (Ptr{Cvoid}, Cint, Ptr{ComplexF64}), A.h, form, Z))
    return Z
end

# ------------------------------------------------------- matrix-free operator
"""Matrix-free EFIE operator on the GPU (`MatrixFreeEFIEOperator`,
src/assembly/EFIE.jl:759-1007): every product recomputes all entries in FP64
on the device; nothing of size N^2 is stored.  `adjoint(A)` multiplies by
Z^H."""
mutable struct HIPMatrixFreeEFIEOperator <: AbstractMatrix{ComplexF64}
    h::Ptr{Cvoid}
    N::Int
    rwg::RWGData
end
Base.size(A::HIPMatrixFreeEFIEOperator) = (A.N, A.N)
function _mf_apply!(y, A::HIPMatrixFreeEFIEOperator, x, adjoint::Integer)
    length(x) == A.N || throw(DimensionMismatch("x length $(length(x)) != $(A.N)"))
    length(y) == A.N || throw(DimensionMismatch("y length $(length(y)) != $(A.N)"))
    xs = (x === y || !(x isa Vector{ComplexF64})) ? Vector{ComplexF64}(x) : x
    yd = y isa Vector{ComplexF64} ? y : Vector{ComplexF64}(undef, A.N)
    GC.@preserve xs yd _check(ccall((:dm_mf_apply, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{ComplexF64}, Ptr{ComplexF64}), context(), A.h, adjoint, xs, yd))
    yd === y || copyto!(y, yd)
    return y
end
LinearAlgebra.mul!(y::AbstractVector, A::HIPMatrixFreeEFIEOperator, x::AbstractVector) = _mf_apply!(y, A, x, 0)
LinearAlgebra.mul!(y::AbstractVector, At::Adjoint{ComplexF64,HIPMatrixFreeEFIEOperator}, x::AbstractVector) =
    _mf_apply!(y, parent(At), x, 1)
Base.:*(A::HIPMatrixFreeEFIEOperator, x::AbstractVector) = mul!(Vector{ComplexF64}(undef, A.N), A, x)
# A[i, j]: column j of the recomputed operator (a full matrix-free product;
# the AbstractMatrix protocol needs it, solvers do not)
function Base.getindex(A::HIPMatrixFreeEFIEOperator, i::Int, j::Int)
    checkbounds(A, i, j)
    e = zeros(ComplexF64, A.N); e[j] = 1
    return (A * e)[i]
end

"""
    matrixfree_efie_operator(mesh, rwg, k; quad_order=3, eta0=376.730313668, mesh_precheck=true,
                             allow_boundary=true, require_closed=false, area_tol_rel=1e-12,
                             max_cache_bytes=2_000_000_000, max_adjacency_pairs=20_000_000)

Drop-in for `DiffMoM.matrixfree_efie_operator` (src/assembly/EFIE.jl:975-1000).
"""
function matrixfree_efie_operator(mesh::TriMesh, rwg::RWGData, k; quad_order::Int=3, eta0=376.730313668,
                                  mesh_precheck::Bool=true, allow_boundary::Bool=true,
                                  require_closed::Bool=false, area_tol_rel::Float64=1e-12,
                                  max_cache_bytes::Integer=DiffMoM._DEFAULT_MAX_EFIE_CACHE_BYTES,
                                  max_adjacency_pairs::Integer=DiffMoM._DEFAULT_MAX_EFIE_ADJACENCY_PAIRS)
    _reference_checks(mesh, rwg, k; quad_order, eta0, mesh_precheck, allow_boundary, require_closed,
                      area_tol_rel, max_cache_bytes, max_adjacency_pairs)
    rh = _rwg_handle(mesh, rwg)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    kc = ComplexF64(k)
    _check(ccall((:dm_mf_create, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cint, Ptr{Ptr{Cvoid}}),
        context(), rh, real(kc), imag(kc), Float64(eta0), quad_order, out))
    A = HIPMatrixFreeEFIEOperator(out[], rwg.nedges, rwg)
    finalizer(a -> ccall((:dm_mf_free, libdiffmom), Cint, (Ptr{Cvoid},), a.h), A)
    return A
end
matrixfree_efie_adjoint_operator(A::HIPMatrixFreeEFIEOperator) = adjoint(A)

# -------------------------------------------------------------- excitation
"""Drop-in for `DiffMoM.assemble_multiple_excitations` with plane waves
(src/assembly/Excitation.jl:2738-2797)."""
function assemble_multiple_excitations(mesh::TriMesh, rwg::RWGData, exc::Vector{PlaneWaveExcitation};
                                       quad_order::Int=3,
                                       max_output_bytes::Integer=DiffMoM._DEFAULT_MAX_DENSE_PAYLOAD_BYTES,
                                       max_work_bytes::Integer=DiffMoM._DEFAULT_MAX_EXCITATION_WORK_BYTES,
                                       max_terms::Integer=DiffMoM._DEFAULT_MAX_EXCITATION_TERMS)
    DiffMoM._validate_mesh_rwg_pair(mesh, rwg)
    foreach(DiffMoM._validate_excitation_model, exc)
    M = length(exc)
    bytes = DiffMoM._checked_array_payload_bytes(ComplexF64, rwg.nedges, M; label="multiple-excitation RHS matrix")
    DiffMoM._enforce_payload_limit(bytes, max_output_bytes, "multiple-excitation RHS matrix", "max_output_bytes")
    DiffMoM._validated_resource_limit("max_work_bytes", max_work_bytes)
    DiffMoM._validated_resource_limit("max_terms", max_terms)
    rh = _rwg_handle(mesh, rwg)
    kv = reduce(hcat, [Vector{Float64}(real.(e.k_vec)) for e in exc])
    pol = reduce(hcat, [Vector{Float64}(real.(e.pol)) for e in exc])
    E0 = Float64[real(e.E0) for e in exc]
    V = Matrix{ComplexF64}(undef, rwg.nedges, M)
    _check(ccall((:dm_excitation_planewave, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Cint, Ptr{ComplexF64}),
        context(), rh, kv, pol, E0, M, quad_order, V))
    return V
end

# ----------------------------------------------------------- preconditioner
"""Near-field preconditioner on the GPU (`NearFieldOperator`'s `mul!`,
src/solver/NearFieldPreconditioner.jl:1034-1060): `mul!(y, P, x)` applies a
sparse approximate inverse of the near-field matrix (`factorization=:spai`),
Jacobi (`:diag`) or block Jacobi (`:block`)."""
mutable struct HIPNearFieldOperator <: AbstractMatrix{ComplexF64}
    h::Ptr{Cvoid}
    N::Int
    A::HIPEFIEMatrix
    kind::Cint
    cutoff::Float64
    block_size::Int
end
Base.size(P::HIPNearFieldOperator) = (P.N, P.N)
Base.getindex(P::HIPNearFieldOperator, i::Int, j::Int) =
    throw(ArgumentError("HIPNearFieldOperator supports mul! only"))
function LinearAlgebra.mul!(y::AbstractVector, P::HIPNearFieldOperator, x::AbstractVector)
    xs = Vector{ComplexF64}(x)
    yd = Vector{ComplexF64}(undef, P.N)
    GC.@preserve xs yd _check(ccall((:dm_precond_apply, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}), context(), P.h, xs, yd))
    return copyto!(y, yd)
end

const _NO_FACTOR = "the GPU path does not factor the near-field matrix; it builds a sparse approximate " *
                   "inverse of it (factorization=:spai), or Jacobi (:diag) / block Jacobi (:block)"

"""
    build_nearfield_preconditioner(A::HIPEFIEMatrix, mesh, rwg, cutoff; factorization=:spai, block_size=128)

GPU counterpart of `DiffMoM.build_nearfield_preconditioner`
(src/solver/NearFieldPreconditioner.jl:732-759).  The reference's `:lu` /
`:ilu` factorisations are refused with ArgumentError.
"""
function build_nearfield_preconditioner(A::HIPEFIEMatrix, mesh::TriMesh, rwg::RWGData, cutoff::Real;
                                        factorization::Symbol=:spai, block_size::Int=128)
    factorization in (:lu, :ilu) && throw(ArgumentError("factorization=:$factorization: " * _NO_FACTOR))
    kind = factorization === :spai ? DM_PRECOND_NEARFIELD : factorization === :diag ? DM_PRECOND_DIAG :
           factorization === :block ? Cint(4) :
           throw(ArgumentError("factorization must be :spai, :diag or :block, got :$factorization"))
    (isfinite(cutoff) && cutoff >= 0) || throw(ArgumentError("cutoff must be finite and nonnegative"))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    _check(ccall((:dm_precond_create, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Cint, Cint, Ptr{Ptr{Cvoid}}),
        context(), A.h, kind, Float64(cutoff), block_size, A.form, out))
    P = HIPNearFieldOperator(out[], A.N, A, kind, Float64(cutoff), block_size)
    finalizer(p -> ccall((:dm_precond_free, libdiffmom), Cint, (Ptr{Cvoid},), p.h), P)
    return P
end

"""
    nearfield_pattern_info(rwg, cutoff) -> (nnz, mean_radius, max_cols, capped_cols)

The pattern the sparse approximate inverse uses for `cutoff` metres: each
column keeps the nearest `max_cols` bases within the cutoff; `capped_cols`
columns had more bases within the cutoff than that.
"""
function nearfield_pattern_info(mesh::TriMesh, rwg::RWGData, cutoff::Real)
    rh = _rwg_handle(mesh, rwg)
    nnz = Ref{Int64}(0); rad = Ref{Cdouble}(0); cap = Ref{Cint}(0); capped = Ref{Int64}(0)
    _check(ccall((:dm_nearfield_pattern_info, libdiffmom), Cint,
        (Ptr{Cvoid}, Cdouble, Ref{Int64}, Ref{Cdouble}, Ref{Cint}, Ref{Int64}),
        rh, Float64(cutoff), nnz, rad, cap, capped))
    return (nnz[], rad[], Int(cap[]), capped[])
end

# ------------------------------------------------------------------- solves
"""
    solve_gmres(A::HIPEFIEMatrix, rhs; preconditioner=nothing, precond_side=:left, tol=1e-8,
                maxiter=200, memory=20, max_workspace_bytes=536_870_912, verbose=false,
                check_gmres_convergence=true)

`DiffMoM.solve_gmres` (src/solver/IterativeSolve.jl:341-392) with the whole
Krylov loop on the GPU: `preconditioner` is `nothing` or a
`HIPNearFieldOperator` built from the same matrix.  Returns `(x, stats)`
with `stats` a NamedTuple (niter, solved, residual).
"""
function solve_gmres(A::HIPEFIEMatrix, rhs::AbstractVector; preconditioner=nothing, precond_side::Symbol=:left,
                     tol::Float64=1e-8, maxiter::Int=200, memory::Int=20,
                     max_workspace_bytes::Integer=DiffMoM._DEFAULT_MAX_GMRES_WORKSPACE_BYTES, verbose::Bool=false,
                     check_gmres_convergence::Bool=true)
    DiffMoM._validate_gmres_options(tol, maxiter, memory, precond_side)
    length(rhs) == A.N || throw(DimensionMismatch("rhs length $(length(rhs)) != $(A.N)"))
    limit = DiffMoM._validated_resource_limit("max_workspace_bytes", max_workspace_bytes)
    DiffMoM._preflight_gmres_workspace(A.N, memory, limit)
    pc, cut, bs = DM_PRECOND_NONE, 0.0, 128
    if preconditioner !== nothing
        (preconditioner isa HIPNearFieldOperator && preconditioner.A === A) ||
            throw(ArgumentError("preconditioner must be a HIPNearFieldOperator built from this matrix"))
        pc, cut, bs = preconditioner.kind, preconditioner.cutoff, preconditioner.block_size
    end
    b = Vector{ComplexF64}(rhs)
    x = Vector{ComplexF64}(undef, A.N)
    st = dm_solve_stats()
    opts = _opts(; method=DM_METHOD_GMRES, precond=pc, precond_side=precond_side === :right ? 1 : 0,
                 maxiter, tol, nf_cutoff=cut, check_convergence=check_gmres_convergence, block_size=bs)
    GC.@preserve b x _check(ccall((:dm_solve, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}, Int64, Int64, Ref{dm_solve_opts}, Ptr{ComplexF64}, Int64,
         Ref{dm_solve_stats}), context(), A.h, b, A.N, 1, opts, x, A.N, st))
    return x, (niter=Int(st.iterations), solved=st.solved == 1, residual=st.residual)
end

"""
    solve_scattering(mesh, freq_hz, excitation; method=:auto, dense_direct_limit=2000,
                     dense_gmres_limit=10000, mlfma_threshold=50000, check_resolution=true,
                     points_per_wavelength=10.0, error_on_underresolved=false, gmres_tol=1e-6,
                     gmres_maxiter=300, check_gmres_convergence=true, nf_cutoff_lambda=1.0,
                     preconditioner=:auto, aca_tol=1e-6, aca_leaf_size=64, aca_eta=1.5,
                     aca_max_rank=50, verbose=true, quad_order=3, c0=299792458.0,
                     max_dense_matrix_bytes=2_000_000_000, precision=:fp64, compress_tol=5e-8)

The dense branches of `DiffMoM.solve_scattering` (src/Workflow.jl:67-317) end
to end on the GPU, with the reference's keywords and checks.

* `:aca_gmres` / `:mlfma` (chosen or selected by `:auto`) are not on the GPU
  path: ArgumentError (force `method=:dense_gmres`).
* `preconditioner=:auto` builds the sparse approximate inverse of the
  near-field matrix over `nf_cutoff_lambda`·λ (each column keeps its nearest
  80 bases within the cutoff; a warning says when that cap cuts the cutoff
  short); `:diag` is Jacobi, `:none` none; `:lu` / `:ilu` raise ArgumentError.
* `precision=:mixed` iterates on the complex64 matrix with FP64 Arnoldi and
  FP64 refinement to the same true residual; with `compress_tol > 0` the inner
  iterations use the cross-approximated complex64 matrix
  (||Z - Z_H||_F <= compress_tol ||Z||_F) while the refinement residual stays
  the dense complex128 product.
* Plane waves are assembled on the GPU; other excitations with the
  reference's `assemble_excitation`, then solved on the GPU.
"""
function solve_scattering(mesh::TriMesh, freq_hz::Real, excitation; method::Symbol=:auto,
                          dense_direct_limit::Int=2000, dense_gmres_limit::Int=10000, mlfma_threshold::Int=50000,
                          check_resolution::Bool=true, points_per_wavelength::Real=10.0,
                          error_on_underresolved::Bool=false, gmres_tol::Float64=1e-6, gmres_maxiter::Int=300,
                          check_gmres_convergence::Bool=true, nf_cutoff_lambda::Float64=1.0,
                          preconditioner::Symbol=:auto, aca_tol::Float64=1e-6, aca_leaf_size::Int=64,
                          aca_eta::Float64=1.5, aca_max_rank::Int=50, verbose::Bool=true, quad_order::Int=3,
                          c0::Real=DiffMoM.C0_DEFAULT,
                          max_dense_matrix_bytes::Integer=DiffMoM._DEFAULT_MAX_DENSE_PAYLOAD_BYTES,
                          precision::Symbol=:fp64, compress_tol::Float64=5e-8)
    frequency = Float64(freq_hz)
    (isfinite(frequency) && frequency > 0) ||
        throw(ArgumentError("solve_scattering: freq_hz must be finite and positive, got $freq_hz"))
    speed = Float64(c0)
    (isfinite(speed) && speed > 0) || throw(ArgumentError("solve_scattering: c0 must be finite and positive, got $c0"))
    method in (:auto, :dense_direct, :dense_gmres, :aca_gmres, :mlfma) ||
        error("solve_scattering: method must be :auto, :dense_direct, :dense_gmres, :aca_gmres, or :mlfma")
    preconditioner in (:auto, :lu, :ilu, :diag, :none) ||
        throw(ArgumentError("solve_scattering: preconditioner must be :auto, :lu, :ilu, :diag, or :none"))
    preconditioner in (:lu, :ilu) &&
        throw(ArgumentError("solve_scattering: preconditioner=:$preconditioner: " * _NO_FACTOR *
                            "; preconditioner=:auto builds the sparse approximate inverse"))
    precision in (:fp64, :mixed) || throw(ArgumentError("precision must be :fp64 or :mixed, got :$precision"))
    (isfinite(nf_cutoff_lambda) && nf_cutoff_lambda > 0) ||
        throw(ArgumentError("solve_scattering: nf_cutoff_lambda must be finite and positive, got $nf_cutoff_lambda"))
    dense_direct_limit >= 0 ||
        throw(ArgumentError("solve_scattering: dense_direct_limit must be nonnegative, got $dense_direct_limit"))
    dense_gmres_limit >= dense_direct_limit ||
        throw(ArgumentError("solve_scattering: dense_gmres_limit ($dense_gmres_limit) must be at least " *
                            "dense_direct_limit ($dense_direct_limit)"))
    mlfma_threshold >= dense_gmres_limit ||
        throw(ArgumentError("solve_scattering: mlfma_threshold ($mlfma_threshold) must be at least " *
                            "dense_gmres_limit ($dense_gmres_limit)"))
    warnings = String[]
    lambda = speed / frequency
    k = 2π * frequency / speed
    excitation isa PlaneWaveExcitation &&
        DiffMoM._validate_plane_wave_wavenumber(excitation, k, "solve_scattering")
    report = DiffMoM.mesh_resolution_report(mesh, frequency; points_per_wavelength=Float64(points_per_wavelength),
                                            c0=speed)
    if check_resolution && !report.meets_target
        msg = "Mesh under-resolved: edge_max/lambda=$(round(report.edge_max_over_lambda, digits=3)), " *
              "target <= $(round(1.0/points_per_wavelength, digits=3)). Results may be inaccurate."
        push!(warnings, msg)
        error_on_underresolved && error("solve_scattering: $msg")
        verbose && println("  WARNING: $msg")
    end
    rwg = build_rwg(mesh)
    N = rwg.nedges
    N >= 1 || throw(ArgumentError("solve_scattering: mesh produces no RWG unknowns; at least two triangles " *
                                  "sharing an edge are required"))
    sel = method
    if method === :auto
        sel = N <= dense_direct_limit ? :dense_direct : N <= dense_gmres_limit ? :dense_gmres :
              N <= mlfma_threshold ? :aca_gmres : :mlfma
    end
    sel in (:aca_gmres, :mlfma) &&
        throw(ArgumentError("solve_scattering: method :$sel (N=$N) is not on the GPU path; " *
                            "force method=:dense_gmres"))
    bytes = DiffMoM._checked_array_payload_bytes(ComplexF64, N, N; label="solve_scattering dense EFIE matrix")
    DiffMoM._enforce_payload_limit(bytes, max_dense_matrix_bytes, "solve_scattering dense EFIE matrix",
                                   "max_dense_matrix_bytes")
    rh = _rwg_handle(mesh, rwg)
    pc = DM_PRECOND_NONE
    cutoff = nf_cutoff_lambda * lambda
    if sel === :dense_gmres
        pc = preconditioner === :diag ? DM_PRECOND_DIAG : preconditioner === :none ? DM_PRECOND_NONE :
             DM_PRECOND_NEARFIELD
        if pc == DM_PRECOND_NEARFIELD
            nnz, rad, cap, capped = nearfield_pattern_info(mesh, rwg, cutoff)
            capped > 0 && push!(warnings, "near-field preconditioner: sparse approximate inverse over " *
                                          "nf_cutoff_lambda=$nf_cutoff_lambda ($(round(cutoff, sigdigits=4)) m) " *
                                          "keeps the $cap nearest bases per column ($capped of $N columns cut " *
                                          "short; mean radius $(round(rad, sigdigits=4)) m, $nnz nonzeros)")
        end
    end
    meth = sel === :dense_direct ? DM_METHOD_LU : (precision === :mixed ? DM_METHOD_MIXED : DM_METHOD_GMRES)
    opts = _opts(; method=meth, precond=pc, maxiter=gmres_maxiter, tol=gmres_tol, nf_cutoff=cutoff,
                 check_convergence=check_gmres_convergence,
                 compress_tol=precision === :mixed ? compress_tol : 0.0)
    X = Vector{ComplexF64}(undef, N)
    st = dm_solve_stats()
    tasm = Ref{Cdouble}(0.0)
    if excitation isa PlaneWaveExcitation
        kv = Vector{Float64}(real.(excitation.k_vec)); pol = Vector{Float64}(real.(excitation.pol))
        E0 = [Float64(real(excitation.E0))]
        _check(ccall((:dm_solve_scattering_pw, libdiffmom), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
             Int64, Ref{dm_solve_opts}, Ptr{ComplexF64}, Ref{dm_solve_stats}, Ref{Cdouble}, Ptr{Ptr{Cvoid}}),
            context(), rh, k, 0.0, 376.730313668, quad_order, kv, pol, E0, 1, opts, X, st, tasm, C_NULL))
    else
        v = excitation isa AbstractVector ? Vector{ComplexF64}(excitation) :
            DiffMoM.assemble_excitation(mesh, rwg, excitation; quad_order=quad_order)
        length(v) == N || throw(DimensionMismatch("solve_scattering: excitation length $(length(v)) != N=$N"))
        all(isfinite, v) || throw(ArgumentError("solve_scattering: excitation vector must contain only finite values"))
        forms = precision === :mixed ? (DM_Z64 | DM_Z32) : DM_Z64
        zh = Ref{Ptr{Cvoid}}(C_NULL)
        tasm[] = @elapsed _check(ccall((:dm_assemble_z, libdiffmom), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cint, Cint, Int64, Ptr{Ptr{Cvoid}}),
            context(), rh, k, 0.0, 376.730313668, quad_order, forms, 0, zh))
        try
            GC.@preserve v X _check(ccall((:dm_solve, libdiffmom), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}, Int64, Int64, Ref{dm_solve_opts}, Ptr{ComplexF64}, Int64,
                 Ref{dm_solve_stats}), context(), zh[], v, N, 1, opts, X, N, st))
        finally
            ccall((:dm_zmat_free, libdiffmom), Cint, (Ptr{Cvoid},), zh[])
        end
    end
    iters = sel === :dense_gmres ? Int(st.iterations) : -1
    resid = sel === :dense_gmres ? st.residual : NaN
    return ScatteringResult(X, sel, N, tasm[], st.solve_time_s, st.precond_time_s, iters, resid, report, warnings)
end

end # module
