"""
    DiffMoMHIP

Julia binding of libdiffmom.so (MI355X / gfx950) for DiffMoM.jl's dense
surface-EFIE path.  Every function is a `ccall` into the C ABI declared in
include/diffmom.h; DiffMoM's own types (`TriMesh`, `RWGData`,
`PlaneWaveExcitation`, `ScatteringResult`) pass through unchanged because the
library takes Julia's in-memory layout (column-major 3xNv / 3xNt, 1-based
Int64 indices, interleaved ComplexF64).

    using DiffMoM, DiffMoMHIP
    mesh = make_rect_plate(0.1, 0.1, 6, 6)
    rwg  = build_rwg(mesh)
    Z    = DiffMoMHIP.assemble_Z_efie(mesh, rwg, 2π/0.1)            # Matrix{ComplexF64}
    A    = DiffMoMHIP.assemble_Z_efie(mesh, rwg, 2π/0.1; device=true) # GPU AbstractMatrix
    x, stats = solve_gmres(A, v)                                      # Krylov.jl calls mul! on the GPU
"""
module DiffMoMHIP

using LinearAlgebra
using DiffMoM

const libdiffmom = joinpath(@__DIR__, "..", "libdiffmom.so")

const DM_Z64 = Cint(1)
const DM_Z32 = Cint(2)
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
end
mutable struct dm_solve_stats
    iterations::Cint; outer_iterations::Cint; solved::Cint
    residual::Cdouble; precond_time_s::Cdouble; solve_time_s::Cdouble
    dm_solve_stats() = new(0, 0, 0, 0.0, 0.0, 0.0)
end

# ------------------------------------------------------------------ handles
const _ctx = Ref{Ptr{Cvoid}}(C_NULL)
const _handles = IdDict{Any,Ptr{Cvoid}}()   # TriMesh / RWGData object -> library handle
const _lock = ReentrantLock()

function context()
    if _ctx[] == C_NULL
        out = Ref{Ptr{Cvoid}}(C_NULL)
        _check(ccall((:dm_create, libdiffmom), Cint, (Cint, Ptr{Ptr{Cvoid}}), 0, out))
        _ctx[] = out[]
    end
    return _ctx[]
end

function _mesh_handle(mesh::TriMesh)
    lock(_lock) do
        get!(_handles, mesh) do
            out = Ref{Ptr{Cvoid}}(C_NULL)
            GC.@preserve mesh _check(ccall((:dm_mesh_create, libdiffmom), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int64}, Int64, Ptr{Ptr{Cvoid}}),
                context(), mesh.xyz, size(mesh.xyz, 2), mesh.tri, size(mesh.tri, 2), out))
            h = out[]
            finalizer(m -> ccall((:dm_mesh_free, libdiffmom), Cint, (Ptr{Cvoid},), h), mesh)
            h
        end
    end
end

function _rwg_handle(mesh::TriMesh, rwg::RWGData)
    rwg.mesh === mesh || throw(ArgumentError("RWG data must be built from the supplied mesh object."))
    eltype(rwg.coeff_plus) <: Real ||
        throw(ArgumentError("Bloch-periodic (complex-coefficient) RWG data is not on the GPU path"))
    mh = _mesh_handle(mesh)
    lock(_lock) do
        get!(_handles, rwg) do
            out = Ref{Ptr{Cvoid}}(C_NULL)
            cp = Vector{Float64}(rwg.coeff_plus); cm = Vector{Float64}(rwg.coeff_minus)
            GC.@preserve rwg cp cm _check(ccall((:dm_rwg_create, libdiffmom), Cint,
                (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Ptr{Cvoid}}),
                mh, rwg.nedges, rwg.tplus, rwg.tminus, rwg.evert, rwg.vplus_opp, rwg.vminus_opp,
                rwg.len, rwg.area_plus, rwg.area_minus, cp, cm, out))
            h = out[]
            finalizer(r -> ccall((:dm_rwg_free, libdiffmom), Cint, (Ptr{Cvoid},), h), rwg)
            h
        end
    end
end

# --------------------------------------------------------- device matrix
"""GPU-resident EFIE matrix: an `AbstractMatrix{ComplexF64}` whose `mul!`
runs on the MI355X (so it drops into `solve_gmres` / `Krylov.gmres`)."""
mutable struct HIPEFIEMatrix <: AbstractMatrix{ComplexF64}
    h::Ptr{Cvoid}
    N::Int
    form::Cint
end
Base.size(A::HIPEFIEMatrix) = (A.N, A.N)
function Base.getindex(A::HIPEFIEMatrix, i::Int, j::Int)
    out = Ref{ComplexF64}()
    _check(ccall((:dm_zmat_entries, libdiffmom), Cint,
        (Ptr{Cvoid}, Cint, Ref{Int64}, Ref{Int64}, Int64, Ref{ComplexF64}), A.h, A.form, i - 1, j - 1, 1, out))
    return out[]
end
function LinearAlgebra.mul!(y::StridedVector{ComplexF64}, A::HIPEFIEMatrix, x::StridedVector{ComplexF64})
    length(x) == A.N || throw(DimensionMismatch("x length $(length(x)) != $(A.N)"))
    length(y) == A.N || throw(DimensionMismatch("y length $(length(y)) != $(A.N)"))
    xs = x === y ? copy(x) : x
    GC.@preserve xs y _check(ccall((:dm_apply, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Int64, Int64, Ptr{ComplexF64}, Int64),
        context(), A.h, A.form, 0, xs, A.N, 1, y, A.N))
    return y
end
function LinearAlgebra.mul!(y::StridedVector{ComplexF64}, At::Adjoint{ComplexF64,HIPEFIEMatrix},
                            x::StridedVector{ComplexF64})
    A = parent(At)
    xs = x === y ? copy(x) : x
    GC.@preserve xs y _check(ccall((:dm_apply, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Int64, Int64, Ptr{ComplexF64}, Int64),
        context(), A.h, A.form, 1, xs, A.N, 1, y, A.N))
    return y
end
Base.:*(A::HIPEFIEMatrix, x::AbstractVector) = mul!(similar(x, ComplexF64, A.N), A, Vector{ComplexF64}(x))

"""
    assemble_Z_efie(mesh, rwg, k; quad_order=3, eta0=376.730313668,
                    max_output_bytes=2_000_000_000, device=false, precision=:fp64)

Drop-in for `DiffMoM.assemble_Z_efie` (src/assembly/EFIE.jl:699-753).
`device=true` keeps Z on the GPU and returns a `HIPEFIEMatrix`;
`precision=:fp32` stores the complex64 form (FP32 far quadrature, FP64
singular terms).
"""
function assemble_Z_efie(mesh::TriMesh, rwg::RWGData, k; quad_order::Int=3, eta0=376.730313668,
                         max_output_bytes::Integer=2_000_000_000, device::Bool=false, precision::Symbol=:fp64)
    rh = _rwg_handle(mesh, rwg)
    form = precision === :fp32 ? DM_Z32 : DM_Z64
    out = Ref{Ptr{Cvoid}}(C_NULL)
    kc = ComplexF64(k)
    _check(ccall((:dm_assemble_z, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cint, Cint, Int64, Ptr{Ptr{Cvoid}}),
        context(), rh, real(kc), imag(kc), Float64(eta0), quad_order, form, max_output_bytes, out))
    A = HIPEFIEMatrix(out[], rwg.nedges, form)
    finalizer(a -> ccall((:dm_zmat_free, libdiffmom), Cint, (Ptr{Cvoid},), a.h), A)
    device && return A
    Z = Matrix{ComplexF64}(undef, A.N, A.N)
    _check(ccall((:dm_zmat_get, libdiffmom), Cint, (Ptr{Cvoid}, Cint, Ptr{ComplexF64}), A.h, form, Z))
    return Z
end

"""Drop-in for `DiffMoM.assemble_multiple_excitations` with plane waves
(src/assembly/Excitation.jl:2738-2797)."""
function assemble_multiple_excitations(mesh::TriMesh, rwg::RWGData, exc::Vector{PlaneWaveExcitation};
                                       quad_order::Int=3)
    rh = _rwg_handle(mesh, rwg)
    M = length(exc)
    kv = reduce(hcat, [Vector(e.k_vec) for e in exc]); pol = reduce(hcat, [Vector(e.pol) for e in exc])
    E0 = [e.E0 for e in exc]
    V = Matrix{ComplexF64}(undef, rwg.nedges, M)
    _check(ccall((:dm_excitation_planewave, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Cint, Ptr{ComplexF64}),
        context(), rh, kv, pol, E0, M, quad_order, V))
    return V
end

"""
    solve_scattering(mesh, freq_hz, pw::PlaneWaveExcitation; method=:dense_gmres,
                     precision=:fp64, gmres_tol=1e-6, gmres_maxiter=300, c0=299792458.0,
                     compress_tol=5e-8)

Dense branches of `DiffMoM.solve_scattering` (src/Workflow.jl:67-317) end to end on
the GPU.  `precision=:mixed` iterates on the complex64 matrix with FP64 Arnoldi
and FP64 refinement to the same true residual; with `compress_tol > 0` the
inner iterations use the cross-approximated complex64 matrix
(||Z - Z_H||_F <= compress_tol ||Z||_F) while the refinement residual stays the
dense complex128 product.  `compress_tol=0` iterates on the dense complex64 matrix.
"""
function solve_scattering(mesh::TriMesh, freq_hz::Real, pw::PlaneWaveExcitation; method::Symbol=:dense_gmres,
                          precision::Symbol=:fp64, gmres_tol::Float64=1e-6, gmres_maxiter::Int=300,
                          quad_order::Int=3, c0::Real=299792458.0, nf_cutoff::Float64=0.0,
                          compress_tol::Float64=5e-8)
    rwg = build_rwg(mesh)
    rh = _rwg_handle(mesh, rwg)
    k = 2π * Float64(freq_hz) / Float64(c0)
    meth = method === :dense_direct ? DM_METHOD_LU : (precision === :mixed ? DM_METHOD_MIXED : DM_METHOD_GMRES)
    opts = dm_solve_opts(meth, method === :dense_direct ? DM_PRECOND_NONE : DM_PRECOND_NEARFIELD, 1,
                         gmres_maxiter, gmres_tol, nf_cutoff, 0, 1, 0.0, 10, 128,
                         0, C_NULL, 0, precision === :mixed ? compress_tol : 0.0)
    X = Vector{ComplexF64}(undef, rwg.nedges)
    st = dm_solve_stats(); tasm = Ref{Cdouble}(0.0)
    kv = Vector(pw.k_vec); pol = Vector(pw.pol); E0 = [pw.E0]
    t = @elapsed _check(ccall((:dm_solve_scattering_pw, libdiffmom), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64,
         Ref{dm_solve_opts}, Ptr{ComplexF64}, Ref{dm_solve_stats}, Ref{Cdouble}, Ptr{Ptr{Cvoid}}),
        context(), rh, k, 0.0, 376.730313668, quad_order, kv, pol, E0, 1, opts, X, st, tasm, C_NULL))
    report = mesh_resolution_report(mesh, freq_hz; c0=c0)
    return ScatteringResult(X, method, rwg.nedges, tasm[], st.solve_time_s, st.precond_time_s,
                            method === :dense_direct ? -1 : Int(st.iterations), st.residual, report, String[])
end

end # module
