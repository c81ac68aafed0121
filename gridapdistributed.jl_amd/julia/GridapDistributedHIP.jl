"""
    GridapDistributedHIP

Thin Julia binding of libgdx.so (C ABI in include/gdx.h): a
`HIPSparseMatrixAssembler <: SparseMatrixAssembler` that lowers the supported
weak-form vocabulary (∫(κ*∇(v)⋅∇(u))dΩ, ∫(f*v)dΩ on a CartesianDiscreteModel
with Lagrangian Q1/Q2, Dirichlet on "boundary" with affine data) to a problem
descriptor, assembles on one MI355X and returns PartitionedArrays
PSparseMatrix/PVector with the reference's row/col PRanges
(src/Algebra.jl:1013-1015, 1240-1257).  Anything else falls back to the
reference DistributedSparseMatrixAssembler unchanged.

Not executed in this repository (no Julia toolchain is available to the build);
the binding mirrors the ctypes binding that the tests exercise
(gridapdistributed.jl_amd/gridap_amd/_lib.py).
"""
module GridapDistributedHIP

using Gridap
using Gridap.FESpaces
using GridapDistributed
using PartitionedArrays
using SparseMatricesCSR
using LinearAlgebra

const libgdx = joinpath(@__DIR__, "..", "libgdx.so")

struct GdxProblemDesc
  dim::Int32; order::Int32; qdeg::Int32; strategy::Int32
  cells::NTuple{3,Int64}; parts::NTuple{3,Int32}; domain::NTuple{6,Float64}
  coef_const::Float64; nmodes::Int32; modes::Ptr{Float64}
  rhs_kind::Int32; rhs_const::Float64; dir_kind::Int32; dir_coef::NTuple{4,Float64}
  op_kind::Int32; flux_p::Float64; flux_alpha::Float64; flux_beta::Float64
  lame_lambda::Float64; lame_mu::Float64; body_force::NTuple{3,Float64}   # op_kind 2: linear elasticity
end

# gdx_newton_opts / gdx_newton_report
struct GdxNewtonOpts
  rtol::Float64; atol::Float64; maxiter::Int32; lin_rtol::Float64; lin_maxiter::Int32; init::Int32
  precond::Ptr{Cvoid}   # gdx_nn* or C_NULL (Jacobi-PCG corrections)
end
struct GdxNewtonReport
  iters::Int32; lin_iters::Int32; converged::Int32; res_ref::Float64; res_final::Float64
  ms_total::Float64; ms_assembly::Float64; ms_solve::Float64; lin_fallbacks::Int32
end

check(rc, f) = rc == 0 || error("$f failed ($rc): ", unsafe_string(ccall((:gdx_last_error, libgdx), Cstring, ())))

mutable struct GdxContext
  h::Ptr{Cvoid}
  function GdxContext(device::Integer=0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:gdx_init, libgdx), Cint, (Int32, Ref{Ptr{Cvoid}}), device, r), :gdx_init)
    c = new(r[])
    finalizer(c -> ccall((:gdx_finalize, libgdx), Cint, (Ptr{Cvoid},), c.h), c)
  end
end

mutable struct GdxProblem
  h::Ptr{Cvoid}
  modes::Vector{Float64}   # kept alive while the descriptor is in use
end

function GdxProblem(ctx::GdxContext, desc::GdxProblemDesc, modes)
  r = Ref{Ptr{Cvoid}}(C_NULL)
  GC.@preserve modes check(ccall((:gdx_problem_create, libgdx), Cint,
    (Ptr{Cvoid}, Ref{GdxProblemDesc}, Ref{Ptr{Cvoid}}), ctx.h, desc, r), :gdx_problem_create)
  p = GdxProblem(r[], modes)
  finalizer(p -> ccall((:gdx_problem_destroy, libgdx), Cint, (Ptr{Cvoid},), p.h), p)
  check(ccall((:gdx_symbolic, libgdx), Cint, (Ptr{Cvoid},), p.h), :gdx_symbolic)
  p
end

numeric!(p::GdxProblem) = check(ccall((:gdx_numeric, libgdx), Cint, (Ptr{Cvoid}, Int32), p.h, 0), :gdx_numeric)

function part_sizes(p::GdxProblem, part::Integer)
  s = zeros(Int64, 5)
  check(ccall((:gdx_part_sizes, libgdx), Cint, (Ptr{Cvoid}, Int32, Ptr{Int64}), p.h, part, s), :gdx_part_sizes)
  s
end

# Local CSR of one part (0-based indices, as SparseMatrixCSR{0,Float64,Int}) and
# the own/ghost row and col indices (converted to the 1-based PArrays convention).
function export_part(p::GdxProblem, part::Integer)
  nl, no, ngr, ngc, nnz = part_sizes(p, part)
  nr, nc = no + ngr, no + ngc
  rg = Vector{Int64}(undef, nr); ro = Vector{Int32}(undef, nr)
  cg = Vector{Int64}(undef, nc); co = Vector{Int32}(undef, nc)
  rowptr = Vector{Int64}(undef, nr + 1); colidx = Vector{Int32}(undef, nnz); vals = Vector{Float64}(undef, nnz)
  b = Vector{Float64}(undef, nr)
  check(ccall((:gdx_export_rows, libgdx), Cint, (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int32}), p.h, part, rg, ro), :rows)
  check(ccall((:gdx_export_cols, libgdx), Cint, (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int32}), p.h, part, cg, co), :cols)
  check(ccall((:gdx_export_csr, libgdx), Cint, (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int32}, Ptr{Float64}),
              p.h, part, rowptr, colidx, vals), :csr)
  check(ccall((:gdx_export_rhs, libgdx), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}), p.h, part, b), :rhs)
  A = SparseMatrixCSR{0}(nr, nc, rowptr, Vector{Int}(colidx), vals)
  (A, rg .+ 1, ro .+ Int32(1), cg .+ 1, co .+ Int32(1), b, no)
end

"""
    assemble_matrix_and_vector_hip(ctx, parts_ids, desc, modes, ngdofs) -> (A::PSparseMatrix, b::PVector)

Builds the PArrays objects exactly as src/Algebra.jl:1013-1015 and :1240-1252 do:
rows/cols are OwnAndGhostIndices(OwnIndices(ngdofs,part,own_to_global),
GhostIndices(ngdofs,ghost_to_global,ghost_to_owner)).
"""
function assemble_matrix_and_vector_hip(p::GdxProblem, ranks, ngdofs)
  numeric!(p)
  data = map(ranks) do part
    A, rg, ro, cg, co, b, no = export_part(p, part - 1)
    rows = OwnAndGhostIndices(OwnIndices(ngdofs, part, rg[1:no]), GhostIndices(ngdofs, rg[no+1:end], ro[no+1:end]))
    cols = OwnAndGhostIndices(OwnIndices(ngdofs, part, cg[1:no]), GhostIndices(ngdofs, cg[no+1:end], co[no+1:end]))
    (A, rows, cols, b)
  end
  values, rows, cols, bs = tuple_of_arrays(data)
  A = PSparseMatrix(values, rows, cols)
  b = PVector(bs, rows)
  A, b
end

"""
    residual_and_jacobian_hip(p, ranks, ngdofs, x) -> (b::PVector, A::PSparseMatrix)

residual_and_jacobian(op, uh) of test/PLaplacianTests.jl:44 for a GDX_OP_FLUX
problem: x holds uh's free values in global-id order (get_free_dof_values
gathered by gid); Dirichlet values come from the descriptor.
"""
function residual_and_jacobian_hip(p::GdxProblem, ranks, ngdofs, x::Vector{Float64})
  GC.@preserve x check(ccall((:gdx_set_state, libgdx), Cint, (Ptr{Cvoid}, Ptr{Float64}), p.h, x), :gdx_set_state)
  A, b = assemble_matrix_and_vector_hip(p, ranks, ngdofs)
  b, A
end

"""
    newton_hip(p, x0; rtol=1e-10, maxiter=50, lin_rtol=1e-8, poisson_init=false, precond=C_NULL) -> (x, report)

solve(FESolver(NLSolver(method=:newton)), op) of test/PLaplacianTests.jl:54-57
with device corrections (Jacobi-PCG, or two-grid FPCG when precond is a gdx_nn
handle on the same context); x is the solution's free values by gid.
"""
function newton_hip(p::GdxProblem, x0::Vector{Float64}; rtol=1e-10, atol=0.0, maxiter=50, lin_rtol=1e-8,
                    lin_maxiter=20000, poisson_init=false, precond::Ptr{Cvoid}=C_NULL)
  x = copy(x0)
  o = GdxNewtonOpts(rtol, atol, maxiter, lin_rtol, lin_maxiter, poisson_init ? 1 : 0, precond)
  rep = Ref{GdxNewtonReport}()
  GC.@preserve x check(ccall((:gdx_newton, libgdx), Cint,
                             (Ptr{Cvoid}, Ref{GdxNewtonOpts}, Ptr{Float64}, Ref{GdxNewtonReport}), p.h, o, x, rep),
                       :gdx_newton)
  x, rep[]
end

"""
    GdxPVector(p, which)   # which = 0: A's row PRange, 1: its column PRange

A PVector in device memory: every part's local vector (own, then ghost
entries, in the order of the exported PRanges) concatenated in part order.
The PArrays operations below run on the device; like PArrays' own
`consistent!(v) |> wait`, they return something to `wait` on (work is queued
on the problem's stream; `Vector(v)` synchronizes).
"""
mutable struct GdxPVector
  h::Ptr{Cvoid}
  p::GdxProblem
  n::Int64
  function GdxPVector(p::GdxProblem, which::Integer)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:gdx_pvector_create, libgdx), Cint, (Ptr{Cvoid}, Int32, Ref{Ptr{Cvoid}}), p.h, which, h),
          :gdx_pvector_create)
    n = Ref{Int64}(0)
    check(ccall((:gdx_pvector_length, libgdx), Cint, (Ptr{Cvoid}, Ref{Int64}), h[], n), :gdx_pvector_length)
    v = new(h[], p, n[])
    finalizer(x -> ccall((:gdx_pvector_destroy, libgdx), Cint, (Ptr{Cvoid},), x.h), v)
  end
end

struct GdxQueued end
Base.wait(::GdxQueued) = nothing

function Base.copyto!(v::GdxPVector, x::Vector{Float64})
  length(x) == v.n || throw(DimensionMismatch("a GdxPVector on this PRange has $(v.n) local entries"))
  GC.@preserve x check(ccall((:gdx_pvector_set, libgdx), Cint, (Ptr{Cvoid}, Ptr{Float64}), v.h, x), :gdx_pvector_set)
  v
end
function Base.Vector(v::GdxPVector)
  x = Vector{Float64}(undef, v.n)
  GC.@preserve x check(ccall((:gdx_pvector_get, libgdx), Cint, (Ptr{Cvoid}, Ptr{Float64}), v.h, x), :gdx_pvector_get)
  x
end
function PartitionedArrays.consistent!(v::GdxPVector)
  check(ccall((:gdx_pvector_consistent, libgdx), Cint, (Ptr{Cvoid},), v.h), :gdx_pvector_consistent)
  GdxQueued()
end
function PartitionedArrays.assemble!(v::GdxPVector)
  check(ccall((:gdx_pvector_assemble, libgdx), Cint, (Ptr{Cvoid},), v.h), :gdx_pvector_assemble)
  GdxQueued()
end
# src/Algebra.jl:42-49
function LinearAlgebra.axpy!(α::Real, x::GdxPVector, y::GdxPVector)
  check(ccall((:gdx_pvector_axpy, libgdx), Cint, (Float64, Ptr{Cvoid}, Ptr{Cvoid}), α, x.h, y.h), :gdx_pvector_axpy)
  y
end
# src/Algebra.jl:270-290: own values of a onto b's PRange
function change_ghost!(b::GdxPVector, a::GdxPVector; make_consistent=false)
  check(ccall((:gdx_pvector_change_ghost, libgdx), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int32), a.h, b.h,
              make_consistent ? 1 : 0), :gdx_pvector_change_ghost)
  b
end
# mul!(y, A, x) of PSparseMatrix (test/PLaplacianTests.jl:50-52): y on A's rows, x on its cols
function LinearAlgebra.mul!(y::GdxPVector, A::GdxProblem, x::GdxPVector)
  check(ccall((:gdx_psparse_mul, libgdx), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), A.h, y.h, x.h), :gdx_psparse_mul)
  y
end

end # module
