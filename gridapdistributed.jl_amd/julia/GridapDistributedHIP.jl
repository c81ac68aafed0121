"""
    GridapDistributedHIP

Julia binding of libgdx.so (C ABI in include/gdx.h).

- `HIPSparseMatrixAssembler <: SparseMatrixAssembler`: a drop-in for the
  reference's `DistributedSparseMatrixAssembler` (src/FESpaces.jl:746-862).
  `get_rows`/`get_cols` are the FE spaces' dof PRanges and
  `get_assembly_strategy` the strategy, as there.  `assemble_matrix_and_vector`,
  `assemble_matrix`, `assemble_vector`, their in-place `!` forms and the
  `allocate_*` functions lower the weak form to a `GdxProblemDesc` and assemble
  on one MI355X, returning PartitionedArrays `PSparseMatrix`/`PVector` with the
  reference's row/col PRanges (src/Algebra.jl:1013-1015, 1240-1257) and local
  matrices of the requested type (`SparseMatrixCSC{Float64,Int}` by default,
  `SparseMatrixCSR{0,Float64,Int}`).
- The lowered vocabulary: forms built with `HIPForm` from `Diffusion(κ)`,
  `Source(f)`, `Elasticity(λ, μ)`, `BodyForce(f)` on a `CartesianDiscreteModel`
  with Lagrangian Q1/Q2; diffusion also on a model built with a `map` (the
  mapped vertices go to `gdx_set_vertex_coords`).  A `HIPForm` is also an ordinary Gridap form (it
  carries the Julia closure), so whatever does not lower (`lower` returns
  `nothing`) is assembled by the reference's `DistributedSparseMatrixAssembler`
  unchanged.
- Lower-level wrappers: `GdxProblem`, `residual_and_jacobian_hip`,
  `newton_hip`, device `GdxPVector`s with `consistent!`, `assemble!`, `axpy!`,
  `change_ghost!` and `mul!`.

Not executed in this repository (no Julia toolchain is available to the
build); the struct layouts and ccall signatures are checked against
include/gdx.h by tests/test_julia_shim.py, and the same C ABI is exercised by
the ctypes binding gridapdistributed.jl_amd/gridap_amd/_lib.py.
"""
module GridapDistributedHIP

using Gridap
using Gridap.FESpaces
using Gridap.Geometry
using GridapDistributed
using PartitionedArrays
using SparseArrays
using SparseMatricesCSR
using LinearAlgebra

export HIPSparseMatrixAssembler, HIPForm, Diffusion, Source, Elasticity, BodyForce
export ConstantCoefficient, FourierExpCoefficient, ConstantSource, ManufacturedSource, AffineFunction

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
  jac_precision::Int32  # 0: FP64 Jacobians, 1: FP32 cell contractions (residual FP64)
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

# Local matrix of one part from its 0-based CSR arrays: SparseMatrixCSR{0} as is,
# SparseMatrixCSC (the reference's default local type, src/FESpaces.jl:860) by
# transposing the CSC of the transpose (columns stay sorted).
local_matrix(::Type{<:SparseMatrixCSR}, nr, nc, rowptr, colidx, vals) =
  SparseMatrixCSR{0}(nr, nc, rowptr, Vector{Int}(colidx), vals)
function local_matrix(::Type{<:SparseMatrixCSC}, nr, nc, rowptr, colidx, vals)
  At = SparseMatrixCSC(nc, nr, rowptr .+ 1, Vector{Int}(colidx) .+ 1, vals)
  SparseMatrixCSC(transpose(At))
end
# position in the CSR values of each stored entry of the local matrix, in its
# own storage order (identity for CSR{0}; the transpose permutation for CSC)
value_order(::Type{<:SparseMatrixCSR}, nr, nc, rowptr, colidx) = collect(1:length(colidx))
function value_order(::Type{<:SparseMatrixCSC}, nr, nc, rowptr, colidx)
  A = local_matrix(SparseMatrixCSC, nr, nc, rowptr, colidx, collect(1.0:length(colidx)))
  Int.(nonzeros(A))
end

# Exports of one part: CSR arrays (0-based), own/ghost row and col indices
# converted to the 1-based PArrays convention, RHS, own count.
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
  (nr, nc, rowptr, colidx, vals, rg .+ 1, ro .+ Int32(1), cg .+ 1, co .+ Int32(1), b, no)
end

"""
    assemble_matrix_and_vector_hip(p, ranks, ngdofs, Tm=SparseMatrixCSC{Float64,Int}) -> (A::PSparseMatrix, b::PVector)

Runs the numeric phase and builds the PArrays objects exactly as
src/Algebra.jl:1013-1015 and :1240-1252 do: rows/cols are
OwnAndGhostIndices(OwnIndices(ngdofs,part,own_to_global),
GhostIndices(ngdofs,ghost_to_global,ghost_to_owner)); local matrices of type Tm.
"""
function assemble_matrix_and_vector_hip(p::GdxProblem, ranks, ngdofs, Tm::Type=SparseMatrixCSC{Float64,Int};
                                        numeric=true)
  numeric && numeric!(p)
  data = map(ranks) do part
    nr, nc, rowptr, colidx, vals, rg, ro, cg, co, b, no = export_part(p, part - 1)
    numeric || (fill!(vals, 0.0); fill!(b, 0.0))
    A = local_matrix(Tm, nr, nc, rowptr, colidx, vals)
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
  A, b = assemble_matrix_and_vector_hip(p, ranks, ngdofs, SparseMatrixCSR{0,Float64,Int})
  b, A
end

"""
    set_cell_lame!(p, lame)

Piecewise-constant material of a GDX_OP_ELASTICITY problem: lame is a
2 x ncells matrix, (lambda_c, mu_c) in column c (cells x fastest), for the
constitutive law lambda_c tr(eps) I + 2 mu_c eps; `nothing` returns to the
descriptor's constants.  Takes effect at the next assembly.
"""
function set_cell_lame!(p::GdxProblem, lame::Union{Nothing,Matrix{Float64}})
  if lame === nothing
    check(ccall((:gdx_set_cell_lame, libgdx), Cint, (Ptr{Cvoid}, Ptr{Float64}), p.h, C_NULL), :gdx_set_cell_lame)
  else
    size(lame, 1) == 2 || error("lame must be 2 x ncells")
    GC.@preserve lame check(ccall((:gdx_set_cell_lame, libgdx), Cint, (Ptr{Cvoid}, Ptr{Float64}), p.h, lame),
                            :gdx_set_cell_lame)
  end
  nothing
end

"""
    set_reaction!(p, diffusion_weight, reaction)

Diffusion-reaction form on the mapped-grid path (after `set_vertex_coords!`):
diffusion_weight kappa grad(u).grad(v) + reaction u v, the terms of a
multi-field block such as u2 v2 or u1 v2 (test/BlockSparseMatrixAssemblersTests.jl:110).
"""
function set_reaction!(p::GdxProblem, diffusion_weight::Real, reaction::Real)
  check(ccall((:gdx_set_reaction, libgdx), Cint, (Ptr{Cvoid}, Float64, Float64), p.h, diffusion_weight, reaction),
        :gdx_set_reaction)
  nothing
end

"""
    newton_hip(p, x0; rtol=1e-10, maxiter=50, lin_rtol=1e-8, poisson_init=false, precond=C_NULL, jac32=false) -> (x, report)

solve(FESolver(NLSolver(method=:newton)), op) of test/PLaplacianTests.jl:54-57
with device corrections (Jacobi-PCG, or two-grid FPCG when precond is a gdx_nn
handle on the same context); x is the solution's free values by gid.
"""
function newton_hip(p::GdxProblem, x0::Vector{Float64}; rtol=1e-10, atol=0.0, maxiter=50, lin_rtol=1e-8,
                    lin_maxiter=20000, poisson_init=false, precond::Ptr{Cvoid}=C_NULL, jac32::Bool=false)
  x = copy(x0)
  o = GdxNewtonOpts(rtol, atol, maxiter, lin_rtol, lin_maxiter, poisson_init ? 1 : 0, precond, jac32 ? 1 : 0)
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

# ---------------------------------------------------------------------------
# The weak-form vocabulary the GPU path lowers (the Python mirror is
# gridapdistributed.jl_amd/gridap_amd/fespaces.py)
# ---------------------------------------------------------------------------
abstract type HIPCoefficient end
"κ = value"
struct ConstantCoefficient <: HIPCoefficient
  value::Float64
end
"κ(x) = scale * exp(Σ_m a_m cos(2π k_m⋅x + φ_m)); modes rows (k1, k2, k3, a, φ)"
struct FourierExpCoefficient <: HIPCoefficient
  modes::Matrix{Float64}
  scale::Float64
end
FourierExpCoefficient(modes::Matrix{Float64}) = FourierExpCoefficient(modes, 1.0)

abstract type HIPSource end
"f = value"
struct ConstantSource <: HIPSource
  value::Float64
end
"f = -∇⋅(κ∇u_D) for the affine Dirichlet data u_D"
struct ManufacturedSource <: HIPSource end

"u_D(x) = c0 + c⋅x"
struct AffineFunction
  c0::Float64
  c::NTuple{3,Float64}
end
(u::AffineFunction)(x) = u.c0 + sum(u.c[i] * x[i] for i in 1:length(x))

"∫( κ*∇(v)⋅∇(u) )dΩ"
struct Diffusion
  kappa::HIPCoefficient
end
"∫( f*v )dΩ"
struct Source
  f::HIPSource
end
"∫( ε(v)⊙(λ tr(ε(u)) I + 2μ ε(u)) )dΩ"
struct Elasticity
  lambda::Float64
  mu::Float64
end
"∫( v⋅f )dΩ"
struct BodyForce
  f::NTuple{3,Float64}
end

"""
    HIPForm(integrand, qdeg, form)

A bilinear (`Diffusion`, `Elasticity`) or linear (`Source`, `BodyForce`) form:
`integrand` and the Measure's degree `qdeg` are what the GPU path lowers,
`form` is the same form as an ordinary Julia closure, e.g.
`HIPForm(Diffusion(κ), 4, (u,v) -> ∫(κf*∇(v)⋅∇(u))dΩ)`.  Calling a HIPForm
evaluates the closure and keeps the lowered description beside it, so a
HIPForm works wherever Gridap expects the form.
"""
struct HIPForm{I,F} <: Function
  integrand::I
  qdeg::Int
  form::F
end
struct HIPContribution{I,C}
  integrand::I
  qdeg::Int
  contribution::C   # the DomainContribution of the Julia closure
end
(h::HIPForm)(args...) = HIPContribution(h.integrand, h.qdeg, h.form(args...))

# collect_cell_*: the lowered descriptions travel beside Gridap's own data
struct HIPData{M,V,D}
  mat::M
  vec::V
  fallback::D       # what collect_cell_* returns for the Julia closures
end
function FESpaces.collect_cell_matrix_and_vector(U::FESpace, V::FESpace, m::HIPContribution, l::HIPContribution, uhd)
  HIPData(m, l, collect_cell_matrix_and_vector(U, V, m.contribution, l.contribution, uhd))
end
FESpaces.collect_cell_matrix(U::FESpace, V::FESpace, m::HIPContribution) =
  HIPData(m, nothing, collect_cell_matrix(U, V, m.contribution))
FESpaces.collect_cell_vector(V::FESpace, l::HIPContribution) =
  HIPData(nothing, l, collect_cell_vector(V, l.contribution))

# ---------------------------------------------------------------------------
# HIPSparseMatrixAssembler
# ---------------------------------------------------------------------------
"""
    HIPSparseMatrixAssembler([Tm, Tv,] U, V, strategy=SubAssembledRows();
                             order, dirichlet=:boundary, uD=AffineFunction(0.0, (0.0, 0.0, 0.0)),
                             ctx=GdxContext(0))

The drop-in for `SparseMatrixAssembler(Tm, Tv, U, V, strategy)`
(src/FESpaces.jl:841-862).  It wraps the reference's assembler (`fallback`) for
everything that does not lower.  The FE spaces' construction parameters that
Julia cannot read back from U and V are given here: the Lagrangian `order`,
`dirichlet` (`:boundary` for dirichlet_tags="boundary", `:clamped_x0` for the
elasticity clamp, `:none`) and the affine Dirichlet data `uD`; `lower` checks
them against U and V (free dof counts) and falls back when they disagree.
"""
struct HIPSparseMatrixAssembler{A,M} <: SparseMatrixAssembler
  fallback::A       # DistributedSparseMatrixAssembler (src/FESpaces.jl:746-753)
  trial::FESpace
  test::FESpace
  order::Int
  dirichlet::Symbol
  uD::AffineFunction
  ctx::GdxContext
  problems::Dict{Any,GdxProblem}   # one symbolic phase per lowered descriptor
end
local_matrix_type(::HIPSparseMatrixAssembler{A,M}) where {A,M} = M

function HIPSparseMatrixAssembler(Tm::Type, Tv::Type, U::FESpace, V::FESpace, strategy=SubAssembledRows();
                                  order::Integer, dirichlet::Symbol=:boundary,
                                  uD::AffineFunction=AffineFunction(0.0, (0.0, 0.0, 0.0)), ctx=GdxContext(0))
  fallback = SparseMatrixAssembler(Tm, Tv, U, V, strategy)
  HIPSparseMatrixAssembler{typeof(fallback),Tm}(fallback, U, V, Int(order), dirichlet, uD, ctx, Dict{Any,GdxProblem}())
end
HIPSparseMatrixAssembler(U::FESpace, V::FESpace, strategy=SubAssembledRows(); kwargs...) =
  HIPSparseMatrixAssembler(SparseMatrixCSC{Float64,Int}, Vector{Float64}, U, V, strategy; kwargs...)

FESpaces.get_rows(a::HIPSparseMatrixAssembler) = get_rows(a.fallback)
FESpaces.get_cols(a::HIPSparseMatrixAssembler) = get_cols(a.fallback)
FESpaces.get_matrix_builder(a::HIPSparseMatrixAssembler) = get_matrix_builder(a.fallback)
FESpaces.get_vector_builder(a::HIPSparseMatrixAssembler) = get_vector_builder(a.fallback)
FESpaces.get_assembly_strategy(a::HIPSparseMatrixAssembler) = get_assembly_strategy(a.fallback)

strategy_code(::SubAssembledRows) = Int32(0)
strategy_code(::FullyAssembledRows) = Int32(1)
strategy_code(::Any) = nothing

# global Cartesian description of a DistributedDiscreteModel made by
# CartesianDiscreteModel(ranks, parts, domain, cells) (src/Geometry.jl:378-411),
# or nothing
function cartesian_info(model)
  pdesc = model isa GridapDistributed.DistributedDiscreteModel ? model.metadata : nothing
  pdesc isa GridapDistributed.DistributedCartesianDescriptor || return nothing
  desc = pdesc.descriptor
  (!any(desc.isperiodic) && all(pdesc.ghost)) || return nothing
  D = length(desc.partition)
  D in (2, 3) || return nothing
  origin = Tuple(desc.origin)
  cells = Tuple(desc.partition)
  domain = ntuple(k -> isodd(k) ? origin[(k + 1) ÷ 2] : origin[k ÷ 2] + desc.sizes[k ÷ 2] * cells[k ÷ 2], 2D)
  (D, cells, Tuple(pdesc.mesh_partition), domain, desc.map)
end

# the descriptor's map applied to every vertex of the uniform box: a D x nv
# matrix, vertices x fastest (the layout of gdx_set_vertex_coords)
function mapped_vertices(D, cells, domain, gmap)
  nv = prod(c + 1 for c in cells)
  xyz = Matrix{Float64}(undef, D, nv)
  for (v, I) in enumerate(CartesianIndices(Tuple(c + 1 for c in cells)))
    x = ntuple(d -> domain[2d - 1] + (domain[2d] - domain[2d - 1]) * (I[d] - 1) / cells[d], D)
    y = gmap(Point(x))
    for d in 1:D
      xyz[d, v] = y[d]
    end
  end
  xyz
end

"""
    set_vertex_coords!(p, xyz)

Mapped grid of a diffusion problem: xyz is a D x nv matrix of vertex
coordinates (vertices x fastest), `nothing` returns to the uniform box.
"""
function set_vertex_coords!(p::GdxProblem, xyz::Union{Nothing,Matrix{Float64}})
  ptr = xyz === nothing ? C_NULL : pointer(xyz)
  GC.@preserve xyz check(ccall((:gdx_set_vertex_coords, libgdx), Cint, (Ptr{Cvoid}, Ptr{Float64}), p.h, ptr),
                         :gdx_set_vertex_coords)
  nothing
end

pad3(t, v) = ntuple(k -> k <= length(t) ? t[k] : v, 3)
pad6(t) = ntuple(k -> k <= length(t) ? Float64(t[k]) : (isodd(k) ? 0.0 : 1.0), 6)

# free dofs the lowered problem has: Q_K scalar (or 3 components) on the Cartesian mesh
function expected_free_dofs(D, cells, K, dirichlet, ncomp)
  if dirichlet == :boundary
    prod(K * c - 1 for c in cells) * ncomp
  elseif dirichlet == :clamped_x0
    K * cells[1] * prod(K * c + 1 for c in cells[2:end]) * ncomp
  else
    prod(K * c + 1 for c in cells) * ncomp
  end
end

"""
    lower(a::HIPSparseMatrixAssembler, data::HIPData) -> Union{Nothing, Tuple{GdxProblemDesc, Vector{Float64}}}

The weak form as a problem descriptor (plus the Fourier modes it points to),
or `nothing` when anything lies outside the vocabulary; the caller then
assembles with the reference's assembler.
"""
function lower(a::HIPSparseMatrixAssembler, data::HIPData)
  m, l = data.mat, data.vec
  (m === nothing || l === nothing) && return nothing   # matrix- or vector-only data: both forms are needed
  m.qdeg == l.qdeg || return nothing
  info = cartesian_info(get_background_model(get_triangulation(a.test)))
  info === nothing && return nothing
  D, cells, parts, domain, gmap = info
  code = strategy_code(get_assembly_strategy(a))
  code === nothing && return nothing
  a.order in (1, 2) || return nothing
  m.qdeg ÷ 2 + 1 <= 4 || return nothing
  modes = Float64[]
  if m.integrand isa Diffusion && l.integrand isa Source
    ncomp = 1
    a.dirichlet in (:boundary, :none) || return nothing
    κ = m.integrand.kappa
    coef, nmodes = κ isa ConstantCoefficient ? (κ.value, 0) : (κ.scale, size(κ.modes, 1))
    κ isa FourierExpCoefficient && (size(κ.modes, 2) == 5 || return nothing; modes = vec(permutedims(κ.modes)))
    f = l.integrand.f
    rhs_kind, rhs_const = f isa ConstantSource ? (Int32(0), f.value) : (Int32(1), 0.0)
    op, lam, mu, bf = Int32(0), 0.0, 0.0, (0.0, 0.0, 0.0)
    dir_kind = a.dirichlet == :boundary ? Int32(1) : Int32(0)
  elseif m.integrand isa Elasticity && l.integrand isa BodyForce
    ncomp = 3
    D == 3 || return nothing
    gmap === identity || return nothing  # mapped grids: diffusion only
    a.dirichlet in (:clamped_x0, :none) || return nothing
    (a.uD.c0 == 0.0 && all(iszero, a.uD.c)) || return nothing
    coef, nmodes, rhs_kind, rhs_const = 1.0, 0, Int32(0), 0.0
    op, lam, mu, bf = Int32(2), m.integrand.lambda, m.integrand.mu, l.integrand.f
    dir_kind = a.dirichlet == :clamped_x0 ? Int32(2) : Int32(0)
  else
    return nothing
  end
  # the spaces must be the ones described: same free dof count as the closed form
  num_free_dofs(a.test) == expected_free_dofs(D, cells, a.order, a.dirichlet, ncomp) || return nothing
  num_free_dofs(a.trial) == num_free_dofs(a.test) || return nothing
  desc = GdxProblemDesc(Int32(D), Int32(a.order), Int32(m.qdeg), code, pad3(Int64.(cells), Int64(1)),
                        pad3(Int32.(parts), Int32(1)), pad6(domain), coef, Int32(nmodes), pointer(modes),
                        rhs_kind, rhs_const, dir_kind, (a.uD.c0, a.uD.c...), op, 0.0, 0.0, 0.0, lam, mu, bf)
  xyz = gmap === identity ? nothing : mapped_vertices(D, cells, domain, gmap)
  desc, modes, xyz
end

# the problem of a lowered descriptor: created and run through gdx_symbolic once
function hip_problem(a::HIPSparseMatrixAssembler, lowered)
  desc, modes, xyz = lowered
  key = (desc.dim, desc.order, desc.qdeg, desc.strategy, desc.cells, desc.parts, desc.domain, desc.coef_const,
         copy(modes), desc.rhs_kind, desc.rhs_const, desc.dir_kind, desc.dir_coef, desc.op_kind,
         desc.lame_lambda, desc.lame_mu, desc.body_force, xyz === nothing ? nothing : copy(xyz))
  get!(a.problems, key) do
    p = GdxProblem(a.ctx, desc, modes)
    xyz === nothing || set_vertex_coords!(p, xyz)
    p
  end
end

ranks_of(a::HIPSparseMatrixAssembler) = linear_indices(partition(get_rows(a)))

function FESpaces.assemble_matrix_and_vector(a::HIPSparseMatrixAssembler, data::HIPData)
  lowered = lower(a, data)
  lowered === nothing && return assemble_matrix_and_vector(a.fallback, data.fallback)
  p = hip_problem(a, lowered)
  assemble_matrix_and_vector_hip(p, ranks_of(a), num_free_dofs(a.test), local_matrix_type(a))
end
FESpaces.assemble_matrix_and_vector(a::HIPSparseMatrixAssembler, data) = assemble_matrix_and_vector(a.fallback, data)

function FESpaces.allocate_matrix_and_vector(a::HIPSparseMatrixAssembler, data::HIPData)
  lowered = lower(a, data)
  lowered === nothing && return allocate_matrix_and_vector(a.fallback, data.fallback)
  p = hip_problem(a, lowered)
  assemble_matrix_and_vector_hip(p, ranks_of(a), num_free_dofs(a.test), local_matrix_type(a); numeric=false)
end
FESpaces.allocate_matrix_and_vector(a::HIPSparseMatrixAssembler, data) = allocate_matrix_and_vector(a.fallback, data)

# in place (src/Algebra.jl:323-393,530-535): the numeric phase refills the same
# structure; the local matrices and b of A/b must come from this assembler
function FESpaces.assemble_matrix_and_vector!(A, b, a::HIPSparseMatrixAssembler, data::HIPData)
  lowered = lower(a, data)
  lowered === nothing && return assemble_matrix_and_vector!(A, b, a.fallback, data.fallback)
  p = hip_problem(a, lowered)
  numeric!(p)
  Tm = local_matrix_type(a)
  map(ranks_of(a), partition(A), partition(b)) do part, Al, bl
    nl, no, ngr, ngc, nnz = part_sizes(p, part - 1)
    (length(nonzeros(Al)) == nnz && length(bl) == no + ngr) || error("A and b were not allocated by this assembler")
    rowptr = Vector{Int64}(undef, no + ngr + 1); colidx = Vector{Int32}(undef, nnz); vals = Vector{Float64}(undef, nnz)
    check(ccall((:gdx_export_csr, libgdx), Cint, (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int32}, Ptr{Float64}),
                p.h, part - 1, rowptr, colidx, vals), :csr)
    check(ccall((:gdx_export_rhs, libgdx), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}), p.h, part - 1, bl), :rhs)
    nonzeros(Al) .= vals[value_order(Tm, no + ngr, no + ngc, rowptr, colidx)]
  end
  A, b
end
FESpaces.assemble_matrix_and_vector!(A, b, a::HIPSparseMatrixAssembler, data) =
  assemble_matrix_and_vector!(A, b, a.fallback, data)

# matrix-only and vector-only data: both forms are needed to lower (the RHS
# carries the Dirichlet lifting), so these go through the reference assembler
# unless the data holds both forms
FESpaces.assemble_matrix(a::HIPSparseMatrixAssembler, data::HIPData) =
  data.vec === nothing ? assemble_matrix(a.fallback, data.fallback) : first(assemble_matrix_and_vector(a, data))
FESpaces.assemble_matrix(a::HIPSparseMatrixAssembler, data) = assemble_matrix(a.fallback, data)
FESpaces.assemble_vector(a::HIPSparseMatrixAssembler, data::HIPData) =
  data.mat === nothing ? assemble_vector(a.fallback, data.fallback) : last(assemble_matrix_and_vector(a, data))
FESpaces.assemble_vector(a::HIPSparseMatrixAssembler, data) = assemble_vector(a.fallback, data)
FESpaces.allocate_matrix(a::HIPSparseMatrixAssembler, data::HIPData) = allocate_matrix(a.fallback, data.fallback)
FESpaces.allocate_matrix(a::HIPSparseMatrixAssembler, data) = allocate_matrix(a.fallback, data)
FESpaces.allocate_vector(a::HIPSparseMatrixAssembler, data::HIPData) = allocate_vector(a.fallback, data.fallback)
FESpaces.allocate_vector(a::HIPSparseMatrixAssembler, data) = allocate_vector(a.fallback, data)
FESpaces.assemble_matrix!(A, a::HIPSparseMatrixAssembler, data::HIPData) = assemble_matrix!(A, a.fallback, data.fallback)
FESpaces.assemble_matrix!(A, a::HIPSparseMatrixAssembler, data) = assemble_matrix!(A, a.fallback, data)
FESpaces.assemble_vector!(b, a::HIPSparseMatrixAssembler, data::HIPData) = assemble_vector!(b, a.fallback, data.fallback)
FESpaces.assemble_vector!(b, a::HIPSparseMatrixAssembler, data) = assemble_vector!(b, a.fallback, data)

# HIPData handed to any other assembler (e.g. the default one that
# AffineFEOperator(a, l, U, V) builds) assembles Gridap's own data
FESpaces.assemble_matrix_and_vector(a::SparseMatrixAssembler, data::HIPData) =
  assemble_matrix_and_vector(a, data.fallback)
FESpaces.allocate_matrix_and_vector(a::SparseMatrixAssembler, data::HIPData) =
  allocate_matrix_and_vector(a, data.fallback)
FESpaces.assemble_matrix_and_vector!(A, b, a::SparseMatrixAssembler, data::HIPData) =
  assemble_matrix_and_vector!(A, b, a, data.fallback)
FESpaces.assemble_matrix(a::SparseMatrixAssembler, data::HIPData) = assemble_matrix(a, data.fallback)
FESpaces.assemble_vector(a::SparseMatrixAssembler, data::HIPData) = assemble_vector(a, data.fallback)

end # module
