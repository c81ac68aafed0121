"""gridap_amd: MI355X-native drop-in for GridapDistributed's assembly hot path.

Host-side mirror of the reference interface (CartesianDiscreteModel,
FESpace/TrialFESpace, SparseMatrixAssembler, assemble_matrix_and_vector,
AffineFEOperator) over the C ABI of libgdx.so (include/gdx.h).
"""
from . import workloads
from ._lib import GdxError
from .algebra import DevicePVector, FullyAssembledRows, LocalIndices, PRange, PSparseMatrix, PVector, SubAssembledRows
from .fespaces import (CLAMPED_X0_TAGS, AffineFEOperator, BodyForce, ConstantCoefficient, ConstantSource, Diffusion,
                       Elasticity, FESpace, Float64, VectorValue,
                       FourierExpCoefficient, GpuProblem, Integral, LinearFunction, ManufacturedSource, ReferenceFE,
                       SparseMatrixAssembler, Source, TestFESpace, TrialFESpace, allocate_matrix_and_vector,
                       assemble_matrix_and_vector, assemble_matrix_and_vector_, collect_cell_matrix_and_vector,
                       lagrangian, problem_desc)
from .geometry import CartesianDiscreteModel, DeviceContext, Measure, Triangulation, with_gpu
from .batch import AssemblyBatch
from .learned import LearnedCoarseSolver, fpcg
from .visualization import interpolate, writevtk
from .multifield import (Block, BlockMultiFieldStyle, BlockPMatrix, BlockPVector, Mass, MultiFieldFESpace,
                         assemble_block_matrix_and_vector)
from .nonlinear import (FEFunction, FEOperator, FESolver, FluxJacobian, FluxResidual, NLSolver, PowerLawFlux,
                        jacobian, residual, residual_and_jacobian, residual_and_jacobian_, solve, zero)


def problem_from_workload(ctx, p, strategy=None):
    """Build model/spaces/forms for a workload dict (gridap_amd.workloads) and
    return (assembler, data) ready for assemble_matrix_and_vector."""
    model = CartesianDiscreteModel(ctx, p["parts"], p["domain"], p["cells"], vertex_coords=p.get("vertex_coords"))
    if strategy is None:
        strategy = FullyAssembledRows() if p.get("strategy") == "full" else SubAssembledRows()
    if p.get("op") == "elasticity":
        reffe = ReferenceFE(lagrangian, VectorValue(3, Float64), p["order"])
        V = TestFESpace(model, reffe, dirichlet_tags="clamped_x0" if p.get("dirichlet") == "clamped_x0" else None)
        U = TrialFESpace(V, LinearFunction())
        dO = Measure(Triangulation(model), p["qdeg"])
        a = Integral(Elasticity(*p["lame"]), dO)
        l = Integral(BodyForce(p["body_force"]), dO)
        assem = SparseMatrixAssembler(U, V, strategy)
        return assem, collect_cell_matrix_and_vector(U, V, a, l, None)
    reffe = ReferenceFE(lagrangian, Float64, p["order"])
    V = TestFESpace(model, reffe, dirichlet_tags="boundary" if p.get("dirichlet", True) else None)
    uD = p.get("uD", (0.0, 0.0, 0.0, 0.0))
    U = TrialFESpace(V, LinearFunction(uD[0], uD[1:]))
    dO = Measure(Triangulation(model), p["qdeg"])
    modes = p.get("modes")
    if modes is not None and len(modes):
        kappa = FourierExpCoefficient(modes, p.get("coef_const", 1.0))
    else:
        kappa = ConstantCoefficient(p.get("coef_const", 1.0))
    f = ManufacturedSource() if p.get("rhs") == "manufactured" else ConstantSource(p.get("rhs_const", 0.0))
    a = Integral(Diffusion(kappa), dO)
    l = Integral(Source(f), dO)
    if strategy is None:
        strategy = FullyAssembledRows() if p.get("strategy") == "full" else SubAssembledRows()
    assem = SparseMatrixAssembler(U, V, strategy)
    return assem, collect_cell_matrix_and_vector(U, V, a, l, None)


def flux_operator_from_workload(ctx, p, strategy=None):
    """FEOperator(r, j, U, V, assem) of a flux workload (op == "flux")."""
    model = CartesianDiscreteModel(ctx, p["parts"], p["domain"], p["cells"])
    reffe = ReferenceFE(lagrangian, Float64, p["order"])
    V = TestFESpace(model, reffe, dirichlet_tags="boundary" if p.get("dirichlet", True) else None)
    uD = p.get("uD", (0.0, 0.0, 0.0, 0.0))
    U = TrialFESpace(V, LinearFunction(uD[0], uD[1:]))
    dO = Measure(Triangulation(model), p["qdeg"])
    fp, fa, fb = p["flux"]
    flux = PowerLawFlux(fp, fa, fb)
    r = Integral(FluxResidual(flux, ConstantSource(p.get("rhs_const", 0.0))), dO)
    j = Integral(FluxJacobian(flux), dO)
    if strategy is None:
        strategy = FullyAssembledRows() if p.get("strategy") == "full" else SubAssembledRows()
    return FEOperator(r, j, U, V, SparseMatrixAssembler(U, V, strategy))
