"""Nonlinear FE operators on the GPU path (mirror of the reference's
FEOperator / residual_and_jacobian / NLSolver usage, test/PLaplacianTests.jl:17-57).

Vocabulary lowered to GDX_OP_FLUX (Q1):

    sigma(g)    = (alpha + beta |g|^2)^((p-2)/2) g                 -> PowerLawFlux(p, alpha, beta)
    r(u,v)      = int( grad(v) . sigma(grad(u)) - v f ) dOmega     -> Integral(FluxResidual(flux, f), dO)
    j(u,du,v)   = int( grad(v) . dsigma(grad(du), grad(u)) ) dOmega -> Integral(FluxJacobian(flux), dO)

The reference test's sigma(g) = (1 + g.g) g is PowerLawFlux(4, 1, 1); the
p-Laplacian with regularisation eps is PowerLawFlux(p, eps**2, 1).
"""
import numpy as np

from . import _lib
from .fespaces import ConstantSource, Integral, SparseMatrixAssembler, TrialFESpace


class PowerLawFlux:
    """sigma(g) = (alpha + beta |g|^2)^((p-2)/2) g; its derivative is
    dsigma(d, g) = kappa d + (p-2) beta (alpha + beta |g|^2)^((p-4)/2) (g.d) g."""

    def __init__(self, p, alpha=0.0, beta=1.0):
        self.p, self.alpha, self.beta = float(p), float(alpha), float(beta)


class FluxResidual:
    """Integrand of r(u,v) = grad(v) . sigma(grad(u)) - v f (test/PLaplacianTests.jl:30)."""

    def __init__(self, flux, f=None):
        self.flux = flux
        self.f = f if f is not None else ConstantSource(0.0)


class FluxJacobian:
    """Integrand of j(u,du,v) = grad(v) . dsigma(grad(du), grad(u)) (test/PLaplacianTests.jl:31)."""

    def __init__(self, flux):
        self.flux = flux


class FEFunction:
    """A function of the trial space given by its free values (global-id order;
    None means zero) and the space's Dirichlet data."""

    def __init__(self, space, free_values=None):
        self.space = space
        self.free_values = None if free_values is None else np.asarray(free_values, dtype=np.float64)

    def get_free_dof_values(self):
        return self.free_values


def zero(U):
    """zero(U) (src/FESpaces.jl:38-42): free values zero, Dirichlet values from U."""
    return FEFunction(U, None)


class FEOperator:
    """FEOperator(r, j, U, V, assem) or FEOperator(r, U, V, assem) (the Jacobian
    then follows from the flux, as the reference's autodiff variant does)."""

    def __init__(self, res, *args):
        if args and isinstance(args[0], TrialFESpace):
            jac, rest = None, args
        else:
            jac, rest = args[0], args[1:]
        U, V = rest[0], rest[1]
        assem = rest[2] if len(rest) > 2 else None
        if not (isinstance(res, Integral) and isinstance(res.integrand, FluxResidual)):
            raise NotImplementedError("residual outside the GPU vocabulary")
        if jac is not None:
            if not (isinstance(jac, Integral) and isinstance(jac.integrand, FluxJacobian)):
                raise NotImplementedError("jacobian outside the GPU vocabulary")
            if jac.integrand.flux is not res.integrand.flux and vars(jac.integrand.flux) != vars(res.integrand.flux):
                raise NotImplementedError("the jacobian must be the derivative of the residual's flux")
            if jac.measure.degree != res.measure.degree:
                raise NotImplementedError("residual and jacobian must use the same Measure")
        self.res, self.jac, self.trial, self.test = res, jac, U, V
        self.assem = assem if assem is not None else SparseMatrixAssembler(U, V)

    def _problem(self):
        return self.assem._problem_for({"flux_op": self})


def flux_problem_desc(op, strategy):
    V, U = op.test, op.trial
    m = V.model
    if V.reffe.order != 1:
        raise NotImplementedError("the flux operator is lowered for Q1 only")
    f = op.res.integrand.f
    if not isinstance(f, ConstantSource):
        raise NotImplementedError("the flux operator takes a constant source")
    fl = op.res.integrand.flux
    d = _lib.ProblemDesc()
    d.dim, d.order, d.qdeg = m.dim, 1, op.res.measure.degree
    d.strategy = strategy.code
    for k in range(3):
        d.cells[k] = m.cells[k] if k < m.dim else 1
        d.parts[k] = m.parts[k] if k < m.dim else 1
    dom = list(m.domain) + [0.0, 1.0] * (3 - m.dim)
    for k in range(6):
        d.domain[k] = dom[k]
    d.coef_const, d.nmodes, d.modes = 1.0, 0, None
    d.rhs_kind, d.rhs_const = _lib.RHS_CONST, f.value
    d.dir_kind = 1 if V.dirichlet else 0
    for k in range(4):
        d.dir_coef[k] = U.uD.coef[k]
    d.op_kind, d.flux_p, d.flux_alpha, d.flux_beta = _lib.OP_FLUX, fl.p, fl.alpha, fl.beta
    return d, None


def _assemble_at(op, uh):
    prob = op._problem()
    prob.set_state(uh.free_values)
    prob.numeric()
    prob.sync()
    return prob


def residual_and_jacobian(op, uh):
    """(b, A) = residual and Jacobian at uh, as PVector / PSparseMatrix."""
    _assemble_at(op, uh)
    A, b = op.assem.export()
    return b, A


def residual_and_jacobian_(b, A, op, uh):
    """In-place variant residual_and_jacobian!(b, A, op, uh)."""
    prob = _assemble_at(op, uh)
    for r in range(prob.nparts):
        P = prob.export_part(r)
        A.matrix_partition[r].data[:] = P["val"]
        b.partition[r][:] = P["b"]
    return b, A


def residual(op, uh):
    return residual_and_jacobian(op, uh)[0]


def jacobian(op, uh):
    return residual_and_jacobian(op, uh)[1]


class NLSolver:
    """NLSolver(method=:newton) with device corrections: Jacobi-PCG, or flexible
    PCG with a LearnedCoarseSolver's multilevel preconditioner (preconditioner=).
    ftol: relative residual tolerance ||r|| <= ftol ||r(u0)||."""

    def __init__(self, method="newton", ftol=1e-10, atol=0.0, iterations=50, linear_rtol=1e-8,
                 linear_iterations=20000, initial_guess="given", show_trace=False, preconditioner=None):
        if method != "newton":
            raise NotImplementedError("only method='newton' runs on the GPU path")
        if initial_guess not in ("given", "poisson"):
            raise ValueError("initial_guess must be 'given' or 'poisson'")
        self.ftol, self.atol, self.iterations = ftol, atol, iterations
        self.linear_rtol, self.linear_iterations = linear_rtol, linear_iterations
        self.initial_guess, self.show_trace = initial_guess, show_trace
        self.preconditioner = preconditioner
        self.report = None


class FESolver:
    def __init__(self, nls):
        self.nls = nls


def solve(solver, op, uh0=None):
    """solve(FESolver(NLSolver(...)), op) -> uh; the Newton report is kept on
    solver.nls.report."""
    nls = solver.nls if isinstance(solver, FESolver) else solver
    prob = op._problem()
    x0 = None if uh0 is None else uh0.free_values
    init = _lib.NEWTON_INIT_POISSON if nls.initial_guess == "poisson" else _lib.NEWTON_INIT_GIVEN
    x, rep = prob.newton(x0, rtol=nls.ftol, atol=nls.atol, maxiter=nls.iterations, lin_rtol=nls.linear_rtol,
                         lin_maxiter=nls.linear_iterations, init=init, precond=nls.preconditioner)
    nls.report = rep
    if nls.show_trace:
        print(f"Newton: {rep['iters']} iterations, ||r|| {rep['res_ref']:.3e} -> {rep['res_final']:.3e}, "
              f"{rep['lin_iters']} PCG iterations ({rep['lin_fallbacks']} fallbacks), {rep['ms_total']:.1f} ms")
    return FEFunction(op.trial, x)
