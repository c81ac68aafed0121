"""FE spaces, weak forms and the assembler (mirror of src/FESpaces.jl:570-862).

The GPU path lowers a restricted weak-form vocabulary:

    a(u,v) = int( kappa * grad(v) . grad(u) ) dOmega      -> Diffusion(kappa)
    l(v)   = int( f * v ) dOmega                          -> Source(f)

and, on ReferenceFE(lagrangian, VectorValue(3, Float64), order), linear elasticity

    a(u,v) = int( eps(v) : (lambda tr(eps(u)) I + 2 mu eps(u)) ) dOmega -> Elasticity(lambda, mu)
    l(v)   = int( v . f ) dOmega                                      -> BodyForce(f)

with kappa a ``ConstantCoefficient`` or ``FourierExpCoefficient`` and f a
``ConstantSource`` or ``ManufacturedSource`` (f = -div(kappa grad uD)).
``SparseMatrixAssembler(U, V, strategy)`` owns one device problem (gdx_problem):
the symbolic phase (numbering, gids, PRanges, CSR pattern, position map) runs
once, every numeric assembly refills the values in HBM.
"""
import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib
from .algebra import DevicePVector, FullyAssembledRows, LocalIndices, PSparseMatrix, PVector, SubAssembledRows


class _Lagrangian:
    pass


lagrangian = _Lagrangian()
Float64 = np.float64


class VectorValue:
    """The value type VectorValue{n,Float64} of a vector-valued reference FE."""

    def __init__(self, n, T=np.float64):
        if T is not np.float64:
            raise NotImplementedError("only Float64 components are lowered to the GPU")
        self.n = int(n)


class ReferenceFE:
    """ReferenceFE(lagrangian, Float64, order) or ReferenceFE(lagrangian, VectorValue(3, Float64), order)."""

    def __init__(self, kind, T, order):
        if not isinstance(kind, _Lagrangian):
            raise NotImplementedError("only Lagrangian reference FEs are lowered to the GPU")
        if isinstance(T, VectorValue):
            if T.n != 3:
                raise NotImplementedError("vector-valued reference FEs are lowered for 3 components")
            self.ncomp = 3
        elif T is np.float64:
            self.ncomp = 1
        else:
            raise NotImplementedError("value type outside the GPU vocabulary")
        if order not in (1, 2):
            raise NotImplementedError("only order 1 and 2 are supported")
        self.order = int(order)


class LinearFunction:
    """u(x) = c0 + c . x (Dirichlet data of the supported vocabulary)."""

    def __init__(self, c0=0.0, c=(0.0, 0.0, 0.0)):
        c = tuple(float(v) for v in c) + (0.0,) * (3 - len(c))
        self.coef = (float(c0),) + c[:3]

    def __call__(self, x):
        x = np.atleast_2d(x)
        return self.coef[0] + sum(self.coef[1 + d] * x[:, d] for d in range(x.shape[1]))


# the entities of CartesianDiscreteModel's face labeling on the face x = x0 of
# a 3D box: face 25, its edges 13, 15, 17, 19 and vertices 1, 3, 5, 7
CLAMPED_X0_TAGS = frozenset(f"tag_{k:02d}" for k in (1, 3, 5, 7, 13, 15, 17, 19, 25))


class FESpace:
    """TestFESpace(model, reffe; dirichlet_tags="boundary") (src/FESpaces.jl:570-614).
    Vector-valued spaces also take the face x = x0 with its closure:
    dirichlet_tags="clamped_x0" or the tag list CLAMPED_X0_TAGS."""

    def __init__(self, model, reffe, dirichlet_tags=None):
        tags = dirichlet_tags
        if isinstance(tags, (list, tuple, set, frozenset)) and frozenset(tags) == CLAMPED_X0_TAGS:
            tags = "clamped_x0"
        if tags not in (None, "boundary", "clamped_x0"):
            raise NotImplementedError("dirichlet_tags outside the GPU vocabulary ('boundary', 'clamped_x0' or none)")
        if tags == "clamped_x0" and (reffe.ncomp == 1 or model.dim != 3):
            raise NotImplementedError("the clamped face is lowered for 3D vector-valued spaces")
        self.model, self.reffe = model, reffe
        self.dirichlet = tags == "boundary"
        self.clamped = tags == "clamped_x0"


TestFESpace = FESpace


class TrialFESpace:
    """TrialFESpace(V, uD) / TrialFESpace(uD, V) (src/FESpaces.jl:396-401)."""

    def __init__(self, V, uD=None):
        if not isinstance(V, FESpace):
            V, uD = uD, V
        self.space = V
        self.uD = uD if uD is not None else LinearFunction()
        if not isinstance(self.uD, LinearFunction):
            raise NotImplementedError("Dirichlet data must be a LinearFunction on the GPU path")

    @property
    def model(self):
        return self.space.model


class ConstantCoefficient:
    def __init__(self, value=1.0):
        self.value = float(value)


class FourierExpCoefficient:
    """kappa(x) = scale * exp(sum_m a_m cos(2 pi k_m . x + phi_m)); modes rows (k1,k2,k3,a,phi)."""

    def __init__(self, modes, scale=1.0):
        self.modes = np.ascontiguousarray(np.asarray(modes, np.float64).reshape(-1, 5))
        self.scale = float(scale)


class ConstantSource:
    def __init__(self, value=0.0):
        self.value = float(value)


class ManufacturedSource:
    """f = -div(kappa grad uD) for the linear Dirichlet data uD."""


class Diffusion:
    def __init__(self, kappa=None):
        self.kappa = kappa if kappa is not None else ConstantCoefficient(1.0)


class Source:
    def __init__(self, f=None):
        self.f = f if f is not None else ConstantSource(0.0)


class Elasticity:
    """eps(v) : (lambda tr(eps(u)) I + 2 mu eps(u)), constant Lame parameters."""

    def __init__(self, lame_lambda, lame_mu):
        self.lame_lambda, self.lame_mu = float(lame_lambda), float(lame_mu)


class BodyForce:
    """v . f with a constant body force f."""

    def __init__(self, f=(0.0, 0.0, 0.0)):
        self.f = tuple(float(v) for v in f)
        if len(self.f) != 3:
            raise ValueError("the body force has 3 components")


class Integral:
    """int(integrand) dOmega -- a DistributedDomainContribution in the reference."""

    def __init__(self, integrand, measure):
        self.integrand, self.measure = integrand, measure


def collect_cell_matrix_and_vector(U, V, a, l, uhd=None):
    """Collect the (lazy) cell data of the bilinear and linear forms
    (src/FESpaces.jl:703-715).  Here the cell data stays symbolic; it is lowered
    to a problem descriptor and integrated on the GPU."""
    vector = V.reffe.ncomp > 1
    if not (isinstance(a, Integral) and isinstance(a.integrand, Elasticity if vector else Diffusion)):
        raise NotImplementedError("bilinear form outside the GPU vocabulary")
    if not (isinstance(l, Integral) and isinstance(l.integrand, BodyForce if vector else Source)):
        raise NotImplementedError("linear form outside the GPU vocabulary")
    if vector and any(U.uD.coef):
        raise NotImplementedError("vector-valued spaces take zero Dirichlet data on the GPU path")
    if a.measure.degree != l.measure.degree:
        raise NotImplementedError("both forms must use the same Measure")
    return dict(U=U, V=V, a=a, l=l)


def problem_desc(data, strategy):
    """Lower collected data to the C descriptor (gdx_problem_desc)."""
    if "flux_op" in data:
        from .nonlinear import flux_problem_desc

        return flux_problem_desc(data["flux_op"], strategy)
    U, V, a, l = data["U"], data["V"], data["a"], data["l"]
    m = V.model
    d = _lib.ProblemDesc()
    d.dim, d.order, d.qdeg = m.dim, V.reffe.order, a.measure.degree
    d.strategy = strategy.code
    for k in range(3):
        d.cells[k] = m.cells[k] if k < m.dim else 1
        d.parts[k] = m.parts[k] if k < m.dim else 1
    dom = list(m.domain) + [0.0, 1.0] * (3 - m.dim)
    for k in range(6):
        d.domain[k] = dom[k]
    if isinstance(a.integrand, Elasticity):
        d.op_kind = _lib.OP_ELASTICITY
        d.lame_lambda, d.lame_mu = a.integrand.lame_lambda, a.integrand.lame_mu
        for k in range(3):
            d.body_force[k] = l.integrand.f[k]
        d.coef_const, d.nmodes, d.modes = 1.0, 0, None
        d.rhs_kind, d.rhs_const = _lib.RHS_CONST, 0.0
        d.dir_kind = _lib.DIR_CLAMPED_X0 if V.clamped else 0
        return d, None
    kap = a.integrand.kappa
    keep = None
    if isinstance(kap, ConstantCoefficient):
        d.coef_const, d.nmodes, d.modes = kap.value, 0, None
    elif isinstance(kap, FourierExpCoefficient):
        keep = kap.modes
        d.coef_const, d.nmodes = kap.scale, keep.shape[0]
        d.modes = keep.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if keep.size else None
    else:
        raise NotImplementedError("coefficient outside the GPU vocabulary")
    f = l.integrand.f
    if isinstance(f, ConstantSource):
        d.rhs_kind, d.rhs_const = _lib.RHS_CONST, f.value
    elif isinstance(f, ManufacturedSource):
        d.rhs_kind, d.rhs_const = _lib.RHS_MANUFACTURED, 0.0
    else:
        raise NotImplementedError("source outside the GPU vocabulary")
    d.dir_kind = 1 if V.dirichlet else 0
    for k in range(4):
        d.dir_coef[k] = U.uD.coef[k]
    return d, keep


class GpuProblem:
    """Owner of one gdx_problem handle."""

    def __init__(self, ctx, desc, keep=None):
        self._keep = keep
        h = ctypes.c_void_p()
        _lib.call("gdx_problem_create", ctx.handle, ctypes.byref(desc), ctypes.byref(h))
        self.handle = h
        self.symbolic_done = False
        self.mapped_coords_set = False  # vertex coordinates on the device (a symbolic phase drops them)

    def symbolic(self):
        _lib.call("gdx_symbolic", self.handle)
        self.symbolic_done = True
        self.mapped_coords_set = False
        n = ctypes.c_int32()
        _lib.call("gdx_num_parts", self.handle, ctypes.byref(n))
        g = ctypes.c_int64()
        _lib.call("gdx_num_global_dofs", self.handle, ctypes.byref(g))
        self.nparts, self.ngdofs = n.value, g.value

    def numeric(self, precision=_lib.PREC_FP64):
        _lib.call("gdx_numeric", self.handle, int(precision))

    def sync(self):
        _lib.call("gdx_sync", self.handle)

    def phase_times(self):
        t = (ctypes.c_double * 4)()
        _lib.call("gdx_phase_times", self.handle, t, 4)
        return list(t)

    def owned_nnz(self):
        n = ctypes.c_int64()
        _lib.call("gdx_owned_nnz", self.handle, ctypes.byref(n))
        return n.value

    def spmv(self, x):
        """y = A x over the owned rows (global-id order), on the device."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.ngdofs)
        _lib.call("gdx_spmv", self.handle, x.ctypes.data, y.ctypes.data)
        return y

    def spmv_device_timing(self, reps=20):
        """(ms, bytes): device time of one SpMV of the assembled matrix (hipEvents,
        averaged over reps) and its algorithmic bytes per launch."""
        ms, nb = ctypes.c_double(), ctypes.c_double()
        _lib.call("gdx_spmv_timed", self.handle, int(reps), ctypes.byref(ms), ctypes.byref(nb))
        return ms.value, nb.value

    def spmv_device_ms(self, reps=20):
        """Device time (ms) of one SpMV of the assembled matrix (hipEvents, averaged over reps)."""
        return self.spmv_device_timing(reps)[0]

    def pvector(self, which=_lib.PRANGE_ROWS):
        """A device PVector on A's row (default) or column PRange."""
        return DevicePVector(self, which)

    def mul(self, y, x):
        """mul!(y, A, x) on the device: x (column PRange) is made consistent first,
        then every part's local CSR times its local x; y is on the row PRange."""
        _lib.call("gdx_psparse_mul", self.handle, y.handle, x.handle)
        return y

    def pcg(self, rtol=1e-8, maxiter=10000, want_x=True):
        """Jacobi-PCG on the assembled system; returns (x, iterations, true relative residual, ms),
        ms the solve's wall time on the device (the copy of x to the host is outside it)."""
        x = np.empty(self.ngdofs) if want_x else None
        it, rr, ms = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
        _lib.call("gdx_pcg", self.handle, float(rtol), int(maxiter), None if x is None else x.ctypes.data,
                  ctypes.byref(it), ctypes.byref(rr), ctypes.byref(ms))
        return x, it.value, rr.value, ms.value

    def set_state(self, x):
        """State of a nonlinear operator: free values by global id (None = zero)."""
        if x is None:
            _lib.call("gdx_set_state", self.handle, None)
            return
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.ngdofs,):
            raise ValueError(f"state must have {self.ngdofs} free values")
        _lib.call("gdx_set_state", self.handle, x.ctypes.data)

    def set_kernel(self, variant, rows_z=32):
        """3D matrix kernel: 0 z-marching columns (rows_z z rows per workgroup), 1 per tile,
        2 per tile with direct stores, 3 per tile with wave-owned rows (Q2)."""
        _lib.call("gdx_problem_set_kernel", self.handle, int(variant), int(rows_z))

    def get_kernel(self):
        v, r = ctypes.c_int32(), ctypes.c_int32()
        _lib.call("gdx_problem_get_kernel", self.handle, ctypes.byref(v), ctypes.byref(r))
        return v.value, r.value

    def set_cell_coefficient(self, kappa):
        """kappa per cell (x fastest) for the next numeric assembly; None returns to the form's kappa."""
        if kappa is None:
            _lib.call("gdx_set_cell_coefficient", self.handle, None)
            return
        kappa = np.ascontiguousarray(kappa, dtype=np.float64).ravel()
        _lib.call("gdx_set_cell_coefficient", self.handle, kappa.ctypes.data)

    def set_vertex_coords(self, xyz):
        """Mapped grid: coordinates of every mesh vertex (shape (nvertices, dim),
        x fastest) for the next numeric assembly; None returns to the uniform box."""
        if xyz is None:
            _lib.call("gdx_set_vertex_coords", self.handle, None)
            self.mapped_coords_set = False
            return
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        _lib.call("gdx_set_vertex_coords", self.handle, xyz.ctypes.data)
        self.mapped_coords_set = True

    def set_reaction(self, diffusion_weight, reaction):
        """Mapped-grid path: assemble diffusion_weight kappa grad(u).grad(v) + reaction u v."""
        _lib.call("gdx_set_reaction", self.handle, float(diffusion_weight), float(reaction))

    def set_cell_lame(self, lame):
        """Elasticity: (lambda, mu) per cell (shape (ncells, 2), x fastest) for the
        next numeric assembly; None returns to the descriptor's constants."""
        if lame is None:
            _lib.call("gdx_set_cell_lame", self.handle, None)
            return
        lame = np.ascontiguousarray(lame, dtype=np.float64).reshape(-1, 2)
        _lib.call("gdx_set_cell_lame", self.handle, lame.ctypes.data)

    def cell_coefficient_grf(self, seed, corr_len, log_scale):
        """kappa = exp(log_scale g), g a unit-variance Gaussian random field drawn on the device."""
        _lib.call("gdx_cell_coefficient_grf", self.handle, int(seed), float(corr_len), float(log_scale))

    def export_cell_coefficient(self, ncells):
        k = np.empty(int(ncells))
        _lib.call("gdx_export_cell_coefficient", self.handle, k.ctypes.data)
        return k

    def newton(self, x0=None, rtol=1e-10, atol=0.0, maxiter=50, lin_rtol=1e-8, lin_maxiter=20000,
               init=_lib.NEWTON_INIT_GIVEN, precond=None, jac_precision=_lib.PREC_FP64):
        """Newton on the device; returns (x, report dict).  Corrections by
        Jacobi-PCG, or by FPCG with precond's multilevel preconditioner (a
        LearnedCoarseSolver on the same context).  jac_precision=PREC_FP32
        assembles the Jacobians with FP32 cell contractions (the residual and
        the Newton test stay FP64)."""
        x = np.zeros(self.ngdofs) if x0 is None else np.array(x0, dtype=np.float64)
        if x.shape != (self.ngdofs,):
            raise ValueError(f"initial guess must have {self.ngdofs} free values")
        o = _lib.NewtonOpts(float(rtol), float(atol), int(maxiter), float(lin_rtol), int(lin_maxiter), int(init),
                            None if precond is None else precond.handle, int(jac_precision))
        r = _lib.NewtonReport()
        _lib.call("gdx_newton", self.handle, ctypes.byref(o), x.ctypes.data, ctypes.byref(r))
        rep = {k: getattr(r, k) for k, _ in _lib.NewtonReport._fields_}
        rep["converged"] = bool(rep["converged"])
        return x, rep

    def stream(self):
        s = ctypes.c_void_p()
        _lib.call("gdx_stream", self.handle, ctypes.byref(s))
        return s.value

    def sizes(self, part):
        s = (ctypes.c_int64 * _lib.SZ_COUNT)()
        _lib.call("gdx_part_sizes", self.handle, int(part), s)
        return list(s)

    def export_fe(self, part):
        """The part's FE dof PRange only: (ldof_gid, ldof_owner, ldof_gnode)."""
        nl = self.sizes(part)[_lib.SZ_NLDOFS]
        gid, owner, gnode = np.empty(nl, np.int64), np.empty(nl, np.int32), np.empty(nl, np.int64)
        _lib.call("gdx_export_fe", self.handle, int(part), gid.ctypes.data, owner.ctypes.data, gnode.ctypes.data)
        return gid, owner, gnode

    def export_part(self, part):
        nl, no, ngr, ngc, nnz = self.sizes(part)
        P = dict(n_ldofs=nl, n_own=no, n_ghost_rows=ngr, n_ghost_cols=ngc, nnz=nnz)
        P["ldof_gid"] = np.empty(nl, np.int64)
        P["ldof_owner"] = np.empty(nl, np.int32)
        P["ldof_gnode"] = np.empty(nl, np.int64)
        _lib.call("gdx_export_fe", self.handle, int(part), P["ldof_gid"].ctypes.data, P["ldof_owner"].ctypes.data,
                  P["ldof_gnode"].ctypes.data)
        nr, nc = no + ngr, no + ngc
        for k, n, t in (("row_gid", nr, np.int64), ("row_owner", nr, np.int32), ("col_gid", nc, np.int64),
                        ("col_owner", nc, np.int32), ("rowptr", nr + 1, np.int64), ("colidx", nnz, np.int32),
                        ("val", nnz, np.float64), ("b", nr, np.float64)):
            P[k] = np.empty(n, t)
        _lib.call("gdx_export_rows", self.handle, int(part), P["row_gid"].ctypes.data, P["row_owner"].ctypes.data)
        _lib.call("gdx_export_cols", self.handle, int(part), P["col_gid"].ctypes.data, P["col_owner"].ctypes.data)
        _lib.call("gdx_export_csr", self.handle, int(part), P["rowptr"].ctypes.data, P["colidx"].ctypes.data,
                  P["val"].ctypes.data)
        _lib.call("gdx_export_rhs", self.handle, int(part), P["b"].ctypes.data)
        return P

    def close(self):
        if self.handle:
            _lib.call("gdx_problem_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SparseMatrixAssembler:
    """SparseMatrixAssembler(U, V[, strategy]) (src/FESpaces.jl:819-862).

    Default strategy SubAssembledRows; local matrices are CSR (0-based)."""

    def __init__(self, U, V, strategy=None):
        self.U, self.V = U, V
        self.strategy = strategy if strategy is not None else SubAssembledRows()
        if not isinstance(self.strategy, (SubAssembledRows, FullyAssembledRows)):
            raise ValueError("strategy must be SubAssembledRows() or FullyAssembledRows()")
        self.problem = None
        self._key = None

    def _problem_for(self, data):
        desc, keep = problem_desc(data, self.strategy)
        key = (bytes(desc), None if keep is None else keep.tobytes())
        if self.problem is None or key != self._key:
            if self.problem is not None:
                self.problem.close()
            self.problem = GpuProblem(self.V.model.ctx, desc, keep)
            self.problem.symbolic()
            if self.V.model.vertex_coords is not None:
                self.problem.set_vertex_coords(self.V.model.vertex_coords)
            self._key = key
        return self.problem

    def export(self):
        prob = self.problem
        mats, rows, cols, bs = [], [], [], []
        for r in range(prob.nparts):
            P = prob.export_part(r)
            no = P["n_own"]
            rows.append(LocalIndices(r, prob.ngdofs, P["row_gid"][:no], P["row_gid"][no:], P["row_owner"][no:]))
            cols.append(LocalIndices(r, prob.ngdofs, P["col_gid"][:no], P["col_gid"][no:], P["col_owner"][no:]))
            nr, nc = len(P["row_gid"]), len(P["col_gid"])
            mats.append(sp.csr_matrix((P["val"], P["colidx"], P["rowptr"]), shape=(nr, nc)))
            bs.append(P["b"])
        return PSparseMatrix(mats, rows, cols), PVector(bs, rows)


def assemble_matrix_and_vector(assem, data, export=True):
    """Fresh assembly: symbolic (once per assembler) + numeric on the GPU."""
    prob = assem._problem_for(data)
    prob.numeric()
    prob.sync()
    return assem.export() if export else None


def allocate_matrix_and_vector(assem, data):
    """Symbolic phase only: matrix/vector with the final structure, zero values."""
    assem._problem_for(data)
    return assem.export()


def assemble_matrix_and_vector_(A, b, assem, data):
    """In-place numeric re-assembly (assemble_matrix_and_vector! in the
    reference, src/Algebra.jl:323-393,530-535): refills the device values and
    updates A and b's local values."""
    prob = assem._problem_for(data)
    prob.numeric()
    prob.sync()
    for r in range(prob.nparts):
        P = prob.export_part(r)
        A.matrix_partition[r].data[:] = P["val"]
        b.partition[r][:] = P["b"]
    return A, b


class AffineFEOperator:
    """AffineFEOperator(a, l, U, V[, assem]): assembles A and b on construction."""

    def __init__(self, a, l, U, V, assem=None):
        self.assem = assem if assem is not None else SparseMatrixAssembler(U, V)
        data = collect_cell_matrix_and_vector(U, V, a, l, None)
        self.matrix, self.vector = assemble_matrix_and_vector(self.assem, data)
        self.trial, self.test = U, V

    def get_matrix(self):
        return self.matrix

    def get_vector(self):
        return self.vector
