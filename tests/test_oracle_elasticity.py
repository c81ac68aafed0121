"""Linear elasticity (vector-valued Lagrangian Q1/Q2, 81 x 81 Q2 cell matrices):
pins of the C oracle and the library's host symbolic phase, all on the CPU.

What pins the oracle (it computes each cell matrix by 3D Gauss quadrature with
the inverse Jacobian, component-major cell dofs):
- an independent restatement (oracle/micro.py: 1D matrices assembled over the
  mesh, combined as Kronecker products of the bilinear form's separable terms,
  numbering-free) gives the same matrix and body-force vector to 1e-13;
- the six rigid-body modes lie in the kernel of the unconstrained matrix (a
  dropped or transposed mu d_b(v) d_a(u) term breaks the rotations);
- a linear field u = G x stores the energy |Omega| (lambda tr(e)^2 + 2 mu e:e),
  e = sym(G) (pins lambda and mu separately; the skew part of G adds nothing);
- the body-force vector sums to |Omega| f per component;
- nnz = 9 prod(8 n_d + 1) for unconstrained Q2 (9 component pairs), the
  closed form of the scope table's elasticity row;
- the owned rows do not depend on the partition.
With Lame parameters given per cell (p["cell_lame"], a piecewise-constant
material): the restatement summed cell by cell gives the same matrix; the
rigid-body modes stay in the kernel; a linear field stores
sum_c |cell| (lambda_c tr(e)^2 + 2 mu_c e:e); a pair that is nonzero on one
cell only couples that cell's nodes and nothing else (pins the x-fastest cell
index).
The library's host symbolic phase is compared field by field with the oracle
(numbering, gids, owners, row/col PRanges with the component-interleaved ghost
columns, CSR pattern).
"""
import numpy as np
import pytest

import micro
import oracle
from _helpers import STRUCT_KEYS, owned_global
from gridap_amd import workloads as W

DOM = (0.0, 1.5, 0.0, 1.0, 0.0, 0.7)


def gid_to_latdof(res):
    """global id -> 3 * lattice node + component (component = local dof % 3)."""
    m = np.zeros(res.ngdofs, np.int64)
    for P in res.parts:
        m[P["ldof_gid"]] = 3 * P["ldof_gnode"] + np.arange(P["n_ldofs"]) % 3
    return m


@pytest.mark.parametrize("cells,parts,clamped,order", [
    ((3, 2, 2), (2, 1, 2), True, 2), ((4, 3, 2), (1, 1, 1), False, 2), ((4, 3, 5), (2, 1, 2), True, 1),
    ((2, 3, 3), (2, 3, 1), True, 2)])
def test_oracle_matches_kronecker_restatement(cells, parts, clamped, order):
    p = W.elasticity3d_q2(cells=cells, parts=parts, clamped=clamped, order=order, domain=DOM,
                          E=2.5, nu=0.27, body_force=(0.3, -0.2, -1.0))
    r = oracle.run(p)
    A, b = owned_global(r.parts, r.ngdofs)
    g2d = gid_to_latdof(r)
    assert r.ngdofs == micro.elasticity_free_dofs(p).sum()
    Am, bm = micro.elasticity_global(p)
    Am, bm = Am[g2d][:, g2d], bm[g2d]
    assert abs(A - Am).max() <= 1e-13 * abs(Am).max()
    assert np.abs(b - bm).max() <= 1e-13 * np.abs(bm).max()
    assert abs(A - A.T).max() <= 1e-14 * abs(A).max()


def test_rigid_modes_energy_and_body_force():
    lam, mu = W.lame_parameters(1.0, 0.3)
    p = W.elasticity3d_q2(cells=(3, 2, 4), parts=(2, 2, 1), clamped=False, domain=DOM)
    r = oracle.run(p)
    A, b = owned_global(r.parts, r.ngdofs)
    g2d = gid_to_latdof(r)
    node, comp = g2d // 3, g2d % 3
    _, X = micro.node_coords(p, node)
    vol = 1.5 * 1.0 * 0.7
    scale = abs(A).max()
    for k in range(3):  # translations
        assert np.abs(A @ (comp == k).astype(float)).max() <= 1e-13 * scale
    for i, j in ((0, 1), (1, 2), (0, 2)):  # infinitesimal rotations u_i = -x_j, u_j = x_i
        rot = np.where(comp == i, -X[:, j], 0.0) + np.where(comp == j, X[:, i], 0.0)
        assert np.abs(A @ rot).max() <= 1e-13 * scale
    G = np.array([[0.3, -0.7, 0.2], [0.5, -0.1, 0.4], [-0.6, 0.9, 0.25]])
    u = np.einsum("nj,nj->n", G[comp], X)  # u_c(x) = sum_j G[c, j] x_j
    e = 0.5 * (G + G.T)
    energy = vol * (lam * np.trace(e) ** 2 + 2.0 * mu * np.sum(e * e))
    assert abs(u @ (A @ u) - energy) <= 1e-12 * energy
    for k in range(3):
        assert abs(b[comp == k].sum() - vol * p["body_force"][k]) <= 1e-13
    assert A.nnz == 9 * np.prod([8 * n + 1 for n in (3, 2, 4)])


def test_partition_invariance():
    base = dict(cells=(4, 3, 3), clamped=True, domain=DOM)
    mats = []
    for parts in ((1, 1, 1), (3, 2, 2), (2, 3, 1)):
        r = oracle.run(W.elasticity3d_q2(parts=parts, **base))
        A, b = owned_global(r.parts, r.ngdofs)
        g2d = gid_to_latdof(r)
        perm = np.argsort(g2d)
        mats.append((A[perm][:, perm], b[perm]))
    for A, b in mats[1:]:
        assert abs(A - mats[0][0]).max() <= 1e-14 * abs(A).max()
        assert np.abs(b - mats[0][1]).max() <= 1e-15


@pytest.mark.parametrize("cells,parts,strategy,clamped,order", [
    ((3, 2, 2), (2, 1, 2), "sub", True, 2), ((4, 3, 3), (2, 2, 1), "full", True, 2),
    ((3, 3, 2), (1, 1, 1), "sub", False, 2), ((5, 3, 4), (3, 1, 2), "sub", True, 1),
    ((4, 4, 3), (2, 2, 2), "full", False, 1), ((5, 4, 3), (3, 2, 1), "sub", True, 2)])
def test_host_symbolic_structure_matches_oracle(host_ctx, cells, parts, strategy, clamped, order):
    from test_host_symbolic import host_problem

    p = W.elasticity3d_q2(cells=cells, parts=parts, clamped=clamped, strategy=strategy, order=order)
    r = oracle.run(p)
    gp = host_problem(host_ctx, p)
    assert gp.ngdofs == r.ngdofs
    for k, P in enumerate(r.parts):
        Q = gp.export_part(k)
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
    gp.close()


def test_descriptor_errors(host_ctx):
    """gdx_problem_create refuses elasticity descriptors outside what is lowered."""
    from gridap_amd import _lib, fespaces, problem_from_workload
    from gridap_amd._lib import GdxError

    p = W.elasticity3d_q2(cells=(2, 2, 2), parts=(1, 1, 1))
    assem, data = problem_from_workload(host_ctx, p)
    for field, value, code in (("dim", 2, -4), ("lame_mu", 0.0, -1), ("dir_kind", 1, -4), ("nmodes", 1, -4)):
        desc, keep = fespaces.problem_desc(data, assem.strategy)
        setattr(desc, field, value)
        if field == "nmodes":
            keep = np.zeros((1, 5))
            desc.modes = keep.ctypes.data_as(_lib.ctypes.POINTER(_lib.ctypes.c_double))
        with pytest.raises(GdxError) as e:
            fespaces.GpuProblem(host_ctx, desc, keep)
        assert e.value.code == code, field
    desc, keep = fespaces.problem_desc(data, assem.strategy)
    desc.dir_coef[0] = 1.0
    with pytest.raises(GdxError):
        fespaces.GpuProblem(host_ctx, desc, keep)
    # the Python vocabulary refuses nonzero Dirichlet data for vector spaces
    from gridap_amd import (BodyForce, CartesianDiscreteModel, Elasticity, Float64, Integral, LinearFunction, Measure,
                            ReferenceFE, TestFESpace, Triangulation, TrialFESpace, VectorValue, lagrangian,
                            collect_cell_matrix_and_vector)
    model = CartesianDiscreteModel(host_ctx, (1, 1, 1), (0, 1, 0, 1, 0, 1), (2, 2, 2))
    V = TestFESpace(model, ReferenceFE(lagrangian, VectorValue(3, Float64), 2), dirichlet_tags="clamped_x0")
    U = TrialFESpace(V, LinearFunction(1.0))
    dO = Measure(Triangulation(model), 4)
    with pytest.raises(NotImplementedError):
        collect_cell_matrix_and_vector(U, V, Integral(Elasticity(1, 1), dO), Integral(BodyForce((0, 0, 1)), dO))


def random_lame(cells, seed):
    rng = np.random.default_rng(seed)
    nc = int(np.prod(cells))
    return np.stack([rng.uniform(0.2, 3.0, nc), rng.uniform(0.1, 2.0, nc)], axis=1)


@pytest.mark.parametrize("cells,parts,clamped,order", [
    ((3, 2, 2), (2, 1, 2), True, 2), ((4, 3, 2), (1, 1, 1), False, 2), ((4, 3, 5), (2, 1, 2), True, 1),
    ((2, 3, 3), (2, 3, 1), False, 1)])
def test_cellwise_lame_matches_restatement(cells, parts, clamped, order):
    p = W.elasticity3d_q2(cells=cells, parts=parts, clamped=clamped, order=order, domain=DOM,
                          body_force=(0.3, -0.2, -1.0))
    p["cell_lame"] = random_lame(cells, 17 + order)
    r = oracle.run(p)
    A, b = owned_global(r.parts, r.ngdofs)
    g2d = gid_to_latdof(r)
    Am, bm = micro.elasticity_global(p)
    Am, bm = Am[g2d][:, g2d], bm[g2d]
    assert abs(A - Am).max() <= 1e-13 * abs(Am).max()
    assert np.abs(b - bm).max() <= 1e-13 * np.abs(bm).max()
    assert abs(A - A.T).max() <= 1e-14 * abs(A).max()
    # the constant pair everywhere is the constant-coefficient form
    q = dict(p, cell_lame=np.tile(p["lame"], (int(np.prod(cells)), 1)))
    A1, _ = owned_global(oracle.run(q).parts, r.ngdofs)
    A0, _ = owned_global(oracle.run(dict(p, cell_lame=None)).parts, r.ngdofs)
    assert abs(A1 - A0).max() <= 1e-14 * abs(A0).max()


def test_cellwise_lame_rigid_modes_energy_and_locality():
    cells = (3, 2, 4)
    p = W.elasticity3d_q2(cells=cells, parts=(2, 2, 1), clamped=False, domain=DOM)
    lame = random_lame(cells, 5)
    r = oracle.run(dict(p, cell_lame=lame))
    A, _ = owned_global(r.parts, r.ngdofs)
    g2d = gid_to_latdof(r)
    node, comp = g2d // 3, g2d % 3
    _, X = micro.node_coords(p, node)
    scale = abs(A).max()
    for k in range(3):
        assert np.abs(A @ (comp == k).astype(float)).max() <= 1e-13 * scale
    for i, j in ((0, 1), (1, 2), (0, 2)):
        rot = np.where(comp == i, -X[:, j], 0.0) + np.where(comp == j, X[:, i], 0.0)
        assert np.abs(A @ rot).max() <= 1e-13 * scale
    G = np.array([[0.3, -0.7, 0.2], [0.5, -0.1, 0.4], [-0.6, 0.9, 0.25]])
    u = np.einsum("nj,nj->n", G[comp], X)
    e = 0.5 * (G + G.T)
    vcell = 1.5 * 1.0 * 0.7 / np.prod(cells)
    energy = vcell * np.sum(lame[:, 0] * np.trace(e) ** 2 + 2.0 * lame[:, 1] * np.sum(e * e))
    assert abs(u @ (A @ u) - energy) <= 1e-12 * energy
    # one cell with a nonzero pair: only that cell's nodes are coupled
    nn = W.lattice_shape(p)
    for cx, cy, cz in ((2, 0, 3), (0, 1, 1)):
        one = np.zeros_like(lame)
        one[cx + cells[0] * (cy + cells[1] * cz)] = (1.3, 0.6)
        A1, _ = owned_global(oracle.run(dict(p, cell_lame=one)).parts, r.ngdofs)
        ix, iy, iz = node % nn[0], (node // nn[0]) % nn[1], node // (nn[0] * nn[1])
        inside = ((ix >= 2 * cx) & (ix <= 2 * cx + 2) & (iy >= 2 * cy) & (iy <= 2 * cy + 2) &
                  (iz >= 2 * cz) & (iz <= 2 * cz + 2))
        rows = np.unique(A1.nonzero()[0])
        assert len(rows) == 81 and inside[rows].all()
