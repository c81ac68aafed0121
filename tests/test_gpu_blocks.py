"""Multi-field block assembly on device 0 (gridap_amd.assemble_block_matrix_and_vector,
the BlockMultiFieldStyle path of src/MultiField.jl:470-560).

The reference's own block test (test/BlockSparseMatrixAssemblersTests.jl:100-120):
2D Q1 on 8 x 8 cells, Dirichlet data x + y on the whole boundary for every
field, quadrature degree 2, FullyAssembledRows, the forms
  biform2((u1,u2),(v1,v2)) = int( grad(u1).grad(v1) + u2 v2 + u1 v2 )
  liform2((v1,v2))          = int( v1 + v2 )
and the three-field biform3 / liform3.  It checks that the block assembly and
the monolithic one agree on b and on A 1, field by field, to 1e-10.  Here every
block is compared with the oracle's single-field assembly of its own
diffusion-reaction form (structure bit-identical, values within 1e-12), and
per field b and A 1 are compared with the oracle's blocks summed to 1e-10:
the reference's per-field invariant, its other side taken from the oracle
(no monolithic multi-field numbering is built here).
Also: the diffusion-reaction form alone (gdx_set_reaction) on mapped and
uniform grids, 2D and 3D, against the oracle; its refusals.
"""
import numpy as np
import pytest

import oracle
from _helpers import STRUCT_KEYS, mapped_vertices, owned_global, rel_err
from gridap_amd import (Block, BlockMultiFieldStyle, CartesianDiscreteModel, ConstantSource, Diffusion, FESpace,
                        Float64, FullyAssembledRows, LinearFunction, Mass, Measure, MultiFieldFESpace, ReferenceFE,
                        Source, Triangulation, TrialFESpace, assemble_block_matrix_and_vector, lagrangian,
                        problem_from_workload)
from gridap_amd import workloads as W
from gridap_amd._lib import GdxError

pytestmark = pytest.mark.gpu


def block_workload(weights, rhs_const):
    """the single-field problem of one block of the reference's block test"""
    p = W.poisson2d_q1(8, (2, 2), strategy="full")
    return dict(p, modes=np.zeros((0, 5)), coef_const=1.0, rhs="const", rhs_const=rhs_const, uD=(0.0, 1.0, 1.0, 0.0),
                dirichlet=True, weights=weights)


def global_of(parts, ngdofs):
    return owned_global(parts, ngdofs)


@pytest.mark.parametrize("nfields", [2, 3])
def test_reference_block_forms(gpu_ctx, nfields):
    model = CartesianDiscreteModel(gpu_ctx, (2, 2), (0, 1, 0, 1), (8, 8))
    V = FESpace(model, ReferenceFE(lagrangian, Float64, 1), dirichlet_tags="boundary")  # TestFESpace
    U = TrialFESpace(V, LinearFunction(0.0, (1.0, 1.0)))  # sol(x) = sum(x)
    dO = Measure(Triangulation(model), 2)
    X = MultiFieldFESpace([U] * nfields, style=BlockMultiFieldStyle())
    Y = MultiFieldFESpace([V] * nfields, style=BlockMultiFieldStyle())
    if nfields == 2:  # grad(u1).grad(v1) + u2 v2 + u1 v2
        terms = [Block(0, 0, Diffusion(), dO), Block(1, 1, Mass(1.0), dO), Block(1, 0, Mass(1.0), dO)]
    else:  # ... + u3 v2 + u2 v3
        terms = [Block(0, 0, Diffusion(), dO), Block(1, 1, Mass(1.0), dO), Block(1, 0, Mass(1.0), dO),
                 Block(1, 2, Mass(1.0), dO), Block(2, 1, Mass(1.0), dO)]
    sources = [Source(ConstantSource(1.0))] * nfields
    A, b, probs = assemble_block_matrix_and_vector(X, Y, terms, sources, FullyAssembledRows())
    ref = {}
    for t in terms:
        w = (1.0, 0.0) if isinstance(t.integrand, Diffusion) else (0.0, 1.0)
        ref[(t.i, t.j)] = oracle.run(block_workload(w, 0.0))
    src = oracle.run(block_workload((1.0, 0.0), 1.0))  # the source integral: b with f = 1 minus b with f = 0
    ng = src.ngdofs
    _, b_f1 = global_of(src.parts, ng)
    _, b_f0 = global_of(ref[(0, 0)].parts, ng)
    source = b_f1 - b_f0
    ones = np.ones(ng)
    for i in range(nfields):
        yi = np.zeros(ng)
        bi = source.copy()
        yg = np.zeros(ng)
        for j in range(nfields):
            if A.blocks[i][j] is None:
                assert (i, j) not in ref
                continue
            r = ref[(i, j)]
            _, prob = probs[(i, j)]
            for k, P in enumerate(r.parts):
                Q = prob.export_part(k)
                for key in STRUCT_KEYS:
                    assert np.array_equal(P[key], Q[key]), (i, j, k, key)
                assert rel_err(Q["val"], P["val"]) <= 1e-12, (i, j, k)
            Ar, br = global_of(r.parts, ng)
            yi += Ar @ ones
            bi += br
            Ag, _ = global_of([prob.export_part(k) for k in range(prob.nparts)], ng)
            yg += Ag @ ones
        # the block vector of field i (owned values by gid)
        bg = np.zeros(ng)
        for part, rows in zip(b.blocks[i].partition, b.blocks[i].index_partition):
            no = len(rows.own_to_global)
            bg[rows.own_to_global] = part[:no]
        assert np.abs(bg - bi).max() <= 1e-10, i
        assert np.abs(yg - yi).max() <= 1e-10, i
    for _, prob in probs.values():
        prob.close()


@pytest.mark.parametrize("dim,order,weights,mapped", [(2, 1, (0.0, 1.0), False), (2, 2, (0.7, 2.5), True),
                                                      (3, 1, (1.0, 3.0), True), (3, 2, (0.0, 1.5), False)])
def test_reaction_form_matches_oracle(gpu_ctx, dim, order, weights, mapped):
    p = W.poisson2d_q1(7, (2, 2)) if dim == 2 else W.poisson3d_q1(5, (2, 1, 2))
    p = dict(p, order=order, qdeg=2 * order, modes=np.zeros((0, 5)), coef_const=1.3, rhs="const", rhs_const=0.8,
             uD=(0.3, 1.0, -0.5, 0.25), weights=weights)
    vc = mapped_vertices(p, amp=0.4, seed=order + 3 * dim) if mapped else None
    print("reaction case", dim, order, weights, mapped, flush=True)
    r = oracle.run(dict(p, vertex_coords=vc))
    q = {k: v for k, v in p.items() if k != "weights"}
    assem, data = problem_from_workload(gpu_ctx, dict(q, vertex_coords=vc))
    prob = assem._problem_for(data)
    if not prob.mapped_coords_set:
        from gridap_amd import workloads

        prob.set_vertex_coords(workloads.mapped_vertices(p, amp=0.0))
    prob.set_reaction(*weights)
    prob.numeric()
    prob.sync()
    for k, P in enumerate(r.parts):
        Q = prob.export_part(k)
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
        assert rel_err(Q["val"], P["val"]) <= 1e-12, k
        assert rel_err(Q["b"], P["b"]) <= 1e-12, k
    # clearing the coordinates returns to the plain diffusion form
    prob.set_vertex_coords(None)
    prob.numeric()
    prob.sync()
    r0 = oracle.run(q)
    for k, P in enumerate(r0.parts):
        assert rel_err(prob.export_part(k)["val"], P["val"]) <= 1e-12, k
    prob.close()


def test_reaction_refusals(gpu_ctx):
    p = W.poisson3d_q1(4, (1, 1, 1))
    assem, data = problem_from_workload(gpu_ctx, p)
    prob = assem._problem_for(data)
    with pytest.raises(GdxError) as e:  # needs the mapped-grid path
        prob.set_reaction(1.0, 1.0)
    assert e.value.code == -3
    from gridap_amd import workloads

    prob.set_vertex_coords(workloads.mapped_vertices(p, amp=0.0))
    for w in ((0.0, 0.0), (-1.0, 1.0), (1.0, np.nan)):
        with pytest.raises(GdxError) as e:
            prob.set_reaction(*w)
        assert e.value.code == -1, w
    prob.close()
