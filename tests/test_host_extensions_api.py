"""The host side of the mapped-grid, reaction, cell-wise material and block
assembly interfaces, CPU only: what the C ABI refuses without a device
context, and what the Python mirror refuses before any problem is made."""
import numpy as np
import pytest

from gridap_amd import (Block, BlockMultiFieldStyle, CartesianDiscreteModel, ConstantSource, Diffusion, Elasticity,
                        FESpace, Float64, LinearFunction, Mass, Measure, MultiFieldFESpace, ReferenceFE, Source,
                        Triangulation, TrialFESpace, assemble_block_matrix_and_vector, lagrangian)
from gridap_amd import workloads as W
from gridap_amd._lib import GdxError
from test_host_symbolic import host_problem


def test_host_only_problems_refuse_device_state(host_ctx):
    """set_vertex_coords, set_reaction and set_cell_lame need a device context
    (GDX_ERR_STATE), after the symbolic phase"""
    p = W.poisson3d_q1(4, (2, 1, 1))
    gp = host_problem(host_ctx, p)
    with pytest.raises(GdxError) as e:
        gp.set_vertex_coords(W.mapped_vertices(p))
    assert e.value.code == -3
    with pytest.raises(GdxError) as e:
        gp.set_reaction(1.0, 1.0)
    assert e.value.code == -3
    with pytest.raises(GdxError) as e:
        gp.set_cell_lame(np.ones((64, 2)))
    assert e.value.code == -3
    gp.close()


def test_mapped_vertices_shape_and_boundary():
    p = W.poisson3d_q1(5, (1, 1, 1))
    X = W.mapped_vertices(p, amp=0.0)
    assert X.shape == (216, 3) and np.allclose(X[1], [0.2, 0.0, 0.0]) and np.allclose(X[6], [0.0, 0.2, 0.0])
    Y = W.mapped_vertices(p, amp=0.5, seed=1, smooth=0.05)
    idx = np.stack(np.meshgrid(*[np.arange(6)] * 3, indexing="ij"), -1).reshape(-1, 3, order="F")
    boundary = np.any((idx == 0) | (idx == 5), axis=1)
    assert np.array_equal(Y[boundary], X[boundary])  # the box's boundary stays in place
    assert np.abs(Y - X).max() <= 0.25 * 0.2 + 0.05 + 1e-15


def test_model_map_and_vertex_coords(host_ctx):
    m = CartesianDiscreteModel(host_ctx, (1, 1), (0, 2, 0, 1), (4, 2), map=lambda X: X * [1.0, 3.0])
    assert m.vertex_coords.shape == (15, 2)
    assert np.allclose(m.vertex_coords[-1], [2.0, 3.0]) and np.allclose(m.vertex_coords[1], [0.5, 0.0])
    with pytest.raises(ValueError):
        CartesianDiscreteModel(host_ctx, (1, 1), (0, 1, 0, 1), (4, 2), vertex_coords=np.zeros((14, 2)))
    with pytest.raises(ValueError):
        CartesianDiscreteModel(host_ctx, (1, 1), (0, 1, 0, 1), (4, 2), map=lambda X: X,
                               vertex_coords=np.zeros((15, 2)))


def test_block_forms_refuse_what_is_not_lowered(host_ctx):
    model = CartesianDiscreteModel(host_ctx, (1, 1), (0, 1, 0, 1), (4, 4))
    V = FESpace(model, ReferenceFE(lagrangian, Float64, 1), dirichlet_tags="boundary")
    U = TrialFESpace(V, LinearFunction(0.0, (1.0, 1.0)))
    dO = Measure(Triangulation(model), 2)
    X, Y = MultiFieldFESpace([U, U]), MultiFieldFESpace([V, V], style=BlockMultiFieldStyle())
    src = [Source(ConstantSource(1.0))] * 2
    with pytest.raises(NotImplementedError):
        MultiFieldFESpace([U, U], style=object())  # only the block style
    with pytest.raises(NotImplementedError):
        Block(0, 0, Elasticity(1.0, 1.0), dO)
    with pytest.raises(ValueError):  # a field index out of range
        assemble_block_matrix_and_vector(X, Y, [Block(0, 2, Mass(1.0), dO)], src)
    with pytest.raises(ValueError):  # a test field without a term
        assemble_block_matrix_and_vector(X, Y, [Block(0, 0, Diffusion(), dO)], src)
    with pytest.raises(ValueError):  # one source per field
        assemble_block_matrix_and_vector(X, Y, [Block(0, 0, Diffusion(), dO)], src[:1])
    with pytest.raises(NotImplementedError):  # two diffusion terms in one block
        assemble_block_matrix_and_vector(X, Y, [Block(0, 0, Diffusion(), dO), Block(0, 0, Diffusion(), dO),
                                                Block(1, 1, Mass(1.0), dO)], src)


def test_vtk_node_coordinates_on_a_mapped_model(host_ctx):
    """VTK points and interpolation use the mapped node coordinates: vertices
    where the map puts them, Q2 midpoints at the mean of their entity's
    vertices (the rule of the oracle's mapped geometry)"""
    import micro
    from gridap_amd.visualization import node_coords

    p = W.poisson3d_q1(3, (1, 1, 1))
    vc = W.mapped_vertices(p, amp=0.5, seed=2, smooth=0.05)
    m = CartesianDiscreteModel(host_ctx, (1, 1, 1), (0, 1, 0, 1, 0, 1), (3, 3, 3), vertex_coords=vc)
    assert np.array_equal(node_coords(m, 1), vc)
    x2 = node_coords(m, 2)
    q = dict(p, order=2, vertex_coords=vc)
    ref = micro.mapped_node_coords(q, np.arange(7 ** 3))
    assert np.abs(x2 - ref[:, :3]).max() <= 1e-15
    m0 = CartesianDiscreteModel(host_ctx, (1, 1, 1), (0, 1, 0, 1, 0, 1), (3, 3, 3))
    assert np.abs(node_coords(m0, 2) - micro.node_coords(dict(p, order=2), np.arange(7 ** 3))[1]).max() <= 1e-15
