"""Diffusion on a mapped Cartesian grid on device 0 (gdx_set_vertex_coords).

- parity with the oracle given the same vertex coordinates: structure
  bit-identical, values and RHS within 1e-12 relative, ghost rows zero
  (2D/3D, Q1/Q2, both strategies, Fourier kappa with the manufactured source
  and Dirichlet lifting, 37 Fourier modes, cell-wise kappa, no Dirichlet);
- repeated assembly is bit-identical; clearing the map, or a new symbolic
  phase, returns to the uniform assembly bit for bit;
- at 3D Q2 32^3, parts (2,2,2): the patch test (nodal values of a linear u
  solve A x = b with f = 0) and the RHS summing to f |box| with the boundary
  left in place, through the exported matrix;
- a hipGraph captured with the mapped problem replays what eager assembly
  writes;
- refusals: FP32, elasticity, non-finite coordinates.
"""
import numpy as np
import pytest

import micro
import oracle
from _helpers import STRUCT_KEYS, mapped_vertices, owned_global, rel_err
from gridap_amd import AssemblyBatch, problem_from_workload
from gridap_amd import workloads as W
from gridap_amd._lib import PREC_FP32, GdxError

pytestmark = pytest.mark.gpu

TOL = 1e-12
MODES = np.array([[1, 2, 0, 0.3, 0.1], [0, 1, 3, -0.2, 0.5], [2, 0, 1, 0.25, -0.4]], np.float64)


def case(dim, order, n, parts, **kw):
    p = W.poisson2d_q1(n, parts) if dim == 2 else W.poisson3d_q1(n, parts)
    p = dict(p, order=order, qdeg=2 * order)
    p.update(kw)
    return p


CASES = [
    case(2, 1, 9, (2, 2), modes=MODES * [1, 1, 0, 1, 1], rhs="manufactured", uD=(0.5, 1.0, -2.0, 0.0)),
    case(2, 2, 7, (3, 1), strategy="full", modes=MODES * [1, 1, 0, 1, 1], rhs_const=1.0),
    case(3, 1, 8, (2, 2, 1), modes=MODES, rhs="manufactured", uD=(0.5, 1.0, -2.0, 0.7)),
    case(3, 1, 7, (1, 2, 3), strategy="full", rhs_const=1.0, dirichlet=False),
    case(3, 2, 5, (2, 1, 2), modes=MODES, rhs="manufactured", uD=(0.2, -1.0, 0.5, 1.5)),
    case(3, 2, 4, (1, 1, 1), rhs_const=2.0, uD=(1.0, 0.0, 0.0, 0.0)),
]


def ident(p):
    return f"{p['dim']}d-q{p['order']}-{p['cells']}-{p['parts']}-{p.get('strategy', 'sub')}"


def gpu_problem(ctx, p, coords=True):
    q = dict(p)
    if not coords:
        q.pop("vertex_coords", None)
    assem, data = problem_from_workload(ctx, q)
    prob = assem._problem_for(data)
    prob.numeric()
    prob.sync()
    return prob


def values(prob):
    prob.sync()
    return [(Q["val"].copy(), Q["b"].copy()) for Q in (prob.export_part(k) for k in range(prob.nparts))]


def same(a, b):
    return all(np.array_equal(v0.view(np.int64), v1.view(np.int64)) and np.array_equal(b0.view(np.int64),
                                                                                       b1.view(np.int64))
               for (v0, b0), (v1, b1) in zip(a, b))


@pytest.mark.parametrize("p", CASES, ids=ident)
def test_parity_with_oracle(gpu_ctx, p):
    print("mapped case", ident(p), flush=True)
    p = dict(p, vertex_coords=mapped_vertices(p, amp=0.5, seed=p["dim"] * 10 + p["order"], smooth=0.06))
    r = oracle.run(p)
    prob = gpu_problem(gpu_ctx, p)
    assert prob.ngdofs == r.ngdofs
    for k, P in enumerate(r.parts):
        Q = prob.export_part(k)
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
        assert rel_err(Q["val"], P["val"]) <= TOL, k
        assert rel_err(Q["b"], P["b"]) <= TOL, k
        no = P["n_own"]
        assert np.all(Q["val"][Q["rowptr"][no]:] == 0.0) and np.all(Q["b"][no:] == 0.0)
    first = values(prob)
    prob.numeric()
    assert same(first, values(prob))
    # back to the uniform box: the unmapped assembly, bit for bit
    ref = values(gpu_problem(gpu_ctx, p, coords=False))
    prob.set_vertex_coords(None)
    prob.numeric()
    assert same(ref, values(prob))
    prob.close()


def test_cellwise_kappa_on_a_mapped_grid(gpu_ctx):
    p = case(3, 1, 9, (2, 2, 1), rhs_const=1.0)
    p["vertex_coords"] = mapped_vertices(p, amp=0.4, seed=5)
    kap = np.exp(np.random.default_rng(2).normal(size=int(np.prod(p["cells"]))))
    r = oracle.run(dict(p, cell_kappa=kap))
    prob = gpu_problem(gpu_ctx, p)
    prob.set_cell_coefficient(kap)
    prob.numeric()
    prob.sync()
    for k, P in enumerate(r.parts):
        Q = prob.export_part(k)
        assert rel_err(Q["val"], P["val"]) <= TOL and rel_err(Q["b"], P["b"]) <= TOL, k
    prob.close()


def test_many_fourier_modes(gpu_ctx):
    """37 modes: the cell kernel stages the first 32 in LDS and reads the rest
    from HBM, the two lanes of a Gauss point continuing their alternation"""
    rng = np.random.default_rng(17)
    modes = np.column_stack([rng.integers(0, 4, (37, 3)), rng.uniform(-0.08, 0.08, 37),
                             rng.uniform(-np.pi, np.pi, 37)]).astype(np.float64)
    p = case(3, 2, 4, (2, 1, 1), modes=modes, rhs="manufactured", uD=(0.5, 1.0, -2.0, 0.7))
    p["vertex_coords"] = mapped_vertices(p, amp=0.4, seed=8)
    print("mapped case 37 modes", flush=True)
    r = oracle.run(p)
    prob = gpu_problem(gpu_ctx, p)
    for k, P in enumerate(r.parts):
        Q = prob.export_part(k)
        assert rel_err(Q["val"], P["val"]) <= TOL and rel_err(Q["b"], P["b"]) <= TOL, k
    prob.close()


def test_new_symbolic_phase_returns_to_the_uniform_box(gpu_ctx):
    """a new symbolic phase drops the vertex coordinates and a cell-wise kappa
    with the other device buffers: the uniform form with the descriptor's kappa"""
    p = case(3, 2, 4, (2, 1, 1), rhs_const=1.0)
    ref = values(gpu_problem(gpu_ctx, p, coords=False))
    prob = gpu_problem(gpu_ctx, dict(p, vertex_coords=mapped_vertices(p, amp=0.5, seed=1)))
    assert not same(ref, values(prob))
    prob.set_cell_coefficient(np.full(int(np.prod(p["cells"])), 2.0))  # dropped with the buffers too
    prob.symbolic()
    prob.numeric()
    assert same(ref, values(prob))
    prob.close()


def test_patch_test_and_volume_at_32(gpu_ctx):
    uD = (0.4, 1.3, -0.6, 0.9)
    p = case(3, 2, 32, (2, 2, 2), coef_const=2.3, rhs_const=0.0, uD=uD)
    p["vertex_coords"] = mapped_vertices(p, amp=0.5, seed=7, smooth=0.05)
    print("mapped patch test", p["cells"], flush=True)
    prob = gpu_problem(gpu_ctx, p)
    parts = [prob.export_part(k) for k in range(prob.nparts)]
    A, b = owned_global(parts, prob.ngdofs)
    g2n = np.zeros(prob.ngdofs, np.int64)
    for P in parts:
        g2n[P["ldof_gid"]] = P["ldof_gnode"]
    x = micro.mapped_node_coords(p, g2n)
    u = uD[0] + x @ np.asarray(uD[1:])
    assert np.abs(A @ u - b).max() <= 1e-11 * np.abs(b).max()
    prob.close()
    q = dict(p, dirichlet=False, rhs_const=1.5)
    prob = gpu_problem(gpu_ctx, q)
    _, b = owned_global([prob.export_part(k) for k in range(prob.nparts)], prob.ngdofs)
    assert abs(b.sum() - 1.5) <= 1e-12
    prob.close()


def test_graph_replays_match_eager(gpu_ctx):
    p = case(3, 2, 6, (2, 2, 1), modes=MODES, rhs="manufactured", uD=(0.5, 1.0, -2.0, 0.7))
    p["vertex_coords"] = mapped_vertices(p, amp=0.5, seed=4)
    prob = gpu_problem(gpu_ctx, p)
    eager = values(prob)
    batch = AssemblyBatch(gpu_ctx)
    batch.add(prob)
    batch.autotune(1)
    batch.capture()
    for rep in range(3):
        batch.launch()
        batch.sync()
        assert same(eager, values(prob)), rep
        prob.numeric()  # eager work between replays
        prob.sync()
    prob.set_vertex_coords(None)  # a mode change drops the graph
    with pytest.raises(GdxError):
        batch.launch()
    batch.close()
    prob.close()


def test_refusals(gpu_ctx):
    p = case(3, 1, 4, (1, 1, 1), rhs_const=1.0)
    p["vertex_coords"] = mapped_vertices(p, amp=0.3, seed=0)
    prob = gpu_problem(gpu_ctx, p)
    with pytest.raises(GdxError) as e:
        prob.numeric(PREC_FP32)
    assert e.value.code == -4
    bad = p["vertex_coords"].copy()
    bad[7, 1] = np.nan
    with pytest.raises(GdxError) as e:
        prob.set_vertex_coords(bad)
    assert e.value.code == -1
    prob.close()
    el = W.elasticity3d_q2(cells=(2, 2, 2), parts=(1, 1, 1))
    assem, data = problem_from_workload(gpu_ctx, el)
    pe = assem._problem_for(data)
    with pytest.raises(GdxError) as e:
        pe.set_vertex_coords(micro.uniform_vertices(el))
    assert e.value.code == -4
    pe.close()


def test_solvers_on_a_mapped_grid(gpu_ctx):
    """The assembled mapped-grid system through the device solvers: Jacobi-PCG
    and FPCG with the multilevel preconditioner (Galerkin levels of this matrix)
    reach 1e-8 and agree with a direct solve of the exported matrix."""
    import scipy.sparse.linalg as spla

    from gridap_amd import LearnedCoarseSolver, fpcg

    p = case(3, 1, 16, (2, 2, 1), modes=MODES, rhs="manufactured", uD=(0.5, 1.0, -2.0, 0.7))
    p["vertex_coords"] = mapped_vertices(p, amp=0.5, seed=9, smooth=0.05)
    prob = gpu_problem(gpu_ctx, p)
    A, b = owned_global([prob.export_part(k) for k in range(prob.nparts)], prob.ngdofs)
    xref = spla.spsolve(A.tocsc(), b)
    x, it, rr, _ = prob.pcg(1e-10, 5000)
    assert rr <= 1e-10 and np.linalg.norm(x - xref) <= 1e-7 * np.linalg.norm(xref)
    net = LearnedCoarseSolver(gpu_ctx, 8, 0)
    net.attach(prob)
    net.set_options(8, use_network=False)
    xf, itf, rf, _, fb = fpcg(prob, net, None, 1e-10, 2000)
    print("mapped 16^3 Q1: Jacobi-PCG", it, "multilevel FPCG", itf, flush=True)
    assert rf <= 1e-10 and not fb and itf < it
    assert np.linalg.norm(xf - xref) <= 1e-7 * np.linalg.norm(xref)
    net.close()
    prob.close()
