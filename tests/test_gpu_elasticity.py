"""Linear elasticity (vector-valued Q1/Q2) on device 0.

- parity with the CPU oracle: structure bit-identical (component-interleaved
  ghost columns included), matrix and body-force vector within 1e-12 relative,
  ghost rows zero;
- at a larger size, where the oracle is slow, size-independent properties
  through the device SpMV: the six rigid-body modes in the kernel of the
  unconstrained matrix and the energy of a linear field
  |Omega| (lambda tr(e)^2 + 2 mu e:e), plus sampled entries against the numpy
  Kronecker restatement;
- repeated assembly is bit-identical; Jacobi-PCG solves the clamped system
  (checked against scipy on the oracle's matrix);
- an elasticity problem and Q1 Poisson problems replayed in one hipGraph give
  the eager result;
- a piecewise-constant material ((lambda_c, mu_c) per cell, or kappa_c scaling
  the constants): parity with the oracle given the same pairs, within 1e-12;
  the kappa form is bit-identical to the pairs (lambda kappa_c, mu kappa_c);
  clearing it returns the constant form bit for bit; a graph captured in the
  cell-wise mode replays what eager assembly writes, and a mode change drops it.
"""
import numpy as np
import pytest

import micro
import oracle
from _helpers import STRUCT_KEYS, owned_global, rel_err
from gridap_amd import AssemblyBatch, problem_from_workload
from gridap_amd._lib import GdxError
from gridap_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-12
DOM = (0.0, 1.5, 0.0, 1.0, 0.0, 0.7)

CASES = [
    dict(cells=(3, 2, 2), parts=(2, 1, 2), clamped=True, order=2, strategy="sub"),
    dict(cells=(6, 5, 4), parts=(3, 2, 2), clamped=True, order=2, strategy="sub"),
    dict(cells=(5, 4, 4), parts=(2, 2, 1), clamped=True, order=2, strategy="full"),
    dict(cells=(4, 3, 3), parts=(1, 1, 1), clamped=False, order=2, strategy="sub"),
    dict(cells=(7, 5, 6), parts=(2, 1, 3), clamped=True, order=1, strategy="sub"),
    dict(cells=(5, 6, 4), parts=(2, 2, 2), clamped=False, order=1, strategy="full"),
]


def elast(case, **kw):
    return W.elasticity3d_q2(domain=DOM, E=2.5, nu=0.27, body_force=(0.3, -0.2, -1.0), **case, **kw)


def gpu_problem(ctx, p):
    assem, data = problem_from_workload(ctx, p)
    prob = assem._problem_for(data)
    prob.numeric()
    prob.sync()
    return assem, prob


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"q{c['order']}-{c['cells']}-{c['parts']}-{c['strategy']}"
                         f"-{'clamped' if c['clamped'] else 'free'}")
def test_parity_with_oracle(gpu_ctx, case):
    print("elasticity case", case, flush=True)
    p = elast(case)
    r = oracle.run(p)
    assem, prob = gpu_problem(gpu_ctx, p)
    assert prob.ngdofs == r.ngdofs
    for k, P in enumerate(r.parts):
        Q = prob.export_part(k)
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
        assert rel_err(Q["val"], P["val"]) <= TOL, k
        assert rel_err(Q["b"], P["b"]) <= TOL, k
        no = P["n_own"]
        assert np.all(Q["val"][Q["rowptr"][no]:] == 0.0) and np.all(Q["b"][no:] == 0.0)
    prob.close()


def latdof(prob):
    m = np.zeros(prob.ngdofs, np.int64)
    for k in range(prob.nparts):
        P = prob.export_part(k)
        m[P["ldof_gid"]] = 3 * P["ldof_gnode"] + np.arange(P["n_ldofs"]) % 3
    return m


def test_invariants_through_device_spmv(gpu_ctx):
    lam, mu = W.lame_parameters(2.5, 0.27)
    p = elast(dict(cells=(24, 20, 18), parts=(3, 2, 2), clamped=False, order=2, strategy="sub"))
    assem, prob = gpu_problem(gpu_ctx, p)
    g2d = latdof(prob)
    node, comp = g2d // 3, g2d % 3
    _, X = micro.node_coords(p, node)
    u1 = np.where(comp == 0, 1.0, 0.0)
    Au1 = prob.spmv(u1)
    ones_scale = np.abs(prob.spmv(np.where(comp == 0, X[:, 0], 0.0))).max()
    assert np.abs(Au1).max() <= 1e-12 * ones_scale
    for i, j in ((0, 1), (1, 2), (0, 2)):
        rot = np.where(comp == i, -X[:, j], 0.0) + np.where(comp == j, X[:, i], 0.0)
        assert np.abs(prob.spmv(rot)).max() <= 1e-12 * ones_scale
    G = np.array([[0.3, -0.7, 0.2], [0.5, -0.1, 0.4], [-0.6, 0.9, 0.25]])
    u = np.einsum("nj,nj->n", G[comp], X)
    e = 0.5 * (G + G.T)
    energy = 1.5 * 1.0 * 0.7 * (lam * np.trace(e) ** 2 + 2.0 * mu * np.sum(e * e))
    assert abs(u @ prob.spmv(u) - energy) <= 1e-11 * energy
    assert prob.owned_nnz() == 9 * np.prod([8 * n + 1 for n in (24, 20, 18)])
    # sampled rows against the Kronecker restatement, keyed by lattice dof
    Am, bm = micro.elasticity_global(p)
    d2g = np.full(3 * int(np.prod(W.lattice_shape(p))), -1, np.int64)
    d2g[g2d] = np.arange(prob.ngdofs)
    rng = np.random.default_rng(3)
    parts = [prob.export_part(k) for k in range(prob.nparts)]
    for k in rng.choice(prob.nparts, 4, replace=False):
        P = parts[k]
        for row in rng.choice(P["n_own"], 20, replace=False):
            gi = P["row_gid"][row]
            cols = P["col_gid"][P["colidx"][P["rowptr"][row]:P["rowptr"][row + 1]]]
            vals = P["val"][P["rowptr"][row]:P["rowptr"][row + 1]]
            ref = Am[g2d[gi]].toarray().ravel()
            mine = np.zeros_like(ref)
            mine[g2d[cols]] = vals
            assert np.abs(mine - ref).max() <= TOL * np.abs(ref).max()
            assert abs(P["b"][row] - bm[g2d[gi]]) <= TOL * np.abs(bm).max()
    prob.close()


def test_repeatable_bitwise_and_solve(gpu_ctx):
    import scipy.sparse.linalg as spla

    p = elast(dict(cells=(8, 6, 5), parts=(2, 2, 1), clamped=True, order=2, strategy="sub"))
    assem, prob = gpu_problem(gpu_ctx, p)
    first = [prob.export_part(k) for k in range(prob.nparts)]
    prob.numeric()
    prob.sync()
    for k, P in enumerate(first):
        Q = prob.export_part(k)
        assert np.array_equal(P["val"].view(np.int64), Q["val"].view(np.int64))
        assert np.array_equal(P["b"].view(np.int64), Q["b"].view(np.int64))
    r = oracle.run(p)
    A, b = owned_global(r.parts, r.ngdofs)
    xref = spla.spsolve(A.tocsc(), b)
    x, it, rr, _ = prob.pcg(1e-10, 20000)
    print("elasticity Jacobi-PCG iterations", it, "relres", rr, flush=True)
    assert rr <= 1e-10
    assert np.linalg.norm(x - xref) <= 1e-7 * np.linalg.norm(xref)
    prob.close()


def test_batch_with_poisson_problems_matches_eager(gpu_ctx):
    """The batch's problems start with zero values (symbolic phase only); the
    graph replays must write what separate eager assemblies give, with eager
    work on the same stream between replays."""
    ws = (elast(dict(cells=(6, 5, 4), parts=(3, 2, 2), clamped=True, order=2, strategy="sub")),
          W.poisson3d_q1(14, (2, 2, 1)), W.poisson3d_q1(9, (1, 1, 1)))
    refs = []
    for p in ws:
        _, pr = gpu_problem(gpu_ctx, p)
        refs.append([(Q["val"].copy(), Q["b"].copy()) for Q in (pr.export_part(k) for k in range(pr.nparts))])
        pr.close()
    probs = []
    for p in ws:
        assem, data = problem_from_workload(gpu_ctx, p)
        probs.append(assem._problem_for(data))
    assert np.all(probs[0].export_part(0)["val"] == 0.0)
    batch = AssemblyBatch(gpu_ctx)
    for pr in probs:
        batch.add(pr)
    batch.autotune(2)
    batch.capture()
    for rep in range(3):
        batch.launch()
        batch.sync()
        for pr, ref in zip(probs, refs):
            for k, (v, bb) in enumerate(ref):
                Q = pr.export_part(k)
                assert np.array_equal(v.view(np.int64), Q["val"].view(np.int64)), (rep, k)
                assert np.array_equal(bb.view(np.int64), Q["b"].view(np.int64)), (rep, k)
        probs[rep % len(probs)].numeric()  # eager work between replays
        probs[rep % len(probs)].sync()
    batch.close()
    for pr in probs:
        pr.close()


def random_lame(cells, seed):
    rng = np.random.default_rng(seed)
    nc = int(np.prod(cells))
    return np.stack([rng.uniform(0.2, 3.0, nc), rng.uniform(0.1, 2.0, nc)], axis=1)


def parts_values(prob):
    prob.sync()
    return [(Q["val"].copy(), Q["b"].copy()) for Q in (prob.export_part(k) for k in range(prob.nparts))]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"q{c['order']}-{c['cells']}-{c['parts']}-{c['strategy']}"
                         f"-{'clamped' if c['clamped'] else 'free'}")
def test_cellwise_lame_parity_with_oracle(gpu_ctx, case):
    print("cell-wise elasticity case", case, flush=True)
    p = elast(case)
    lame = random_lame(case["cells"], 7)
    r = oracle.run(dict(p, cell_lame=lame))
    assem, prob = gpu_problem(gpu_ctx, p)
    const = parts_values(prob)
    prob.set_cell_lame(lame)
    prob.numeric()
    prob.sync()
    for k, P in enumerate(r.parts):
        Q = prob.export_part(k)
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
        assert rel_err(Q["val"], P["val"]) <= TOL, k
        assert rel_err(Q["b"], P["b"]) <= TOL, k
        no = P["n_own"]
        assert np.all(Q["val"][Q["rowptr"][no]:] == 0.0) and np.all(Q["b"][no:] == 0.0)
    # kappa scaling the constants: the pairs (lambda kappa_c, mu kappa_c), bit for bit
    kap = np.exp(np.random.default_rng(8).normal(size=len(lame)))
    prob.set_cell_coefficient(kap)
    prob.numeric()
    via_kappa = parts_values(prob)
    prob.set_cell_lame(np.stack([p["lame"][0] * kap, p["lame"][1] * kap], axis=1))
    prob.numeric()
    for (v0, b0), (v1, b1) in zip(via_kappa, parts_values(prob)):
        assert np.array_equal(v0.view(np.int64), v1.view(np.int64)) and np.array_equal(b0, b1)
    r2 = oracle.run(dict(p, cell_lame=np.stack([p["lame"][0] * kap, p["lame"][1] * kap], axis=1)))
    for P, (v, bb) in zip(r2.parts, via_kappa):
        assert rel_err(v, P["val"]) <= TOL
    # back to the constants
    prob.set_cell_lame(None)
    prob.numeric()
    for (v0, b0), (v1, b1) in zip(const, parts_values(prob)):
        assert np.array_equal(v0.view(np.int64), v1.view(np.int64)) and np.array_equal(b0, b1)
    prob.close()


def test_cellwise_lame_refusals(gpu_ctx):
    p = elast(dict(cells=(3, 2, 2), parts=(1, 1, 1), clamped=True, order=2, strategy="sub"))
    assem, prob = gpu_problem(gpu_ctx, p)
    for bad in ((1.0, 0.0), (-1.0, 1.0), (np.nan, 1.0), (1.0, np.inf)):
        lame = np.tile([1.0, 0.5], (12, 1))
        lame[5] = bad
        with pytest.raises(GdxError) as e:
            prob.set_cell_lame(lame)
        assert e.value.code == -1, bad
    prob.close()
    q = W.poisson3d_q1(4, (1, 1, 1))
    _, pq = gpu_problem(gpu_ctx, q)
    with pytest.raises(GdxError) as e:
        pq.set_cell_lame(np.ones((64, 2)))
    assert e.value.code == -4
    pq.close()


def test_cellwise_lame_graph_replays(gpu_ctx):
    case = dict(cells=(6, 5, 4), parts=(3, 2, 2), clamped=True, order=2, strategy="sub")
    p = elast(case)
    lames = [random_lame(case["cells"], s) for s in (1, 2)]
    _, prob = gpu_problem(gpu_ctx, p)
    eager = []
    for lame in lames:
        prob.set_cell_lame(lame)
        prob.numeric()
        eager.append(parts_values(prob))
    batch = AssemblyBatch(gpu_ctx)
    batch.add(prob)
    batch.autotune(1)
    batch.capture()
    for rep in range(4):
        prob.set_cell_lame(lames[rep % 2])  # same device buffer, new contents
        batch.launch()
        batch.sync()
        for (v0, b0), (v1, b1) in zip(eager[rep % 2], parts_values(prob)):
            assert np.array_equal(v0.view(np.int64), v1.view(np.int64)) and np.array_equal(b0, b1), rep
        prob.numeric()  # eager work between replays
        prob.sync()
    prob.set_cell_lame(None)  # a mode change drops the graph
    with pytest.raises(GdxError):
        batch.launch()
    batch.close()
    prob.close()


def test_cellwise_lame_solve(gpu_ctx):
    """Jacobi-PCG on the clamped system with a piecewise-constant material
    reaches 1e-10 and agrees with scipy's direct solve of the oracle's matrix."""
    import scipy.sparse.linalg as spla

    case = dict(cells=(8, 6, 5), parts=(2, 2, 1), clamped=True, order=2, strategy="sub")
    p = elast(case)
    lame = random_lame(case["cells"], 3)
    _, prob = gpu_problem(gpu_ctx, p)
    prob.set_cell_lame(lame)
    prob.numeric()
    prob.sync()
    r = oracle.run(dict(p, cell_lame=lame))
    A, b = owned_global(r.parts, r.ngdofs)
    xref = spla.spsolve(A.tocsc(), b)
    x, it, rr, _ = prob.pcg(1e-10, 20000)
    print("cell-wise elasticity Jacobi-PCG iterations", it, "relres", rr, flush=True)
    assert rr <= 1e-10
    assert np.linalg.norm(x - xref) <= 1e-7 * np.linalg.norm(xref)
    prob.close()
