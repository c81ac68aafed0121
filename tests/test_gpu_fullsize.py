"""The configurations the bench measures, at their full sizes, on device 0.

The C oracle is too slow to assemble these whole, so the checks are ones that
hold at any size, plus sampled rows against an independent restatement and,
for the nonlinear solve, one oracle residual:

- linear elasticity Q2, 96 x 80 x 72 cells, parts (3,2,2), clamped at x = 0:
  free dof count; the rigid-body modes in the kernel of every row away from the
  clamp; the energy of a linear field that vanishes on the clamp,
  |Omega| (lambda tr(e)^2 + 2 mu e:e); sampled rows (extracted with the device
  SpMV) against the Kronecker restatement of the 1D assembled factors;
- the same with a lognormal Young's modulus field (kappa_c scaling lambda and
  mu on each cell, drawn on the device): rigid-body modes in the kernel away
  from the clamp and the linear field's energy
  sum_c |cell| kappa_c (lambda tr(e)^2 + 2 mu e:e) with the exported field;
- learned-preconditioner workload, Q1 128^3, lognormal kappa (held-out field):
  FPCG with the multilevel preconditioner and Jacobi-PCG both reach the FP64 true
  residual 1e-8, FPCG in at most half the Jacobi-PCG time (the workload's
  target); nnz and free dofs in closed form;
- p-Laplacian Q1 96^3, parts (4,2,2): Newton with multilevel FPCG corrections
  and with Jacobi-PCG corrections reach 1e-10 and agree; the oracle's residual
  at the returned state is below 1e-10 of the initial one.
"""
import numpy as np
import pytest

import micro
import oracle
from _helpers import owned_global
from gridap_amd import LearnedCoarseSolver, flux_operator_from_workload, fpcg, problem_from_workload
from gridap_amd import workloads as W
from gridap_amd._lib import NEWTON_INIT_POISSON

pytestmark = pytest.mark.gpu


def gid_map(prob):
    """global id -> lattice node of every free dof (and the component, by order)"""
    g2n = np.full(prob.ngdofs, -1, np.int64)
    comp = np.zeros(prob.ngdofs, np.int64)
    for k in range(prob.nparts):
        gid, _, gnode = prob.export_fe(k)
        c = np.zeros(len(gid), np.int64)
        if len(gid) > 1:  # the components of a node are consecutive local dofs
            same = np.concatenate([[False], gnode[1:] == gnode[:-1]])
            run = np.cumsum(~same) - 1
            first = np.flatnonzero(~same)
            c = np.arange(len(gid)) - first[run]
        g2n[gid] = gnode
        comp[gid] = c
    return g2n, comp


def elasticity_row(F, lam, mu, nn, node, a):
    """Row (node, component a) of the unconstrained elasticity matrix over the
    whole lattice from the 1D assembled factors F[d] = (K, M, P, Q, m):
    {(node j, component b): value} for the nonzero entries."""
    i = (node % nn[0], (node // nn[0]) % nn[1], node // (nn[0] * nn[1]))
    idx = {"K": 0, "M": 1, "P": 2, "Q": 3}
    out = {}
    for b in range(3):
        terms = []
        if a == b:
            for e in range(3):
                kinds = ["M", "M", "M"]
                kinds[e] = "K"
                terms.append((mu + (lam + mu if e == a else 0.0), kinds))
        else:
            k1 = ["M", "M", "M"]
            k1[a], k1[b] = "P", "Q"
            k2 = ["M", "M", "M"]
            k2[a], k2[b] = "Q", "P"
            terms += [(lam, k1), (mu, k2)]
        rows = [[F[d][idx[kk]][i[d]] for kk in "KMPQ"] for d in range(3)]
        js = [np.flatnonzero(np.abs(np.stack(rows[d])).sum(axis=0) > 0) for d in range(3)]
        for jz in js[2]:
            for jy in js[1]:
                for jx in js[0]:
                    v = 0.0
                    for c, kinds in terms:
                        v += c * (F[0][idx[kinds[0]]][i[0], jx] * F[1][idx[kinds[1]]][i[1], jy] *
                                  F[2][idx[kinds[2]]][i[2], jz])
                    if v != 0.0:
                        out[(int(jx + nn[0] * (jy + nn[1] * jz)), b)] = v
    return out


def test_c3_elasticity_full_size(gpu_ctx):
    p = W.elasticity3d_q2()  # 96 x 80 x 72 cells, parts (3,2,2), E = 1, nu = 0.3, clamped at x = 0
    print("full-size elasticity", p["cells"], p["parts"], flush=True)
    lam, mu = p["lame"]
    assem, data = problem_from_workload(gpu_ctx, p)
    prob = assem._problem_for(data)
    prob.numeric()
    prob.sync()
    nn = W.lattice_shape(p)
    assert prob.ngdofs == 3 * (nn[0] - 1) * nn[1] * nn[2]
    g2n, comp = gid_map(prob)
    X = micro.node_coords(p, g2n)[1]
    away = (g2n % nn[0]) >= 3  # rows whose stencil does not reach the clamped nodes (Q2: 2 nodes)
    scale = np.abs(prob.spmv(np.where(comp == 0, X[:, 0], 0.0))).max()
    for a in range(3):  # translations
        r = prob.spmv(np.where(comp == a, 1.0, 0.0))
        assert np.abs(r[away]).max() <= 1e-12 * scale
    for i, j in ((0, 1), (1, 2), (0, 2)):  # rotations
        rot = np.where(comp == i, -X[:, j], 0.0) + np.where(comp == j, X[:, i], 0.0)
        assert np.abs(prob.spmv(rot)[away]).max() <= 1e-12 * scale
    # energy of u = G x with only G[:, 0] nonzero: zero on the clamp x = 0
    g = np.array([0.3, -0.5, 0.7])
    u = g[comp] * X[:, 0]
    e = np.zeros((3, 3))
    e[0, 0] = g[0]
    e[0, 1] = e[1, 0] = g[1] / 2
    e[0, 2] = e[2, 0] = g[2] / 2
    energy = lam * np.trace(e) ** 2 + 2.0 * mu * np.sum(e * e)  # |Omega| = 1
    assert abs(u @ prob.spmv(u) - energy) <= 1e-10 * energy
    # sampled rows: A e_i (the matrix is symmetric) against the restatement
    F = [micro.assembled_factors_1d(p, d) for d in range(3)]
    rng = np.random.default_rng(11)
    picks = list(rng.choice(prob.ngdofs, 12, replace=False))
    for gi in picks:
        ei = np.zeros(prob.ngdofs)
        ei[gi] = 1.0
        col = prob.spmv(ei)
        nz = np.flatnonzero(col)
        got = {(int(g2n[j]), int(comp[j])): col[j] for j in nz}
        ref = elasticity_row(F, lam, mu, nn, int(g2n[gi]), int(comp[gi]))
        ref = {k: v for k, v in ref.items() if k[0] % nn[0] != 0}  # clamped columns are not stored
        sc = max(abs(v) for v in ref.values())
        for k in set(ref) | set(got):  # entries the restatement finds zero must be zero to rounding
            assert abs(got.get(k, 0.0) - ref.get(k, 0.0)) <= 1e-12 * sc, (gi, k)
    prob.close()


def test_c3_elasticity_cellwise_full_size(gpu_ctx):
    p = W.elasticity3d_q2()
    print("full-size elasticity, cell-wise Young's modulus", p["cells"], p["parts"], flush=True)
    lam, mu = p["lame"]
    assem, data = problem_from_workload(gpu_ctx, p)
    prob = assem._problem_for(data)
    prob.cell_coefficient_grf(31, 0.1, 1.0)
    kap = prob.export_cell_coefficient(int(np.prod(p["cells"])))
    assert kap.min() > 0.0 and kap.max() / kap.min() > 10.0
    prob.numeric()
    prob.sync()
    nn = W.lattice_shape(p)
    g2n, comp = gid_map(prob)
    X = micro.node_coords(p, g2n)[1]
    away = (g2n % nn[0]) >= 3
    scale = np.abs(prob.spmv(np.where(comp == 0, X[:, 0], 0.0))).max()
    for a in range(3):
        assert np.abs(prob.spmv(np.where(comp == a, 1.0, 0.0))[away]).max() <= 1e-12 * scale
    for i, j in ((0, 1), (1, 2), (0, 2)):
        rot = np.where(comp == i, -X[:, j], 0.0) + np.where(comp == j, X[:, i], 0.0)
        assert np.abs(prob.spmv(rot)[away]).max() <= 1e-12 * scale
    g = np.array([0.3, -0.5, 0.7])
    u = g[comp] * X[:, 0]
    e = np.zeros((3, 3))
    e[0, 0] = g[0]
    e[0, 1] = e[1, 0] = g[1] / 2
    e[0, 2] = e[2, 0] = g[2] / 2
    energy = (lam * np.trace(e) ** 2 + 2.0 * mu * np.sum(e * e)) * kap.sum() / len(kap)  # |Omega| = 1
    assert abs(u @ prob.spmv(u) - energy) <= 1e-10 * energy
    prob.close()


def test_c4_learned_fpcg_full_size(gpu_ctx):
    cells = 128
    p = W.poisson3d_q1(cells, (2, 2, 2), rhs_const=0.0)
    print("full-size learned FPCG", cells, flush=True)
    assem, data = problem_from_workload(gpu_ctx, p)
    prob = assem._problem_for(data)
    assert prob.owned_nnz() == W.nnz_free(p) == 379 ** 3 and prob.ngdofs == 127 ** 3
    prob.cell_coefficient_grf(900000, 0.1, 2.0)  # a held-out field of the workload
    prob.numeric()
    prob.sync()
    net = LearnedCoarseSolver(gpu_ctx, cells // 2, 1)
    net.attach(prob)
    net.set_multigrid(True, 1)
    net.set_options(8, use_network=False)
    rhs = np.random.default_rng(5000).normal(size=prob.ngdofs)
    fpcg(prob, None, rhs, 1e-8, 50000)  # first solves build their graphs and buffers
    fpcg(prob, net, rhs, 1e-8, 5000)
    xj, itj, rj, msj, _ = fpcg(prob, None, rhs, 1e-8, 50000)
    xl, itl, rl, msl, fb = fpcg(prob, net, rhs, 1e-8, 5000)
    print("jacobi", itj, rj, msj, "multilevel", itl, rl, msl, flush=True)
    assert rj <= 1e-8 and rl <= 1e-8 and not fb
    assert msl <= 0.5 * msj
    assert np.linalg.norm(xl - xj) <= 1e-6 * np.linalg.norm(xj)
    net.close()
    prob.close()


def test_c5_newton_full_size(gpu_ctx):
    q = W.plaplacian3d_q1(96, (4, 2, 2), p=3.0, eps=1e-2)
    print("full-size p-Laplacian Newton", q["cells"], q["parts"], flush=True)
    op = flux_operator_from_workload(gpu_ctx, q)
    prob = op._problem()
    assert prob.owned_nnz() == W.nnz_free(q) == 283 ** 3
    net = LearnedCoarseSolver(gpu_ctx, 48, 0)
    net.attach(prob)
    net.set_options(8, use_network=False)
    kw = dict(rtol=1e-10, maxiter=40, lin_rtol=1e-8, lin_maxiter=20000, init=NEWTON_INIT_POISSON)
    x_ml, rep_ml = prob.newton(None, precond=net, **kw)
    x_j, rep_j = prob.newton(None, **kw)
    print("newton", rep_ml, rep_j, flush=True)
    assert rep_ml["converged"] and rep_j["converged"]
    assert rep_ml["res_final"] <= 1e-10 * rep_ml["res_ref"]
    assert np.linalg.norm(x_ml - x_j) <= 1e-7 * np.linalg.norm(x_j)
    # the oracle's residual at the returned state (and at zero, the reference)
    g2n, _ = gid_map(prob)
    u = np.zeros(int(np.prod(W.lattice_shape(q))))
    u[g2n] = x_ml
    q1 = dict(q, state=u)
    _, r1 = owned_global(oracle.run(q1).parts, prob.ngdofs)
    assert np.linalg.norm(r1) <= 1e-10 * rep_ml["res_ref"]
    assert np.all(x_ml > 0.0)  # f = 1 > 0, u = 0 on the boundary
    net.close()
    prob.close()
