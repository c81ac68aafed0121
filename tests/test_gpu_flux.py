"""Parity of the nonlinear flux operator (GDX_OP_FLUX) and of Newton's method
on device 0 with the CPU oracle.

Bar, as for the affine path: structure bit-identical, Jacobian values and
residual within 1e-12 relative in FP64, ghost rows zero, repeated assemblies
bit-identical.  Newton (test/PLaplacianTests.jl:54-62): the reference test's
problem recovers u = x + y to 1e-9; the p-Laplacian's returned state has an
oracle-recomputed residual below the requested tolerance.
"""
import numpy as np
import pytest

import oracle
from _helpers import STRUCT_KEYS, owned_global, rel_err
from gridap_amd import (FESolver, NLSolver, flux_operator_from_workload, residual_and_jacobian,
                        residual_and_jacobian_, solve, zero)
from gridap_amd.nonlinear import FEFunction
from gridap_amd import workloads as W
from gridap_amd._lib import NEWTON_INIT_POISSON

pytestmark = pytest.mark.gpu

TOL = 1e-12


def flux3d(n, parts, p=3.0, alpha=0.05, beta=1.0, f=1.0, strategy="sub", uD=(0.3, 0.5, -0.2, 0.4)):
    q = W.plaplacian3d_q1(n, parts, p=p, eps=0.0, strategy=strategy)
    q["cells"] = tuple(n) if not np.isscalar(n) else (n, n, n)
    q["flux"] = (p, alpha, beta)
    q["rhs_const"] = f
    q["uD"] = uD
    return q


def flux2d(n, parts, p=4.0, alpha=1.0, beta=1.0, f=0.5, strategy="sub", qdeg=2):
    q = W.flux2d_reference_test(parts, strategy)
    q["cells"] = (n, n) if np.isscalar(n) else tuple(n)
    q["flux"] = (p, alpha, beta)
    q["rhs_const"] = f
    q["uD"] = (0.1, 0.3, -0.2, 0.0)
    q["qdeg"] = qdeg
    return q


CASES = [
    flux3d(4, (2, 2, 1)),
    flux3d((5, 4, 6), (2, 1, 3), p=4.0, alpha=1.0),
    flux3d(5, (2, 2, 2), strategy="full"),
    flux3d(6, (3, 2, 1), p=2.5, alpha=0.3, beta=0.7),
    dict(flux3d(4, (1, 1, 1)), qdeg=4),
    flux2d(6, (2, 2)),
    flux2d(9, (3, 2), p=3.0, alpha=0.01, strategy="full"),
    flux2d(7, (1, 2), p=2.0, alpha=1.0, beta=0.0),
]


def smooth_state(q, seed):
    rng = np.random.default_rng(seed)
    x = W.lattice_coords(q)
    a = rng.normal(size=(3, x.shape[1]))
    return 0.2 * np.sin(x @ a[0] * 2.0 + 1.0) + 0.3 * np.cos(x @ a[1]) + 0.1 * (x @ a[2])


def gpu_flux(ctx, q, u_lattice):
    op = flux_operator_from_workload(ctx, q)
    prob = op._problem()
    parts0 = [prob.export_part(k) for k in range(prob.nparts)]
    g2n = np.full(prob.ngdofs, -1, np.int64)
    for P in parts0:
        g2n[P["ldof_gid"]] = P["ldof_gnode"]
    uh = FEFunction(op.trial, u_lattice[g2n])
    b, A = residual_and_jacobian(op, uh)
    parts = [prob.export_part(k) for k in range(prob.nparts)]
    return op, prob, uh, parts, b, A


@pytest.mark.parametrize("q", CASES, ids=lambda q: f"{q['dim']}d-{q['cells']}-{q['parts']}-{q['strategy']}-"
                                                   f"p{q['flux'][0]}-qd{q['qdeg']}")
def test_flux_parity_with_oracle(gpu_ctx, q):
    print("flux case", q["dim"], q["cells"], q["parts"], q["strategy"], q["flux"], q["qdeg"], flush=True)
    u = smooth_state(q, 11)
    qs = dict(q)
    qs["state"] = u
    r = oracle.run(qs)
    op, prob, uh, parts, b, A = gpu_flux(gpu_ctx, q, u)
    assert prob.ngdofs == r.ngdofs
    for k, (P, Q) in enumerate(zip(r.parts, parts)):
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
        assert rel_err(Q["val"], P["val"]) <= TOL, k
        assert rel_err(Q["b"], P["b"]) <= TOL, k
        no = P["n_own"]
        assert np.all(Q["val"][Q["rowptr"][no]:] == 0.0) and np.all(Q["b"][no:] == 0.0)


def test_inplace_equals_fresh_and_repeatable(gpu_ctx):
    """residual_and_jacobian! equals residual_and_jacobian (test/PLaplacianTests.jl:44-52)."""
    q = flux3d(6, (2, 2, 2))
    u = smooth_state(q, 3)
    op, prob, uh, parts, b, A = gpu_flux(gpu_ctx, q, u)
    b2 = b.copy()
    A2 = A.copy()
    for m in A2.matrix_partition:
        m.data[:] = 0.0
    residual_and_jacobian_(b2, A2, op, uh)
    for k in range(prob.nparts):
        assert np.array_equal(A.matrix_partition[k].data, A2.matrix_partition[k].data)
        assert np.array_equal(b.partition[k], b2.partition[k])
    x = np.ones(prob.ngdofs)
    assert np.array_equal(prob.spmv(x), prob.spmv(x))


@pytest.mark.parametrize("strategy", ["sub", "full"])
def test_newton_reference_problem(gpu_ctx, strategy):
    """test/PLaplacianTests.jl:20-62: sigma = (1+|g|^2) g, u = x + y on (0,4)^2, 4x4 cells."""
    q = W.flux2d_reference_test((2, 2), strategy)
    op = flux_operator_from_workload(gpu_ctx, q)
    nls = NLSolver(method="newton", ftol=1e-14, atol=1e-13, iterations=20, linear_rtol=1e-14)
    uh = solve(FESolver(nls), op)
    prob = op._problem()
    g2n = np.full(prob.ngdofs, -1, np.int64)
    for k in range(prob.nparts):
        P = prob.export_part(k)
        g2n[P["ldof_gid"]] = P["ldof_gnode"]
    x = W.lattice_coords(q)[g2n]
    err = uh.free_values - (x[:, 0] + x[:, 1])
    assert np.abs(err).max() < 1e-9, nls.report
    assert nls.report["converged"]


def test_newton_plaplacian_residual_checked_by_oracle(gpu_ctx):
    """p = 3 regularised p-Laplacian, f = 1, u = 0 on the boundary: the state
    Newton returns has an oracle residual below 1e-10 of the initial one."""
    q = W.plaplacian3d_q1(12, (4, 2, 2), p=3.0, eps=1e-2)
    op = flux_operator_from_workload(gpu_ctx, q)
    nls = NLSolver(method="newton", ftol=1e-10, iterations=30, linear_rtol=1e-11, initial_guess="poisson")
    uh = solve(FESolver(nls), op)
    rep = nls.report
    print("newton report", rep, flush=True)
    assert rep["converged"] and rep["res_final"] <= 1e-10 * rep["res_ref"]
    prob = op._problem()
    g2n = np.full(prob.ngdofs, -1, np.int64)
    for k in range(prob.nparts):
        P = prob.export_part(k)
        g2n[P["ldof_gid"]] = P["ldof_gnode"]
    u = np.zeros(int(np.prod(W.lattice_shape(q))))
    u[g2n] = uh.free_values
    q0, q1 = dict(q), dict(q)
    q0["state"], q1["state"] = np.zeros_like(u), u
    _, r0 = owned_global(oracle.run(q0).parts, prob.ngdofs)
    _, r1 = owned_global(oracle.run(q1).parts, prob.ngdofs)
    assert abs(np.linalg.norm(r0) - rep["res_ref"]) <= 1e-12 * np.linalg.norm(r0)
    assert np.linalg.norm(r1) <= 1e-10 * np.linalg.norm(r0)
    assert np.all(uh.free_values > 0.0)  # f = 1 > 0 with u = 0 on the boundary: maximum principle


@pytest.mark.parametrize("use_network", [False, True], ids=["two-grid", "two-grid+network"])
def test_newton_with_two_grid_fpcg_corrections(gpu_ctx, use_network):
    """Newton whose corrections are solved by FPCG with the two-grid
    preconditioner (rebuilt from every Jacobian) reaches the same solution as
    Newton with Jacobi-PCG corrections, with fewer linear iterations when the
    coarse network is off; the returned state's residual is checked by the
    oracle."""
    from gridap_amd import LearnedCoarseSolver

    cells = 32
    q = W.plaplacian3d_q1(cells, (2, 2, 2), p=3.0, eps=1e-2)
    op = flux_operator_from_workload(gpu_ctx, q)
    prob = op._problem()
    kw = dict(rtol=1e-10, maxiter=30, lin_rtol=1e-10, init=NEWTON_INIT_POISSON)
    x_j, rep_j = prob.newton(None, **kw)
    net = LearnedCoarseSolver(gpu_ctx, cells // 2, 3)
    net.attach(prob)
    net.set_options(8, use_network=use_network)
    x_t, rep_t = prob.newton(None, precond=net, **kw)
    print("jacobi", rep_j, "\ntwo-grid", rep_t, flush=True)
    assert rep_j["converged"] and rep_t["converged"]
    assert rep_t["lin_fallbacks"] == 0 or use_network
    assert np.linalg.norm(x_t - x_j) <= 1e-8 * np.linalg.norm(x_j)
    if not use_network:
        assert rep_t["lin_iters"] < rep_j["lin_iters"]
    g2n = np.full(prob.ngdofs, -1, np.int64)
    for k in range(prob.nparts):
        P = prob.export_part(k)
        g2n[P["ldof_gid"]] = P["ldof_gnode"]
    u = np.zeros(int(np.prod(W.lattice_shape(q))))
    u[g2n] = x_t
    q0, q1 = dict(q), dict(q)
    q0["state"], q1["state"] = np.zeros_like(u), u
    _, r0 = owned_global(oracle.run(q0).parts, prob.ngdofs)
    _, r1 = owned_global(oracle.run(q1).parts, prob.ngdofs)
    assert np.linalg.norm(r1) <= 1e-10 * np.linalg.norm(r0)
    net.close()


MIXED_CASES = [flux3d(6, (2, 2, 1)), flux3d((9, 5, 7), (2, 1, 3), p=4.0, alpha=1.0),
               flux3d(5, (2, 2, 2), strategy="full"), flux3d(8, (2, 2, 2), p=2.5, alpha=0.3, beta=0.7)]


@pytest.mark.parametrize("q", MIXED_CASES, ids=lambda q: f"{q['cells']}-{q['parts']}-{q['strategy']}-p{q['flux'][0]}")
def test_flux_mixed_jacobian(gpu_ctx, q):
    """gdx_numeric(GDX_PREC_FP32) on a flux problem: cell Jacobians in FP32,
    rows summed in FP64.  Sparsity and the FP64 residual are bit-identical to
    the FP64 form; the Jacobian is within 1e-6 of the oracle's (relative to
    its largest entry)."""
    from gridap_amd._lib import PREC_FP32

    print("mixed flux case", q["cells"], q["parts"], q["strategy"], q["flux"], flush=True)
    u = smooth_state(q, 5)
    qs = dict(q)
    qs["state"] = u
    r = oracle.run(qs)
    op, prob, uh, parts64, b, A = gpu_flux(gpu_ctx, q, u)
    prob.numeric(PREC_FP32)
    prob.sync()
    parts32 = [prob.export_part(k) for k in range(prob.nparts)]
    worst = 0.0
    for k, (P, Q, O) in enumerate(zip(parts64, parts32, r.parts)):
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
        assert np.array_equal(P["b"].view(np.int64), Q["b"].view(np.int64)), k
        worst = max(worst, rel_err(Q["val"], O["val"]))
        assert rel_err(P["val"], O["val"]) <= TOL, k
    print("mixed Jacobian max rel error", worst, flush=True)
    assert 0.0 < worst <= 1e-6
    prob.numeric()  # back to FP64: the same values as before, bit for bit
    prob.sync()
    for P, Q in zip(parts64, (prob.export_part(k) for k in range(prob.nparts))):
        assert np.array_equal(P["val"].view(np.int64), Q["val"].view(np.int64))


def test_flux_mixed_jacobian_unsupported_forms(gpu_ctx):
    """The mixed form lives in the fused 3D kernel: 2D and 3 Gauss points per
    direction are refused (GDX_ERR_UNSUPPORTED), for gdx_numeric and Newton."""
    from gridap_amd._lib import GdxError, PREC_FP32

    for q in (flux2d(6, (2, 2)), dict(flux3d(4, (1, 1, 1)), qdeg=4)):
        op = flux_operator_from_workload(gpu_ctx, q)
        prob = op._problem()
        with pytest.raises(GdxError) as e:
            prob.numeric(PREC_FP32)
        assert e.value.code == -4
        with pytest.raises(GdxError) as e:
            prob.newton(None, jac_precision=PREC_FP32)
        assert e.value.code == -4
        prob.close()


@pytest.mark.parametrize("multilevel", [False, True], ids=["jacobi-pcg", "multilevel-fpcg"])
def test_newton_with_mixed_jacobians(gpu_ctx, multilevel):
    """Newton with FP32-contracted Jacobians (residual, test and solves FP64)
    reaches the FP64 Newton's solution within 1e-8 in at most one more step;
    the returned state's residual is checked by the oracle."""
    from gridap_amd import LearnedCoarseSolver
    from gridap_amd._lib import PREC_FP32

    cells = 32
    q = W.plaplacian3d_q1(cells, (2, 2, 2), p=3.0, eps=1e-2)
    op = flux_operator_from_workload(gpu_ctx, q)
    prob = op._problem()
    net = None
    if multilevel:
        net = LearnedCoarseSolver(gpu_ctx, cells // 2, 3)
        net.attach(prob)
        net.set_options(8, use_network=False)
    kw = dict(rtol=1e-10, maxiter=30, lin_rtol=1e-10, init=NEWTON_INIT_POISSON, precond=net)
    x64, rep64 = prob.newton(None, **kw)
    x32, rep32 = prob.newton(None, jac_precision=PREC_FP32, **kw)
    print("fp64", rep64, "\nfp32 jacobians", rep32, flush=True)
    assert rep64["converged"] and rep32["converged"]
    assert rep32["iters"] <= rep64["iters"] + 1
    assert np.linalg.norm(x32 - x64) <= 1e-8 * np.linalg.norm(x64)
    g2n = np.full(prob.ngdofs, -1, np.int64)
    for k in range(prob.nparts):
        P = prob.export_part(k)
        g2n[P["ldof_gid"]] = P["ldof_gnode"]
    u = np.zeros(int(np.prod(W.lattice_shape(q))))
    u[g2n] = x32
    q0, q1 = dict(q), dict(q)
    q0["state"], q1["state"] = np.zeros_like(u), u
    _, r0 = owned_global(oracle.run(q0).parts, prob.ngdofs)
    _, r1 = owned_global(oracle.run(q1).parts, prob.ngdofs)
    assert np.linalg.norm(r1) <= 1e-10 * np.linalg.norm(r0)
    if net is not None:
        net.close()
    prob.close()
