"""Pins of the oracle's nonlinear flux operator (residual and Jacobian of
test/PLaplacianTests.jl:24-31), checked against what the mathematics fixes:

- p = 2, beta = 0 reduces it to the Poisson operator with kappa = 1, whose
  oracle path is pinned by the closed-form tests in test_oracle_pins.py;
- the residual is the gradient of an energy functional (computed by the oracle
  without the residual's formula), checked by central differences;
- the Jacobian is the derivative of the residual, checked by central differences;
- the reference test's manufactured solution u = x + y (constant gradient, f = 0)
  makes the residual vanish;
- the regularised p = 3 Jacobian is symmetric positive definite.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from gridap_amd import workloads as W


def global_system(res):
    I, J, V = oracle.global_coo(res)
    A = sp.csr_matrix((V, (I, J)), shape=(res.ngdofs, res.ngdofs))
    b = np.zeros(res.ngdofs)
    for P in res.parts:
        no = P["n_own"]
        b[P["row_gid"][:no]] = P["b"][:no]
    return A, b


def small_flux3d(n=(3, 4, 3), parts=(2, 2, 1), p=3.0, alpha=0.1, beta=1.0, f=1.0, strategy="sub"):
    q = W.plaplacian3d_q1(n, parts, p=p, eps=0.0, strategy=strategy)
    q["cells"] = tuple(n)
    q["flux"] = (p, alpha, beta)
    q["rhs_const"] = f
    q["uD"] = (0.3, 0.5, -0.2, 0.4)
    return q


def smooth_state(q, seed):
    """A smooth random lattice state (Dirichlet nodes are overridden by uD)."""
    rng = np.random.default_rng(seed)
    x = W.lattice_coords(q)
    a = rng.normal(size=(4, x.shape[1]))
    u = 0.2 * np.sin(x @ a[0] * 2.0 + 1.0) + 0.3 * np.cos(x @ a[1]) + 0.1 * (x @ a[2])
    return u


def run_state(q, u):
    q = dict(q)
    q["state"] = u
    return oracle.run(q)


def test_p2_reduces_to_poisson():
    q = small_flux3d(p=2.0, alpha=1.0, beta=0.0, f=0.7)
    u = smooth_state(q, 1)
    rf = run_state(q, u)
    qp = dict(q)
    qp["op"] = "diffusion"
    rp = oracle.run(qp)
    Af, bf = global_system(rf)
    Ap, bp = global_system(rp)
    assert abs(Af - Ap).max() <= 1e-13 * abs(Ap).max()
    # linear operator: r(u) = A u_free - b (b carries the source and the Dirichlet lifting)
    g2n = oracle.gid_to_gnode(rf)
    uf = u[g2n]
    r = Ap @ uf - bp
    assert np.abs(bf - r).max() <= 1e-12 * np.abs(r).max()


@pytest.mark.parametrize("p,alpha,beta", [(3.0, 0.1, 1.0), (4.0, 1.0, 1.0), (2.5, 0.3, 0.7)])
def test_residual_is_energy_gradient(p, alpha, beta):
    q = small_flux3d(p=p, alpha=alpha, beta=beta, f=0.9)
    u = smooth_state(q, 2)
    res = run_state(q, u)
    _, r = global_system(res)
    g2n = oracle.gid_to_gnode(res)
    rng = np.random.default_rng(3)
    v = np.zeros_like(u)
    v[g2n] = rng.normal(size=len(g2n))  # free directions only
    h = 1e-5
    qp, qm = dict(q), dict(q)
    qp["state"], qm["state"] = u + h * v, u - h * v
    dE = (oracle.energy(qp) - oracle.energy(qm)) / (2 * h)
    rv = r @ v[g2n]
    assert abs(dE - rv) <= 1e-7 * max(abs(rv), np.abs(r).sum() * np.abs(v).max() * 1e-3)


@pytest.mark.parametrize("p,alpha,beta", [(3.0, 0.1, 1.0), (4.0, 1.0, 1.0)])
def test_jacobian_is_residual_derivative(p, alpha, beta):
    q = small_flux3d(p=p, alpha=alpha, beta=beta)
    u = smooth_state(q, 4)
    res = run_state(q, u)
    A, _ = global_system(res)
    g2n = oracle.gid_to_gnode(res)
    rng = np.random.default_rng(5)
    v = np.zeros_like(u)
    v[g2n] = rng.normal(size=len(g2n))
    h = 1e-5
    _, rp = global_system(run_state(q, u + h * v))
    _, rm = global_system(run_state(q, u - h * v))
    fd = (rp - rm) / (2 * h)
    Jv = A @ v[g2n]
    assert np.abs(fd - Jv).max() <= 1e-7 * np.abs(Jv).max()


@pytest.mark.parametrize("strategy", ["sub", "full"])
def test_reference_manufactured_solution_zero_residual(strategy):
    q = W.flux2d_reference_test(strategy=strategy)
    x = W.lattice_coords(q)
    u = x[:, 0] + x[:, 1]
    res = run_state(q, u)
    A, r = global_system(res)
    assert np.abs(r).max() <= 1e-12 * np.abs(A).max()
    # constant gradient g = (1,1): A = (1+|g|^2) K + 2 (g.grad)(g.grad) assembled, symmetric
    assert abs(A - A.T).max() <= 1e-13 * abs(A).max()


def test_plaplacian_jacobian_spd_and_partition_invariant():
    q1 = small_flux3d(parts=(1, 1, 1), p=3.0, alpha=1e-4, beta=1.0)
    q8 = small_flux3d(parts=(2, 2, 1), p=3.0, alpha=1e-4, beta=1.0)
    u = smooth_state(q1, 6)
    r1, r8 = run_state(q1, u), run_state(q8, u)
    A1, b1 = global_system(r1)
    A8, b8 = global_system(r8)
    # gids differ between partitions: compare through the lattice node key
    n1, n8 = oracle.gid_to_gnode(r1), oracle.gid_to_gnode(r8)
    p = np.empty_like(n8)
    pos = {g: i for i, g in enumerate(n1)}
    for i, g in enumerate(n8):
        p[i] = pos[g]
    P = sp.csr_matrix((np.ones(len(p)), (np.arange(len(p)), p)), shape=(len(p), len(p)))
    assert abs(P @ A1 @ P.T - A8).max() <= 1e-13 * abs(A1).max()
    assert np.abs(P @ b1 - b8).max() <= 1e-13 * np.abs(b1).max()
    Ad = A1.toarray()
    assert np.abs(Ad - Ad.T).max() <= 1e-13 * np.abs(Ad).max()
    assert np.linalg.eigvalsh(0.5 * (Ad + Ad.T)).min() > 0.0
