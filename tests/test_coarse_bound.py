"""The ceiling of the coarse-grid network (CPU): with an exact coarse solve in
place of the V-cycle the multilevel FCG saves only a few iterations, so a
network that at best makes the coarse correction exact can save no more.  The
restated Q1 matrix is first pinned against the oracle's assembly."""
import numpy as np

import coarse_bound as CB
import oracle
from gridap_amd import workloads as W


def test_restated_matrix_is_the_oracles():
    N = 6
    kappa = np.exp(2.0 * np.random.default_rng(3).normal(size=N ** 3))
    r = oracle.run(dict(W.poisson3d_q1(N, (2, 1, 2)), cell_kappa=kappa))
    I, J, V = oracle.global_coo(r)
    g2n = oracle.gid_to_gnode(r)
    n1 = N + 1
    i, j, k = g2n % n1, (g2n // n1) % n1, g2n // (n1 * n1)
    lex = (i - 1) + (N - 1) * ((j - 1) + (N - 1) * (k - 1))  # interior lattice order
    ref = np.zeros(((N - 1) ** 3,) * 2)
    np.add.at(ref, (lex[I], lex[J]), V)
    A = CB.q1_matrix(N, kappa).toarray()
    assert np.abs(A - ref).max() <= 1e-12 * np.abs(ref).max()


def test_prolongation_is_trilinear():
    N = 8
    P = CB.prolong3(N)
    # a trilinear function on the coarse interior nodes is reproduced on the fine interior nodes
    xc = np.arange(1, N // 2) * 2.0 / N
    xf = np.arange(1, N) * 1.0 / N
    f = lambda x, y, z: 1.0 + 2 * x - y + 0.5 * z + x * y * z
    Xc = np.meshgrid(xc, xc, xc, indexing="ij")
    Xf = np.meshgrid(xf, xf, xf, indexing="ij")
    vc = f(*Xc).transpose(2, 1, 0).ravel()
    vf = f(*Xf).transpose(2, 1, 0).ravel()
    # interior fine nodes next to the boundary interpolate towards the (zero) boundary: compare away from it
    inner = np.zeros((N - 1,) * 3, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    assert np.allclose((P @ vc)[inner.ravel()], vf[inner.ravel()], atol=1e-13)


def test_exact_coarse_solve_bounds_the_network():
    it = CB.iterations(32)
    print("FCG iterations at 32^3:", it)
    assert it["exact"] < it["vcycle"] <= it["exact"] + 5
