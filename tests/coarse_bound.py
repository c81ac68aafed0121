"""What a coarse-grid network can gain at most in the learned multilevel
preconditioner (test infrastructure; run as a script for larger sizes:
``python tests/coarse_bound.py 128``).

The network only predicts the coarse-grid error, so an exact coarse solve is
the best it can reach.  This module restates the library's preconditioner in
scipy (damped Jacobi, full weighting R = P^T with trilinear P, a coarse
correction, the energy-optimal step along P ec, damped Jacobi again; the
coarse correction either the Galerkin V-cycle with nu sweeps per level and 24
sweeps on the coarsest grid, or an exact solve) and counts the iterations of
flexible CG with one stored direction to 1e-8 on the lognormal-kappa Poisson
problem of the learned workload.  The Q1 matrix is assembled here from cell
kappa (pinned against the oracle's assembly by test_coarse_bound.py)."""
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def q1_reference_stiffness(h):
    """8 x 8 Q1 stiffness of a cube of side h (2-point Gauss: exact); local node a = x | 2 y | 4 z"""
    g = np.array([-1.0, 1.0]) / np.sqrt(3.0)
    K = np.zeros((8, 8))
    for qx in g:
        for qy in g:
            for qz in g:
                t = ((1 + qx) / 2, (1 + qy) / 2, (1 + qz) / 2)
                G = np.empty((8, 3))
                for a in range(8):
                    b = (a & 1, (a >> 1) & 1, (a >> 2) & 1)
                    val = [t[d] if b[d] else 1.0 - t[d] for d in range(3)]
                    der = [1.0 if b[d] else -1.0 for d in range(3)]
                    G[a] = [der[0] * val[1] * val[2], val[0] * der[1] * val[2], val[0] * val[1] * der[2]]
                G /= h
                K += G @ G.T * (h ** 3 / 8.0)
    return K


def q1_matrix(N, kappa):
    """3D Q1 stiffness on N^3 cells of (0,1)^3 with cell kappa (x fastest), on the
    interior nodes in lattice order (x fastest)."""
    n1 = N + 1
    K = q1_reference_stiffness(1.0 / N)
    c = np.arange(N ** 3)
    cx, cy, cz = c % N, (c // N) % N, c // (N * N)
    nodes = np.stack([(cx + (a & 1)) + n1 * ((cy + ((a >> 1) & 1)) + n1 * (cz + ((a >> 2) & 1))) for a in range(8)], 1)
    I = np.repeat(nodes, 8, axis=1).ravel()
    J = np.tile(nodes, (1, 8)).ravel()
    V = (kappa[c][:, None] * K.ravel()[None, :]).ravel()
    A = sp.coo_matrix((V, (I, J)), shape=(n1 ** 3, n1 ** 3)).tocsr()
    i = np.arange(n1)
    inter = (i > 0) & (i < N)
    free = np.flatnonzero((inter[None, None, :] & inter[None, :, None] & inter[:, None, None]).ravel())
    return A[free][:, free].tocsr()


def prolong3(N):
    """trilinear P from the interior nodes of the N/2 grid (even fine indices) to those of the N grid"""
    nf, nc = N - 1, N // 2 - 1
    J = np.arange(nc)
    f = 2 * J + 1  # fine index (0-based over interior nodes) of coarse node J
    rows = np.concatenate([f, f - 1, f + 1])
    cols = np.concatenate([J, J, J])
    vals = np.concatenate([np.ones(nc), np.full(nc, 0.5), np.full(nc, 0.5)])
    P1 = sp.csr_matrix((vals, (rows, cols)), shape=(nf, nc))
    return sp.kron(P1, sp.kron(P1, P1)).tocsr()  # x fastest


class Multilevel:
    """the preconditioner; coarse = "vcycle" or "exact" """

    def __init__(self, A, N, nu=2, coarse="vcycle", coarsest_cells=4, coarsest_sweeps=24):
        self.A, self.nu, self.coarse, self.cs = A, nu, coarse, coarsest_sweeps
        self.Dinv = 1.0 / A.diagonal()
        self.P = prolong3(N)
        self.levels = []
        Ac, m = (self.P.T @ A @ self.P).tocsr(), N // 2
        while True:
            lev = dict(A=Ac, Dinv=1.0 / Ac.diagonal())
            self.levels.append(lev)
            if m % 2 or m <= coarsest_cells:
                break
            lev["P"] = prolong3(m)
            Ac, m = (lev["P"].T @ Ac @ lev["P"]).tocsr(), m // 2
        self.lu = spla.splu(self.levels[0]["A"].tocsc()) if coarse == "exact" else None

    def vcycle(self, l, rhs):
        L, w = self.levels[l], 2.0 / 3.0
        if l + 1 == len(self.levels):
            e = np.zeros_like(rhs)
            for _ in range(self.cs):
                e = e + w * L["Dinv"] * (rhs - L["A"] @ e)
            return e
        e = w * L["Dinv"] * rhs
        for _ in range(self.nu - 1):
            e = e + w * L["Dinv"] * (rhs - L["A"] @ e)
        e = e + L["P"] @ self.vcycle(l + 1, L["P"].T @ (rhs - L["A"] @ e))
        for _ in range(self.nu):
            e = e + w * L["Dinv"] * (rhs - L["A"] @ e)
        return e

    def coarse_solve(self, rc):
        if self.lu is not None:
            return self.lu.solve(rc)
        if self.coarse == "vcycle":
            return self.vcycle(0, rc)
        # "exact" by PCG on Ac with the V-cycle to 1e-12 (grids too large for a sparse LU)
        Ac, n0 = self.levels[0]["A"], np.linalg.norm(rc)
        e, r = np.zeros_like(rc), rc.copy()
        z = self.vcycle(0, r)
        p, rz = z.copy(), r @ z
        for _ in range(200):
            q = Ac @ p
            a = rz / (p @ q)
            e += a * p
            r -= a * q
            if np.linalg.norm(r) <= 1e-12 * n0:
                break
            z = self.vcycle(0, r)
            rz, rz_old = r @ z, rz
            p = z + (rz / rz_old) * p
        return e

    def __call__(self, r):
        w = 2.0 / 3.0
        z = w * self.Dinv * r
        s = r - self.A @ z
        c = self.P @ self.coarse_solve(self.P.T @ s)
        Ac = self.A @ c
        a = (c @ s) / (c @ Ac)
        return z + a * c + w * self.Dinv * (s - a * Ac)


def fcg(A, b, M, rtol=1e-8, maxit=1000):
    """flexible CG with one stored direction; iterations to the true residual rtol"""
    x, r, nb = np.zeros_like(b), b.copy(), np.linalg.norm(b)
    p_old = q_old = None
    for it in range(1, maxit + 1):
        z = M(r)
        p = z if p_old is None else z - (z @ q_old) / (p_old @ q_old) * p_old
        q = A @ p
        a = (p @ r) / (p @ q)
        x += a * p
        r -= a * q
        p_old, q_old = p, q
        if np.linalg.norm(b - A @ x) <= rtol * nb:
            return x, it
    return x, maxit


def iterations(N, seed=900000, nu=2):
    """FCG iterations with the V-cycle and with an exact coarse solve, on the
    learned workload's problem (kappa = exp(2 g), correlation length 0.1)"""
    import grf_oracle

    kappa = grf_oracle.lognormal_cells((N, N, N), [1.0 / N] * 3, seed, 0.1, 2.0)
    A = q1_matrix(N, kappa)
    b = np.random.default_rng(5000).normal(size=A.shape[0])
    exact = "exact" if N <= 64 else "exact_pcg"
    return {mode: fcg(A, b, Multilevel(A, N, nu, mode))[1] for mode in ("vcycle", exact)}


if __name__ == "__main__":
    import os

    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    N = int(sys.argv[1]) if len(sys.argv) > 1 else 32
    print(N, iterations(N))
