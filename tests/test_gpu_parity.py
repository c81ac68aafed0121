"""Parity of the HIP assembly (libgdx.so on device 0) with the CPU oracle.

Bar: structure (FE dof gids/owners/lattice nodes, row and col PRanges, CSR
rowptr/colidx) bit-identical; values and RHS within 1e-12 relative
(max|dA| / max|A|) in FP64.  At the full 3D Q2 64^3 size, where the C oracle
is too slow, sampled rows are checked against the numpy brute-force row
computation and size-independent invariants.
"""
import os

import numpy as np
import pytest

import micro
import oracle
from _helpers import STRUCT_KEYS, gid_to_gnode, owned_global, rel_err
from gridap_amd import (AffineFEOperator, assemble_matrix_and_vector, assemble_matrix_and_vector_,
                        problem_from_workload)
from gridap_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-12

CASES = [
    W.poisson2d_q1(8, (2, 2)),
    W.poisson2d_q1(37, (2, 3)),
    W.poisson2d_q1(20, (3, 1), strategy="full"),
    dict(W.poisson2d_q1(9, (2, 2)), order=2, qdeg=4, modes=W.fourier_modes(5, 6, 4), rhs="manufactured",
         uD=(0.2, 1.0, -1.0, 0.0)),
    W.poisson3d_q2_fourier(4, (2, 2, 2)),
    W.poisson3d_q2_fourier(9, (2, 1, 3)),
    W.poisson3d_q2_fourier(10, (1, 2, 2), strategy="full"),
    W.poisson3d_q2_fourier(3, (1, 1, 1)),
    dict(W.poisson3d_q2_fourier(7, (2, 2, 1)), qdeg=2),
    W.poisson3d_q1(11, (2, 2, 1)),
    W.poisson3d_q1(13, (3, 2, 2), strategy="full", modes=W.fourier_modes(9, 8, 5)),
    dict(W.poisson3d_q1((9, 6, 10), (2, 1, 3)), dirichlet=False),
]


def gpu_parts(ctx, p):
    assem, data = problem_from_workload(ctx, p)
    prob = assem._problem_for(data)
    prob.numeric()
    prob.sync()
    return assem, prob, [prob.export_part(k) for k in range(prob.nparts)]


@pytest.mark.parametrize("p", CASES, ids=lambda p: f"{p['dim']}d-q{p['order']}-{p['cells']}-{p['parts']}-{p['strategy']}")
def test_parity_with_oracle(gpu_ctx, p):
    print("case", p["dim"], p["order"], p["cells"], p["parts"], p["strategy"], flush=True)
    r = oracle.run(p)
    assem, prob, parts = gpu_parts(gpu_ctx, p)
    assert prob.ngdofs == r.ngdofs
    for k, (P, Q) in enumerate(zip(r.parts, parts)):
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
        assert rel_err(Q["val"], P["val"]) <= TOL, k
        assert rel_err(Q["b"], P["b"]) <= TOL, k
        no = P["n_own"]
        assert np.all(Q["val"][Q["rowptr"][no]:] == 0.0) and np.all(Q["b"][no:] == 0.0)  # ghost rows zero
    prob.close()


def test_repeatable_bitwise(gpu_ctx):
    # no float atomics: two numeric assemblies give identical bits
    p = W.poisson3d_q2_fourier(12, (2, 2, 2))
    assem, prob, parts1 = gpu_parts(gpu_ctx, p)
    prob.numeric()
    prob.sync()
    for k, P in enumerate(parts1):
        Q = prob.export_part(k)
        assert np.array_equal(P["val"].view(np.int64), Q["val"].view(np.int64))
        assert np.array_equal(P["b"].view(np.int64), Q["b"].view(np.int64))
    prob.close()


def test_inplace_reassembly_matches_fresh(gpu_ctx):
    # assemble_matrix_and_vector! (test/PLaplacianTests.jl:44-52, test/FESpacesTests.jl:33-39)
    p = W.poisson3d_q1(10, (2, 1, 2))
    assem, data = problem_from_workload(gpu_ctx, p)
    A, b = assemble_matrix_and_vector(assem, data)
    vals = [M.data.copy() for M in A.matrix_partition]
    for M in A.matrix_partition:
        M.data[:] = 0.0
    assemble_matrix_and_vector_(A, b, assem, data)
    for v, M in zip(vals, A.matrix_partition):
        assert np.array_equal(v, M.data)
    x = np.ones(A.shape[1])
    assert np.allclose(A @ x, A.to_global() @ x)


def test_affine_operator_recovers_linear_solution(gpu_ctx):
    # u = x + y on (0,1)^2, Q1, f = 0: A\b reproduces u at the nodes (test/FESpacesTests.jl:21-27)
    import scipy.sparse.linalg as spla
    from gridap_amd import (CartesianDiscreteModel, ConstantSource, Diffusion, Float64, Integral, LinearFunction,
                            Measure, ReferenceFE, Source, SubAssembledRows, FullyAssembledRows, TestFESpace,
                            Triangulation, TrialFESpace, lagrangian, SparseMatrixAssembler)
    for strategy in (SubAssembledRows(), FullyAssembledRows()):
        model = CartesianDiscreteModel(gpu_ctx, (2, 2), (0, 1, 0, 1), (32, 32))
        V = TestFESpace(model, ReferenceFE(lagrangian, Float64, 1), dirichlet_tags="boundary")
        U = TrialFESpace(V, LinearFunction(0.0, (1.0, 1.0)))
        dO = Measure(Triangulation(model), 2)
        op = AffineFEOperator(Integral(Diffusion(), dO), Integral(Source(ConstantSource(0.0)), dO), U, V,
                              SparseMatrixAssembler(U, V, strategy))
        A, b = op.get_matrix(), op.get_vector()
        for M, rows, cols in zip(A.matrix_partition, A.row_partition, A.col_partition):
            assert M.shape == (rows.local_length(), cols.local_length())
        Ag = A.to_global()
        x = spla.spsolve(Ag.tocsc(), b.to_global())
        gnode = np.zeros(A.shape[0], np.int64)
        for k in range(len(A.matrix_partition)):
            P = op.assem.problem.export_part(k)
            gnode[P["ldof_gid"]] = P["ldof_gnode"]
        xs, ys = (gnode % 33) / 32.0, (gnode // 33) / 32.0
        assert np.abs(x - (xs + ys)).max() < 1e-9


def test_golden_fixture(gpu_ctx):
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_small.npz"), allow_pickle=False)
    assem, prob, parts = gpu_parts(gpu_ctx, W.poisson2d_q1(6, (2, 2)))
    for k, Q in enumerate(parts):
        for key in ("row_gid", "col_gid", "rowptr", "colidx"):
            assert np.array_equal(Q[key], g[f"p{k}_{key}"])
        assert rel_err(Q["val"], g[f"p{k}_val"]) <= TOL
        assert rel_err(Q["b"], g[f"p{k}_b"]) <= TOL
    prob.close()


@pytest.mark.parametrize("kernel", [(2, 8, 0, TOL), (0, 32, 0, TOL), (0, 16, 1, 1e-6), (2, 8, 1, 1e-6)],
                         ids=["direct-store-fp64", "z-marching-fp64", "z-marching-fp32", "direct-store-fp32"])
def test_full_size_q2_64(gpu_ctx, kernel):
    """3D Q2 64^3, parts (2,2,2), kappa = exp(16 Fourier modes), run by the
    kernel the bench times (direct store), the z-marching kernel and the FP32
    mixed form: size-independent checks and 47 sampled rows against the numpy
    brute-force row integral (1e-12 in FP64, 1e-6 for FP32 contractions)."""
    variant, rows_z, precision, tol = kernel
    p = W.poisson3d_q2_fourier(64, (2, 2, 2))
    print("full-size case 3D Q2 64^3 parts (2,2,2), kernel", kernel, flush=True)
    assem, data = problem_from_workload(gpu_ctx, p)
    prob = assem._problem_for(data)
    prob.set_kernel(variant, rows_z)
    prob.numeric(precision)
    prob.sync()
    parts = [prob.export_part(k) for k in range(prob.nparts)]
    assert prob.owned_nnz() == W.nnz_free(p) == 127263527
    assert prob.ngdofs == W.free_dofs(p) == 2048383
    g2n = gid_to_gnode(parts, prob.ngdofs)
    n2g = {int(n): g for g, n in enumerate(g2n)}
    # locate owned rows by gid
    where = np.zeros((prob.ngdofs, 2), np.int64)
    for k, P in enumerate(parts):
        where[P["row_gid"][: P["n_own"]], 0] = k
        where[P["row_gid"][: P["n_own"]], 1] = np.arange(P["n_own"])
    rng = np.random.default_rng(7)
    picks = list(rng.choice(prob.ngdofs, 40, replace=False))
    # rows touching part interfaces and the boundary layer
    nn = 129
    for ijk in [(64, 64, 64), (1, 1, 1), (64, 1, 127), (63, 64, 65), (2, 64, 64), (127, 127, 127), (64, 63, 64)]:
        picks.append(n2g[ijk[0] + nn * (ijk[1] + nn * ijk[2])])
    for gid in picks:
        k, o = where[gid]
        P = parts[k]
        lo, hi = P["rowptr"][o], P["rowptr"][o + 1]
        got = dict(zip(g2n[P["col_gid"][P["colidx"][lo:hi]]].tolist(), P["val"][lo:hi]))
        ref, bref = micro.row(p, int(g2n[gid]))
        assert set(got) == set(ref)
        scale = max(abs(v) for v in ref.values())
        assert max(abs(got[j] - ref[j]) for j in ref) <= tol * scale
        assert abs(P["b"][o] - bref) <= tol * max(abs(bref), scale)
    # ghost rows stay zero
    for P in parts:
        no = P["n_own"]
        assert np.all(P["val"][P["rowptr"][no]:] == 0.0) and np.all(P["b"][no:] == 0.0)
    # symmetry on the sampled rows (A_ij == A_ji up to round-off)
    for gid in picks[:10]:
        k, o = where[gid]
        P = parts[k]
        lo, hi = P["rowptr"][o], P["rowptr"][o + 1]
        for c, v in zip(P["col_gid"][P["colidx"][lo:hi]], P["val"][lo:hi]):
            k2, o2 = where[c]
            P2 = parts[k2]
            lo2, hi2 = P2["rowptr"][o2], P2["rowptr"][o2 + 1]
            cols2 = P2["col_gid"][P2["colidx"][lo2:hi2]]
            v2 = P2["val"][lo2:hi2][np.nonzero(cols2 == gid)[0][0]]
            assert abs(v - v2) <= (1e-14 if precision == 0 else tol) * abs(P["val"][lo:hi]).max()
    prob.close()
