"""Solve side on the GPU: CSR SpMV over the owned rows and Jacobi-PCG, on the
matrix the device assembled (the reference applies A*x and solves A\\b,
test/FESpacesTests.jl:21-27, README.md:76)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from _helpers import gid_to_gnode, owned_global, rel_err
from gridap_amd import problem_from_workload
from gridap_amd import workloads as W

pytestmark = pytest.mark.gpu


def _prob(ctx, p):
    assem, data = problem_from_workload(ctx, p)
    prob = assem._problem_for(data)
    prob.numeric()
    prob.sync()
    return prob


@pytest.mark.parametrize("p", [W.poisson3d_q2_fourier(6, (2, 2, 1)), W.poisson3d_q1(9, (1, 2, 3)),
                               W.poisson2d_q1(20, (2, 2))], ids=["q2", "q1", "2d"])
def test_spmv_matches_oracle_matrix(gpu_ctx, p):
    r = oracle.run(p)
    A, b = owned_global(r.parts, r.ngdofs)
    prob = _prob(gpu_ctx, p)
    x = np.random.default_rng(1).standard_normal(r.ngdofs)
    y = prob.spmv(x)
    assert rel_err(y, A @ x) < 1e-13
    prob.close()


def test_spmv_mixed_column_code_blocks(gpu_ctx):
    """64^3 Q1 on parts (2,2,2): 250,047 rows span more than 16384 ids, so the
    block SpMV reads most 128-row blocks as 16-bit (window, offset) codes and
    the few whose columns need more than four windows (part corners) as 32-bit
    ids; the product matches the assembled matrix's either way."""
    p = W.poisson3d_q1(64, (2, 2, 2))
    prob = _prob(gpu_ctx, p)
    parts = [prob.export_part(k) for k in range(prob.nparts)]
    A, _ = owned_global(parts, prob.ngdofs)
    x = np.random.default_rng(5).standard_normal(prob.ngdofs)
    assert rel_err(prob.spmv(x), A @ x) < 1e-13
    _, nbytes = prob.spmv_device_timing(2)
    nnz, n = prob.owned_nnz(), prob.ngdofs
    nblk = -(-n // 128)
    lo = 10 * nnz + 28 * n + 16 * nblk  # every entry 16-bit
    assert lo < nbytes < lo + 0.05 * (2 * nnz)  # some entries 32-bit, at most 5%
    prob.close()


def test_pcg_linear_solution_2d(gpu_ctx):
    # u = x + y is in the Q1 space: the solve reproduces it at the nodes
    p = W.poisson2d_q1(64, (2, 2))
    prob = _prob(gpu_ctx, p)
    x, it, rr, ms = prob.pcg(rtol=1e-12, maxiter=5000)
    assert rr <= 1e-11
    parts = [prob.export_part(k) for k in range(prob.nparts)]
    g2n = gid_to_gnode(parts, prob.ngdofs)
    u = (g2n % 65) / 64 + (g2n // 65) / 64
    assert np.abs(x - u).max() < 1e-9
    prob.close()


def test_pcg_matches_direct_solve_and_repeats(gpu_ctx):
    p = W.poisson3d_q2_fourier(10, (2, 1, 2))
    prob = _prob(gpu_ctx, p)
    parts = [prob.export_part(k) for k in range(prob.nparts)]
    A, b = owned_global(parts, prob.ngdofs)
    x, it, rr, ms = prob.pcg(rtol=1e-10, maxiter=5000)
    assert rr <= 1e-10 and it > 0
    xd = spla.spsolve(A.tocsc(), b)
    assert rel_err(x, xd) < 1e-7
    x2, it2, rr2, _ = prob.pcg(rtol=1e-10, maxiter=5000)
    assert it2 == it and np.array_equal(x.view(np.int64), x2.view(np.int64))
    prob.close()


def test_spmv_timed_runs_the_solver_kernel(gpu_ctx):
    """gdx_spmv_timed: a positive device time per SpMV, bytes per launch between
    the all-16-bit and all-32-bit column id forms, and the SpMV it times (rows
    in 128-row blocks, 16-bit column offsets inside a part, 32-bit ids across
    part boundaries) still gives the oracle's product afterwards."""
    p = W.poisson3d_q1(20, (2, 2, 1))
    prob = _prob(gpu_ctx, p)
    ms, nbytes = prob.spmv_device_timing(10)
    assert 0.0 < ms < 100.0
    nnz, n = prob.owned_nnz(), prob.ngdofs
    nblk = -(-n // 128)
    assert 10 * nnz + 28 * n + 16 * nblk <= nbytes < 12 * nnz + 28 * n + 16 * nblk
    r = oracle.run(p)
    A, _ = owned_global(r.parts, r.ngdofs)
    x = np.random.default_rng(2).standard_normal(prob.ngdofs)
    assert rel_err(prob.spmv(x), A @ x) <= 1e-13
    with pytest.raises(Exception):
        prob.spmv_device_ms(0)
    prob.close()
