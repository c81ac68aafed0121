"""Failed device allocations in the symbolic phase, on device 0.

GDX_FAULT_ALLOC=n makes the n-th device allocation of one gdx_symbolic call
fail as exhausted HBM would.  For every n up to the phase's allocation count:
the call returns GDX_ERR_ALLOC, the problem refuses a numeric phase
(GDX_ERR_STATE), and a second gdx_symbolic on the same problem succeeds and
assembles exactly what a problem that never failed assembles.  Covers the
device numbering, CSR, pattern, row-table and tile stages, the buffers
gdx_symbolic allocates after them, and the flux and elasticity set-ups.
"""
import os

import numpy as np
import pytest

from gridap_amd import fespaces, problem_from_workload
from gridap_amd import workloads as W
from gridap_amd._lib import GdxError

pytestmark = pytest.mark.gpu

ERR_STATE, ERR_ALLOC = -3, -5

CASES = [
    W.poisson3d_q2_fourier(4, (2, 2, 1)),
    W.plaplacian3d_q1(6, (2, 1, 2)),
    W.elasticity3d_q2((4, 3, 3), (2, 1, 2)),
]


def fresh(ctx, p):
    assem, data = problem_from_workload(ctx, p)
    desc, keep = fespaces.problem_desc(data, assem.strategy)
    return fespaces.GpuProblem(ctx, desc, keep)


def assembled(gp):
    gp.numeric()
    gp.sync()
    return [gp.export_part(k) for k in range(gp.nparts)]


@pytest.mark.parametrize("p", CASES, ids=lambda p: p.get("op", "diffusion"))
def test_every_failed_allocation_is_reported_and_recoverable(gpu_ctx, p):
    print("allocation faults", p.get("op", "diffusion"), p["cells"], p["parts"], flush=True)
    ref_gp = fresh(gpu_ctx, p)
    ref_gp.symbolic()
    ref = assembled(ref_gp)
    ref_gp.close()
    n = 0
    while True:
        gp = fresh(gpu_ctx, p)
        os.environ["GDX_FAULT_ALLOC"] = str(n)
        try:
            gp.symbolic()
            failed = False
        except GdxError as e:
            assert e.code == ERR_ALLOC, (n, e)
            failed = True
        finally:
            os.environ.pop("GDX_FAULT_ALLOC", None)
        if failed:
            with pytest.raises(GdxError) as e:
                gp.numeric()
            assert e.value.code == ERR_STATE
            gp.symbolic()
        for P, Q in zip(ref, assembled(gp)):
            for key in ("ldof_gid", "row_gid", "col_gid", "rowptr", "colidx"):
                assert np.array_equal(P[key], Q[key]), (n, key)
            assert np.array_equal(P["val"].view(np.int64), Q["val"].view(np.int64)), n
            assert np.array_equal(P["b"].view(np.int64), Q["b"].view(np.int64)), n
        gp.close()
        if not failed:
            break
        n += 1
        assert n < 500
    print("device allocations in the symbolic phase:", n, flush=True)
    assert n >= 20
