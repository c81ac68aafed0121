"""gdx_nn_apply_noalloc against gdx_nn_apply, called through the C ABI on device 0.

The no-alloc entry point stages r and z in the library's static device array
instead of per-call allocations; it runs the same kernels, so z must equal
gdx_nn_apply's bit for bit.  Two calls in a row share the array: the second,
on a smaller problem, must not see the first one's contents.  A problem above
the staging limit is refused with GDX_ERR_ARG.
"""
import ctypes

import numpy as np
import pytest

from gridap_amd import LearnedCoarseSolver, _lib, problem_from_workload
from gridap_amd import workloads as W
from gridap_amd._lib import GdxError

pytestmark = pytest.mark.gpu

ERR_ARG = -1


def attached(ctx, cells, seed):
    p = W.poisson3d_q1(cells, (2, 2, 2), rhs_const=0.0)
    assem, data = problem_from_workload(ctx, p)
    prob = assem._problem_for(data)
    prob.cell_coefficient_grf(seed, 0.1, 2.0)
    prob.numeric()
    prob.sync()
    net = LearnedCoarseSolver(ctx, cells // 2, seed)
    net.attach(prob)
    return prob, net


def apply_both(net, r):
    z1 = np.empty_like(r)
    z2 = np.full_like(r, np.nan)
    _lib.call("gdx_nn_apply", net.handle, r.ctypes.data, z1.ctypes.data)
    _lib.call("gdx_nn_apply_noalloc", net.handle, r.ctypes.data, z2.ctypes.data)
    return z1, z2


def test_noalloc_apply_matches_apply(gpu_ctx):
    rng = np.random.default_rng(3)
    big, netb = attached(gpu_ctx, 32, 11)
    small, nets = attached(gpu_ctx, 16, 12)
    for net, prob in ((netb, big), (nets, small)):  # the second call has the smaller input
        print("no-alloc apply, ngdofs", prob.ngdofs, flush=True)
        r = rng.normal(size=prob.ngdofs)
        z1, z2 = apply_both(net, r)
        assert np.all(np.isfinite(z1)) and np.abs(z1).max() > 0
        assert np.array_equal(z1.view(np.int64), z2.view(np.int64))
    # and back to the larger problem after the smaller one
    r = rng.normal(size=big.ngdofs)
    z1, z2 = apply_both(netb, r)
    assert np.array_equal(z1.view(np.int64), z2.view(np.int64))
    for x in (netb, nets, big, small):
        x.close()


def test_noalloc_apply_refuses_oversize(gpu_ctx):
    prob, net = attached(gpu_ctx, 136, 5)  # 135^3 = 2,460,375 free dofs > 2,097,152
    print("no-alloc apply oversize, ngdofs", prob.ngdofs, flush=True)
    r = np.zeros(prob.ngdofs)
    z = np.zeros(prob.ngdofs)
    with pytest.raises(GdxError) as e:
        _lib.call("gdx_nn_apply_noalloc", net.handle, r.ctypes.data, z.ctypes.data)
    assert e.value.code == ERR_ARG
    net.close()
    prob.close()
