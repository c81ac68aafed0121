"""State handling of batch graphs on device 0.

A captured batch graph holds the device pointers and kernels of its problems
at capture time.  Whatever re-allocates those buffers or changes the launches
(a kernel variant, a coefficient mode, a precision, a new symbolic phase) must
drop the graph, so that the next gdx_batch_launch returns GDX_ERR_STATE
instead of replaying freed memory.  Flux problems get their device state in
gdx_symbolic, so a batch captured right after it records only kernel launches
and replays the eager numeric phase bit for bit.
"""
import numpy as np
import pytest

from gridap_amd import AssemblyBatch, flux_operator_from_workload, problem_from_workload
from gridap_amd import workloads as W
from gridap_amd._lib import GdxError

pytestmark = pytest.mark.gpu

ERR_STATE = -3


def make(ctx, p):
    assem, data = problem_from_workload(ctx, p)
    return assem._problem_for(data)


def values(prob):
    prob.sync()
    return [(P["val"].copy(), P["b"].copy()) for P in (prob.export_part(k) for k in range(prob.nparts))]


def launch_state_error(b):
    with pytest.raises(GdxError) as e:
        b.launch()
    assert e.value.code == ERR_STATE


def test_kernel_change_drops_the_graph(gpu_ctx):
    prob = make(gpu_ctx, W.poisson3d_q2_fourier(6, (2, 1, 1)))
    other = make(gpu_ctx, W.poisson3d_q1(7, (1, 2, 1)))
    b1, b2 = AssemblyBatch(gpu_ctx), AssemblyBatch(gpu_ctx)
    b1.add(prob)
    b1.add(other)
    b2.add(other)
    b1.capture()
    b2.capture()
    b1.launch()
    b1.sync()
    prob.set_kernel(2)  # re-uploads the tiles and row table of prob
    launch_state_error(b1)
    b2.launch()  # a batch without prob keeps its graph
    b2.sync()
    b1.capture()
    b1.launch()
    b1.sync()
    ref = values(prob)
    prob.numeric()
    for (v0, c0), (v1, c1) in zip(ref, values(prob)):
        assert np.array_equal(v0, v1) and np.array_equal(c0, c1)
    for x in (b1, b2, prob, other):
        x.close()


def test_coefficient_mode_and_precision_drop_the_graph(gpu_ctx):
    q = W.poisson3d_q1(8, (2, 1, 1), modes=np.zeros((0, 5)))
    prob = make(gpu_ctx, q)
    b = AssemblyBatch(gpu_ctx)
    b.add(prob)
    b.capture()
    kap = np.exp(np.random.default_rng(1).normal(size=int(np.prod(q["cells"]))))
    prob.set_cell_coefficient(kap)  # analytic -> cell-wise kappa: another coefficient launch
    launch_state_error(b)
    b.capture()
    prob.set_cell_coefficient(2.0 * kap)  # same mode, same buffer: the graph stays valid
    b.launch()
    b.sync()
    prob.set_cell_coefficient(None)
    launch_state_error(b)
    b.capture()
    prob.set_kernel(0, 8)
    b.capture()
    prob.numeric(1)  # FP32 contractions: other kernels than the captured FP64 ones
    launch_state_error(b)
    prob.numeric(0)
    b.capture()
    prob.symbolic()  # device buffers re-made
    launch_state_error(b)
    b.close()
    prob.close()


def test_flux_batch_captured_right_after_symbolic(gpu_ctx):
    ops = [flux_operator_from_workload(gpu_ctx, dict(W.plaplacian3d_q1(n, parts), flux=(3.0, 0.01, 1.0)))
           for n, parts in ((6, (2, 1, 1)), (5, (1, 2, 2)))]
    probs = [op._problem() for op in ops]  # symbolic phase done, no numeric yet
    poisson = make(gpu_ctx, W.poisson3d_q1(6, (2, 2, 1)))
    b = AssemblyBatch(gpu_ctx)
    for prob in probs + [poisson]:
        b.add(prob)
    b.capture()
    got = []
    for rep in range(3):
        b.launch()
        b.sync()
        got.append([values(prob) for prob in probs])
        probs[rep % 2].numeric()  # eager work between replays
        probs[rep % 2].sync()
    for prob, g in zip(probs, zip(*got)):
        prob.numeric()
        ref = values(prob)
        for rep in g:
            for (v0, c0), (v1, c1) in zip(ref, rep):
                assert np.array_equal(v0.view(np.int64), v1.view(np.int64))
                assert np.array_equal(c0.view(np.int64), c1.view(np.int64))
        assert any(np.abs(v).max() > 0 for v, _ in ref)
    b.close()
    for prob in probs + [poisson]:
        prob.close()
