"""Cell-wise kappa on device 0: the device random field against its CPU
restatement (bit-level agreement up to the device's log/cos rounding), and the
assembly with per-cell kappa against the oracle (structure identical, values
within 1e-12)."""
import numpy as np
import pytest

import grf_oracle
import oracle
from _helpers import STRUCT_KEYS, rel_err
from gridap_amd import problem_from_workload
from gridap_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cells,corr", [((16, 16, 16), 0.1), ((12, 10, 14), 0.2), ((20, 24), 0.15)])
def test_grf_matches_cpu_restatement(gpu_ctx, cells, corr):
    dim = len(cells)
    p = W.poisson3d_q1(cells) if dim == 3 else W.poisson2d_q1(cells[0])
    p["cells"] = tuple(cells)
    p["parts"] = (2,) * dim
    assem, data = problem_from_workload(gpu_ctx, p)
    prob = assem._problem_for(data)
    prob.cell_coefficient_grf(123, corr, 2.0)
    k = prob.export_cell_coefficient(int(np.prod(cells)))
    h = [(p["domain"][2 * d + 1] - p["domain"][2 * d]) / cells[d] for d in range(dim)]
    ref = grf_oracle.lognormal_cells(cells, h, 123, corr, 2.0)
    assert np.abs(np.log(k) - np.log(ref)).max() < 1e-12
    prob.close()


CASES = [W.poisson3d_q1(9, (2, 2, 1)), W.poisson3d_q1((8, 6, 10), (1, 2, 2), strategy="full"),
         dict(W.poisson3d_q2_fourier(5, (2, 1, 2)), modes=np.zeros((0, 5)), rhs="const", rhs_const=1.0),
         dict(W.poisson2d_q1(12, (2, 2)), rhs_const=0.5)]


@pytest.mark.parametrize("p", CASES, ids=lambda p: f"{p['dim']}d-q{p['order']}-{p['cells']}-{p['strategy']}")
def test_cell_kappa_assembly_matches_oracle(gpu_ctx, p):
    rng = np.random.default_rng(4)
    nc = int(np.prod(p["cells"]))
    kappa = np.exp(2.0 * rng.normal(size=nc))
    q = dict(p)
    q["cell_kappa"] = kappa
    r = oracle.run(q)
    assem, data = problem_from_workload(gpu_ctx, p)
    prob = assem._problem_for(data)
    prob.set_cell_coefficient(kappa)
    prob.numeric()
    prob.sync()
    for k, P in enumerate(r.parts):
        Q = prob.export_part(k)
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
        assert rel_err(Q["val"], P["val"]) <= 1e-12, k
        assert rel_err(Q["b"], P["b"]) <= 1e-12, k
    prob.close()
