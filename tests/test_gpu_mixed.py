"""The mixed form (GDX_PREC_FP32: sum-factorised contractions in FP32, CSR
values in FP64) on device 0, with the z-marching kernel and the direct-store
kernel (the one the FP64 bench times): same structure as the oracle,
max relative matrix error <= 1e-6 (max|dA| / max|A|) and Frobenius relative
error <= 1e-6, RHS likewise; the FP64 form afterwards is unchanged."""
import numpy as np
import pytest

import oracle
from _helpers import STRUCT_KEYS, rel_err
from gridap_amd import problem_from_workload
from gridap_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [W.poisson3d_q2_fourier(8, (2, 2, 2)), W.poisson3d_q2_fourier(9, (1, 2, 3), strategy="full"),
         W.poisson3d_q1(12, (2, 2, 1)), dict(W.poisson3d_q2_fourier(6, (2, 1, 1)), qdeg=2)]


@pytest.mark.parametrize("variant", [0, 2], ids=["z-marching", "direct-store"])
@pytest.mark.parametrize("p", CASES, ids=["q2", "q2-full", "q1", "q2-qd2"])
def test_fp32_form_error_bound(gpu_ctx, p, variant):
    print("fp32 form, kernel", variant, p["order"], p["cells"], p["parts"], flush=True)
    r = oracle.run(p)
    assem, data = problem_from_workload(gpu_ctx, p)
    prob = assem._problem_for(data)
    prob.set_kernel(variant, 16)
    prob.numeric(precision=1)
    prob.sync()
    for k, P in enumerate(r.parts):
        Q = prob.export_part(k)
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
        assert rel_err(Q["val"], P["val"]) <= 1e-6
        assert np.linalg.norm(Q["val"] - P["val"]) <= 1e-6 * np.linalg.norm(P["val"])
        assert rel_err(Q["b"], P["b"]) <= 1e-6
        assert rel_err(Q["val"], P["val"]) > 0.0 or np.all(P["val"] == 0)  # it is a different computation
    prob.numeric()  # FP64 again
    prob.sync()
    for k, P in enumerate(r.parts):
        Q = prob.export_part(k)
        assert rel_err(Q["val"], P["val"]) <= 1e-12
    prob.close()


def test_fp32_refused_where_not_built(gpu_ctx):
    from gridap_amd._lib import GdxError

    assem, data = problem_from_workload(gpu_ctx, W.poisson2d_q1(8, (2, 2)))
    prob = assem._problem_for(data)
    with pytest.raises(GdxError) as e:
        prob.numeric(precision=1)
    assert e.value.code == -4
    prob.close()
