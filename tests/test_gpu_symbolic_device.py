"""The device stage of the symbolic phase (local CSR, position patterns, row
table and tiles built in HBM) against the host stage, on device 0.

Every case is set up twice on the device: once through the device stage (the
default for problems on a context) and once with GDX_HOST_SYMBOLIC=1, which
makes the host stage build the same structure.  The exported structure
(numbering, PRanges, rowptr, colidx) must be identical, and the numeric phase
through every 3D kernel variant must give bit-identical values and RHS (the
kernels read the position patterns, the row table and the tiles, so a
different pattern id order or tile list would change what they write or
fail).  The oracle comparison of the same structure runs in the parity tests.
"""
import os

import numpy as np
import pytest

import oracle
from _helpers import STRUCT_KEYS
from gridap_amd import fespaces, problem_from_workload
from gridap_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [
    W.poisson2d_q1(9, (2, 2)),
    W.poisson2d_q1(8, (3, 1), strategy="full"),
    dict(W.poisson2d_q1(7, (2, 2)), order=2, qdeg=4),
    W.poisson3d_q2_fourier(6, (2, 2, 2)),
    W.poisson3d_q2_fourier(7, (2, 1, 3), strategy="full"),
    W.poisson3d_q2_fourier(3, (1, 1, 1)),
    W.poisson3d_q1(9, (2, 2, 1)),
    W.poisson3d_q1(11, (3, 2, 2), strategy="full"),
    dict(W.poisson3d_q1((6, 5, 7), (2, 1, 3)), dirichlet=False),
    W.plaplacian3d_q1(8, (2, 2, 1)),
    W.elasticity3d_q2((6, 5, 4), (3, 2, 2)),
    W.elasticity3d_q2((5, 4, 4), (2, 2, 1), clamped=False, strategy="full"),
    W.elasticity3d_q2((5, 6, 4), (2, 2, 2), order=1),
]


def ident(p):
    return f"{p.get('op', 'diffusion')}-{p['dim']}d-q{p['order']}-{p['cells']}-{p['parts']}-{p['strategy']}"


def build(ctx, p, host):
    assem, data = problem_from_workload(ctx, p)
    desc, keep = fespaces.problem_desc(data, assem.strategy)
    gp = fespaces.GpuProblem(ctx, desc, keep)
    if host:
        os.environ["GDX_HOST_SYMBOLIC"] = "1"
    try:
        gp.symbolic()
    finally:
        os.environ.pop("GDX_HOST_SYMBOLIC", None)
    return gp


def exported(gp):
    gp.sync()
    return [gp.export_part(k) for k in range(gp.nparts)]


@pytest.mark.parametrize("p", CASES, ids=ident)
def test_device_stage_equals_host_stage(gpu_ctx, p):
    print("device symbolic case", ident(p), flush=True)
    dev, host = build(gpu_ctx, p, False), build(gpu_ctx, p, True)
    assert dev.ngdofs == host.ngdofs and dev.nparts == host.nparts
    assert dev.owned_nnz() == host.owned_nnz()
    variants = [(0, 32)]
    if p["dim"] == 3 and p.get("op", "diffusion") == "diffusion":
        variants += [(0, 16), (1, 32), (2, 32)] + ([(3, 32)] if p["order"] == 2 else [])
    first = None
    for v in variants:
        for gp in (dev, host):
            if len(variants) > 1:
                gp.set_kernel(*v)
            gp.numeric()
        A, B = exported(dev), exported(host)
        for k, (Pa, Pb) in enumerate(zip(A, B)):
            for key in STRUCT_KEYS:
                assert np.array_equal(Pa[key], Pb[key]), (v, k, key)
            assert np.array_equal(Pa["val"].view(np.int64), Pb["val"].view(np.int64)), (v, k)
            assert np.array_equal(Pa["b"].view(np.int64), Pb["b"].view(np.int64)), (v, k)
        if first is None:
            first = A
    # and the structure is the oracle's
    r = oracle.run(p)
    for k, P in enumerate(r.parts):
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], first[k][key]), (k, key)
    dev.close()
    host.close()


def test_device_stage_twice_and_after_host_copies(gpu_ctx):
    """A second symbolic phase on the same problem re-makes the device
    structure; host copies made in between (an export, a solve's tables) do
    not go stale."""
    p = W.poisson3d_q2_fourier(5, (2, 2, 1))
    gp = build(gpu_ctx, p, False)
    gp.numeric()
    ref = exported(gp)
    x, it, rr, _ = gp.pcg(1e-10, 2000)
    assert rr <= 1e-10
    gp.symbolic()
    gp.numeric()
    for Pa, Pb in zip(ref, exported(gp)):
        for key in STRUCT_KEYS + ("val", "b"):
            assert np.array_equal(Pa[key], Pb[key]), key
    x2, it2, rr2, _ = gp.pcg(1e-10, 2000)
    assert it2 == it and np.array_equal(x, x2)
    gp.close()


@pytest.mark.parametrize("p", [CASES[3], CASES[9], CASES[10]], ids=ident)
def test_host_stage_takes_over_after_a_hash_collision(gpu_ctx, p):
    """GDX_FAULT_HASH=1 reports a collision in the device pattern table: the
    host CSR stage then builds the structure from the device numbering's host
    copies, with the same result."""
    dev = build(gpu_ctx, p, False)
    os.environ["GDX_FAULT_HASH"] = "1"
    try:
        fb = build(gpu_ctx, p, False)
    finally:
        os.environ.pop("GDX_FAULT_HASH", None)
    for gp in (dev, fb):
        gp.numeric()
    for Pa, Pb in zip(exported(dev), exported(fb)):
        for key in STRUCT_KEYS:
            assert np.array_equal(Pa[key], Pb[key]), key
        assert np.array_equal(Pa["val"].view(np.int64), Pb["val"].view(np.int64))
        assert np.array_equal(Pa["b"].view(np.int64), Pb["b"].view(np.int64))
    dev.close()
    fb.close()


def test_resymbolic_detaches_networks(gpu_ctx):
    """A second symbolic phase re-makes the solver's tables, which an attached
    network refers to: the network is detached (FPCG with it is refused until
    it is attached again), and works after re-attachment."""
    from gridap_amd import LearnedCoarseSolver, fpcg
    from gridap_amd._lib import GdxError

    p = W.poisson3d_q1(8, (2, 1, 1))
    gp = build(gpu_ctx, p, False)
    gp.numeric()
    net = LearnedCoarseSolver(gpu_ctx, 4, 0)
    net.attach(gp)
    net.set_options(8, use_network=False)
    x1, it1, r1, _, _ = fpcg(gp, net, None, 1e-10, 500)
    gp.symbolic()
    gp.numeric()
    with pytest.raises(GdxError):
        fpcg(gp, net, None, 1e-10, 500)
    net.attach(gp)
    net.set_options(8, use_network=False)  # options belong to the attachment
    x2, it2, r2, _, _ = fpcg(gp, net, None, 1e-10, 500)
    assert r1 <= 1e-10 and r2 <= 1e-10 and it1 == it2 and np.array_equal(x1, x2)
    net.close()
    gp.close()
