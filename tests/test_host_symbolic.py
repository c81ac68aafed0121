"""The library's host symbolic phase (numbering, gids, PRanges, CSR pattern)
against the oracle, without a device (host-only problems), plus the C-ABI
surface: every symbol of include/gdx.h is exported and argument errors come
back as status codes.  CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from _helpers import STRUCT_KEYS
from gridap_amd import _lib, fespaces, problem_from_workload
from gridap_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASES = [
    W.poisson2d_q1(4, (2, 2)),
    W.poisson2d_q1(7, (2, 3)),
    W.poisson2d_q1(6, (3, 1), strategy="full"),
    W.poisson2d_q1(5, (1, 1)),
    dict(W.poisson2d_q1(6, (2, 2)), order=2, qdeg=4),
    W.poisson3d_q2_fourier(4, (2, 2, 2)),
    W.poisson3d_q2_fourier(5, (2, 1, 3)),
    W.poisson3d_q2_fourier(6, (1, 2, 2), strategy="full"),
    W.poisson3d_q2_fourier(3, (1, 1, 1)),
    W.poisson3d_q1(5, (2, 2, 1)),
    W.poisson3d_q1(7, (3, 2, 2), strategy="full"),
    dict(W.poisson3d_q1((5, 4, 6), (2, 1, 3)), dirichlet=False),
]


def host_problem(host_ctx, p):
    assem, data = problem_from_workload(host_ctx, p)
    desc, keep = fespaces.problem_desc(data, assem.strategy)
    gp = fespaces.GpuProblem(host_ctx, desc, keep)
    gp.symbolic()
    return gp


@pytest.mark.parametrize("p", CASES, ids=lambda p: f"{p['dim']}d-q{p['order']}-{p['cells']}-{p['parts']}-{p['strategy']}")
def test_structure_matches_oracle(host_ctx, p):
    r = oracle.run(p)
    gp = host_problem(host_ctx, p)
    assert gp.ngdofs == r.ngdofs
    for k, P in enumerate(r.parts):
        Q = gp.export_part(k)
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
        assert np.all(Q["val"] == 0.0)  # host-only problem exports an allocated (zero) matrix
    gp.close()


def test_owned_nnz_closed_form(host_ctx):
    p = W.poisson3d_q2_fourier(12, (2, 2, 2))
    gp = host_problem(host_ctx, p)
    assert gp.owned_nnz() == W.nnz_free(p)
    assert gp.ngdofs == W.free_dofs(p)
    gp.close()


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "gdx.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(gdx_\w+)\s*\(", hdr, re.M)))
    assert names and set(names) == set(_lib.EXPORTED)
    L = _lib.lib()
    for n in names:
        assert hasattr(L, n), n


def test_version_and_errors(host_ctx):
    L = _lib.lib()
    a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    assert L.gdx_version(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 0
    assert (a.value, b.value) == (0, 1)
    assert L.gdx_version(None, None, None) == -1
    # invalid descriptors come back as status codes with a message (no abort)
    assem, data = problem_from_workload(host_ctx, W.poisson2d_q1(4, (2, 2)))
    desc, _ = fespaces.problem_desc(data, assem.strategy)
    h = ctypes.c_void_p()
    bad = _lib.ProblemDesc.from_buffer_copy(desc)
    bad.dim = 4
    assert L.gdx_problem_create(None, ctypes.byref(bad), ctypes.byref(h)) == -1
    assert b"dim" in L.gdx_last_error()
    bad = _lib.ProblemDesc.from_buffer_copy(desc)
    bad.order = 3
    assert L.gdx_problem_create(None, ctypes.byref(bad), ctypes.byref(h)) == -4
    bad = _lib.ProblemDesc.from_buffer_copy(desc)
    bad.cells[0] = 1
    bad.parts[0] = 2
    assert L.gdx_problem_create(None, ctypes.byref(bad), ctypes.byref(h)) == -1
    # numeric on a host-only problem is refused, not crashed
    gp = fespaces.GpuProblem(host_ctx, desc)
    with pytest.raises(_lib.GdxError) as e:
        gp.numeric()
    assert e.value.code == -3  # GDX_ERR_STATE: before symbolic
    gp.symbolic()
    with pytest.raises(_lib.GdxError) as e:
        gp.numeric()
    assert e.value.code == -3  # host-only
    with pytest.raises(_lib.GdxError):
        gp.sizes(99)
    gp.close()


FLUX_CASES = [
    W.plaplacian3d_q1(5, (2, 2, 1)),
    W.plaplacian3d_q1(6, (1, 2, 2), strategy="full"),
    W.flux2d_reference_test((2, 2)),
    W.flux2d_reference_test((2, 1), "full"),
]


@pytest.mark.parametrize("p", FLUX_CASES, ids=lambda p: f"flux-{p['dim']}d-{p['cells']}-{p['parts']}-{p['strategy']}")
def test_flux_structure_matches_oracle(host_ctx, p):
    """The nonlinear operator shares the affine path's symbolic phase: same
    PRanges and CSR pattern as the oracle's Jacobian."""
    from gridap_amd import flux_operator_from_workload
    from gridap_amd.nonlinear import flux_problem_desc

    r = oracle.run(p)
    op = flux_operator_from_workload(host_ctx, p)
    desc, _ = flux_problem_desc(op, op.assem.strategy)
    assert desc.op_kind == _lib.OP_FLUX
    gp = fespaces.GpuProblem(host_ctx, desc)
    gp.symbolic()
    assert gp.ngdofs == r.ngdofs == W.free_dofs(p)
    for k, P in enumerate(r.parts):
        Q = gp.export_part(k)
        for key in STRUCT_KEYS:
            assert np.array_equal(P[key], Q[key]), (k, key)
    # a host-only problem has no device state
    with pytest.raises(_lib.GdxError) as e:
        gp.set_state(None)
    assert e.value.code == -3
    gp.close()


def test_flux_descriptor_errors(host_ctx):
    from gridap_amd import flux_operator_from_workload
    from gridap_amd.nonlinear import flux_problem_desc

    L = _lib.lib()
    op = flux_operator_from_workload(host_ctx, W.plaplacian3d_q1(4, (1, 1, 1)))
    desc, _ = flux_problem_desc(op, op.assem.strategy)
    h = ctypes.c_void_p()
    bad = _lib.ProblemDesc.from_buffer_copy(desc)
    bad.order = 2
    assert L.gdx_problem_create(None, ctypes.byref(bad), ctypes.byref(h)) == -4
    bad = _lib.ProblemDesc.from_buffer_copy(desc)
    bad.flux_alpha = -1.0
    assert L.gdx_problem_create(None, ctypes.byref(bad), ctypes.byref(h)) == -1
    bad = _lib.ProblemDesc.from_buffer_copy(desc)
    bad.op_kind = 7
    assert L.gdx_problem_create(None, ctypes.byref(bad), ctypes.byref(h)) == -1
    # Newton on an affine problem is refused
    assem, data = problem_from_workload(host_ctx, W.poisson2d_q1(4, (2, 2)))
    d2, _ = fespaces.problem_desc(data, assem.strategy)
    gp = fespaces.GpuProblem(host_ctx, d2)
    gp.symbolic()
    o = _lib.NewtonOpts(1e-10, 0.0, 10, 1e-8, 100, 0)
    rep = _lib.NewtonReport()
    assert L.gdx_newton(gp.handle, ctypes.byref(o), None, ctypes.byref(rep)) == -3
    gp.close()


@pytest.mark.parametrize("p", [W.poisson3d_q2_fourier(6, (2, 2, 2)), W.poisson3d_q1(7, (3, 2, 2), strategy="full"),
                               W.poisson2d_q1(9, (3, 2)),
                               W.elasticity3d_q2(cells=(5, 4, 3), parts=(3, 2, 1), clamped=True)],
                         ids=["q2-sub", "q1-full", "2d", "elasticity"])
def test_structure_independent_of_host_threads(host_ctx, p, monkeypatch):
    """The per-part loops of the symbolic phase run on several host threads;
    one thread and several give the same exports (pattern ids are merged in
    part order, offsets come from prefix sums)."""
    out = []
    for threads in ("1", "3"):
        monkeypatch.setenv("GDX_HOST_THREADS", threads)
        gp = host_problem(host_ctx, p)
        out.append([gp.export_part(k) for k in range(gp.nparts)])
        gp.close()
    for A, B in zip(*out):
        for key in STRUCT_KEYS:
            assert np.array_equal(A[key], B[key]), key
