"""Device PVectors (row and column PRanges) on device 0 against the host
semantics of gridap_amd.algebra.PVector (PArrays' consistent! and assemble!):

- consistent!, assemble!, change_ghost and copy_rhs are bit-identical to the
  host results (same additions in the same order);
- mul!(y, A, x) equals every part's local CSR times its consistent local x
  (scipy, 1e-13 relative) and, on owned rows, the global device SpMV;
- axpy! is y = a x + y followed by consistent!;
- after the problem is destroyed its PVectors report GDX_ERR_STATE.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from gridap_amd import _lib, problem_from_workload
from gridap_amd import workloads as W
from gridap_amd._lib import GdxError
from gridap_amd.algebra import LocalIndices, PVector

pytestmark = pytest.mark.gpu

CASES = [W.poisson3d_q2_fourier(6, (2, 2, 2)), W.poisson3d_q1(9, (3, 2, 2), strategy="full"),
         W.poisson2d_q1(12, (2, 3)), W.elasticity3d_q2(cells=(5, 4, 3), parts=(2, 2, 1), clamped=True)]
IDS = ["q2", "q1-full", "2d", "elasticity"]


def setup(ctx, p):
    assem, data = problem_from_workload(ctx, p)
    prob = assem._problem_for(data)
    prob.numeric()
    prob.sync()
    parts = [prob.export_part(k) for k in range(prob.nparts)]
    return prob, parts


def ranges(prob, parts, which):
    out = []
    for k, P in enumerate(parts):
        no = P["n_own"]
        g, o = (P["row_gid"], P["row_owner"]) if which == _lib.PRANGE_ROWS else (P["col_gid"], P["col_owner"])
        out.append(LocalIndices(k, prob.ngdofs, g[:no], g[no:], o[no:]))
    return out


def host_vector(flat, ips):
    cuts = np.cumsum([0] + [ip.local_length() for ip in ips])
    return PVector([flat[cuts[i]:cuts[i + 1]].copy() for i in range(len(ips))], ips)


@pytest.mark.parametrize("p", CASES, ids=IDS)
@pytest.mark.parametrize("which", [_lib.PRANGE_ROWS, _lib.PRANGE_COLS], ids=["rows", "cols"])
def test_consistent_assemble_bitwise(gpu_ctx, p, which):
    print("pvector case", p["dim"], p["order"], p["cells"], p["parts"], which, flush=True)
    prob, parts = setup(gpu_ctx, p)
    ips = ranges(prob, parts, which)
    v = prob.pvector(which)
    assert v.n == sum(ip.local_length() for ip in ips)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(v.n)
    got = v.set(x).consistent().get()
    assert np.array_equal(got, host_vector(x, ips).consistent().concatenated())
    got = v.set(x).assemble().get()
    assert np.array_equal(got, host_vector(x, ips).assemble().concatenated())
    v.close()
    prob.close()


@pytest.mark.parametrize("p", CASES, ids=IDS)
def test_mul_axpy_change_ghost_rhs(gpu_ctx, p):
    print("pvector ops", p["dim"], p["order"], p["cells"], p["parts"], flush=True)
    prob, parts = setup(gpu_ctx, p)
    rows, cols = ranges(prob, parts, _lib.PRANGE_ROWS), ranges(prob, parts, _lib.PRANGE_COLS)
    rng = np.random.default_rng(3)
    x = prob.pvector(_lib.PRANGE_COLS)
    y = prob.pvector(_lib.PRANGE_ROWS)
    xv = rng.standard_normal(x.n)  # ghost entries are garbage until mul! makes x consistent
    prob.mul(y, x.set(xv))
    xh = host_vector(xv, cols).consistent()
    assert np.array_equal(x.get(), xh.concatenated())
    yh = []
    for P, xl in zip(parts, xh.partition):
        nr = len(P["rowptr"]) - 1
        yh.append(sp.csr_matrix((P["val"], P["colidx"], P["rowptr"]), shape=(nr, len(xl))) @ xl)
    yh = np.concatenate(yh)
    yd = y.get()
    scale = max(np.abs(yh).max(), 1e-300)
    assert np.abs(yd - yh).max() <= 1e-13 * scale
    # owned rows against the global device SpMV
    g = xh._own_global()
    ref = prob.spmv(g)
    cuts = np.cumsum([0] + [ip.local_length() for ip in rows])
    for k, ip in enumerate(rows):
        no = ip.own_length()
        if no:
            assert np.abs(yd[cuts[k]:cuts[k] + no] - ref[ip.own_to_global]).max() <= 1e-13 * scale
    # axpy!(a, x, y) on the column PRange: y = a x + y, then consistent!
    z = prob.pvector(_lib.PRANGE_COLS)
    zv = rng.standard_normal(z.n)
    got = z.set(zv).axpy(0.75, x).get()
    zh = host_vector(0.75 * x.get() + zv, cols).consistent().concatenated()
    assert np.abs(got - zh).max() <= 1e-15 * np.abs(zh).max()
    # change_ghost: y's own values onto the column PRange
    w = prob.pvector(_lib.PRANGE_COLS)
    got = w.fill(7.0).change_ghost_from(y.set(yh), make_consistent=False).get()
    exp = host_vector(yh, rows)
    want = [np.concatenate([pv[: ip.own_length()], np.zeros(cp.ghost_length())])
            for pv, ip, cp in zip(exp.partition, rows, cols)]
    assert np.array_equal(got, np.concatenate(want))
    got = w.change_ghost_from(y, make_consistent=True).get()
    assert np.array_equal(got, host_vector(np.concatenate(want), cols).consistent().concatenated())
    # copy_rhs: the assembled b on the row PRange
    assert np.array_equal(y.copy_rhs().get(), np.concatenate([P["b"] for P in parts]))
    prob.close()
    with pytest.raises(GdxError) as e:
        x.get()
    assert e.value.code == -3
    for v in (x, y, z, w):
        v.close()


def test_argument_errors(gpu_ctx):
    prob, parts = setup(gpu_ctx, W.poisson3d_q1(5, (2, 1, 1)))
    x, y = prob.pvector(_lib.PRANGE_COLS), prob.pvector(_lib.PRANGE_ROWS)
    with pytest.raises(GdxError) as e:
        prob.mul(x, y)  # y and x swapped
    assert e.value.code == -1
    with pytest.raises(GdxError):
        x.copy_rhs()  # b lives on the row PRange
    with pytest.raises(GdxError):
        prob.pvector(7)
    other, _ = setup(gpu_ctx, W.poisson3d_q1(5, (2, 1, 1)))
    z = other.pvector(_lib.PRANGE_COLS)
    with pytest.raises(GdxError):
        z.axpy(1.0, x)
    for v in (x, y, z):
        v.close()
    prob.close()
    other.close()
