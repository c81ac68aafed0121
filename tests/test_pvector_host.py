"""PArrays semantics of PVector.consistent() / assemble() (the host mirror the
device PVectors are checked against), on the oracle's row and column PRanges.
CPU only.

Pins: consistent! makes every copy of a global id equal to its owner's value
and is idempotent; assemble! zeroes the ghosts and conserves the sum over all
local entries; on vectors with one nonzero copy per global id, assemble! of
the copies equals the global vector exactly; the local CSR of every part times
its consistent local x equals the global matrix times x on the owned rows."""
import numpy as np
import pytest

import oracle
from _helpers import owned_global
from gridap_amd import workloads as W
from gridap_amd.algebra import LocalIndices, PVector


def ranges(res, which):
    out = []
    for k, P in enumerate(res.parts):
        no = P["n_own"]
        g, o = (P["row_gid"], P["row_owner"]) if which == "rows" else (P["col_gid"], P["col_owner"])
        out.append(LocalIndices(k, res.ngdofs, g[:no], g[no:], o[no:]))
    return out


CASES = [W.poisson3d_q2_fourier(4, (2, 2, 2)), W.poisson3d_q1(5, (3, 2, 1), strategy="full"),
         W.poisson2d_q1(7, (2, 3)), W.elasticity3d_q2(cells=(3, 2, 2), parts=(2, 1, 2), clamped=True)]


@pytest.mark.parametrize("p", CASES, ids=["q2", "q1-full", "2d", "elasticity"])
@pytest.mark.parametrize("which", ["rows", "cols"])
def test_consistent_and_assemble(p, which):
    res = oracle.run(p)
    ips = ranges(res, which)
    rng = np.random.default_rng(5)
    v = PVector([rng.standard_normal(ip.local_length()) for ip in ips], ips)
    c = v.copy().consistent()
    own = c._own_global()
    for x, ip in zip(c.partition, ips):
        assert np.array_equal(x[ip.own_length():], own[ip.ghost_to_global])
    assert all(np.array_equal(a, b) for a, b in zip(c.copy().consistent().partition, c.partition))
    total = sum(x.sum() for x in v.partition)
    a = v.copy().assemble()
    assert all(np.all(x[ip.own_length():] == 0.0) for x, ip in zip(a.partition, ips))
    assert abs(sum(x.sum() for x in a.partition) - total) <= 1e-12 * max(1.0, abs(total)) * len(ips)
    # one nonzero copy per global id: assemble! recovers the global vector exactly
    gv = rng.standard_normal(res.ngdofs)
    seen = np.zeros(res.ngdofs, bool)
    parts = []
    for ip in ips:
        x = np.zeros(ip.local_length())
        for lid, g in enumerate(ip.local_to_global()):
            if not seen[g] and lid >= ip.own_length():
                x[lid] = gv[g]
                seen[g] = True
        parts.append(x)
    for x, ip in zip(parts, ips):
        for lid in range(ip.own_length()):
            g = ip.own_to_global[lid]
            if not seen[g]:
                x[lid] = gv[g]
                seen[g] = True
    assert np.array_equal(PVector(parts, ips).assemble()._own_global(), gv)


@pytest.mark.parametrize("p", CASES, ids=["q2", "q1-full", "2d", "elasticity"])
def test_local_products_equal_global_product(p):
    import scipy.sparse as sp

    res = oracle.run(p)
    A, _ = owned_global(res.parts, res.ngdofs)
    cols = ranges(res, "cols")
    rng = np.random.default_rng(9)
    xg = rng.standard_normal(res.ngdofs)
    x = PVector([rng.standard_normal(ip.local_length()) for ip in cols], cols)
    for v, ip in zip(x.partition, cols):
        v[: ip.own_length()] = xg[ip.own_to_global]
    x.consistent()
    ref = A @ xg
    for P, xl in zip(res.parts, x.partition):
        nr = len(P["rowptr"]) - 1
        Al = sp.csr_matrix((P["val"], P["colidx"], P["rowptr"]), shape=(nr, len(xl)))
        y = Al @ xl
        no = P["n_own"]
        if no:
            assert np.abs(y[:no] - ref[P["row_gid"][:no]]).max() <= 1e-13 * np.abs(ref).max()
