"""Shared comparison helpers for the parity tests."""
import numpy as np
import scipy.sparse as sp

STRUCT_KEYS = ("ldof_gid", "ldof_owner", "ldof_gnode", "row_gid", "row_owner", "col_gid", "col_owner", "rowptr",
               "colidx")


def owned_global(parts, ngdofs):
    """Owned rows of all parts as a global CSR (gid-indexed) and the RHS."""
    I, J, V, b = [], [], [], np.zeros(ngdofs)
    for P in parts:
        no, rp = P["n_own"], P["rowptr"]
        cnt = np.diff(rp[: no + 1])
        I.append(np.repeat(P["row_gid"][:no], cnt))
        J.append(P["col_gid"][P["colidx"][: rp[no]]])
        V.append(P["val"][: rp[no]])
        b[P["row_gid"][:no]] = P["b"][:no]
    A = sp.csr_matrix((np.concatenate(V), (np.concatenate(I), np.concatenate(J))), shape=(ngdofs, ngdofs))
    return A, b


def gid_to_gnode(parts, ngdofs):
    m = np.full(ngdofs, -1, np.int64)
    for P in parts:
        m[P["ldof_gid"]] = P["ldof_gnode"]
    return m


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale) if a.size else 0.0



def mapped_vertices(p, amp=0.2, seed=0, smooth=0.0):
    """a mapped grid's vertex coordinates (the workload definition's)"""
    from gridap_amd import workloads as W

    return W.mapped_vertices(p, amp=amp, seed=seed, smooth=smooth)
