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
    """Vertex coordinates of a mapped grid for workload p: the uniform box's
    vertices, interior ones moved by up to amp (in units of the local cell size)
    at random, plus a smooth displacement of amplitude smooth that vanishes on
    the box boundary (the domain stays the box)."""
    import micro

    D = p["dim"]
    cells = list(p["cells"])[:D]
    X = micro.uniform_vertices(p)
    dom = np.asarray(list(p["domain"])[:2 * D], np.float64).reshape(D, 2)
    h = (dom[:, 1] - dom[:, 0]) / np.asarray(cells)
    rng = np.random.default_rng(seed)
    idx = np.stack(np.meshgrid(*[np.arange(c + 1) for c in cells], indexing="ij"), -1).reshape(-1, D, order="F")
    interior = np.all((idx > 0) & (idx < np.asarray(cells)), axis=1)
    Y = X.copy()
    Y[interior] += amp * h * rng.uniform(-0.5, 0.5, size=(interior.sum(), D))
    if smooth:
        s = (X - dom[:, 0]) / (dom[:, 1] - dom[:, 0])
        bump = np.prod(np.sin(np.pi * s), axis=1)
        Y += smooth * bump[:, None] * (dom[:, 1] - dom[:, 0]) * np.cos(np.pi * np.arange(1, D + 1) / 3.0)
    return Y
