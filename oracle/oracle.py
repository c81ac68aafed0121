"""ctypes front-end of the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (built from ``gdx_oracle.c`` by ``oracle/Makefile``)
and returns the per-part PSparseMatrix/PVector data the reference's
``assemble_matrix_and_vector`` would return (src/Algebra.jl:593-646,
src/Algebra.jl:1013-1015), with 0-based indices.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcDesc(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("order", ctypes.c_int), ("qdeg", ctypes.c_int), ("strategy", ctypes.c_int),
        ("cells", ctypes.c_longlong * 3), ("parts", ctypes.c_int * 3), ("domain", ctypes.c_double * 6),
        ("nmodes", ctypes.c_int), ("modes", ctypes.POINTER(ctypes.c_double)), ("coef_const", ctypes.c_double),
        ("rhs_kind", ctypes.c_int), ("rhs_const", ctypes.c_double),
        ("dir_kind", ctypes.c_int), ("dir_coef", ctypes.c_double * 4),
        ("op_kind", ctypes.c_int), ("flux_p", ctypes.c_double), ("flux_alpha", ctypes.c_double),
        ("flux_beta", ctypes.c_double), ("state", ctypes.POINTER(ctypes.c_double)),
        ("cell_kappa", ctypes.POINTER(ctypes.c_double)),
        ("lame_lambda", ctypes.c_double), ("lame_mu", ctypes.c_double), ("body_force", ctypes.c_double * 3),
        ("cell_lame", ctypes.POINTER(ctypes.c_double)),
        ("vertex_coords", ctypes.POINTER(ctypes.c_double)),
        ("weights_set", ctypes.c_int), ("diffusion_weight", ctypes.c_double), ("reaction", ctypes.c_double),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)
        L.orc_run.restype = vp
        L.orc_run.argtypes = [ctypes.POINTER(OrcDesc), ctypes.c_int]
        L.orc_nparts.argtypes = [vp]
        L.orc_ngdofs.restype = ctypes.c_longlong
        L.orc_ngdofs.argtypes = [vp]
        L.orc_times.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.orc_part_sizes.argtypes = [vp, ctypes.c_int, i64p]
        L.orc_part_fe.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.orc_part_matrix.argtypes = [vp, ctypes.c_int] + [vp] * 8
        L.orc_free.argtypes = [vp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = None
        L.orc_energy.restype = ctypes.c_double
        L.orc_energy.argtypes = [vp]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def make_desc(p):
    """p: problem dict (keys as produced by gridap_amd.workloads)."""
    d = OrcDesc()
    d.dim, d.order, d.qdeg = p["dim"], p["order"], p["qdeg"]
    d.strategy = 1 if p.get("strategy", "sub") in ("full", 1) else 0
    cells = list(p["cells"]) + [1] * (3 - len(p["cells"]))
    parts = list(p["parts"]) + [1] * (3 - len(p["parts"]))
    dom = list(p["domain"]) + [0.0, 1.0] * ((6 - len(p["domain"])) // 2)
    for i in range(3):
        d.cells[i], d.parts[i] = cells[i], parts[i]
    for i in range(6):
        d.domain[i] = dom[i]
    modes = np.ascontiguousarray(np.asarray(p.get("modes", np.zeros((0, 5))), dtype=np.float64).reshape(-1, 5))
    d.nmodes = modes.shape[0]
    d._modes_keep = modes
    d.modes = modes.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if modes.size else None
    d.coef_const = p.get("coef_const", 1.0)
    d.rhs_kind = {"const": 0, "manufactured": 1}[p.get("rhs", "const")]
    d.rhs_const = p.get("rhs_const", 0.0)
    dk = p.get("dirichlet", True)
    d.dir_kind = 2 if dk == "clamped_x0" else (1 if dk else 0)
    dc = list(p.get("uD", (0.0, 0.0, 0.0, 0.0)))
    for i in range(4):
        d.dir_coef[i] = dc[i]
    # nonlinear flux operator sigma(g) = (alpha + beta|g|^2)^((p-2)/2) g at a lattice state
    d.op_kind = {"diffusion": 0, "flux": 1, "elasticity": 2}[p.get("op", "diffusion")]
    # linear elasticity: Lame parameters and a constant body force
    lam, mu = p.get("lame", (0.0, 0.0))
    d.lame_lambda, d.lame_mu = lam, mu
    bf = list(p.get("body_force", (0.0, 0.0, 0.0)))
    for i in range(3):
        d.body_force[i] = bf[i]
    fp, fa, fb = p.get("flux", (2.0, 1.0, 0.0))
    d.flux_p, d.flux_alpha, d.flux_beta = fp, fa, fb
    st = p.get("state")
    if st is not None:
        st = np.ascontiguousarray(st, dtype=np.float64).ravel()
        d._state_keep = st
        d.state = st.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    else:
        d.state = None
    ck = p.get("cell_kappa")
    if ck is not None:
        ck = np.ascontiguousarray(ck, dtype=np.float64).ravel()
        d._ck_keep = ck
        d.cell_kappa = ck.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    else:
        d.cell_kappa = None
    cl = p.get("cell_lame")  # elasticity: (lambda, mu) per cell, x fastest
    if cl is not None:
        cl = np.ascontiguousarray(cl, dtype=np.float64).reshape(-1, 2).ravel()
        d._cl_keep = cl
        d.cell_lame = cl.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    else:
        d.cell_lame = None
    vc = p.get("vertex_coords")  # mapped grid: dim coordinates per mesh vertex, x fastest
    if vc is not None:
        vc = np.ascontiguousarray(vc, dtype=np.float64).ravel()
        d._vc_keep = vc
        d.vertex_coords = vc.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    else:
        d.vertex_coords = None
    # diffusion-reaction: a(u,v) = int( w_d kappa grad(u).grad(v) + r u v )
    if "weights" in p:
        d.weights_set = 1
        d.diffusion_weight, d.reaction = (float(v) for v in p["weights"])
    else:
        d.weights_set, d.diffusion_weight, d.reaction = 0, 1.0, 0.0
    return d


class OracleResult:
    def __init__(self, parts, ngdofs, times, energy=None):
        self.parts, self.ngdofs, self.times, self.energy = parts, ngdofs, times, energy


def energy(p):
    """Energy functional of the flux operator at p['state'] (see orc_energy)."""
    L = lib()
    d = make_desc(p)
    h = L.orc_run(ctypes.byref(d), 0)
    try:
        return L.orc_energy(h)
    finally:
        L.orc_free(h)


def run(p, inplace_reps=0, threads=1):
    """threads: OpenMP threads of the in-place re-assembly (one part per thread;
    1 = the parts one after another, as in the reference's debug mode)."""
    L = lib()
    L.orc_set_threads(int(threads))
    d = make_desc(p)
    h = L.orc_run(ctypes.byref(d), inplace_reps)
    try:
        n = L.orc_nparts(h)
        t = (ctypes.c_double * 3)()
        L.orc_times(h, t)
        parts = []
        for r in range(n):
            s = (ctypes.c_longlong * 5)()
            L.orc_part_sizes(h, r, s)
            nl, no, ngr, ngc, nnz = list(s)
            P = dict(n_ldofs=nl, n_own=no, n_ghost_rows=ngr, n_ghost_cols=ngc, nnz=nnz)
            P["ldof_gid"] = np.empty(nl, np.int64)
            P["ldof_owner"] = np.empty(nl, np.int32)
            P["ldof_gnode"] = np.empty(nl, np.int64)
            L.orc_part_fe(h, r, _ptr(P["ldof_gid"]), _ptr(P["ldof_owner"]), _ptr(P["ldof_gnode"]))
            nr, nc = no + ngr, no + ngc
            P["row_gid"] = np.empty(nr, np.int64)
            P["row_owner"] = np.empty(nr, np.int32)
            P["col_gid"] = np.empty(nc, np.int64)
            P["col_owner"] = np.empty(nc, np.int32)
            P["rowptr"] = np.empty(nr + 1, np.int64)
            P["colidx"] = np.empty(nnz, np.int32)
            P["val"] = np.empty(nnz, np.float64)
            P["b"] = np.empty(nr, np.float64)
            L.orc_part_matrix(h, r, *[_ptr(P[k]) for k in
                                      ("row_gid", "row_owner", "col_gid", "col_owner", "rowptr", "colidx", "val", "b")])
            parts.append(P)
        return OracleResult(parts, L.orc_ngdofs(h), tuple(t))
    finally:
        L.orc_free(h)


def global_coo(res):
    """Owned rows of every part as global (gid_row, gid_col, value) triplets."""
    I, J, V = [], [], []
    for P in res.parts:
        rp = P["rowptr"]
        no = P["n_own"]
        counts = np.diff(rp[: no + 1])
        rows = np.repeat(P["row_gid"][:no], counts)
        cols = P["col_gid"][P["colidx"][: rp[no]]]
        I.append(rows)
        J.append(cols)
        V.append(P["val"][: rp[no]])
    return np.concatenate(I), np.concatenate(J), np.concatenate(V)


def gid_to_gnode(res):
    """Global dof id -> global lattice node id (coordinate key)."""
    m = np.full(res.ngdofs, -1, np.int64)
    for P in res.parts:
        m[P["ldof_gid"]] = P["ldof_gnode"]
    return m
