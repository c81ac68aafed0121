"""Brute-force numpy micro-oracle (test infrastructure only).

A second, independent CPU restatement of the cell integrals of the assembly
path, used to pin the C oracle (gdx_oracle.c) and, row by row, the GPU result
at full size.  It shares no code with the C oracle: Gauss points come from
numpy.polynomial.legendre.leggauss, the Lagrange basis from numpy polynomials,
the geometry from the affine box map, and assembly is keyed by global lattice
node (no partitioning, no dof numbering), so it is insensitive to Gridap's
numbering.

It follows the integrand of the reference's drivers
(test/FESpacesTests.jl:16-17, test/PoissonTests.jl:34-37):
  a(u,v) = int kappa grad(u).grad(v),  l(v) = int f v,  b -= A u_D (lifting,
  src/FESpaces.jl:703-715).
Only tests/ may import it.
"""
import numpy as np
from numpy.polynomial import legendre as npleg
from numpy.polynomial import polynomial as nppoly


def gauss01(n):
    x, w = npleg.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def lagrange01(K):
    """Coefficient arrays of the 1D Lagrange basis on nodes j/K (tensor order)."""
    nodes = np.arange(K + 1) / K
    basis = []
    for a in range(K + 1):
        c = np.array([1.0])
        for j in range(K + 1):
            if j != a:
                c = nppoly.polymul(c, np.array([-nodes[j], 1.0]) / (nodes[a] - nodes[j]))
        basis.append(c)
    return basis


def kappa_f(p, X):
    """kappa and f at points X (n, 3)."""
    modes = np.asarray(p.get("modes", np.zeros((0, 5))), np.float64).reshape(-1, 5)
    th = 2 * np.pi * X @ modes[:, :3].T + modes[:, 4]
    g = np.cos(th) @ modes[:, 3] if len(modes) else np.zeros(len(X))
    kap = p.get("coef_const", 1.0) * np.exp(g)
    if p.get("rhs") == "manufactured":
        c = np.asarray(p["uD"][1:4])
        # f = -div(kappa grad uD) = -kappa grad(g).c
        dg = -(np.sin(th) * modes[:, 3]) * 2 * np.pi
        f = -kap * (dg @ (modes[:, :3] @ c)) if len(modes) else np.zeros(len(X))
    else:
        f = np.full(len(X), float(p.get("rhs_const", 0.0)))
    return kap, f


def _geometry(p):
    D, K = p["dim"], p["order"]
    cells = np.array(list(p["cells"]) + [1] * (3 - D))
    dom = np.array(list(p["domain"]) + [0.0, 1.0] * (3 - D)).reshape(3, 2)
    h = (dom[:, 1] - dom[:, 0]) / cells
    nn = np.where(np.arange(3) < D, K * cells + 1, 1)
    return D, K, cells, dom, h, nn


def element(p, cell):
    """Element matrix/vector of one cell in tensor node order (x fastest),
    plus the global node ids of the cell nodes."""
    D, K, cells, dom, h, nn = _geometry(p)
    nq = p["qdeg"] // 2 + 1
    xq, wq = gauss01(nq)
    B = lagrange01(K)
    phi = np.array([nppoly.polyval(xq, c) for c in B])             # [a, q]
    dphi = np.array([nppoly.polyval(xq, nppoly.polyder(c)) for c in B])
    n1 = K + 1
    # tensor node list
    idx = np.array(np.meshgrid(*[np.arange(n1) if d < D else np.array([0]) for d in range(3)], indexing="ij"))
    idx = idx.reshape(3, -1).T  # (i,j,k) with i fastest after reordering below
    order = np.lexsort((idx[:, 0], idx[:, 1], idx[:, 2]))
    idx = idx[order]
    qidx = np.array(np.meshgrid(*[np.arange(nq) if d < D else np.array([0]) for d in range(3)], indexing="ij"))
    qidx = qidx.reshape(3, -1).T
    qidx = qidx[np.lexsort((qidx[:, 0], qidx[:, 1], qidx[:, 2]))]
    # quadrature points in physical space and weights (affine box map)
    X = np.zeros((len(qidx), 3))
    w = np.ones(len(qidx))
    for d in range(D):
        X[:, d] = dom[d, 0] + h[d] * (cell[d] + xq[qidx[:, d]])
        w *= wq[qidx[:, d]] * h[d]
    kap, f = kappa_f(p, X)
    # basis values and physical gradients at quadrature points
    Phi = np.ones((len(idx), len(qidx)))
    G = np.ones((D, len(idx), len(qidx)))
    for d in range(D):
        Phi *= phi[idx[:, d]][:, qidx[:, d]]
        for e in range(D):
            G[e] *= (dphi[idx[:, d]][:, qidx[:, d]] / h[d]) if d == e else phi[idx[:, d]][:, qidx[:, d]]
    A = np.einsum("eiq,ejq,q->ij", G, G, kap * w)
    bvec = Phi @ (f * w)
    gnode = (K * cell[0] + idx[:, 0]) + nn[0] * ((K * cell[1] + idx[:, 1]) + nn[1] * (K * cell[2] + idx[:, 2]))
    return A, bvec, gnode


def node_coords(p, gnode):
    D, K, cells, dom, h, nn = _geometry(p)
    g = np.atleast_1d(gnode)
    ijk = np.stack([g % nn[0], (g // nn[0]) % nn[1], g // (nn[0] * nn[1])], 1)
    X = np.zeros((len(g), 3))
    for d in range(D):
        X[:, d] = dom[d, 0] + h[d] * ijk[:, d] / K
    return ijk, X


def is_dirichlet(p, gnode):
    D, K, cells, dom, h, nn = _geometry(p)
    if not p.get("dirichlet", True):
        return np.zeros(len(np.atleast_1d(gnode)), bool)
    ijk, _ = node_coords(p, gnode)
    out = np.zeros(len(ijk), bool)
    for d in range(D):
        out |= (ijk[:, d] == 0) | (ijk[:, d] == nn[d] - 1)
    return out


def uD(p, gnode):
    _, X = node_coords(p, gnode)
    c = p.get("uD", (0, 0, 0, 0))
    return c[0] + X @ np.asarray(c[1:4], np.float64)


def row(p, gnode_i):
    """Assembled row of the free node gnode_i: ({gnode_j: A_ij} over free j, b_i)."""
    D, K, cells, dom, h, nn = _geometry(p)
    ijk, _ = node_coords(p, gnode_i)
    ijk = ijk[0]
    ranges = []
    for d in range(3):
        if d >= D:
            ranges.append([0])
            continue
        cmin, cmax = max(-(-ijk[d] // K) - 1, 0), min(ijk[d] // K, cells[d] - 1)
        ranges.append(list(range(cmin, cmax + 1)))
    vals, bi = {}, 0.0
    for cz in ranges[2]:
        for cy in ranges[1]:
            for cx in ranges[0]:
                A, bv, gn = element(p, (cx, cy, cz))
                a = int(np.nonzero(gn == gnode_i)[0][0])
                dirm = is_dirichlet(p, gn)
                ud = np.where(dirm, uD(p, gn), 0.0)
                bi += bv[a] - A[a] @ ud
                for j, g in enumerate(gn):
                    if not dirm[j]:
                        vals[int(g)] = vals.get(int(g), 0.0) + A[a, j]
    return vals, bi


def assemble_global(p):
    """Global free x free matrix keyed by lattice node (dense dict of rows) and RHS."""
    D, K, cells, dom, h, nn = _geometry(p)
    I, J, V, bI, bV = [], [], [], [], []
    for cz in range(cells[2]):
        for cy in range(cells[1]):
            for cx in range(cells[0]):
                A, bv, gn = element(p, (cx, cy, cz))
                dirm = is_dirichlet(p, gn)
                ud = np.where(dirm, uD(p, gn), 0.0)
                free = ~dirm
                bv = bv - A @ ud
                ii, jj = np.meshgrid(gn[free], gn[free], indexing="ij")
                I.append(ii.ravel())
                J.append(jj.ravel())
                V.append(A[np.ix_(free, free)].ravel())
                bI.append(gn[free])
                bV.append(bv[free])
    return np.concatenate(I), np.concatenate(J), np.concatenate(V), np.concatenate(bI), np.concatenate(bV)


def assembled_factors_1d(p, d):
    """1D matrices assembled over the mesh along direction d (dense, nn_d x nn_d):
    K = int phi_i' phi_j', M = int phi_i phi_j, P = int phi_i' phi_j,
    Q = int phi_i phi_j', and m = int phi_i, with the problem's Gauss rule."""
    D, K, cells, dom, h, nn = _geometry(p)
    nq = p["qdeg"] // 2 + 1
    xq, wq = gauss01(nq)
    B = lagrange01(K)
    phi = np.array([nppoly.polyval(xq, c) for c in B])
    dphi = np.array([nppoly.polyval(xq, nppoly.polyder(c)) for c in B]) / h[d]
    n = nn[d]
    Km, Mm, Pm, Qm, m = (np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n)), np.zeros(n))
    w = wq * h[d]
    for e in range(cells[d]):
        s = slice(K * e, K * e + K + 1)
        Km[s, s] += np.einsum("aq,bq,q->ab", dphi, dphi, w)
        Mm[s, s] += np.einsum("aq,bq,q->ab", phi, phi, w)
        Pm[s, s] += np.einsum("aq,bq,q->ab", dphi, phi, w)
        Qm[s, s] += np.einsum("aq,bq,q->ab", phi, dphi, w)
        m[s] += phi @ w
    return Km, Mm, Pm, Qm, m


def cell_factors_1d(p, d):
    """The 1D matrices of assembled_factors_1d before the sum over cells: for
    each cell e along direction d, (K, M, P, Q) of that cell alone (nn_d x nn_d,
    nonzero on the cell's K + 1 nodes)."""
    D, K, cells, dom, h, nn = _geometry(p)
    nq = p["qdeg"] // 2 + 1
    xq, wq = gauss01(nq)
    B = lagrange01(K)
    phi = np.array([nppoly.polyval(xq, c) for c in B])
    dphi = np.array([nppoly.polyval(xq, nppoly.polyder(c)) for c in B]) / h[d]
    w = wq * h[d]
    out = []
    for e in range(cells[d]):
        s = slice(K * e, K * e + K + 1)
        f = [np.zeros((nn[d], nn[d])) for _ in range(4)]
        f[0][s, s] = np.einsum("aq,bq,q->ab", dphi, dphi, w)
        f[1][s, s] = np.einsum("aq,bq,q->ab", phi, phi, w)
        f[2][s, s] = np.einsum("aq,bq,q->ab", dphi, phi, w)
        f[3][s, s] = np.einsum("aq,bq,q->ab", phi, dphi, w)
        out.append(f)
    return out


def elasticity_global(p):
    """Linear elasticity on the whole lattice, dof = 3 * lattice node + component
    (x-fastest nodes, no Dirichlet elimination): the matrix as a scipy sparse
    matrix and the body-force vector.  Built from 1D factors as the sum of
    Kronecker products of the bilinear form's separable terms,
      (a, b) block = lambda d_a(v) d_b(u) + mu d_b(v) d_a(u) + mu [a == b] grad(v).grad(u),
    an independent route from the C oracle's per-cell 3D quadrature.  With
    constant Lame parameters the factors are assembled over the 1D mesh; with
    p["cell_lame"] ((lambda, mu) per cell, x fastest) every cell adds its own
    Kronecker terms, weighted by its pair."""
    F = [assembled_factors_1d(p, d) for d in range(3)]
    m = np.kron(F[2][4], np.kron(F[1][4], F[0][4]))
    bf = np.asarray(p["body_force"], np.float64)
    bvec = (m[:, None] * bf[None, :]).ravel()
    if p.get("cell_lame") is None:
        lam, mu = p["lame"]
        KMPQ = [dict(K=F[d][0], M=F[d][1], P=F[d][2], Q=F[d][3]) for d in range(3)]
        return _elasticity_blocks(KMPQ, lam, mu).tocsr(), bvec
    cells = [int(c) for c in p["cells"]]
    lame = np.asarray(p["cell_lame"], np.float64).reshape(-1, 2)
    C = [cell_factors_1d(p, d) for d in range(3)]
    A = None
    for ez in range(cells[2]):
        for ey in range(cells[1]):
            for ex in range(cells[0]):
                lam, mu = lame[ex + cells[0] * (ey + cells[1] * ez)]
                KMPQ = [dict(zip("KMPQ", C[d][e])) for d, e in enumerate((ex, ey, ez))]
                t = _elasticity_blocks(KMPQ, lam, mu)
                A = t if A is None else A + t
    return A.tocsr(), bvec


def _elasticity_blocks(KMPQ, lam, mu):
    """sum over the 3 x 3 component blocks of the separable terms, with the 1D
    factors KMPQ[d] = {"K", "M", "P", "Q"}; dof = 3 * node + component"""
    import scipy.sparse as sp

    def kron3(f0, f1, f2):  # x fastest: node = x + nx (y + ny z)
        return sp.kron(sp.csr_matrix(f2), sp.kron(sp.csr_matrix(f1), sp.csr_matrix(f0)))

    def term(kinds):  # kinds[d] in K, M, P, Q
        return kron3(KMPQ[0][kinds[0]], KMPQ[1][kinds[1]], KMPQ[2][kinds[2]])

    A = None
    for a in range(3):
        for b in range(3):
            if a == b:
                blk = None
                for e in range(3):
                    kinds = ["M", "M", "M"]
                    kinds[e] = "K"
                    t = term(kinds) * (mu + (lam + mu if e == a else 0.0))
                    blk = t if blk is None else blk + t
            else:
                k1 = ["M", "M", "M"]
                k1[a], k1[b] = "P", "Q"  # lambda d_a(phi_i) d_b(phi_j)
                k2 = ["M", "M", "M"]
                k2[a], k2[b] = "Q", "P"  # mu d_b(phi_i) d_a(phi_j)
                blk = lam * term(k1) + mu * term(k2)
            E = sp.csr_matrix(([1.0], ([a], [b])), shape=(3, 3))
            t = sp.kron(blk, E)
            A = t if A is None else A + t
    return A


def elasticity_free_dofs(p):
    """Lattice dofs (3 * node + component) that are free: all, or those off the
    clamped face x = x0."""
    D, K, cells, dom, h, nn = _geometry(p)
    nodes = np.arange(int(np.prod(nn)))
    free = np.ones(len(nodes), bool) if p.get("dirichlet") != "clamped_x0" else (nodes % nn[0]) != 0
    return np.repeat(free, 3)


def uniform_vertices(p):
    """The mesh vertices of the uniform box, (nvertices, dim), x fastest."""
    D, K, cells, dom, h, nn = _geometry(p)
    axes = [dom[d, 0] + h[d] * np.arange(cells[d] + 1) for d in range(D)]
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.ravel(order="F") for g in grids], axis=1)


def mapped_node_coords(p, gnode):
    """Coordinates of lattice nodes on a mapped grid (p["vertex_coords"]): the
    image of the node's reference position under the multilinear map of a cell
    holding it, i.e. the mean of the vertices of the edge, face or cell whose
    midpoint it is (a vertex itself)."""
    D, K, cells, dom, h, nn = _geometry(p)
    vc = np.asarray(p["vertex_coords"], np.float64).reshape(-1, D)
    nv = [c + 1 for c in cells]
    ijk, _ = node_coords(p, gnode)
    X = np.zeros((len(ijk), 3))
    for n, g in enumerate(ijk):
        corners = [[g[d] // K] if g[d] % K == 0 else [g[d] // K, g[d] // K + 1] for d in range(D)]
        pts = []
        for c in np.ndindex(*[len(cs) for cs in corners]):
            v = [corners[d][c[d]] for d in range(D)]
            idx = v[0] + nv[0] * (v[1] + (nv[1] * v[2] if D == 3 else 0))
            pts.append(vc[idx])
        X[n, :D] = np.mean(pts, axis=0)
    return X


def affine_mapped_global(p, B):
    """Diffusion with constant kappa = coef_const and constant f = rhs_const on the
    uniform box lattice mapped by x = B X + t (no Dirichlet elimination): the
    matrix keyed by lattice node and the RHS.  With grad_x = B^-T grad_X,
      a(phi_i, phi_j) = sum_de C_de int d_d(phi_i) d_e(phi_j) dX,
      C = kappa |det B| B^-1 B^-T,
    and int d_d(phi_i) d_e(phi_j) is a Kronecker product of the 1D assembled
    factors (K along d when d == e; P along d and Q along e otherwise; M along
    the rest).  With p["weights"] = (w_d, r) the form is
    w_d kappa grad(u).grad(v) + r u v, the second term r |det B| M x M (x M)."""
    import scipy.sparse as sp

    D = p["dim"]
    B = np.asarray(B, np.float64)
    F = [assembled_factors_1d(p, d) for d in range(D)]
    Binv = np.linalg.inv(B)
    C = p.get("coef_const", 1.0) * abs(np.linalg.det(B)) * Binv @ Binv.T

    def kron(fs):  # x fastest
        out = sp.csr_matrix(fs[D - 1])
        for d in range(D - 2, -1, -1):
            out = sp.kron(out, sp.csr_matrix(fs[d]))
        return out

    wd, rw = p.get("weights", (1.0, 0.0))
    A = rw * abs(np.linalg.det(B)) * kron([F[k][1] for k in range(D)])
    for d in range(D):
        for e in range(D):
            fs = [F[k][1] for k in range(D)]  # M
            if d == e:
                fs[d] = F[d][0]  # K
            else:
                fs[d], fs[e] = F[d][2], F[e][3]  # P along d, Q along e
            A = A + wd * C[d, e] * kron(fs)
    m = F[D - 1][4]
    for d in range(D - 2, -1, -1):
        m = np.kron(m, F[d][4])
    return A.tocsr(), p.get("rhs_const", 0.0) * abs(np.linalg.det(B)) * m
