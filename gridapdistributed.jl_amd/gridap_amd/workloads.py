"""Problem descriptors for the assembly workloads.

Each descriptor is a plain dict describing a Poisson problem on a partitioned
Cartesian mesh, in the vocabulary of the reference's drivers
(test/PoissonTests.jl:10-45, test/FESpacesTests.jl:73-88):

  dim, order       Lagrangian Q_order reference FE (ReferenceFE(lagrangian,Float64,order))
  qdeg             Measure(Omega, qdeg): div(qdeg,2)+1 Gauss points per direction
  cells, domain    CartesianDiscreteModel(ranks, parts, domain, cells)
  parts            mesh partition (x-fastest rank numbering, src/PArraysExtras.jl:23)
  strategy         "sub" (SubAssembledRows) or "full" (FullyAssembledRows)
  coef_const,modes kappa(x) = coef_const * exp(sum_m a_m cos(2 pi k_m.x + phi_m))
  rhs, rhs_const   "const": f = rhs_const; "manufactured": f = -div(kappa grad uD)
  dirichlet, uD    Dirichlet on the whole boundary with uD = c0 + c1 x + c2 y + c3 z
  op, flux         "flux": the nonlinear operator of test/PLaplacianTests.jl:24-31 with
                   sigma(g) = (alpha + beta |g|^2)^((p-2)/2) g, flux = (p, alpha, beta)
  op, lame,        "elasticity": ReferenceFE(lagrangian, VectorValue{3,Float64}, order),
  body_force       a(u,v) = int( eps(v) : (lambda tr(eps(u)) I + 2 mu eps(u)) ),
                   l(v) = int( v . body_force ), lame = (lambda, mu); dirichlet
                   "clamped_x0" (u = 0 on the face x = x0 with its edges and
                   vertices) or False
"""
import numpy as np


def fourier_modes(seed, nmodes=16, kmax=8, sigma=0.5):
    """Seeded random Fourier modes (k1,k2,k3,a,phi): k_i uniform in 1..kmax,
    a ~ N(0, sigma^2), phi ~ U[0, 2 pi)."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, kmax + 1, size=(nmodes, 3)).astype(np.float64)
    a = rng.normal(0.0, sigma, size=nmodes)
    phi = rng.uniform(0.0, 2.0 * np.pi, size=nmodes)
    return np.column_stack([k, a, phi])


def poisson2d_q1(n=256, parts=(2, 2), strategy="sub"):
    """2D Poisson on (0,1)^2, Q1, quadrature degree 2, u = x + y, f = 0."""
    return dict(dim=2, order=1, qdeg=2, cells=(n, n), parts=tuple(parts), domain=(0.0, 1.0, 0.0, 1.0),
                strategy=strategy, coef_const=1.0, modes=np.zeros((0, 5)), rhs="const", rhs_const=0.0,
                dirichlet=True, uD=(0.0, 1.0, 1.0, 0.0))


def poisson3d_q2_fourier(n=64, parts=(2, 2, 2), seed=2024, strategy="sub"):
    """3D Poisson on (0,1)^3, Q2, quadrature degree 4, kappa = exp(g) with g a
    seeded sum of 16 Fourier modes, uD = x + y + z, f = -div(kappa grad uD)."""
    return dict(dim=3, order=2, qdeg=4, cells=(n, n, n), parts=tuple(parts),
                domain=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), strategy=strategy, coef_const=1.0,
                modes=fourier_modes(seed), rhs="manufactured", rhs_const=0.0, dirichlet=True,
                uD=(0.0, 1.0, 1.0, 1.0))


def poisson3d_q1(n, parts=(1, 1, 1), strategy="sub", modes=None, rhs_const=1.0):
    """3D Poisson Q1 on (0,1)^3, quadrature degree 2, homogeneous Dirichlet, f = rhs_const."""
    return dict(dim=3, order=1, qdeg=2, cells=(n, n, n) if np.isscalar(n) else tuple(n), parts=tuple(parts),
                domain=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), strategy=strategy, coef_const=1.0,
                modes=np.zeros((0, 5)) if modes is None else modes, rhs="const", rhs_const=rhs_const,
                dirichlet=True, uD=(0.0, 0.0, 0.0, 0.0))


def mapped_vertices(p, amp=0.2, seed=0, smooth=0.0):
    """Vertex coordinates of a mapped grid for workload p ((nvertices, dim), x
    fastest): the uniform box's vertices, interior ones moved at random by up
    to amp / 2 of the cell size per direction (numpy default_rng(seed)), plus a
    smooth displacement of amplitude smooth that vanishes on the box boundary,
    so the domain stays the box."""
    D = p["dim"]
    cells = list(p["cells"])[:D]
    dom = np.asarray(list(p["domain"])[:2 * D], np.float64).reshape(D, 2)
    h = (dom[:, 1] - dom[:, 0]) / np.asarray(cells)
    idx = np.stack(np.meshgrid(*[np.arange(c + 1) for c in cells], indexing="ij"), -1).reshape(-1, D, order="F")
    X = dom[:, 0] + idx * h
    rng = np.random.default_rng(seed)
    interior = np.all((idx > 0) & (idx < np.asarray(cells)), axis=1)
    Y = X.copy()
    Y[interior] += amp * h * rng.uniform(-0.5, 0.5, size=(int(interior.sum()), D))
    if smooth:
        s = (X - dom[:, 0]) / (dom[:, 1] - dom[:, 0])
        bump = np.prod(np.sin(np.pi * s), axis=1)
        Y += smooth * bump[:, None] * (dom[:, 1] - dom[:, 0]) * np.cos(np.pi * np.arange(1, D + 1) / 3.0)
    return Y


def lame_parameters(E, nu):
    """(lambda, mu) of Young's modulus E and Poisson's ratio nu."""
    return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))


def elasticity3d_q2(cells=(96, 80, 72), parts=(3, 2, 2), E=1.0, nu=0.3, body_force=(0.0, 0.0, -1.0),
                    clamped=True, order=2, domain=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), strategy="sub"):
    """3D linear elasticity, vector-valued Lagrangian Q2 (81 x 81 cell matrices),
    quadrature degree 4, E = 1, nu = 0.3, clamped at x = 0, body force (0,0,-1)."""
    return dict(dim=3, order=order, qdeg=2 * order, cells=tuple(cells), parts=tuple(parts), domain=tuple(domain),
                strategy=strategy, coef_const=1.0, modes=np.zeros((0, 5)), rhs="const", rhs_const=0.0,
                dirichlet="clamped_x0" if clamped else False, uD=(0.0, 0.0, 0.0, 0.0), op="elasticity",
                lame=lame_parameters(E, nu), body_force=tuple(float(v) for v in body_force))


def plaplacian3d_q1(n=96, parts=(4, 2, 2), p=3.0, eps=1e-2, strategy="sub"):
    """p-Laplacian on (0,1)^3, Q1, quadrature degree 2, f = 1, u = 0 on the boundary,
    sigma(g) = (eps^2 + |g|^2)^((p-2)/2) g (regularised at grad u = 0)."""
    return dict(dim=3, order=1, qdeg=2, cells=(n, n, n), parts=tuple(parts),
                domain=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), strategy=strategy, coef_const=1.0,
                modes=np.zeros((0, 5)), rhs="const", rhs_const=1.0, dirichlet=True, uD=(0.0, 0.0, 0.0, 0.0),
                op="flux", flux=(float(p), float(eps) ** 2, 1.0))


def flux2d_reference_test(parts=(2, 2), strategy="sub"):
    """The nonlinear problem of test/PLaplacianTests.jl:20-36: domain (0,4)^2,
    4x4 cells, Q1, Measure degree 2, sigma(g) = (1 + |g|^2) g, u = x + y, f = 0."""
    return dict(dim=2, order=1, qdeg=2, cells=(4, 4), parts=tuple(parts), domain=(0.0, 4.0, 0.0, 4.0),
                strategy=strategy, coef_const=1.0, modes=np.zeros((0, 5)), rhs="const", rhs_const=0.0,
                dirichlet=True, uD=(0.0, 1.0, 1.0, 0.0), op="flux", flux=(4.0, 1.0, 1.0))


def even_parts(k):
    """k parts factored as evenly as possible over (x, y, z)."""
    best = None
    for px in range(1, k + 1):
        for py in range(1, k // px + 1):
            if k % (px * py):
                continue
            pz = k // (px * py)
            f = tuple(sorted((px, py, pz), reverse=True))
            if best is None or max(f) - min(f) < max(best) - min(best):
                best = f
    return best


def batched_q1(seed=7, count=32, lo=16, hi=48):
    """Independent 3D Poisson Q1 problems of the batched workload: cube sizes
    uniform in lo..hi cells, 1..8 parts each (factored evenly), kappa from 8
    seeded Fourier modes, f = 1, homogeneous Dirichlet."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(lo, hi + 1))
        k = int(rng.integers(1, 9))
        q = poisson3d_q1(n, even_parts(k), modes=fourier_modes(seed * 1000 + i, 8, 4, 0.5), rhs_const=1.0)
        out.append(q)
    return out


def lattice_shape(p):
    """Global node lattice extents (x fastest): order * cells + 1 per direction."""
    K = p["order"]
    return tuple(K * c + 1 for c in p["cells"])


def lattice_coords(p):
    """Coordinates of every global lattice node, shape (nnodes, dim), x fastest."""
    shp = lattice_shape(p)
    axes = [np.linspace(p["domain"][2 * d], p["domain"][2 * d + 1], shp[d]) for d in range(p["dim"])]
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.ravel(order="F") for g in grids], axis=1)


def free_dofs(p):
    """Closed form for the number of free dofs with Dirichlet on the whole boundary."""
    K = p["order"]
    n = 1
    for c in p["cells"]:
        n *= K * c - 1
    return n


def nnz_free(p):
    """Closed form for nnz(free x free) with Dirichlet on the whole boundary:
    Q1: prod(3 n_d - 5), Q2: prod(8 n_d - 9)."""
    K = p["order"]
    n = 1
    for c in p["cells"]:
        n *= (3 * c - 5) if K == 1 else (8 * c - 9)
    return n
