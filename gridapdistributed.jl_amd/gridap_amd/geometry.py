"""Distributed Cartesian geometry (mirror of src/Geometry.jl:378-411, 660-682).

``with_gpu`` plays the role of ``with_debug``/``with_mpi`` (README.md:23-29):
all parts of the partition are subdomains on one GPU.
"""
import ctypes

from . import _lib


class DeviceContext:
    """One HIP device with the library's stream (gdx_init/gdx_finalize)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _lib.call("gdx_init", int(device), ctypes.byref(h))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            _lib.call("gdx_finalize", self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def with_gpu(f, device=0):
    """Run ``f(ctx)`` with a device context, like ``with_debug() do distribute ... end``."""
    with DeviceContext(device) as ctx:
        return f(ctx)


class CartesianDiscreteModel:
    """CartesianDiscreteModel(ranks, parts, domain, cells; map) (src/Geometry.jl:378-388).

    ``domain`` = (x0, x1, y0, y1[, z0, z1]); ``cells`` and ``parts`` per direction.
    Cells are split with the reference's uniform partition (src/PArraysExtras.jl:5-84).
    ``map`` (optional, the CartesianDescriptor's map) takes an (n, dim) array of
    points of the uniform box to their images; ``vertex_coords`` gives the mapped
    vertices directly ((cells_0 + 1) (cells_1 + 1) [(cells_2 + 1)] rows, x fastest).
    Cells then take the multilinear map of their vertices (diffusion only).
    """

    def __init__(self, ctx, parts, domain, cells, map=None, vertex_coords=None):
        import numpy as np

        parts, cells = tuple(int(p) for p in parts), tuple(int(c) for c in cells)
        domain = tuple(float(v) for v in domain)
        if len(parts) != len(cells) or len(domain) != 2 * len(cells):
            raise ValueError("parts, cells and domain must describe the same dimension")
        if any(c < p for c, p in zip(cells, parts)):
            raise ValueError("every part needs at least one cell per direction")
        self.ctx, self.parts, self.domain, self.cells = ctx, parts, domain, cells
        self.dim = len(cells)
        if map is not None and vertex_coords is not None:
            raise ValueError("give map or vertex_coords, not both")
        if map is not None:
            vertex_coords = map(self.uniform_vertices())
        self.vertex_coords = None
        if vertex_coords is not None:
            vc = np.ascontiguousarray(vertex_coords, dtype=np.float64)
            if vc.shape != (int(np.prod([c + 1 for c in cells])), self.dim):
                raise ValueError("vertex_coords must have one row of dim coordinates per mesh vertex")
            self.vertex_coords = vc

    def uniform_vertices(self):
        """The vertices of the uniform box, (nvertices, dim), x fastest."""
        import numpy as np

        axes = [np.linspace(self.domain[2 * d], self.domain[2 * d + 1], self.cells[d] + 1) for d in range(self.dim)]
        grids = np.meshgrid(*axes, indexing="ij")
        return np.stack([g.ravel(order="F") for g in grids], axis=1)

    def num_parts(self):
        n = 1
        for p in self.parts:
            n *= p
        return n


class Triangulation:
    """Integration domain: owned cells (SubAssembledRows) or owned+ghost cells
    (FullyAssembledRows), src/Geometry.jl:660-682, 797-847."""

    def __init__(self, model, portion=None):
        self.model, self.portion = model, portion


class Measure:
    """Measure(Omega, degree): tensor Gauss rule with degree//2+1 points per direction."""

    def __init__(self, trian, degree):
        self.trian, self.degree = trian, int(degree)
