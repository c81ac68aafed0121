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
    """CartesianDiscreteModel(ranks, parts, domain, cells) (src/Geometry.jl:378-388).

    ``domain`` = (x0, x1, y0, y1[, z0, z1]); ``cells`` and ``parts`` per direction.
    Cells are split with the reference's uniform partition (src/PArraysExtras.jl:5-84).
    """

    def __init__(self, ctx, parts, domain, cells):
        parts, cells, domain = tuple(int(p) for p in parts), tuple(int(c) for c in cells), tuple(map(float, domain))
        if len(parts) != len(cells) or len(domain) != 2 * len(cells):
            raise ValueError("parts, cells and domain must describe the same dimension")
        if any(c < p for c, p in zip(cells, parts)):
            raise ValueError("every part needs at least one cell per direction")
        self.ctx, self.parts, self.domain, self.cells = ctx, parts, domain, cells
        self.dim = len(cells)

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
