"""writevtk of distributed triangulations, models and FE functions (mirror of
src/Visualization.jl and test/VisualizationTests.jl: uh = interpolate(x+y) on a
partitioned 8x8 model, written per part plus a .pvtu index).  The files are
parsed back: pieces, point and cell counts, VTK cell types and connectivity,
the nodal values against the function at the written coordinates, and the
model's gid / part cell data against the partition.  CPU only (host-only
problems for the dof numbering)."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from gridap_amd import (CartesianDiscreteModel, FESpace, Float64, LinearFunction, ReferenceFE, Triangulation,
                        TrialFESpace, VectorValue, interpolate, lagrangian, writevtk)
from gridap_amd.visualization import part_cell_boxes


class HostOnly:
    handle = None


def arrays(piece):
    out = {}
    for sec in ("PointData", "CellData", "Points", "Cells"):
        for da in piece.find(sec).findall("DataArray"):
            vals = np.array(da.text.split(), dtype=float)
            nc = int(da.get("NumberOfComponents", "1"))
            out[(sec, da.get("Name"))] = vals.reshape(-1, nc) if nc > 1 else vals
    return out


def read_pvtu(path):
    root = ET.parse(path).getroot()
    assert root.get("type") == "PUnstructuredGrid"
    srcs = [p.get("Source") for p in root.find("PUnstructuredGrid").findall("Piece")]
    pieces = []
    for s in srcs:
        r = ET.parse(os.path.join(os.path.dirname(path), s)).getroot()
        assert r.get("type") == "UnstructuredGrid"
        pc = r.find("UnstructuredGrid").find("Piece")
        pieces.append((s, int(pc.get("NumberOfPoints")), int(pc.get("NumberOfCells")), arrays(pc)))
    return pieces


@pytest.mark.parametrize("dim,order,parts,cells", [(2, 1, (2, 2), (8, 8)), (2, 2, (3, 1), (7, 5)),
                                                    (3, 1, (2, 1, 2), (4, 5, 6)), (3, 2, (1, 2, 2), (3, 4, 4))])
def test_writevtk_triangulation_with_fe_function(tmp_path, dim, order, parts, cells):
    domain = (0.0, 1.0, 0.0, 2.0, -1.0, 1.0)[:2 * dim]
    model = CartesianDiscreteModel(HostOnly(), parts, domain, cells)
    V = FESpace(model, ReferenceFE(lagrangian, Float64, order), dirichlet_tags="boundary")
    f = LinearFunction(0.5, (1.0, 1.0, -2.0)[:dim])
    U = TrialFESpace(V, f)
    uh = interpolate(f, U)
    path = writevtk(Triangulation(model), str(tmp_path / "t1"), cellfields={"uh": uh})
    pieces = read_pvtu(path)
    assert [p[0] for p in pieces] == [os.path.join("t1", f"t1_{r + 1}.vtu") for r in range(model.num_parts())]
    assert sum(p[2] for p in pieces) == int(np.prod(cells))  # owned cells only
    boxes = part_cell_boxes(model, ghost=False)
    nv = 4 if dim == 2 else 8
    for r, (src, npts, ncell, A) in enumerate(pieces):
        ext = [b[1] - b[0] for b in boxes[r]]
        assert ncell == int(np.prod(ext)) and npts == int(np.prod([e + 1 for e in ext]))
        assert np.all(A[("CellData", "piece")] == r + 1)
        assert np.all(A[("Cells", "types")] == (9 if dim == 2 else 12))
        assert np.array_equal(A[("Cells", "offsets")], nv * np.arange(1, ncell + 1))
        pts = A[("Points", None)]
        # nodal values = the interpolated affine function at the written points (free and Dirichlet nodes)
        expect = 0.5 + pts[:, :dim] @ np.array((1.0, 1.0, -2.0)[:dim])
        assert np.abs(A[("PointData", "uh")] - expect).max() < 1e-12
        # every cell's vertices span a box of the mesh size (VTK vertex order)
        conn = A[("Cells", "connectivity")].astype(np.int64).reshape(ncell, nv)
        h = np.array([(domain[2 * d + 1] - domain[2 * d]) / cells[d] for d in range(dim)])
        d01 = pts[conn[:, 1], :dim] - pts[conn[:, 0], :dim]
        d03 = pts[conn[:, 3], :dim] - pts[conn[:, 0], :dim]
        assert np.allclose(d01, np.eye(dim)[0] * h) and np.allclose(d03, np.eye(dim)[1] * h)
        if dim == 3:
            assert np.allclose(pts[conn[:, 4], :dim] - pts[conn[:, 0], :dim], np.eye(3)[2] * h)


def test_writevtk_model_cell_data(tmp_path):
    model = CartesianDiscreteModel(HostOnly(), (2, 3), (0, 1, 0, 1), (6, 7))
    pieces = read_pvtu(writevtk(model, str(tmp_path / "model")))
    owned = part_cell_boxes(model, ghost=False)
    seen = np.zeros(6 * 7, int)
    for r, (src, npts, ncell, A) in enumerate(pieces):
        gid, part = A[("CellData", "gid")].astype(int), A[("CellData", "part")].astype(int)
        own = part == r + 1
        seen[gid[own] - 1] += 1
        for g, q in zip(gid, part):  # the owner's box holds the cell
            cx, cy = (g - 1) % 6, (g - 1) // 6
            (x0, x1), (y0, y1) = owned[q - 1]
            assert x0 <= cx < x1 and y0 <= cy < y1
    assert np.all(seen == 1)  # every cell owned by exactly one part


def test_writevtk_vector_field(tmp_path):
    model = CartesianDiscreteModel(HostOnly(), (2, 1, 1), (0, 1, 0, 1, 0, 1), (3, 2, 2))
    V = FESpace(model, ReferenceFE(lagrangian, VectorValue(3, Float64), 1), dirichlet_tags="clamped_x0")
    from gridap_amd.nonlinear import FEFunction
    from gridap_amd.visualization import dof_numbering, node_coords
    gid = dof_numbering(V)
    x = node_coords(model, 1)
    free = np.zeros(int(gid.max()) + 1)
    for c in range(3):
        sel = gid[:, c] >= 0
        free[gid[sel, c]] = (c + 1) * x[sel, 0]  # u_c = (c+1) x, zero on the clamped face x = 0
    uh = FEFunction(TrialFESpace(V, LinearFunction()), free)
    pieces = read_pvtu(writevtk(Triangulation(model), str(tmp_path / "u"), cellfields={"u": uh}))
    for src, npts, ncell, A in pieces:
        u, pts = A[("PointData", "u")], A[("Points", None)]
        assert u.shape == (npts, 3)
        assert np.abs(u - pts[:, :1] * np.array([1.0, 2.0, 3.0])).max() < 1e-12
