"""VTK output of distributed triangulations, models and FE functions (mirror of
src/Visualization.jl:1-250: writevtk(trian, filebase; cellfields), writevtk(model,
filebase)).

One XML UnstructuredGrid piece per part, written as WriteVTK's pvtk_grid names
them (``filebase/<name>_<part>.vtu``, parts 1-based), and a PUnstructuredGrid
index ``filebase.pvtu`` listing them:

- ``writevtk(Triangulation(model), filebase, cellfields={"uh": uh})``: each part's
  owned cells (the SubAssembledRows triangulation, src/Geometry.jl:813-847) with
  the cell data ``piece`` = part id (src/Visualization.jl:78) and the nodal values
  of each FE function at the cells' vertices;
- ``writevtk(model, filebase)``: each part's local cells (owned and one ghost
  layer) with the cell data ``gid`` (global cell id, 1-based) and ``part``
  (owner), as src/Visualization.jl:31-36 does for the model's cells.

FE functions are given by their free values in global-id order (FEFunction,
interpolate) and the space's Dirichlet data; vector-valued functions write
NumberOfComponents = 3.  Data arrays are ASCII (the reference's ``ascii=true``
form of the same files).
"""
import os

import numpy as np

from .fespaces import (ConstantCoefficient, ConstantSource, Diffusion, Integral, LinearFunction, Source, TrialFESpace,
                       problem_desc)
from .geometry import CartesianDiscreteModel, Measure, Triangulation
from .nonlinear import FEFunction

VTK_QUAD, VTK_HEXAHEDRON = 9, 12
# VTK vertex order of a quad / hex as offsets of the cell's lower corner
_VTK_CORNERS = {2: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)],
                3: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]}


def local_range(p, np_, n, ghost):
    """Cells [lo, hi) of part p (0-based) of np_ along a direction of n cells,
    src/PArraysExtras.jl:72-84: the last n % np_ parts get one more cell;
    ghost adds one cell on interior sides."""
    l, rem = divmod(n, np_)
    pp = p + 1
    offset = l * (pp - 1)
    if rem >= np_ - pp + 1:
        l += 1
        offset += pp - (np_ - rem) - 1
    return max(offset - ghost, 0), min(l + offset + ghost, n)


def part_cell_boxes(model, ghost):
    """Per part (x-fastest rank order, src/PArraysExtras.jl:23): its cell box
    (lo, hi) per direction, owned cells only or with the one-cell ghost layer."""
    boxes = []
    D = model.dim
    for r in range(model.num_parts()):
        pc, rr = [], r
        for d in range(D):
            pc.append(rr % model.parts[d])
            rr //= model.parts[d]
        boxes.append([local_range(pc[d], model.parts[d], model.cells[d], 1 if ghost else 0) for d in range(D)])
    return boxes


class _HostOnly:
    handle = None


def dof_numbering(space):
    """Global node -> free global id (-1 on Dirichlet nodes) and, for vector
    spaces, the id of each (node, component): the numbering gdx_symbolic gives
    the space's dofs (the same as any assembler over this space), from a
    host-only problem.  Cached on the space."""
    V = space.space if isinstance(space, TrialFESpace) else space
    cached = getattr(V, "_gdx_numbering", None)
    if cached is not None:
        return cached
    from .fespaces import Elasticity, BodyForce, GpuProblem, SubAssembledRows

    m = V.model
    dO = Measure(Triangulation(m), 2 * V.reffe.order)
    if V.reffe.ncomp > 1:
        a, l = Integral(Elasticity(1.0, 1.0), dO), Integral(BodyForce((0.0, 0.0, 0.0)), dO)
    else:
        a, l = Integral(Diffusion(ConstantCoefficient(1.0)), dO), Integral(Source(ConstantSource(0.0)), dO)
    U = TrialFESpace(V, LinearFunction())
    desc, keep = problem_desc(dict(U=U, V=V, a=a, l=l), SubAssembledRows())
    gp = GpuProblem(_HostOnly(), desc, keep)
    gp.symbolic()
    nc = V.reffe.ncomp
    K = V.reffe.order
    nnodes = int(np.prod([K * c + 1 for c in m.cells]))
    gid = np.full((nnodes, nc), -1, np.int64)
    for r in range(gp.nparts):
        P = gp.export_part(r)
        gn, g = P["ldof_gnode"], P["ldof_gid"]
        # the components of a node are its local dofs in order
        comp = _component_of(gn) if nc > 1 else 0
        gid[gn, comp] = g
    gp.close()
    V._gdx_numbering = gid
    return gid


def _component_of(gnode):
    out = np.zeros(len(gnode), np.int64)
    seen = {}
    for i, n in enumerate(gnode.tolist()):
        out[i] = seen.get(n, 0)
        seen[n] = out[i] + 1
    return out


def node_coords(model, order):
    """Coordinates (x fastest) of the global node lattice of order `order`.  On a
    mapped model a node sits at its cells' multilinear map of its reference
    position: a vertex at its mapped coordinates, a Q2 edge, face or cell
    midpoint at the mean of that entity's mapped vertices."""
    D = model.dim
    shp = [order * c + 1 for c in model.cells]
    if model.vertex_coords is None:
        axes = [np.linspace(model.domain[2 * d], model.domain[2 * d + 1], shp[d]) for d in range(D)]
        grids = np.meshgrid(*axes, indexing="ij")
        return np.stack([g.ravel(order="F") for g in grids], axis=1)
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in shp], indexing="ij"), -1).reshape(-1, D, order="F")
    v0, two = idx // order, (idx % order) != 0
    nv = [c + 1 for c in model.cells]
    acc = np.zeros((len(idx), D))
    cnt = np.zeros(len(idx))
    for s in np.ndindex(*([2] * D)):
        s = np.asarray(s)
        ok = np.all((s == 0) | two, axis=1)
        v = np.minimum(v0 + s, np.asarray(nv) - 1)
        flat = v[:, 0] + (v[:, 1] * nv[0] if D > 1 else 0) + (v[:, 2] * nv[0] * nv[1] if D > 2 else 0)
        acc += np.where(ok[:, None], model.vertex_coords[flat], 0.0)
        cnt += ok
    return acc / cnt[:, None]


def interpolate(f, space):
    """interpolate(f, V) for an affine f (LinearFunction): the FE function whose
    free values are f at the free nodes (exact for Q1 and Q2)."""
    if not isinstance(f, LinearFunction):
        raise NotImplementedError("interpolation is lowered for affine functions")
    V = space.space if isinstance(space, TrialFESpace) else space
    gid = dof_numbering(V)
    x = node_coords(V.model, V.reffe.order)
    vals = f.coef[0] + x @ np.asarray(f.coef[1:1 + V.model.dim])
    free = np.zeros(int(gid.max()) + 1)
    for c in range(gid.shape[1]):
        sel = gid[:, c] >= 0
        free[gid[sel, c]] = vals[sel]
    U = space if isinstance(space, TrialFESpace) else TrialFESpace(V, f)
    return FEFunction(U, free)


def nodal_values(uh):
    """Values of an FE function at every global lattice node: free values by
    global id, Dirichlet nodes from the trial space's data (zero for a test space)."""
    S = uh.space
    V = S.space if isinstance(S, TrialFESpace) else S
    gid = dof_numbering(V)
    x = node_coords(V.model, V.reffe.order)
    D = V.model.dim
    if isinstance(S, TrialFESpace):
        uD = S.uD.coef[0] + x @ np.asarray(S.uD.coef[1:1 + D])
    else:
        uD = np.zeros(len(x))
    fv = uh.free_values if uh.free_values is not None else np.zeros(int(gid.max()) + 1)
    out = np.empty(gid.shape)
    for c in range(gid.shape[1]):
        sel = gid[:, c] >= 0
        out[sel, c] = fv[gid[sel, c]]
        out[~sel, c] = uD[~sel] if gid.shape[1] == 1 else 0.0
    return out


def _fmt(a):
    a = np.asarray(a)
    if a.dtype.kind == "f":
        return " ".join(repr(float(v)) for v in a.ravel())
    return " ".join(str(int(v)) for v in a.ravel())


def _vtk_type(a):
    a = np.asarray(a)
    if a.dtype.kind == "f":
        return "Float64"
    if a.dtype == np.uint8:
        return "UInt8"
    return "Int64"


def _piece(path, points, conn, offsets, types, pointdata, celldata):
    with open(path, "w") as fh:
        fh.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="1.0" '
                 'byte_order="LittleEndian" header_type="UInt64">\n<UnstructuredGrid>\n')
        fh.write(f'<Piece NumberOfPoints="{len(points)}" NumberOfCells="{len(offsets)}">\n<PointData>\n')
        for k, v in pointdata.items():
            nc = v.shape[1] if v.ndim == 2 else 1
            fh.write(f'<DataArray type="{_vtk_type(v)}" Name="{k}" NumberOfComponents="{nc}" format="ascii">\n'
                     f'{_fmt(v)}\n</DataArray>\n')
        fh.write('</PointData>\n<CellData>\n')
        for k, v in celldata.items():
            fh.write(f'<DataArray type="{_vtk_type(v)}" Name="{k}" format="ascii">\n{_fmt(v)}\n</DataArray>\n')
        fh.write('</CellData>\n<Points>\n<DataArray type="Float64" NumberOfComponents="3" format="ascii">\n'
                 f'{_fmt(points)}\n</DataArray>\n</Points>\n<Cells>\n')
        fh.write(f'<DataArray type="Int64" Name="connectivity" format="ascii">\n{_fmt(conn)}\n</DataArray>\n')
        fh.write(f'<DataArray type="Int64" Name="offsets" format="ascii">\n{_fmt(offsets)}\n</DataArray>\n')
        fh.write(f'<DataArray type="UInt8" Name="types" format="ascii">\n{_fmt(types)}\n</DataArray>\n')
        fh.write('</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n')


def _write(model, filebase, boxes, cellfields, extra_celldata):
    D = model.dim
    name = os.path.basename(filebase)
    os.makedirs(filebase, exist_ok=True)
    fields = {k: nodal_values(v) for k, v in (cellfields or {}).items()}
    orders = {k: (v.space.space if isinstance(v.space, TrialFESpace) else v.space).reffe.order
              for k, v in (cellfields or {}).items()}
    nn1 = [c + 1 for c in model.cells]  # vertex lattice
    xv = node_coords(model, 1)
    sources = []
    for r, box in enumerate(boxes):
        lo, hi = [b[0] for b in box], [b[1] for b in box]
        ext = [h - l for h, l in zip(hi, lo)]
        ncell = int(np.prod(ext))
        if ncell == 0:  # empty parts are left out, as filter_empty_parts does (src/Visualization.jl:182-190)
            continue
        vlo, vext = lo, [e + 1 for e in ext]
        # the part's vertices, x fastest, and their global vertex ids
        grids = np.meshgrid(*[np.arange(vlo[d], vlo[d] + vext[d]) for d in range(D)], indexing="ij")
        vc = np.stack([g.ravel(order="F") for g in grids], axis=1)
        gv = vc[:, 0] + (vc[:, 1] * nn1[0] if D > 1 else 0) + (vc[:, 2] * nn1[0] * nn1[1] if D > 2 else 0)
        pts = np.zeros((len(gv), 3))
        pts[:, :D] = xv[gv]
        cgrids = np.meshgrid(*[np.arange(ext[d]) for d in range(D)], indexing="ij")
        cc = np.stack([g.ravel(order="F") for g in cgrids], axis=1)
        conn = []
        for off in _VTK_CORNERS[D]:
            v = cc[:, 0] + off[0] + (cc[:, 1] + off[1]) * vext[0]
            if D == 3:
                v = v + (cc[:, 2] + off[2]) * vext[0] * vext[1]
            conn.append(v)
        conn = np.stack(conn, axis=1)
        nv = len(_VTK_CORNERS[D])
        offsets = nv * np.arange(1, ncell + 1)
        types = np.full(ncell, VTK_HEXAHEDRON if D == 3 else VTK_QUAD, np.uint8)
        pdata = {}
        for k, vals in fields.items():
            K = orders[k]
            nnK = [K * c + 1 for c in model.cells]
            gK = K * vc[:, 0] + (K * vc[:, 1] * nnK[0] if D > 1 else 0) + \
                (K * vc[:, 2] * nnK[0] * nnK[1] if D > 2 else 0)
            pdata[k] = vals[gK] if vals.shape[1] > 1 else vals[gK, 0]
        gcell = cc + np.asarray(lo)
        cdata = extra_celldata(r, gcell)
        rel = os.path.join(name, f"{name}_{r + 1}.vtu")
        _piece(os.path.join(os.path.dirname(filebase) or ".", rel), pts, conn, offsets, types, pdata, cdata)
        sources.append((rel, pdata, cdata))
    with open(filebase + ".pvtu", "w") as fh:
        fh.write('<?xml version="1.0"?>\n<VTKFile type="PUnstructuredGrid" version="1.0" '
                 'byte_order="LittleEndian" header_type="UInt64">\n<PUnstructuredGrid GhostLevel="0">\n<PPointData>\n')
        pd, cd = (sources[0][1], sources[0][2]) if sources else ({}, {})
        for k, v in pd.items():
            nc = v.shape[1] if v.ndim == 2 else 1
            fh.write(f'<PDataArray type="{_vtk_type(v)}" Name="{k}" NumberOfComponents="{nc}"/>\n')
        fh.write('</PPointData>\n<PCellData>\n')
        for k, v in cd.items():
            fh.write(f'<PDataArray type="{_vtk_type(v)}" Name="{k}"/>\n')
        fh.write('</PCellData>\n<PPoints>\n<PDataArray type="Float64" NumberOfComponents="3"/>\n</PPoints>\n')
        for rel, _, _ in sources:
            fh.write(f'<Piece Source="{rel}"/>\n')
        fh.write('</PUnstructuredGrid>\n</VTKFile>\n')
    return filebase + ".pvtu"


def writevtk(obj, filebase, cellfields=None):
    """writevtk(trian_or_model, filebase; cellfields) of src/Visualization.jl:192-203.
    Returns the path of the .pvtu index."""
    if isinstance(obj, Triangulation):
        model = obj.model
        boxes = part_cell_boxes(model, ghost=False)
        return _write(model, filebase, boxes, cellfields,
                      lambda r, gcell: {"piece": np.full(len(gcell), r + 1, np.int64)})
    if isinstance(obj, CartesianDiscreteModel):
        if cellfields:
            raise NotImplementedError("cellfields are written with a Triangulation")
        model = obj
        boxes = part_cell_boxes(model, ghost=True)
        owned = part_cell_boxes(model, ghost=False)
        D = model.dim

        def owner_of(gcell):
            own = np.zeros(len(gcell), np.int64)
            for q, ob in enumerate(owned):
                inside = np.ones(len(gcell), bool)
                for d in range(D):
                    inside &= (gcell[:, d] >= ob[d][0]) & (gcell[:, d] < ob[d][1])
                own[inside] = q + 1
            return own

        def cdata(r, gcell):
            g = gcell[:, 0].copy()
            stride = 1
            for d in range(1, D):
                stride *= model.cells[d - 1]
                g += gcell[:, d] * stride
            return {"gid": g + 1, "part": owner_of(gcell)}

        return _write(model, filebase, boxes, None, cdata)
    raise NotImplementedError("writevtk takes a Triangulation or a CartesianDiscreteModel")
