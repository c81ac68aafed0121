"""End to end on device 0: assemble the affine problem of test/FESpacesTests.jl:21-27
(u = x + y, Q1, Dirichlet on the boundary, f = 0) with the GPU path, solve it
with the device PCG, and write the solution with writevtk (the output step of
test/VisualizationTests.jl); the written nodal values reproduce x + y."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from gridap_amd import (AffineFEOperator, CartesianDiscreteModel, ConstantSource, Diffusion, FEFunction, FESpace,
                        Float64, Integral, LinearFunction, Measure, ReferenceFE, Source, Triangulation, TrialFESpace,
                        lagrangian, writevtk)

pytestmark = pytest.mark.gpu


def test_solution_written_to_vtk(gpu_ctx, tmp_path):
    model = CartesianDiscreteModel(gpu_ctx, (2, 2), (0, 1, 0, 1), (16, 16))
    V = FESpace(model, ReferenceFE(lagrangian, Float64, 1), dirichlet_tags="boundary")
    U = TrialFESpace(V, LinearFunction(0.0, (1.0, 1.0)))
    dO = Measure(Triangulation(model), 2)
    op = AffineFEOperator(Integral(Diffusion(), dO), Integral(Source(ConstantSource(0.0)), dO), U, V)
    x, it, relres, _ = op.assem.problem.pcg(rtol=1e-12, maxiter=2000)
    assert relres <= 1e-10
    path = writevtk(Triangulation(model), str(tmp_path / "uh"), cellfields={"uh": FEFunction(U, x)})
    root = ET.parse(path).getroot()
    srcs = [p.get("Source") for p in root.iter("Piece")]
    assert len(srcs) == 4
    for s in srcs:
        piece = ET.parse(os.path.join(tmp_path, s)).getroot().find("UnstructuredGrid").find("Piece")
        vals = {da.get("Name"): np.array(da.text.split(), float) for da in piece.iter("DataArray")}
        pts = vals[None].reshape(-1, 3)
        assert np.abs(vals["uh"] - (pts[:, 0] + pts[:, 1])).max() < 1e-9
