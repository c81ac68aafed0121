"""Writes tests/golden/c1_small.npz: the oracle's PSparseMatrix/PVector parts for
2D Poisson Q1 on (0,1)^2, 6x6 cells, parts (2,2), u = x + y.  The reference
ships no golden data; this fixture pins the oracle against regressions and is
also compared against the GPU result by the parity tests.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "gridapdistributed.jl_amd"))

import oracle  # noqa: E402
from gridap_amd import workloads as W  # noqa: E402

r = oracle.run(W.poisson2d_q1(6, (2, 2)))
out = {}
for k, P in enumerate(r.parts):
    for key in ("row_gid", "col_gid", "rowptr", "colidx", "val", "b"):
        out[f"p{k}_{key}"] = P[key]
np.savez(os.path.join(HERE, "c1_small.npz"), **out)
print("wrote", os.path.join(HERE, "c1_small.npz"))
