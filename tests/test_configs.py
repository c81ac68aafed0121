"""configs/*.json record every seed and builder-fixed choice of the five
workloads; the workload definitions the tests and bench.py use must agree
with them.  CPU only."""
import json
import os
import re

import numpy as np

from gridap_amd import workloads as W

CFG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "configs")


def load(name):
    with open(os.path.join(CFG, name)) as fh:
        return json.load(fh)


def test_c1_c2_match_the_workloads():
    c1, c2 = load("c1.json"), load("c2.json")
    p1 = W.poisson2d_q1(**c1["args"])
    assert list(p1["cells"]) == c1["cells"] and list(p1["parts"]) == c1["parts"] and p1["qdeg"] == c1["qdeg"]
    assert W.free_dofs(p1) == c1["free_dofs"] and W.nnz_free(p1) == c1["nnz"]
    p2 = W.poisson3d_q2_fourier(**c2["args"])
    assert list(p2["cells"]) == c2["cells"] and list(p2["parts"]) == c2["parts"] and p2["qdeg"] == c2["qdeg"]
    assert np.array_equal(p2["modes"], W.fourier_modes(c2["seeds"]["fourier_modes"]))
    assert W.free_dofs(p2) == c2["free_dofs"] and W.nnz_free(p2) == c2["nnz"]
    k = p2["modes"][:, :3]
    assert k.min() >= 1 and k.max() <= 8 and len(k) == 16


def test_c3_c5_match_the_workloads():
    c3, c5 = load("c3.json"), load("c5.json")
    a = c3["args"]["batched_q1"]
    ws = W.batched_q1(**a)
    assert len(ws) == 32 and all(a["lo"] <= q["cells"][0] <= a["hi"] for q in ws)
    assert np.array_equal(ws[3]["modes"], W.fourier_modes(7000 + 3, 8, 4, 0.5))
    e = W.elasticity3d_q2(**{k: tuple(v) if isinstance(v, list) else v for k, v in c3["args"]["elasticity3d_q2"].items()})
    assert e["cells"] == (96, 80, 72) and e["parts"] == (3, 2, 2) and e["qdeg"] == 4
    p5 = W.plaplacian3d_q1(**{k: tuple(v) if isinstance(v, list) else v for k, v in c5["args"].items()})
    assert p5["flux"] == (3.0, 1e-4, 1.0) and p5["cells"] == (96, 96, 96) and p5["rhs_const"] == 1.0


def test_c5_network_training_matches_bench():
    """the network-on form's training length in the bench is the recorded one"""
    c5 = load("c5.json")
    src = open(os.path.join(os.path.dirname(CFG), "bench.py")).read()
    steps = int(re.search(r"^C5_TRAIN_STEPS = (\d+)", src, flags=re.M).group(1))
    assert f"{steps} Adam steps" in c5["builder_fixed"]["network_on_training"]
    assert f"k = 0 .. {steps - 1}" in c5["seeds"]["training_step_k"]
    assert "31 * k + 7" in src and "LearnedCoarseSolver(ctx, n // 2, 0)" in src


def test_c3_cellwise_seed_matches_bench():
    c3 = load("c3.json")
    src = open(os.path.join(os.path.dirname(CFG), "bench.py")).read()
    seed = int(re.search(r"^ELAST_CELLWISE_SEED = (\d+)", src, flags=re.M).group(1))
    assert c3["seeds"]["elasticity_cellwise_young_modulus"] == seed


def test_c2_mapped_seed_matches_bench():
    c2 = load("c2.json")
    src = open(os.path.join(os.path.dirname(CFG), "bench.py")).read()
    seed = int(re.search(r"^MAPPED_SEED = (\d+)", src, flags=re.M).group(1))
    assert c2["seeds"]["mapped_grid_vertices"] == seed
    assert "amp=0.5, seed=MAPPED_SEED, smooth=0.05" in src
