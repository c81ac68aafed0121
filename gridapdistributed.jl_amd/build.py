"""Build libgdx.so (HIP for gfx950) in-tree.

    python gridapdistributed.jl_amd/build.py

Compiles every source under csrc/ with hipcc into ``gridapdistributed.jl_amd/libgdx.so``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = ["csrc/gdx_capi.hip", "csrc/gdx_kernels.hip", "csrc/gdx_march3d.hip", "csrc/gdx_tile3d.hip",
       "csrc/gdx_solve.hip",
       "csrc/gdx_flux.hip", "csrc/gdx_coef.hip",
       "csrc/gdx_nn.hip", "csrc/gdx_twogrid.hip", "csrc/gdx_batch.hip", "csrc/gdx_elast.hip", "csrc/gdx_mapped.hip",
       "csrc/gdx_pvector.hip",
       "csrc/gdx_symbolic.cpp", "csrc/gdx_symbolic_dev.hip"]
OUT = os.path.join(HERE, "libgdx.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    # no packed-FP32 instructions in generated device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    srcs = [os.path.join(HERE, s) for s in SRC]
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgdx.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
