"""The host code built with AddressSanitizer and UndefinedBehaviorSanitizer.

Two programs are compiled into a temporary folder and run; any sanitizer report
fails them:
- the library's host symbolic stages (gridapdistributed.jl_amd/csrc/gdx_symbolic.cpp,
  the path of problems created without a device) on 2D/3D, Q1/Q2, both
  strategies, the flux operator and elasticity, then a tile rebuild
  (tests/sanitize/symbolic_driver.cpp; three host threads);
- the C oracle on the same kinds of cases, fresh and in-place assembly, with
  one and three OpenMP threads (tests/sanitize/oracle_driver.c).
CPU only; skipped where the compilers lack the sanitizer runtimes."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]


def build_and_run(cmd, exe, env=None):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        if "asan" in (r.stderr + r.stdout).lower() and "cannot find" in (r.stderr + r.stdout):
            pytest.skip("sanitizer runtime not installed")
        raise AssertionError(r.stdout + r.stderr)
    e = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", **(env or {}))
    out = subprocess.run([exe], capture_output=True, text=True, env=e, timeout=300)
    text = out.stdout + out.stderr
    assert out.returncode == 0, text[-4000:]
    assert "AddressSanitizer" not in text and "runtime error" not in text and "LeakSanitizer" not in text, text[-4000:]
    return out.stdout


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"), reason="needs g++ and HIP headers")
def test_host_symbolic_stages_under_sanitizers(tmp_path):
    exe = str(tmp_path / "symbolic_driver")
    src = os.path.join(ROOT, "gridapdistributed.jl_amd", "csrc")
    cmd = ["g++", "-std=c++17", *SAN, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{src}",
           os.path.join(ROOT, "tests", "sanitize", "symbolic_driver.cpp"), os.path.join(src, "gdx_symbolic.cpp"),
           "-o", exe, "-lpthread"]
    out = build_and_run(cmd, exe, env={"GDX_HOST_THREADS": "3"})
    assert "all host symbolic cases clean" in out and out.count("case ") == 8


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_driver")
    cmd = ["gcc", "-std=gnu11", *SAN, "-fopenmp", f"-I{os.path.join(ROOT, 'oracle')}",
           os.path.join(ROOT, "tests", "sanitize", "oracle_driver.c"), "-o", exe, "-lm"]
    out = build_and_run(cmd, exe)
    assert "all oracle cases clean" in out and out.count("case ") == 16
