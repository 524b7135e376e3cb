"""The C++ drop-in (include/vsdmo.hpp) end to end on the GPU: the example
shaped like the reference CLI's dmoVolumeTet (reference src/main.cu:51-73)
is compiled with g++ against libvsdmo.so and run; it optimises a Kuhn cube
through DMO::DMOTetClass, then records training pairs, trains the network
and runs the learned proposal with the deferred fallback, then a batch of
two cubes through vsdmo_cpp::Batch with the network as its tet proposal."""
import os
import shutil
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "volumesurfacediscretemeshoptimization_amd")


def test_cpp_dropin_example_runs(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "dropin_example")
    r = subprocess.run([gxx, "-std=c++17", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "dropin_example.cpp"), "-o", exe, "-L", PKG, "-lvsdmo",
                        "-Wl,-rpath," + PKG], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout, r.stderr)
    assert "min quality" in r.stdout and "q_min" in r.stdout
    assert "learned:" in r.stdout, r.stdout
    assert "batch:" in r.stdout, r.stdout
