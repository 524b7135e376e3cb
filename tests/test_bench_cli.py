"""Host-side contract of bench.py (no GPU needed): one-GPU only, and the
learned defaults it measures are the package's defaults."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_learned_defaults_match_the_package():
    sys.path.insert(0, ROOT)
    import bench
    from volumesurfacediscretemeshoptimization_amd import DMOTetClass, _lib
    w = DMOTetClass.LEARN_WINDOW
    assert tuple(bench.LEARN_WINDOW["tet"]) == tuple(w[_lib.VSDMO_TET])
    assert tuple(bench.LEARN_WINDOW["hex"]) == tuple(w[_lib.VSDMO_HEX])
    assert bench.SPEC_WINDOW == (1, 0.25)
