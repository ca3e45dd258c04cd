"""Changed-node tracking of the host broadphase tree (what the device query's
incremental image update relies on), checked by a C++ program compiled
against the product's host_geometry.cpp: see tests/cpp/bp_dirty_check.cpp."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_incremental_image_equals_the_moved_tree(tmp_path):
    csrc = os.path.join(ROOT, "kizuriphysics3d_amd", "csrc")
    exe = str(tmp_path / "bp_dirty_check")
    build = ["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-pthread", "-D__HIP_PLATFORM_AMD__",
             "-I/opt/rocm/include", "-I" + csrc, os.path.join(HERE, "cpp", "bp_dirty_check.cpp"),
             os.path.join(csrc, "host_geometry.cpp"), "-o", exe]
    r = subprocess.run(build, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "identical" in r.stdout, r.stdout + r.stderr
