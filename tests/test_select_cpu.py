"""The GPU BVH build's selection (csrc/pale_select.h) against libstdc++'s
std::nth_element on the CPU: the sequential restatement, heap selection at
the depth limit, and the rank form of Hoare's partition that the workgroup
kernel uses, element for element (tests/cpu/select_check.cpp, built with
g++ here).  BVH.ixx:49-133 orders the surfels with std::nth_element; this is
what lets the GPU build reproduce the reference's node numbering and
permutation."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "differentiablepointrendering_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_selection_equals_libstdcxx(tmp_path):
    exe = str(tmp_path / "select_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "cpu", "select_check.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok "), r.stdout
