"""The reference's own test drivers (tests/treecode.cpp, tests/treecode_mp.cpp,
tests/directsum_mp.cpp of the reference checkout) compile and link unchanged
against the drop-in headers (bit_amd/src) and libraries (libtree3d.so,
libbit_amd.so).  Compile and link only: running them needs a GPU, and their
recipes run on the GPU through tests/cpp/dropin_treecode.cpp."""
import os
import subprocess

import pytest

from conftest import ROOT

REF_TESTS = "/root/reference/tests"
PKG = os.path.join(ROOT, "bit_amd")


@pytest.mark.parametrize("name", ["treecode", "treecode_mp", "directsum_mp"])
def test_reference_driver_links_against_dropin(name, tmp_path):
    src = os.path.join(REF_TESTS, name + ".cpp")
    if not os.path.exists(src):
        pytest.skip("reference checkout not present")
    if not os.path.exists(os.path.join(PKG, "libtree3d.so")):
        pytest.fail("bit_amd/libtree3d.so not built: run python -m bit_amd.build")
    exe = tmp_path / name
    cmd = ["g++", "-O1", "-std=c++17", "-fopenmp", "-I" + os.path.join(PKG, "src"), src, "-o",
           str(exe), "-L" + PKG, "-ltree3d", "-lbit_amd", "-Wl,-rpath," + PKG]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert exe.exists()
