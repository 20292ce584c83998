"""The C++ header with the reference's API compiles on its own (host C++17, no
HIP headers) and the main.cpp-shaped example is built with the library."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.join(os.path.dirname(__file__), "..")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "rtnurbs.hpp"\nint main() { return BezierPatchOps::getPatch == nullptr; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_example_built():
    assert os.path.exists(os.path.join(ROOT, "rt-nurbs_amd", "lib", "drop_in_render"))
