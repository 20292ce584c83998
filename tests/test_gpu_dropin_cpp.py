"""The reference-named C++ shims (include/rtnurbs.hpp) driven by a
reference-style C++ program (rt-nurbs_amd/examples/drop_in_render.cpp:
getPatch, buildBVH, calculateRaysWrapper, calculateIntersectWrapper on
managed memory) reproduce the exact plain form of the Python host bit for
bit on config 1."""
import os
import subprocess

import numpy as np
import pytest

from rtnurbs import scenes

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(__file__), "..", "rt-nurbs_amd", "lib", "drop_in_render")


def test_cpp_drop_in_matches_exact_form(rtl, tmp_path):
    assert os.path.exists(EXE), "build() compiles the C++ example"
    s = scenes.bicubic_grid_surface(0)
    cam = scenes.c1_camera(96, 80)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([s.n, s.p, s.m, s.q, cam.W, cam.H], np.int32).tobytes())
        for a in (s.U, s.V, s.Pw):
            f.write(np.ascontiguousarray(a, np.float32).tobytes())
        f.write(np.array(list(cam.eye) + list(cam.look) + list(cam.up) + [cam.fov_deg], np.float32).tobytes())
    out = tmp_path / "out.bin"
    print("drop-in C++ render", flush=True)
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.uint8)
    W, H = np.frombuffer(raw[:8].tobytes(), np.int32)
    rec = np.frombuffer(raw[8:].tobytes(), np.float32).reshape(W * H, 4)
    t, u, v = rec[:, 0], rec[:, 1], rec[:, 2]
    idx = rec[:, 3].view(np.int32)
    # the same path through the Python host: exact form, reference constants
    scene = rtl.Scene([s]).build()
    rays = rtl.generate_rays(cam)
    h, _ = rtl.trace(scene, rays, rtl.params(mode=rtl.MODE_EXACT))
    g = h.get()
    hit = g["hit"] >= 0
    assert hit.sum() > 0
    assert np.array_equal(idx, np.where(hit, np.arange(W * H), -1))
    for a, b in ((t, g["t"]), (u, g["u"]), (v, g["v"])):
        assert np.array_equal(a[hit].view(np.uint32), b[hit].view(np.uint32))
