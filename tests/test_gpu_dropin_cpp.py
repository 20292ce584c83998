"""A driver in the shape of the reference's main.cpp, written against the
reference's own API names (rt-nurbs_amd/examples/drop_in_render.cpp over
include/rtnurbs.hpp: lookAt, init_camera, getPerspectiveCamera /
getOrthographicCamera, CudaDevAlloc / CudaAllocManaged, curvature refinement,
getPatch, buildBVH with the re-base from patches[0].heap_pointer,
calculateRaysWrapper(ray_queue, projection, camera), patchDelta,
calculateIntersectWrapper), reproduces on the GPU:

- the reference's rays bit for bit (the oracle's restatement of
  calculateRay<perspective_projection> / <orthographic_projection>, pinned by
  tests/golden/camera.npz);
- the exact plain form of the Python host on the same rays bit for bit
  (records, idx_interact), with and without refinement and flat-skip."""
import os
import subprocess

import numpy as np
import pytest

from rtnurbs import scenes

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(__file__), "..", "rt-nurbs_amd", "lib", "drop_in_render")


def run_driver(tmp_path, s, cam, proj, C=0.0, flat=0.0):
    inp = tmp_path / f"in{proj}.bin"
    with open(inp, "wb") as f:
        f.write(np.array([s.n, s.p, s.m, s.q, cam.W, cam.H, proj], np.int32).tobytes())
        for a in (s.U, s.V, s.Pw):
            f.write(np.ascontiguousarray(a, np.float32).tobytes())
        f.write(np.array(list(cam.eye) + list(cam.look) + list(cam.up) + [cam.fov_deg, C, flat],
                         np.float32).tobytes())
    out = tmp_path / f"out{proj}.bin"
    r = subprocess.run([EXE, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.uint8)
    W, H, npatch = np.frombuffer(raw[:12].tobytes(), np.int32)
    rec = np.frombuffer(raw[12:].tobytes(), np.float32).reshape(W * H, 10)
    newton = int(r.stdout.split("newton_calls")[1].split()[0])
    return rec, int(npatch), newton


@pytest.mark.parametrize("case", ["persp", "ortho", "persp_refined_flat"])
def test_main_shaped_driver_matches(rtl, oracle, tmp_path, case):
    assert os.path.exists(EXE), "build() compiles the C++ example"
    print(f"drop-in C++ driver {case}", flush=True)
    s = scenes.bicubic_grid_surface(0)
    cam = scenes.c1_camera(96, 80)
    if case == "ortho":  # an orthographic view square onto the unit grid
        cam = scenes.Camera((0.5, 0.5, 3.0), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 60.0, 96, 80)
    proj = 0 if case == "ortho" else 1
    C, flat = (4.0, 2.0) if case == "persp_refined_flat" else (0.0, 0.0)
    rec, npatch, newton = run_driver(tmp_path, s, cam, proj, C, flat)
    rays = [np.ascontiguousarray(rec[:, k]) for k in range(6)]
    t, u, v = rec[:, 6], rec[:, 7], rec[:, 8]
    idx = rec[:, 9].view(np.int32)
    # the rays are the reference's (oracle restatement, pinned by the camera fixture)
    wfc = oracle.lookat(cam.eye, cam.look, cam.up)
    cfg = oracle.camera_config(wfc, cam.W, cam.H, proj, cam.fov_deg, 0.01 if proj else 0.0, 1000.0)
    cfr = cfg[6 * 64:7 * 64].view(np.float32)
    want = oracle.gen_rays(wfc, cfr, cam.W, cam.H) if proj else oracle.gen_rays_ortho(wfc, cfr, cam.W, cam.H)
    for k in range(6):
        assert np.array_equal(rays[k].view(np.uint32), want[k].view(np.uint32)), k
    # the same rays through the Python host's exact form (reference constants)
    surf = rtl.refine(s, C, C) if C > 0 else s
    scene = rtl.Scene([surf]).build()
    assert scene.stats()["n_patches"] == npatch
    R = rtl.Rays.from_host(rays)
    # MathUtils::radians in FP32, then cam_k as calculateIntersectWrapper forms it (intersect.cu:212)
    vfov = np.float32(np.float32(3.14159265) / np.float32(180)) * np.float32(cam.fov_deg)
    cam_k = np.float32(0.5 * cam.H / np.tan(np.float64(np.float32(vfov / np.float32(2))))) if flat > 0 else 0.0
    h, tot = rtl.trace(scene, R, rtl.params(mode=rtl.MODE_EXACT, flat_threshold=flat, cam_k=cam_k))
    g = h.get()
    hit = g["hit"] >= 0
    assert hit.sum() > 0
    assert np.array_equal(idx, np.where(hit, np.arange(cam.W * cam.H), -1))
    for a, b in ((t, g["t"]), (u, g["u"]), (v, g["v"])):
        assert np.array_equal(a[hit].view(np.uint32), b[hit].view(np.uint32))
    assert newton == tot["newton_calls"]
    if flat > 0:
        assert tot["skips"] > 0
