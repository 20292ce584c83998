"""GPU pins of the reference-API ray generator and of the per-patch flatness:

- rtn_rays_generate (calculateRaysWrapper(ray_queue, projection, camera)) on
  the reference's camera_config writes the reference's perspective and
  orthographic rays bit for bit (tests/golden/camera.npz) and idx_interact = -1;
- the scene build's delta (patchDelta, main.cpp:18-43) by closed forms: 0 on
  nets that lie on the bilinear interpolant of their corners, |v| when one
  interior control point (w = 1) is moved by v."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("tag", ["land", "port"])
@pytest.mark.parametrize("proj", ["persp", "ortho"])
def test_rays_generate_matches_reference(rtl, tag, proj):
    print(f"rays {tag} {proj}", flush=True)
    g = dict(np.load(os.path.join(GOLD, "camera.npz")))
    cfg = np.ascontiguousarray(g[f"{tag}_{proj}"])
    x = g[f"{tag}_in"]
    W, H = int(x[9]), int(x[10])
    rays = rtl.Rays(W * H)
    rays.idx.fill_bytes(0)
    q = rays.queue()
    code = 1 if proj == "persp" else 0
    rtl.check(rtl.load().rtn_rays_generate(C.byref(q), code, cfg.ctypes.data, None))
    got = rays.get()
    for k in range(6):
        assert np.array_equal(got[k].view(np.uint32), g[f"{tag}_{proj}_rays"][k].view(np.uint32)), k
    assert np.all(rays.idx.get() == -1)


def _bilinear_nets(rng, k, p, q):
    nets, moved = [], []
    for i in range(k):
        c = rng.normal(size=(2, 2, 3)) + 3.0 * i
        pts = np.zeros((p + 1, q + 1, 3))
        for a in range(p + 1):
            for b in range(q + 1):
                u, v = a / p, b / q
                pts[a, b] = (1 - u) * (1 - v) * c[0, 0] + u * (1 - v) * c[1, 0] + (1 - u) * v * c[0, 1] + u * v * c[1, 1]
        vec = np.zeros(3)
        if i % 2:  # odd patches: one interior point moved by a known vector
            vec = rng.normal(size=3) * 0.2
            pts[p // 2, q // 2] += vec
        w = np.ones((p + 1, q + 1, 1))
        nets.append(np.concatenate([pts * w, w], -1).reshape(-1, 4))
        moved.append(np.linalg.norm(vec))
    return np.stack(nets).astype(np.float32), np.array(moved)


@pytest.mark.parametrize("p,q", [(2, 2), (3, 3), (4, 2), (5, 5)])
def test_scene_delta_closed_forms(rtl, p, q):
    print(f"delta closed forms ({p},{q})", flush=True)
    rng = np.random.default_rng(p * 7 + q)
    nets, want = _bilinear_nets(rng, 16, p, q)
    scene = rtl.Scene(patch_groups=[(len(nets), p, q, nets)]).build()
    d = scene.download()
    pid = d["pid"]  # original patch of each BVH position
    got = d["delta"]
    # FP32 rounding of the bilinear values: a few ulps of the coordinates' magnitude
    scale = np.abs(nets[:, :, :3]).max(axis=(1, 2))[pid]
    assert np.all(np.abs(got - want[pid]) <= 1e-5 * want[pid] + 4e-7 * (1 + scale)), (got, want[pid])
