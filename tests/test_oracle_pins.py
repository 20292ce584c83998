"""Pins of the oracle functions the per-pixel fixtures leave open.

- camera_config and orthographic / perspective rays against the reference's
  own init_camera, getPerspectiveCamera, getOrthographicCamera and
  calculateRay (tests/golden/camera.npz, compiled reference headers);
- curvature refinement of bicubic and degree-4 surfaces at C = 1, 4, 16
  against the reference (tests/golden/refine_p3.npz, refine_p4.npz), for the
  oracle and for the library's host refinement (rtn_refine);
- knot insertion leaves the surface unchanged, in FP64, up to degree 5 (the
  reference caps the degree at 4, so degree 5 has no reference output);
- patchDelta (main.cpp:18-43) by closed forms: a net on the bilinear
  interpolant of its corners has delta 0; moving one interior control point
  (w = 1) by v gives delta |v|; moving several gives the largest move."""
import hashlib
import os

import numpy as np
import pytest

from helpers import assert_bits_equal
from rtnurbs import scenes
from test_oracle_math import nurbs_point

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, f"{name}.npz")))


# ------------------------------------------------------------------ camera
@pytest.mark.parametrize("tag", ["land", "port"])
def test_camera_config_and_rays(oracle, tag):
    g = load("camera")
    x = g[f"{tag}_in"]
    W, H, fov, zn, zf = int(x[9]), int(x[10]), float(x[11]), float(x[12]), float(x[13])
    wfc = oracle.lookat(tuple(x[0:3]), tuple(x[3:6]), tuple(x[6:9]))
    assert_bits_equal(wfc, g[f"{tag}_wfc"], "lookAt")
    for pname, proj in (("init", -1), ("persp", 1), ("ortho", 0)):
        got = oracle.camera_config(wfc, W, H, proj, fov, zn, zf)
        assert np.array_equal(got, g[f"{tag}_{pname}"]), pname
    for pname, fn in (("persp", lambda w, c: oracle.gen_rays(w, c, W, H)),
                      ("ortho", lambda w, c: oracle.gen_rays_ortho(w, c, W, H))):
        cfr = g[f"{tag}_{pname}"][6 * 64:7 * 64].view(np.float32)
        rays = fn(wfc, cfr)
        for k in range(6):
            assert_bits_equal(rays[k], g[f"{tag}_{pname}_rays"][k], f"{pname} ray component {k}")


def test_orthographic_rays_closed_form(oracle):
    """Orthographic rays share the camera's z axis as direction, and their
    origins lie on the near plane's raster grid: neighbouring pixels differ
    by the camera's x / y axis scaled by the screen window over the raster."""
    wfc = oracle.lookat((1.0, 2.0, -3.0), (0.0, 0.0, 5.0), (0.0, 1.0, 0.0))
    W, H = 16, 10
    c = oracle.camera_config(wfc, W, H, 0, 60.0, 0.0, 100.0)
    cfr = c[6 * 64:7 * 64].view(np.float32)
    ox, oy, oz, dx, dy, dz = (a.reshape(H, W).astype(np.float64) for a in oracle.gen_rays_ortho(wfc, cfr, W, H))
    M = wfc.reshape(4, 4).astype(np.float64)  # column major: M[k] is column k
    assert np.all(dx == M[2, 0]) and np.all(dy == M[2, 1]) and np.all(dz == M[2, 2])
    o = np.stack([ox, oy, oz], -1)
    aspect = W / H
    step = 2 * aspect / W  # screen window [-aspect, aspect] over W pixels
    np.testing.assert_allclose(o[:, 1:] - o[:, :-1], np.broadcast_to(step * M[0, :3], (H, W - 1, 3)), atol=1e-5)
    np.testing.assert_allclose(o[1:] - o[:-1], np.broadcast_to(-step * M[1, :3], (H - 1, W, 3)), atol=1e-5)
    # the centre of the raster is the eye
    np.testing.assert_allclose((o[H // 2 - 1, W // 2 - 1] + o[H // 2, W // 2]) / 2, M[3, :3], atol=1e-5)


# -------------------------------------------------------------- refinement
def _check_net(got, g, key):
    if f"{key}_Pw" in g:
        assert_bits_equal(got, g[f"{key}_Pw"], key)
    else:
        assert hashlib.sha256(np.ascontiguousarray(got, np.float32).tobytes()).digest() == \
            g[f"{key}_Pw_sha256"].tobytes(), key
        assert_bits_equal(got[::16], g[f"{key}_Pw_every16"], key)


@pytest.mark.parametrize("name", ["p3", "p4"])
@pytest.mark.parametrize("C", [1, 4, 16])
def test_refinement_golden_oracle_and_host(oracle, name, C):
    from rtnurbs import lib as rtl
    g = load(f"refine_{name}")
    n, m, p, q = (int(x) for x in g["nmpq"])
    s = scenes.Nurbs(n, m, p, q, g["U"], g["V"], g["Pw"])
    key = f"C{C}"
    for label, refine in (("oracle", oracle.refine), ("rtn_refine", rtl.refine)):
        mid = refine(s, float(C), None)
        assert mid.n == g[f"{key}_nm"][0], label
        assert_bits_equal(mid.U, g[f"{key}_U"], f"{label} U")
        if f"{key}_Pw_mid" in g:
            assert_bits_equal(mid.Pw, g[f"{key}_Pw_mid"], f"{label} U net")
        out = refine(mid, None, float(C))
        assert out.m == g[f"{key}_nm"][1], label
        assert_bits_equal(out.V, g[f"{key}_V"], f"{label} V")
        _check_net(out.Pw, g, key)


@pytest.mark.parametrize("deg", [(2, 3), (4, 4), (5, 5), (5, 2)])
def test_knot_insertion_keeps_the_surface(oracle, deg):
    """Curvature refinement only inserts knots (A5.1), so the refined surface is
    the same surface: FP64 Cox-de Boor points of both agree at random
    parameters, to the FP32 rounding of the inserted control points."""
    from rtnurbs import lib as rtl
    rng = np.random.default_rng(sum(deg) * 7)
    if deg[0] == deg[1]:  # random interior knot multiplicities
        s = scenes.random_surface(rng, deg_range=deg, cp_range=(deg[0] + 2, deg[0] + 4), box=4.0)
    else:  # mixed degrees on a uniform-knot grid
        s = scenes.grid_surface(rng, 12, *deg)
    for refine in (oracle.refine, rtl.refine):
        r = refine(s, 2.0, 2.0)
        assert r.n > s.n and r.m > s.m
        for u, v in rng.uniform(0, 1, size=(12, 2)):
            a, b = nurbs_point(s, u, v), nurbs_point(r, u, v)
            np.testing.assert_allclose(b, a, rtol=0, atol=2e-5)


# -------------------------------------------------------------- patchDelta
def _bilinear_net(rng, p, q, w_rand=True):
    """A (p+1)x(q+1) homogeneous net whose dehomogenised points lie on the
    bilinear interpolant of four random corners at their nodes (i/p, j/q)."""
    c = rng.normal(size=(2, 2, 3))
    pts = np.zeros((p + 1, q + 1, 3))
    for i in range(p + 1):
        for j in range(q + 1):
            u, v = i / p, j / q
            pts[i, j] = (1 - u) * (1 - v) * c[0, 0] + u * (1 - v) * c[1, 0] + (1 - u) * v * c[0, 1] + u * v * c[1, 1]
    w = rng.uniform(0.5, 2.0, size=(p + 1, q + 1, 1)) if w_rand else np.ones((p + 1, q + 1, 1))
    return np.concatenate([pts * w, w], -1).reshape(-1, 4).astype(np.float32), pts


@pytest.mark.parametrize("p,q", [(1, 1), (2, 2), (3, 3), (2, 5), (4, 3), (5, 5)])
def test_patch_delta_closed_forms(oracle, p, q):
    L = oracle.lib()
    rng = np.random.default_rng(p * 10 + q)
    net, pts = _bilinear_net(rng, p, q)
    scale = np.abs(pts).max()
    assert L.orc_patch_delta(net, p, q) <= 1e-6 * (1 + scale)  # on the interpolant: delta 0
    if p < 2 and q < 2:
        return  # no interior node
    net1, _ = _bilinear_net(rng, p, q, w_rand=False)
    moves = []
    for _ in range(3):  # move interior points (w = 1) by known vectors
        i, j = int(rng.integers(1, max(2, p))), int(rng.integers(0, q + 1))
        if p < 2:
            i, j = int(rng.integers(0, p + 1)), int(rng.integers(1, q))
        vec = rng.normal(size=3) * rng.uniform(0.01, 0.5)
        net1.reshape(p + 1, q + 1, 4)[i, j, :3] += vec.astype(np.float32)
        moves.append((i, j, vec))
        got = float(L.orc_patch_delta(net1, p, q))
        # delta = the largest distance of a moved point from its node's bilinear value
        disp = {}
        for (a, b, vv) in moves:
            disp[(a, b)] = disp.get((a, b), 0) + vv
        want = max(np.linalg.norm(vv) for vv in disp.values())
        assert abs(got - want) <= 1e-5 * (1 + want), (got, want)


def test_patch_delta_is_zero_for_a_flat_bilinear_bezier_from_decomposition(oracle):
    """A degree-(1,1) NURBS is its own bilinear patch: every patch of its
    decomposition has delta 0 (corners give 0 by construction)."""
    rng = np.random.default_rng(5)
    n = m = 4
    P = np.concatenate([rng.normal(size=(n + 1, m + 1, 3)), np.ones((n + 1, m + 1, 1))], -1).astype(np.float32)
    U = np.array([0, 0, 0.25, 0.5, 0.75, 1, 1], np.float32)
    s = scenes.Nurbs(n, m, 1, 1, U, U.copy(), P.reshape(-1, 4))
    d = oracle.decompose(s)
    for net in d["cp"]:
        assert oracle.lib().orc_patch_delta(np.ascontiguousarray(net, np.float32), 1, 1) == 0.0
