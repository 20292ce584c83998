"""Scene files (CPU): rtn_scene_save / rtn_scene_load round-trip the surfaces,
patch groups and camera exactly, the bytes follow the documented layout, and
damaged files are refused with an error instead of being half-read."""
import os
import struct

import numpy as np
import pytest

from rtnurbs import scenes


def _read_layout(path):
    """Independent reader of the documented layout (scene_io.hip header)."""
    b = open(path, "rb").read()
    assert b[:8] == b"RTNSCENE"
    version, ns, ng, has_cam = struct.unpack_from("<4I", b, 8)
    at = 24
    cam = None
    if has_cam:
        cf = struct.unpack_from("<10f", b, at)
        wh = struct.unpack_from("<2i", b, at + 40)
        cam = (cf, wh)
        at += 48
    surfs = []
    for _ in range(ns):
        n, p, m, q = struct.unpack_from("<4i", b, at)
        at += 16
        U = np.frombuffer(b, np.float32, n + p + 2, at); at += 4 * (n + p + 2)
        V = np.frombuffer(b, np.float32, m + q + 2, at); at += 4 * (m + q + 2)
        P = np.frombuffer(b, np.float32, 4 * (n + 1) * (m + 1), at); at += 16 * (n + 1) * (m + 1)
        surfs.append((n, p, m, q, U, V, P.reshape(-1, 4)))
    groups = []
    for _ in range(ng):
        np_, p, q = struct.unpack_from("<3i", b, at)
        at += 12
        k = 4 * np_ * (p + 1) * (q + 1)
        groups.append((np_, p, q, np.frombuffer(b, np.float32, k, at)))
        at += 4 * k
    assert at == len(b)
    return version, cam, surfs, groups


def test_round_trip(rtl, tmp_path):
    sc = scenes.config1(3)
    extra = scenes.torus()
    cp = np.random.default_rng(0).normal(size=(2 * 4 * 3, 4)).astype(np.float32)
    cp[:, 3] = np.abs(cp[:, 3]) + 0.5
    s = rtl.Scene([sc.surfaces[0], extra], patch_groups=[(2, 3, 2, cp)])
    f1 = tmp_path / "a.rtns"
    s.save(f1, sc.camera)
    version, cam, surfs, groups = _read_layout(f1)
    assert version == 1
    want = np.float32(list(sc.camera.eye) + list(sc.camera.look) + list(sc.camera.up) + [sc.camera.fov_deg])
    assert np.array_equal(np.float32(cam[0]), want)
    assert cam[1] == (sc.camera.W, sc.camera.H)
    for (n, p, m, q, U, V, P), ref in zip(surfs, (sc.surfaces[0], extra)):
        assert (n, p, m, q) == (ref.n, ref.p, ref.m, ref.q)
        for a, r in ((U, ref.U), (V, ref.V), (P, ref.Pw)):
            assert np.array_equal(a.view(np.uint32), np.ascontiguousarray(r, np.float32).view(np.uint32))
    assert len(groups) == 1 and groups[0][:3] == (2, 3, 2)
    assert np.array_equal(groups[0][3], cp.ravel())
    # load and save again: the same bytes; the camera comes back
    s2, cam2 = rtl.Scene.load(f1)
    f2 = tmp_path / "b.rtns"
    s2.save(f2, cam2)
    assert open(f1, "rb").read() == open(f2, "rb").read()
    assert (cam2.W, cam2.H) == (sc.camera.W, sc.camera.H) and cam2.fov_deg == np.float32(sc.camera.fov_deg)
    # no camera
    f3 = tmp_path / "c.rtns"
    s.save(f3)
    _, cam3 = rtl.Scene.load(f3)
    assert cam3 is None


def test_damaged_files_are_refused(rtl, tmp_path):
    sc = scenes.config1(4)
    s = rtl.Scene(sc.surfaces)
    good = tmp_path / "g.rtns"
    s.save(good, sc.camera)
    b = open(good, "rb").read()
    cases = {
        "truncated": b[:-7],
        "trailing": b + b"\0\0\0\0",
        "magic": b"RTNSCENX" + b[8:],
        "version": b[:8] + struct.pack("<I", 9) + b[12:],
        "huge count": b[:12] + struct.pack("<I", 1 << 30) + b[16:],
        # the first surface's u degree set to 0: rtn_scene_add_nurbs refuses it
        "bad degree": b[:76] + struct.pack("<i", 0) + b[80:],
    }
    for name, data in cases.items():
        f = tmp_path / f"{name.replace(' ', '_')}.rtns"
        f.write_bytes(data)
        with pytest.raises(rtl.RtnError):
            rtl.Scene.load(f)
    with pytest.raises(rtl.RtnError):
        rtl.Scene.load(tmp_path / "missing.rtns")
