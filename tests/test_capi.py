"""C ABI surface (CPU): the library loads, exports every function declared in
include/rtnurbs.h, and its host-side entry points (camera, refinement,
argument validation) behave like the reference without touching a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import assert_bits_equal
from rtnurbs import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "rtnurbs.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rtn_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported(rtl):
    L = rtl.load()
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(rtl.EXPORTS) == names


def test_camera_matches_oracle(oracle, rtl):
    for cam in (scenes.c1_camera(), scenes.reference_camera(), scenes.c2_camera(), scenes.c1_camera(300, 200)):
        gw, gc = rtl.camera_matrices(cam)
        wfc, cfr = oracle.camera(cam)
        assert_bits_equal(gw, wfc, "world_from_camera")
        assert_bits_equal(gc, cfr, "camera_from_raster")


@pytest.mark.parametrize("tag", ["land", "port"])
def test_camera_config_matches_reference(rtl, tag):
    """rtn_lookat, rtn_camera_init, rtn_camera_get_perspective and
    rtn_camera_get_orthographic (host-side) give the reference's camera_config
    bytes (tests/golden/camera.npz, from the reference's camera.h)."""
    L = rtl.load()
    g = dict(np.load(os.path.join(ROOT, "tests", "golden", "camera.npz")))
    x = g[f"{tag}_in"]
    W, H, fov, zn, zf = int(x[9]), int(x[10]), float(x[11]), float(x[12]), float(x[13])
    f3 = C.c_float * 3
    wfc = np.zeros(16, np.float32)
    assert L.rtn_lookat(f3(*x[0:3]), f3(*x[3:6]), f3(*x[6:9]), wfc.ctypes.data) == 0
    assert_bits_equal(wfc, g[f"{tag}_wfc"], "lookAt")
    init = np.zeros(472, np.uint8)
    assert L.rtn_camera_init(wfc.ctypes.data, W, H, init.ctypes.data) == 0
    assert np.array_equal(init, g[f"{tag}_init"])
    for name, fn, args in (("persp", L.rtn_camera_get_perspective, (fov, zn, zf)),
                           ("ortho", L.rtn_camera_get_orthographic, (zn, zf))):
        out = np.zeros(472, np.uint8)
        assert fn(init.ctypes.data, *[C.c_float(a) for a in args], out.ctypes.data) == 0
        assert np.array_equal(out, g[f"{tag}_{name}"]), name


@pytest.mark.parametrize("C_", [1.0, 4.0, 16.0])
def test_refinement_matches_oracle(oracle, rtl, C_):
    for s in (scenes.torus(), scenes.sphere(), scenes.bicubic_grid_surface(2)):
        a = rtl.refine(s, C_, C_)
        b = oracle.refine(s, C_, C_)
        assert (a.n, a.m) == (b.n, b.m)
        assert np.array_equal(a.U, b.U) and np.array_equal(a.V, b.V)
        assert_bits_equal(a.Pw, b.Pw, "refined net")


def test_config2_patch_count(rtl):
    """Config 2 is frozen at 10,018 patches (closed-form count per surface)."""
    sc = scenes.config2(0)
    assert len(sc.surfaces) == 64
    assert sum(scenes.patch_count(s) for s in sc.surfaces) == scenes.C2_PATCHES == 10018
    degs = {(s.p, s.q) for s in sc.surfaces}
    assert max(max(d) for d in degs) == 5 and min(min(d) for d in degs) == 2


def test_invalid_inputs_rejected(rtl):
    """Host-side validation (no GPU involved): degree above the cap, unsorted
    knots, interior multiplicity above the degree, null pointers."""
    L = rtl.load()
    h = C.c_void_p()
    assert L.rtn_scene_create(C.byref(h)) == 0
    s = scenes.bicubic_grid_surface(0)
    U, V, Pw = s.U.copy(), s.V.copy(), s.Pw.copy()
    args = lambda U_, p_: (h, s.n, p_, U_.ctypes.data, s.m, s.q, V.ctypes.data, Pw.ctypes.data)
    assert L.rtn_scene_add_nurbs(*args(U, 6)) == rtl.RTN_EUNSUPPORTED
    bad = U.copy()
    bad[5], bad[6] = bad[6], bad[5]
    assert L.rtn_scene_add_nurbs(*args(bad, 3)) == rtl.RTN_EINVAL
    mult = np.array([0, 0, 0, 0, .5, .5, .5, .5, 1, 1, 1, 1], np.float32)
    assert L.rtn_scene_add_nurbs(*args(mult, 3)) == rtl.RTN_EINVAL
    assert b"multiplicity" in L.rtn_last_error()
    unclamped = np.array([0, 0, 0, .1, .2, .4, .6, .8, 1, 1, 1, 1], np.float32)
    assert L.rtn_scene_add_nurbs(*args(unclamped, 3)) == rtl.RTN_EINVAL
    assert L.rtn_scene_add_nurbs(h, s.n, 3, None, s.m, 3, V.ctypes.data, Pw.ctypes.data) == rtl.RTN_EINVAL
    assert L.rtn_scene_add_nurbs(*args(U, 3)) == 0
    assert L.rtn_scene_build(None) == rtl.RTN_EINVAL
    assert L.rtn_scene_destroy(h) == 0
    ins = C.c_int()
    assert L.rtn_refine(2, 7, 3, U.ctypes.data, 7, Pw.ctypes.data, 1.0, C.byref(ins), None, None) == rtl.RTN_EINVAL


def test_trace_params_defaults(rtl):
    p = rtl.params()
    assert (p.mode, p.max_iter) == (rtl.MODE_EXACT, 7)
    assert np.float32(p.thr_det) == np.float32(1e-5) and np.float32(p.thr_res) == np.float32(1e-4)
    assert p.seed_mode == rtl.SEED_PLAIN and p.precision == rtl.PREC_FP32


def test_product_does_not_import_the_oracle():
    """No product source refers to the oracle directory."""
    pkg = os.path.join(ROOT, "rt-nurbs_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "pyoracle" not in txt and "oracle/" not in txt, f
