"""A scene read back from its file builds and traces exactly like the scene
it was written from (mixed degrees, the plain method and a fast form)."""
import numpy as np
import pytest

from helpers import c2_small
from rtnurbs import scenes

pytestmark = pytest.mark.gpu

CFG = dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)


def test_loaded_scene_traces_identically(rtl, tmp_path):
    print("case c2_small(8, 2) through a scene file", flush=True)
    sc = scenes.Scene(c2_small(8, 2), scenes.c2_camera(240, 135))
    a = rtl.Scene(sc.surfaces)
    f = tmp_path / "scene.rtns"
    a.save(f, sc.camera)
    a.build()
    b, cam = rtl.Scene.load(f)
    b.build()
    da, db = a.download(), b.download()
    for k in ("cp", "off", "pq", "nodes"):
        assert np.array_equal(da[k].view(np.uint8), db[k].view(np.uint8)), k
    rays = rtl.generate_rays(cam)
    for mode in (rtl.MODE_EXACT, rtl.MODE_WIDE):
        ha, _ = rtl.trace(a, rays, rtl.params(mode=mode, **CFG))
        hb, _ = rtl.trace(b, rays, rtl.params(mode=mode, **CFG))
        ga, gb = ha.get(), hb.get()
        assert np.array_equal(ga["hit"], gb["hit"])
        m = ga["hit"] >= 0
        assert m.any()
        assert np.array_equal(ga["t"][m].view(np.uint32), gb["t"][m].view(np.uint32))
