"""rtn_calculate_intersect_noalloc (the drop-in trace with its totals in a
file-scope device array instead of a per-call allocation) equals
rtn_calculate_intersect bit for bit: records, idx_interact, Newton and skip
counters, on two calls in a row, the second on a smaller raster (the shared
scratch keeps the first call's contents and must be cleared)."""
import ctypes as C

import numpy as np
import pytest

from rtnurbs import scenes
from test_gpu_exact import _dropin_buffers

pytestmark = pytest.mark.gpu


def _run(rtl, fn, d_nodes, d_hdr, d_delta, cam, flat, start_count):
    rays = rtl.generate_rays(cam)
    inter = rtl.DeviceArray((cam.W * cam.H,), rtl.INTERACT_DTYPE)
    inter.fill_bytes(0)
    cnt = rtl.DeviceArray((2,), np.uint64)
    cnt.set(np.array(start_count, np.uint64))
    q = rays.queue()
    v_fov = np.float32(np.float32(3.14159265) / np.float32(180) * np.float32(60.0))
    rtl.check(fn(C.byref(q), d_nodes.ptr, d_hdr.ptr, inter.ptr, cam.W, cam.H, flat, float(v_fov), d_delta.ptr,
                 cnt.ptr, cnt.ptr + 8))
    return inter.get(), rays.idx.get(), cnt.get()


@pytest.mark.parametrize("flat", [0.0, 2.0])
def test_noalloc_dropin_equals_dropin(oracle, rtl, flat):
    print(f"no-alloc drop-in, flat {flat}", flush=True)
    s = rtl.refine(scenes.torus(), 4.0, 4.0)
    d = oracle.decompose(s)
    nets = list(d["cp"])
    bnd = np.stack([oracle.patch_bounds(n, s.p, s.q) for n in nets])
    nodes, perm, used = oracle.build_bvh(bnd)
    d_cp, d_hdr = _dropin_buffers(rtl, nets, s.p, s.q, perm)
    delta = np.array([oracle.lib().orc_patch_delta(oracle.f32(nets[i]), s.p, s.q) for i in perm], np.float32)
    d_nodes = rtl.DeviceArray.from_host(nodes)
    d_delta = rtl.DeviceArray.from_host(delta)
    L = rtl.load()
    for cam, start in ((scenes.reference_camera(320, 180), [7, 3]), (scenes.reference_camera(160, 90), [0, 0])):
        a = _run(rtl, L.rtn_calculate_intersect, d_nodes, d_hdr, d_delta, cam, flat, start)
        b = _run(rtl, L.rtn_calculate_intersect_noalloc, d_nodes, d_hdr, d_delta, cam, flat, start)
        assert np.array_equal(a[0].view(np.uint8), b[0].view(np.uint8)), "surfaceInteraction records"
        assert np.array_equal(a[1], b[1]), "idx_interact"
        assert np.array_equal(a[2], b[2]), "counters"
        assert (a[1] >= 0).sum() > 0 and a[2][0] > start[0]
        if flat > 0:
            assert a[2][1] > start[1]
