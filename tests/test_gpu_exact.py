"""GPU parity of the plain method, exact form, against the CPU oracle.

Bit-exact comparisons (FP32 bit patterns, integer indices) on seeded scenes:
rays, GPU decomposition, GPU exhaustive-SAH BVH, and the per-ray trace,
through the C ABI.  The oracle itself is pinned against the reference's own
outputs (tests/golden) and closed forms (test_oracle_*.py).
"""
import ctypes as C

import numpy as np
import pytest

from helpers import assert_bits_equal, c2_small, jitter_np, nodes_equal
from rtnurbs import scenes

pytestmark = pytest.mark.gpu


def _oracle_patchset(oracle, surfaces):
    """Oracle decomposition of several surfaces, concatenated in getPatch order."""
    nets, degs = [], []
    for s in surfaces:
        d = oracle.decompose(s)
        nets += list(d["cp"])
        degs += [(s.p, s.q)] * len(d["cp"])
    return nets, degs


def _oracle_scene(oracle, surfaces):
    nets, degs = _oracle_patchset(oracle, surfaces)
    bnd = np.stack([oracle.patch_bounds(n, p, q) for n, (p, q) in zip(nets, degs)])
    nodes, perm, used = oracle.build_bvh(bnd)
    ps = oracle.PatchSet([nets[i] for i in perm], [degs[i] for i in perm])
    return nets, degs, bnd, nodes[:used], perm, ps


def test_rays_pixel_centres_bitexact(oracle, rtl):
    for cam in (scenes.c1_camera(256, 256), scenes.reference_camera(320, 180), scenes.c2_camera(192, 108)):
        wfc, cfr = oracle.camera(cam)
        gw, gc = rtl.camera_matrices(cam)
        assert_bits_equal(gw, wfc, "world_from_camera")
        assert_bits_equal(gc, cfr, "camera_from_raster")
        rays = rtl.generate_rays(cam)
        ref = oracle.gen_rays(wfc, cfr, cam.W, cam.H)
        for k, (a, b) in enumerate(zip(rays.get(), ref)):
            assert_bits_equal(a, b, f"ray component {k}")
        assert np.all(rays.idx.get() == -1)


def test_rays_jittered_spp_bitexact(oracle, rtl):
    cam = scenes.c2_camera(64, 36)
    spp, seed = 4, 7
    n = cam.W * cam.H * spp
    wfc, cfr = oracle.camera(cam)
    ref = oracle.gen_rays(wfc, cfr, cam.W, cam.H, spp, jitter_np(n, seed, 0), jitter_np(n, seed, 1))
    rays = rtl.generate_rays(cam, spp=spp, jitter=True, seed=seed)
    for k, (a, b) in enumerate(zip(rays.get(), ref)):
        assert_bits_equal(a, b, f"jittered ray component {k}")


SURFACE_SETS = {
    "c1-seed0": lambda: [scenes.bicubic_grid_surface(0)],
    "c1-seed5": lambda: [scenes.bicubic_grid_surface(5)],
    "torus": lambda: [scenes.torus()],
    "sphere": lambda: [scenes.sphere()],
    "mixed-degree": lambda: c2_small(6, 1),
}


@pytest.mark.parametrize("name", list(SURFACE_SETS))
def test_decomposition_and_bvh_bitexact(oracle, rtl, name):
    surfaces = SURFACE_SETS[name]()
    nets, degs, bnd, nodes, perm, ps = _oracle_scene(oracle, surfaces)
    sc = rtl.Scene(surfaces).build()
    st = sc.stats()
    assert st["n_patches"] == len(nets) == sum(scenes.patch_count(s) for s in surfaces)
    g = sc.download()
    # BVH order and node array equal the reference build
    assert np.array_equal(g["pid"], perm.astype(np.int32)), "primitive order differs"
    nodes_equal(g["nodes"], nodes)
    # nets bit-exact in BVH order
    for i in range(len(nets)):
        p, q = g["pq"][i]
        assert (p, q) == degs[perm[i]]
        net = g["cp"][g["off"][i]: g["off"][i] + (p + 1) * (q + 1)]
        assert_bits_equal(net, nets[perm[i]], f"patch {i} net")
    # knot span indices per patch equal the oracle's
    base = 0
    for s in surfaces:
        d = oracle.decompose(s)
        for k in range(d["nbU"] * d["nbV"]):
            i, j = divmod(k, d["nbV"])
            pos = int(np.nonzero(g["pid"] == base + k)[0][0])
            assert tuple(g["knot"][pos]) == (d["knot_u"][i], d["knot_v"][j])
            assert np.array_equal(g["uv"][pos], d["uv"][k])
        base += d["nbU"] * d["nbV"]
    # patchDelta in BVH order
    for i in range(len(nets)):
        p, q = degs[perm[i]]
        assert_bits_equal(g["delta"][i:i + 1], np.float32([oracle.lib().orc_patch_delta(
            oracle.f32(nets[perm[i]]), p, q)]), "delta")


def test_get_patch_and_bounds_dropin(oracle, rtl):
    for s in (scenes.torus(), scenes.bicubic_grid_surface(2), c2_small(1, 4)[0]):
        cp, uv, nbU, nbV = rtl.get_patch(s)
        d = oracle.decompose(s)
        assert (nbU, nbV) == (d["nbU"], d["nbV"])
        assert_bits_equal(cp, d["cp"], "getPatch nets")
        assert np.array_equal(uv, d["uv"])
        b = rtl.patch_bounds(cp, s.p, s.q)
        bo = np.stack([oracle.patch_bounds(n, s.p, s.q) for n in d["cp"]])
        assert np.array_equal(b, bo)
        nodes, perm, used = rtl.build_bvh(b)
        onodes, operm, oused = oracle.build_bvh(bo)
        assert used == oused and np.array_equal(perm, operm)
        nodes_equal(nodes[:used], onodes[:used])


def _compare_trace(g, r, what):
    assert np.array_equal(g["hit"], r["hit"]), f"{what}: hit patch differs on {(g['hit'] != r['hit']).sum()} rays"
    m = r["hit"] >= 0
    for k in ("t", "u", "v", "dpdu", "dpdv"):
        assert_bits_equal(g[k][m], r[k][m], f"{what}: {k}")


TRACE_CASES = [
    ("c1-seed0-cfg", lambda: scenes.config1(0), dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)),
    ("c1-seed1-ref", lambda: scenes.config1(1), dict(max_iter=7, thr_res=1e-4, thr_det=1e-5)),
    ("c1-seed2-cfg", lambda: scenes.config1(2), dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)),
    ("torus-ref", lambda: scenes.Scene([scenes.torus()], scenes.reference_camera(480, 270)),
     dict(max_iter=7, thr_res=1e-4, thr_det=1e-5)),
    ("sphere-ref", lambda: scenes.Scene([scenes.sphere()], scenes.reference_camera(480, 270)),
     dict(max_iter=7, thr_res=1e-4, thr_det=1e-5)),
    ("mixed-degree", lambda: scenes.Scene(c2_small(8, 2), scenes.c2_camera(240, 135)),
     dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)),
]


@pytest.mark.parametrize("name,mk,kw", TRACE_CASES, ids=[c[0] for c in TRACE_CASES])
def test_trace_exact_bitexact(oracle, rtl, name, mk, kw):
    sc = mk()
    print(f"case {name}", flush=True)
    nets, degs, bnd, nodes, perm, ps = _oracle_scene(oracle, sc.surfaces)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera)
    prm = rtl.params(mode=rtl.MODE_EXACT, count=1, **kw)
    hits, tot = rtl.trace(scene, rays, prm)
    g = hits.get()
    r = oracle.trace_f32(nodes, ps, rays.get(), **kw)
    _compare_trace(g, r, name)
    # per-ray work counters: Newton calls, evaluations, interior/leaf visits, patch tests
    for k in (0, 1, 3, 4, 5):
        assert np.array_equal(g["counters"][:, k].astype(np.int64), r["counters"][:, k]), f"counter {k}"
    assert tot["newton_evals"] == int(r["counters"][:, 1].sum())
    assert tot["overflow"] == 0
    assert (r["hit"] >= 0).sum() > 0


def _dropin_buffers(rtl, nets, p, q, perm):
    """Reference-layout device buffers: a CP block and bezierPatch headers
    whose control_point pointers point into it (main.cpp:117-134)."""
    ps = (p + 1) * (q + 1)
    block = np.ascontiguousarray(np.stack([nets[i] for i in perm]).reshape(-1, 4), np.float32)
    d_cp = rtl.DeviceArray.from_host(block)
    hdr = np.zeros(len(perm), dtype=np.dtype([("cp", np.uint64), ("u0", np.float32), ("u1", np.float32),
                                              ("v0", np.float32), ("v1", np.float32), ("p", np.int32),
                                              ("q", np.int32), ("heap", np.uint64)]))
    assert hdr.dtype.itemsize == 40
    hdr["cp"] = d_cp.ptr + np.arange(len(perm), dtype=np.uint64) * np.uint64(16 * ps)
    hdr["p"], hdr["q"], hdr["heap"] = p, q, d_cp.ptr
    return d_cp, rtl.DeviceArray.from_host(hdr)


@pytest.mark.parametrize("flat", [0.0, 2.0])
def test_calculate_intersect_dropin(oracle, rtl, flat):
    """rtn_calculate_intersect on reference-layout buffers equals the oracle,
    including surfaceInteraction fields, idx_interact, wo and the Newton /
    skip counters (intersect.cu:189-193)."""
    cam = scenes.reference_camera(320, 180)
    s = rtl.refine(scenes.torus(), 4.0, 4.0)
    d = oracle.decompose(s)
    nets = list(d["cp"])
    bnd = np.stack([oracle.patch_bounds(n, s.p, s.q) for n in nets])
    nodes, perm, used = oracle.build_bvh(bnd)
    d_cp, d_hdr = _dropin_buffers(rtl, nets, s.p, s.q, perm)
    ps = oracle.PatchSet([nets[i] for i in perm], [(s.p, s.q)] * len(perm))
    delta = np.array([oracle.lib().orc_patch_delta(oracle.f32(n), s.p, s.q) for n in ps.cp.reshape(len(perm), -1, 4)],
                     np.float32)
    d_nodes = rtl.DeviceArray.from_host(nodes)
    d_delta = rtl.DeviceArray.from_host(delta)
    rays = rtl.generate_rays(cam)
    inter = rtl.DeviceArray((cam.W * cam.H,), rtl.INTERACT_DTYPE)
    inter.fill_bytes(0)
    cnt = rtl.DeviceArray((2,), np.uint64)
    cnt.set(np.array([5, 0], np.uint64))
    q = rays.queue()
    v_fov = np.float32(np.float32(3.14159265) / np.float32(180) * np.float32(60.0))
    rtl.check(rtl.load().rtn_calculate_intersect(C.byref(q), d_nodes.ptr, d_hdr.ptr, inter.ptr, cam.W, cam.H, flat,
                                                 float(v_fov), d_delta.ptr, cnt.ptr, cnt.ptr + 8))
    cam_k = np.float32(0.5 * np.float64(np.float32(cam.H)) / np.tan(np.float64(v_fov / np.float32(2.0))))
    r = oracle.trace_f32(nodes, ps, rays.get(), max_iter=7, thr_res=1e-4, thr_det=1e-5, flat=flat,
                         cam_k=float(cam_k), delta=delta)
    gi = inter.get()
    idx = rays.idx.get()
    hit = r["hit"] >= 0
    assert np.array_equal(idx, np.where(hit, np.arange(len(idx)), -1))
    assert_bits_equal(gi["t_hit"][hit], r["t"][hit], "t_hit")
    assert_bits_equal(gi["uv"][hit], np.stack([r["u"], r["v"]], 1)[hit], "uv")
    assert_bits_equal(gi["dpdu"][hit], r["dpdu"][hit], "dpdu")
    assert_bits_equal(gi["dpdv"][hit], r["dpdv"][hit], "dpdv")
    dirs = np.stack(rays.get()[3:], 1)
    assert_bits_equal(gi["wo"][hit], -dirs[hit], "wo")
    assert np.all(gi["t_hit"][~hit] == 0)
    c = cnt.get()
    assert c[0] == 5 + r["counters"][:, 0].sum()
    assert c[1] == r["counters"][:, 2].sum()
    if flat > 0:
        assert c[1] > 0
