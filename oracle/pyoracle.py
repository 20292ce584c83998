"""ctypes front-end of the CPU oracle (oracle/oracle.cpp) and, where it was
built in this container, of the reference harness (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and by
the benchmark's CPU-baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
vp = C.c_void_p

NODE_DTYPE = np.dtype([("b", np.float32, 6), ("left_first", np.uint32), ("prim_count", np.uint32)])


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_lookat.argtypes = [f32p, f32p, f32p, f32p]
        L.orc_camera_perspective.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, f32p]
        L.orc_gen_rays.argtypes = [f32p, f32p, C.c_int, C.c_int, C.c_int, vp, vp] + [f32p] * 6
        L.orc_camera_config.argtypes = [f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, vp]
        L.orc_gen_rays_ortho.argtypes = [f32p, f32p, C.c_int, C.c_int] + [f32p] * 6
        L.orc_segs.argtypes = [f32p, C.c_int, C.c_int, f32p]
        L.orc_segs.restype = C.c_int
        L.orc_count_segments.argtypes = [f32p, C.c_int, C.c_int]
        L.orc_count_segments.restype = C.c_int
        L.orc_decompose.argtypes = [C.c_int, C.c_int, f32p, C.c_int, C.c_int, f32p, f32p, f32p, f32p, i32p, i32p]
        L.orc_decompose.restype = C.c_int
        L.orc_find_span.argtypes = [C.c_int, C.c_int, C.c_float, f32p]
        L.orc_find_span.restype = C.c_int
        L.orc_surface_point_f64.argtypes = [C.c_int, C.c_int, f32p, C.c_int, C.c_int, f32p, f32p,
                                            C.c_double, C.c_double, f64p]
        L.orc_refine.argtypes = [C.c_int, C.c_int, C.c_int, f32p, C.c_int, f32p, C.c_float, i32p, vp, vp]
        L.orc_refine.restype = C.c_int
        L.orc_patch_bounds.argtypes = [f32p, C.c_int, C.c_int, f32p]
        L.orc_patch_delta.argtypes = [f32p, C.c_int, C.c_int]
        L.orc_patch_delta.restype = C.c_float
        L.orc_build_bvh.argtypes = [f32p, C.c_int, vp, u32p]
        L.orc_build_bvh.restype = C.c_int
        L.orc_trace_f32.argtypes = [vp, f32p, i32p, i32p, i32p, vp, C.c_int, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_uint32, C.c_long] + [f32p] * 6 + \
                                   [i32p, f32p, f32p, f32p, f32p, f32p, i64p, C.c_int]
        L.orc_trace_f64.argtypes = [vp, f32p, C.c_long, i32p, i32p, i32p, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_uint32, C.c_long] + [f32p] * 6 + \
                                   [i32p, f64p, f64p, f64p, f64p, f64p, f64p, i64p, C.c_int]
        L.orc_trace_f64_roots.argtypes = [vp, f32p, C.c_long, i32p, i32p, i32p, C.c_int, C.c_double, C.c_double,
                                          C.c_double, C.c_int, C.c_uint32, C.c_long] + [f32p] * 6 + \
                                         [i32p, f64p, f64p, f64p, f64p, f64p, f64p, i64p, C.c_int]
        L.orc_polish_f64.argtypes = [f32p, C.c_int, C.c_int, f32p, f32p, C.c_double, C.c_double, f64p]
        L.orc_polish_f64.restype = C.c_int
        L.orc_patch_eval_f32.argtypes = [f32p, C.c_int, C.c_int, C.c_float, C.c_float, f32p]
        L.orc_ray_planes_f32.argtypes = [f32p, f32p, f32p]
        L.orc_slab_f32.argtypes = [f32p, f32p, f32p]
        L.orc_slab_f32.restype = C.c_float
        L.orc_bilinear_seed_f32.argtypes = [f32p, C.c_int, C.c_int, f32p, f32p, f32p]
        L.orc_pcg_hash.argtypes = [C.c_uint32]
        L.orc_pcg_hash.restype = C.c_uint32
        L.orc_newton_f32.argtypes = [f32p, C.c_int, C.c_int, f32p, f32p, C.c_float, C.c_float, C.c_float,
                                     C.c_float, C.c_uint32, C.c_int, f32p]
        L.orc_newton_f32.restype = C.c_int
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The compiled reference harness; only present where /root/reference was."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_lookat.argtypes = [f32p, f32p, f32p, f32p]
        R.ref_camera.argtypes = [f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, f32p]
        R.ref_gen_rays.argtypes = [f32p, f32p, C.c_int, C.c_int] + [f32p] * 6
        R.ref_camera_config.argtypes = [f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, vp]
        R.ref_camera_config_size.restype = C.c_int
        R.ref_gen_rays_ortho.argtypes = [f32p, f32p, C.c_int, C.c_int] + [f32p] * 6
        R.ref_get_patch.argtypes = [C.c_int, C.c_int, f32p, C.c_int, C.c_int, f32p, f32p, vp, vp,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]
        R.ref_get_patch.restype = C.c_int
        R.ref_refine_u.argtypes = [f32p, C.c_int, C.c_int, C.c_int, C.c_float, f32p, vp, vp]
        R.ref_refine_u.restype = C.c_int
        R.ref_refine_v.argtypes = [f32p, C.c_int, C.c_int, C.c_int, C.c_float, f32p, vp, vp]
        R.ref_refine_v.restype = C.c_int
        R.ref_build_bvh.argtypes = [f32p, C.c_int, C.c_int, C.c_int, vp, u32p]
        R.ref_build_bvh.restype = C.c_int
        R.ref_patch_bounds.argtypes = [f32p, C.c_int, C.c_int, f32p]
        R.ref_trace.argtypes = [vp, f32p, C.c_int, C.c_int, C.c_int, vp, C.c_float, C.c_float, C.c_long] + \
                               [f32p] * 6 + [i32p, f32p, f32p, f32p, f32p, f32p, i32p]
        R.ref_bilinear_seed.argtypes = [f32p, C.c_int, C.c_int, f32p, f32p, f32p]
        R.ref_newton.argtypes = [f32p, C.c_int, C.c_int, f32p, f32p, C.c_float, C.c_float, C.c_uint32, f32p]
        R.ref_newton.restype = C.c_int
        R.ref_ray_planes.argtypes = [f32p, f32p, f32p]
        R.ref_patch_eval.argtypes = [f32p, C.c_int, C.c_int, C.c_float, C.c_float, f32p]
        R.ref_pcg_hash.argtypes = [C.c_uint32]
        R.ref_pcg_hash.restype = C.c_uint32
        _ref = R
    return _ref


# ------------------------------------------------------------ conveniences
def f32(a):
    return np.ascontiguousarray(a, np.float32)


def lookat(eye, look, up):
    out = np.zeros(16, np.float32)
    lib().orc_lookat(f32(eye), f32(look), f32(up), out)
    return out


def camera(cam):
    """(world_from_camera, camera_from_raster) of a scenes.Camera."""
    wfc = lookat(cam.eye, cam.look, cam.up)
    cfr = np.zeros(16, np.float32)
    lib().orc_camera_perspective(cam.W, cam.H, cam.fov_deg, cam.znear, cam.zfar, cfr)
    return wfc, cfr


def gen_rays(wfc, cfr, W, H, spp=1, jx=None, jy=None):
    n = W * H * spp
    out = [np.zeros(n, np.float32) for _ in range(6)]
    jxp = f32(jx).ctypes.data if jx is not None else None
    jyp = f32(jy).ctypes.data if jy is not None else None
    keep = (jx, jy)
    lib().orc_gen_rays(f32(wfc), f32(cfr), W, H, spp, jxp, jyp, *out)
    del keep
    return out


CAMERA_CONFIG_BYTES = 472  # camera_config (camera.h:13-27): 7 matrices, 2 ints, bounds2f


def camera_config(wfc, W, H, proj, fov_deg=60.0, znear=0.01, zfar=1000.0):
    """init_camera then getPerspectiveCamera (proj 1) / getOrthographicCamera
    (proj 0) / nothing (proj -1): the camera_config bytes as uint8[472]."""
    out = np.zeros(CAMERA_CONFIG_BYTES, np.uint8)
    lib().orc_camera_config(f32(wfc), W, H, proj, fov_deg, znear, zfar, out.ctypes.data)
    return out


def gen_rays_ortho(wfc, cfr, W, H):
    out = [np.zeros(W * H, np.float32) for _ in range(6)]
    lib().orc_gen_rays_ortho(f32(wfc), f32(cfr), W, H, *out)
    return out


def decompose(s):
    """getPatch: returns dict(cp[np,(p+1)(q+1),4], uv[np,4], nbU, nbV, knot_u, knot_v)."""
    L = lib()
    nbU = L.orc_count_segments(s.U, s.n, s.p)
    nbV = L.orc_count_segments(s.V, s.m, s.q)
    npch = nbU * nbV
    ps = (s.p + 1) * (s.q + 1)
    cp = np.zeros((npch, ps, 4), np.float32)
    uv = np.zeros((npch, 4), np.float32)
    ku = np.zeros(nbU, np.int32)
    kv = np.zeros(npch, np.int32)
    rc = L.orc_decompose(s.n, s.p, s.U, s.m, s.q, s.V, f32(s.Pw), cp, uv, ku, kv)
    if rc != 0:
        raise ValueError("decompose failed")
    return dict(cp=cp, uv=uv, nbU=nbU, nbV=nbV, knot_u=ku, knot_v=kv.reshape(nbU, nbV)[0].copy(),
                p=s.p, q=s.q)


def refine(s, C_u, C_v):
    """curvatureRefinementU then V (main.cpp:100-102) on a scenes.Nurbs."""
    from dataclasses import replace
    L = lib()
    cur = s
    for d, Cc in ((0, C_u), (1, C_v)):
        if Cc is None or Cc <= 0:
            continue
        nn = cur.n if d == 0 else cur.m
        deg = cur.p if d == 0 else cur.q
        K = cur.U if d == 0 else cur.V
        add = np.zeros(nn - deg + 1, np.int32)
        total = L.orc_refine(d, cur.n, deg, K, cur.m, f32(cur.Pw), Cc, add, None, None)
        nk = np.zeros(nn + total + deg + 2, np.float32)
        if d == 0:
            P2 = np.zeros(((cur.n + total + 1) * (cur.m + 1), 4), np.float32)
        else:
            P2 = np.zeros(((cur.n + 1) * (cur.m + total + 1), 4), np.float32)
        rc = L.orc_refine(d, cur.n, deg, K, cur.m, f32(cur.Pw), Cc, add, nk.ctypes.data, P2.ctypes.data)
        if rc < 0:
            raise ValueError("knot insertion rejected")
        if d == 0:
            cur = replace(cur, n=cur.n + total, U=nk, Pw=P2)
        else:
            cur = replace(cur, m=cur.m + total, V=nk, Pw=P2)
    return cur


def patch_bounds(cp, p, q):
    out = np.zeros(6, np.float32)
    lib().orc_patch_bounds(f32(cp), p, q, out)
    return out


def build_bvh(bounds):
    """Exhaustive-SAH buildBVH over prim bounds [N,6]: (nodes structured array, perm)."""
    n = len(bounds)
    nodes = np.zeros(max(2 * n - 1, 1), NODE_DTYPE)
    perm = np.zeros(n, np.uint32)
    used = lib().orc_build_bvh(f32(bounds), n, nodes.ctypes.data, perm)
    return nodes, perm, used


class PatchSet:
    """Patches in BVH order with per-patch degree (possibly mixed)."""

    def __init__(self, nets, degs):
        # nets: list of arrays [(p+1)(q+1),4]; degs: list of (p,q)
        self.p = np.array([d[0] for d in degs], np.int32)
        self.q = np.array([d[1] for d in degs], np.int32)
        sizes = (self.p + 1) * (self.q + 1)
        self.off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
        self.cp = f32(np.concatenate([np.asarray(a, np.float32).reshape(-1, 4) for a in nets], axis=0))


def trace_f32(nodes, ps: PatchSet, rays, max_iter=7, thr_det=1e-5, thr_res=1e-4, flat=0.0, cam_k=0.0,
              delta=None, seed_base=0, threads=os.cpu_count() or 1):
    n = len(rays[0])
    hit = np.full(n, -1, np.int32)
    t = np.zeros(n, np.float32); u = np.zeros(n, np.float32); v = np.zeros(n, np.float32)
    du = np.zeros((n, 3), np.float32); dv = np.zeros((n, 3), np.float32)
    cnt = np.zeros((n, 8), np.int64)
    dl = f32(delta) if delta is not None else None
    lib().orc_trace_f32(nodes.ctypes.data, ps.cp, ps.off, ps.p, ps.q, dl.ctypes.data if dl is not None else None,
                        max_iter, thr_det, thr_res, flat, cam_k, seed_base, n, *[f32(a) for a in rays],
                        hit, t, u, v, du, dv, cnt, threads)
    return dict(hit=hit, t=t, u=u, v=v, dpdu=du, dpdv=dv, counters=cnt)


def trace_f64(nodes, ps: PatchSet, rays, max_iter=7, thr_det=1e-5, thr_res=1e-4, prune_slack=1.0, seed_base=0,
              threads=os.cpu_count() or 1, grid=0):
    """FP64 oracle trace; grid > 0 is the closest-root reference (Newton also
    from the centres of a grid x grid partition of each tested patch, its
    nearest converged root kept)."""
    n = len(rays[0])
    hit = np.full(n, -1, np.int32)
    t = np.zeros(n); u = np.zeros(n); v = np.zeros(n); t2 = np.zeros(n)
    du = np.zeros((n, 3)); dv = np.zeros((n, 3))
    cnt = np.zeros((n, 8), np.int64)
    lib().orc_trace_f64_roots(nodes.ctypes.data, ps.cp, len(ps.cp), ps.off, ps.p, ps.q, max_iter, thr_det,
                              thr_res, prune_slack, grid, seed_base, n, *[f32(a) for a in rays], hit, t, u, v, du, dv,
                              t2, cnt, threads)
    return dict(hit=hit, t=t, u=u, v=v, dpdu=du, dpdv=dv, t2=t2, counters=cnt)


def polish_f64(net, p, q, o, d, u, v):
    out = np.zeros(7)
    it = lib().orc_polish_f64(f32(net), p, q, f32(o), f32(d), float(u), float(v), out)
    return it, out


def c1_pipeline(scene):
    """Full host pipeline of the reference for a single-surface scene:
    decompose, bounds, exhaustive SAH, rays.  Returns a dict used by tests."""
    assert len(scene.surfaces) == 1
    s = scene.surfaces[0]
    dec = decompose(s)
    bnd = np.stack([patch_bounds(net, s.p, s.q) for net in dec["cp"]])
    nodes, perm, used = build_bvh(bnd)
    nets = dec["cp"][perm]
    ps = PatchSet(list(nets), [(s.p, s.q)] * len(nets))
    wfc, cfr = camera(scene.camera)
    rays = gen_rays(wfc, cfr, scene.camera.W, scene.camera.H)
    return dict(dec=dec, bounds=bnd, nodes=nodes, perm=perm, used=used, ps=ps, nets=nets, wfc=wfc, cfr=cfr,
                rays=rays)
