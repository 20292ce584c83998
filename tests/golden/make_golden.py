#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs only where the reference checkout is present: it builds
oracle/_ref/libref.so (oracle/build_ref.sh, compiling the reference headers
where they lie) and records, for the scenes below, the reference's outputs:
camera matrices, primary rays, getPatch nets and parameter ranges, the
buildBVH node array and primitive order, curvature-refinement results, and the
per-pixel trace (hit patch, t, u, v, dpdu, dpdv, Newton calls) of the
intersect.cu:33-193 loop.  Inputs are stored beside the outputs so the tests
need nothing else.  The files are data only (inputs and expected outputs).

Usage: python tests/golden/make_golden.py [camera,refine34]   (no argument: every fixture)
"""
import ctypes as C
import hashlib
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "rt-nurbs_amd"), os.path.join(ROOT, "oracle")]

import pyoracle as O  # noqa: E402
from rtnurbs import scenes  # noqa: E402


def ref_scene(R, s, cam, crop=None):
    """Run the reference host pipeline + per-pixel loop for one surface."""
    out = {}
    eye, look, up = (np.asarray(x, np.float32) for x in (cam.eye, cam.look, cam.up))
    wfc = np.zeros(16, np.float32)
    cfr = np.zeros(16, np.float32)
    R.ref_lookat(eye, look, up, wfc)
    R.ref_camera(wfc, cam.W, cam.H, cam.fov_deg, cam.znear, cam.zfar, cfr)
    n = cam.W * cam.H
    rays = [np.zeros(n, np.float32) for _ in range(6)]
    R.ref_gen_rays(wfc, cfr, cam.W, cam.H, *rays)
    a, b = C.c_int(), C.c_int()
    np_ = R.ref_get_patch(s.n, s.p, s.U, s.m, s.q, s.V, O.f32(s.Pw), None, None, C.byref(a), C.byref(b))
    ps = (s.p + 1) * (s.q + 1)
    cp = np.zeros((np_, ps, 4), np.float32)
    uv = np.zeros((np_, 4), np.float32)
    R.ref_get_patch(s.n, s.p, s.U, s.m, s.q, s.V, O.f32(s.Pw), cp.ctypes.data, uv.ctypes.data, C.byref(a),
                    C.byref(b))
    bounds = np.zeros((np_, 6), np.float32)
    for i in range(np_):
        R.ref_patch_bounds(cp[i], s.p, s.q, bounds[i])
    nodes = np.zeros(2 * np_ - 1, O.NODE_DTYPE)
    perm = np.zeros(np_, np.uint32)
    R.ref_build_bvh(cp, np_, s.p, s.q, nodes.ctypes.data, perm)
    nets = np.ascontiguousarray(cp[perm])
    if crop is not None:
        x0, y0, w, h = crop
        yy, xx = np.mgrid[y0:y0 + h, x0:x0 + w]
        sel = (yy * cam.W + xx).ravel()
    else:
        sel = np.arange(n)
    m = len(sel)
    hit = np.full(n, -1, np.int32)
    t = np.zeros(n, np.float32); u = np.zeros(n, np.float32); v = np.zeros(n, np.float32)
    du = np.zeros((n, 3), np.float32); dv = np.zeros((n, 3), np.float32); nc = np.zeros(n, np.int32)
    R.ref_trace(nodes.ctypes.data, nets, np_, s.p, s.q, None, 0.0, 0.0, n, *rays, hit, t, u, v, du, dv, nc)
    out.update(U=s.U, V=s.V, Pw=s.Pw, nmpq=np.array([s.n, s.m, s.p, s.q], np.int32),
               cam=np.array([*cam.eye, *cam.look, *cam.up, cam.fov_deg, cam.W, cam.H, cam.znear, cam.zfar],
                            np.float32),
               wfc=wfc, cfr=cfr, cp=cp, uv=uv, nbUV=np.array([a.value, b.value], np.int32), bounds=bounds,
               nodes=nodes.view(np.uint32).reshape(-1, 8), perm=perm, sel=sel.astype(np.int32),
               rays=np.stack([r[sel] for r in rays]), hit=hit[sel], t=t[sel], u=u[sel], v=v[sel], dpdu=du[sel],
               dpdv=dv[sel], newton_calls=nc[sel])
    assert m == len(sel)
    return out


def ref_refine(R, s, C_u, C_v, room=4096):
    """curvatureRefinementU then V (nurbs.h:290-486) in the reference; the
    output buffers have `room` spare knots, and the reference's returned last
    index trims them."""
    Uo = np.zeros(s.n + room + s.p + 2, np.float32)
    Po = np.zeros(((s.n + room + 1) * (s.m + 1), 4), np.float32)
    nn = R.ref_refine_u(O.f32(s.U), s.p, s.n, s.m, C_u, O.f32(s.Pw), Uo.ctypes.data, Po.ctypes.data)
    Uo, Po = Uo[:nn + s.p + 2].copy(), Po[:(nn + 1) * (s.m + 1)].copy()
    Vo = np.zeros(s.m + room + s.q + 2, np.float32)
    Po2 = np.zeros(((nn + 1) * (s.m + room + 1), 4), np.float32)
    mm = R.ref_refine_v(O.f32(s.V), s.q, s.m, nn, C_v, O.f32(Po), Vo.ctypes.data, Po2.ctypes.data)
    return nn, mm, Uo, Po, Vo[:mm + s.q + 2].copy(), Po2[:(nn + 1) * (mm + 1)].copy()


def make_refine34(R):
    """Refinement of a bicubic and a degree-4 surface (the reference's degree
    cap, constants.h:31) at C = 1, 4 and 16."""
    surfs = {"p3": scenes.random_surface(np.random.default_rng(43), deg_range=(3, 3), cp_range=(5, 6), box=4.0),
             "p4": scenes.random_surface(np.random.default_rng(44), deg_range=(4, 4), cp_range=(6, 7), box=4.0)}
    for name, s in surfs.items():
        d = dict(U=s.U, V=s.V, Pw=s.Pw, nmpq=np.array([s.n, s.m, s.p, s.q], np.int32))
        for Cc in (1.0, 4.0, 16.0):
            nn, mm, Uo, Po, Vo, Po2 = ref_refine(R, s, Cc, Cc)
            k = f"C{int(Cc)}"
            d.update({f"{k}_nm": np.array([nn, mm], np.int32), f"{k}_U": Uo, f"{k}_V": Vo})
            if Po2.size <= 65536:  # small nets in full
                d[f"{k}_Pw"] = Po2
            else:  # large nets: SHA-256 of the bytes and every 16th control point
                d[f"{k}_Pw_sha256"] = np.frombuffer(hashlib.sha256(np.ascontiguousarray(Po2).tobytes()).digest(),
                                                    np.uint8)
                d[f"{k}_Pw_every16"] = Po2[::16].copy()
            if Cc == 1.0:
                d[f"{k}_Pw_mid"] = Po
            print(f"refine {name} C={Cc:g}: n {s.n} -> {nn}, m {s.m} -> {mm}")
        np.savez_compressed(os.path.join(HERE, f"refine_{name}.npz"), **d)


def make_camera(R):
    """The reference's camera_config after init_camera, getPerspectiveCamera
    and getOrthographicCamera, and its perspective and orthographic rays, for
    a landscape and a portrait raster."""
    assert R.ref_camera_config_size() == O.CAMERA_CONFIG_BYTES
    d = {}
    for tag, (eye, look, up, W, H, fov, zn, zf) in {
            "land": ((3.0, 3.0, 0.0), (0.0, 0.0, 6.0), (0.0, 0.5, 0.5), 40, 24, 60.0, 0.01, 1000.0),
            "port": ((0.5, -2.0, 1.5), (0.5, 0.5, 0.0), (0.0, 0.0, 1.0), 18, 30, 45.0, 0.5, 80.0)}.items():
        wfc = np.zeros(16, np.float32)
        R.ref_lookat(*(np.asarray(x, np.float32) for x in (eye, look, up)), wfc)
        cfgs = {}
        for pname, proj in (("init", -1), ("persp", 1), ("ortho", 0)):
            b = np.zeros(O.CAMERA_CONFIG_BYTES, np.uint8)
            R.ref_camera_config(wfc, W, H, proj, fov, zn, zf, b.ctypes.data)
            cfgs[pname] = b
            d[f"{tag}_{pname}"] = b
        rays = {}
        for pname, fn in (("persp", R.ref_gen_rays), ("ortho", R.ref_gen_rays_ortho)):
            cfr = cfgs[pname][6 * 64:7 * 64].view(np.float32).copy()
            r = [np.zeros(W * H, np.float32) for _ in range(6)]
            fn(wfc, cfr, W, H, *r)
            d[f"{tag}_{pname}_rays"] = np.stack(r)
        d[f"{tag}_in"] = np.array([*eye, *look, *up, W, H, fov, zn, zf], np.float32)
        d[f"{tag}_wfc"] = wfc
    np.savez_compressed(os.path.join(HERE, "camera.npz"), **d)
    print("camera: configs and rays written")


def main():
    subprocess.run(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")], check=True)
    if not O.ref_available():
        sys.exit("the reference checkout is not present; nothing written")
    R = O.ref()
    only = sys.argv[1].split(",") if len(sys.argv) > 1 else None
    if only is not None:  # regenerate only the named groups (camera, refine34)
        if "camera" in only:
            make_camera(R)
        if "refine34" in only:
            make_refine34(R)
        return
    make_camera(R)
    make_refine34(R)
    cases = {
        "c1_seed0": (scenes.bicubic_grid_surface(0), scenes.c1_camera(256, 256), (96, 96, 64, 64)),
        "c1_seed7": (scenes.bicubic_grid_surface(7), scenes.c1_camera(256, 256), (64, 128, 64, 64)),
        "torus": (scenes.torus(), scenes.reference_camera(480, 270), (160, 80, 96, 64)),
        "sphere": (scenes.sphere(), scenes.reference_camera(480, 270), (170, 90, 96, 64)),
    }
    for name, (s, cam, crop) in cases.items():
        d = ref_scene(R, s, cam, crop)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **d)
        print(name, "patches", len(d["cp"]), "hits in crop", int((d["hit"] >= 0).sum()), "/", len(d["hit"]))
    # curvature refinement (nurbs.h:290-486), p = q = 2 torus at C = 4
    s = scenes.torus()
    Cc = 4.0
    nn = O.refine(s, Cc, None).n
    Uo = np.zeros(nn + s.p + 2, np.float32)
    Po = np.zeros(((nn + 1) * (s.m + 1), 4), np.float32)
    R.ref_refine_u(O.f32(s.U), s.p, s.n, s.m, Cc, O.f32(s.Pw), Uo.ctypes.data, Po.ctypes.data)
    mid = scenes.Nurbs(nn, s.m, s.p, s.q, Uo, s.V.copy(), Po)
    mm = O.refine(mid, None, Cc).m
    Vo = np.zeros(mm + s.q + 2, np.float32)
    Po2 = np.zeros(((nn + 1) * (mm + 1), 4), np.float32)
    R.ref_refine_v(O.f32(mid.V), mid.q, mid.m, mid.n, Cc, O.f32(mid.Pw), Vo.ctypes.data, Po2.ctypes.data)
    np.savez_compressed(os.path.join(HERE, "refine_torus_C4.npz"), U=s.U, V=s.V, Pw=s.Pw,
                        nmpq=np.array([s.n, s.m, s.p, s.q], np.int32), C=np.float32(Cc), U_out=Uo, V_out=Vo,
                        Pw_mid=Po, Pw_out=Po2, nm_out=np.array([nn, mm], np.int32))
    print("refine: n", s.n, "->", nn, "m", s.m, "->", mm)
    # single-call known answers: pcg_hash, bilinear seed, Newton
    xs = np.array([0, 1, 2, 3, 12345, 65535, 2 ** 31, 2 ** 32 - 1], np.uint32)
    hs = np.array([R.ref_pcg_hash(int(x)) for x in xs], np.uint32)
    rng = np.random.default_rng(11)
    c1 = ref_scene(R, scenes.bicubic_grid_surface(3), scenes.c1_camera(64, 64))
    net = c1["cp"][12]
    uvs = rng.uniform(0, 1, size=(64, 2)).astype(np.float32)
    evals = np.zeros((64, 12), np.float32)
    for i in range(64):
        R.ref_patch_eval(net, 3, 3, uvs[i, 0], uvs[i, 1], evals[i])
    # rays from random origins towards surface points (half of them nudged off)
    origins = rng.uniform(-0.5, 1.5, size=(64, 3)).astype(np.float32) + np.float32([0, 0, 1.5])
    targets = (evals[:, :3] / evals[:, 3:4]).astype(np.float32)
    targets[32:] += rng.normal(0, 0.05, size=(32, 3)).astype(np.float32)
    dirs = targets - origins
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    seeds = np.zeros((64, 2), np.float32)
    newton = np.zeros((64, 11), np.float32)
    nhit = np.zeros(64, np.int32)
    planes = np.zeros((64, 8), np.float32)
    for i in range(64):
        R.ref_bilinear_seed(net, 3, 3, origins[i], dirs[i], seeds[i])
        nhit[i] = R.ref_newton(net, 3, 3, origins[i], dirs[i], seeds[i, 0], seeds[i, 1], i, newton[i])
        R.ref_ray_planes(origins[i], dirs[i], planes[i])
    np.savez_compressed(os.path.join(HERE, "units.npz"), pcg_in=xs, pcg_out=hs, net=net, origins=origins,
                        dirs=dirs, seeds=seeds, newton=newton, newton_hit=nhit, planes=planes, uvs=uvs, evals=evals)
    print("units: newton hits", int(nhit.sum()), "/ 64")


if __name__ == "__main__":
    main()
