"""Curvature-refinement x flat-skip sweep of the plain method, as a CSV.

The reference's benchmark driver (benchmark.cpp:122-251) renders the torus and
the sphere at 1920x1080 from its benchmark camera, refines each surface with
curvatureRefinementU/V at C in {0, 1, 2, 4, 8, 16}, and for every flat-skip
threshold eps (pixels) in {0.25 .. 8} traces the plain method against a
ground truth at eps = 0.  This harness produces the same table through this
library: GPU refinement (rtn_scene_refine), GPU decomposition and BVH (rtn_scene_build),
the exact trace form with `flat_threshold` and `cam_k` (the reference's
flat-skip, intersect.cu:104-123), HIP-event timing.

Columns, as the reference prints them (benchmark.cpp:166):
  surface, C, patches, eps_px, time_ms, hit_pct, newton_per_ray, skip_pct,
  max_pix_err, mean_pix_err, max_norm_deg
Errors are over pixels hit in both runs: the hit points' distance times cam_k
over the ground-truth t (pixels), and the angle between the normals
dpdu x dpdv, sign ignored (degrees).

    python -m rtnurbs.flat_sweep [--res 1920x1080] [--reps 5] [--cs 0,1,2,4,8,16] [--eps 0.25,0.5,1,2,4,8]
"""
from __future__ import annotations

import argparse
import math
import sys

import numpy as np

from . import lib as L
from . import scenes

HEADER = "surface,C,patches,eps_px,time_ms,hit_pct,newton_per_ray,skip_pct,max_pix_err,mean_pix_err,max_norm_deg"


def _normals(g, m):
    n = np.cross(g["dpdu"].reshape(-1, 3)[m].astype(np.float64), g["dpdv"].reshape(-1, 3)[m].astype(np.float64))
    return n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-300)


def _errors(ref, test, o, d, cam_k):
    both = (ref["hit"] >= 0) & (test["hit"] >= 0)
    if not both.any():
        return 0.0, 0.0, 0.0
    tr, tt = ref["t"][both].astype(np.float64), test["t"][both].astype(np.float64)
    dd = d[both]
    # t = (S - o).d with unit d (camera.h:167-170): the hit points are o + t d
    w = np.abs(tr - tt) * np.linalg.norm(dd, axis=1)
    px = np.where(tr > 0, w * cam_k / np.where(tr > 0, tr, 1.0), 0.0)
    c = np.abs(np.sum(_normals(ref, both) * _normals(test, both), axis=1)).clip(0.0, 1.0)
    return float(px.max()), float(px.mean()), float(np.degrees(np.arccos(c)).max())


def sweep(res=(1920, 1080), reps=5, cs=(0, 1, 2, 4, 8, 16), epss=(0.25, 0.5, 1.0, 2.0, 4.0, 8.0), out=sys.stdout):
    """Runs the sweep; writes CSV rows to `out` and returns them as dicts."""
    W, H = res
    cam = scenes.reference_camera(W, H)
    cam_k = 0.5 * H / math.tan(math.radians(cam.fov_deg) / 2.0)
    rays = L.generate_rays(cam)
    rh = [np.asarray(a, np.float64) for a in rays.get()]
    o, d = np.stack(rh[:3], 1), np.stack(rh[3:6], 1)
    n = rays.n
    rows = []
    print(HEADER, file=out, flush=True)
    for name, make in (("torus", scenes.torus), ("sphere", scenes.sphere)):
        for Cc in cs:
            scene = L.Scene([make()])
            if Cc > 0:
                scene.refine(Cc, Cc)
            scene.build()
            patches = scene.stats()["n_patches"]
            base = dict(mode=L.MODE_EXACT, cam_k=cam_k)  # reference constants: 7 iterations, 1e-5 / 1e-4
            ref, _ = L.trace(scene, rays, L.params(**base, flat_threshold=0.0))
            gref = ref.get()
            for eps in epss:
                h, tot = L.trace(scene, rays, L.params(**base, flat_threshold=eps, count=1))
                g = h.get()
                prm = L.params(**base, flat_threshold=eps)
                hits = L.HitBuffers(n)
                L.trace(scene, rays, prm, hits, totals=False)  # warm-up
                L.synchronize()
                e0, e1 = L.Event(), L.Event()
                e0.record()
                for _ in range(reps):
                    L.trace(scene, rays, prm, hits, totals=False)
                e1.record()
                L.synchronize()
                ms = e0.elapsed_ms(e1) / reps
                mx, mean, deg = _errors(gref, g, o, d, cam_k)
                calls, skips = tot["newton_calls"], tot["skips"]
                row = dict(surface=name, C=Cc, patches=int(patches), eps_px=eps, time_ms=ms,
                           hit_pct=100.0 * float(np.mean(g["hit"] >= 0)), newton_per_ray=calls / n,
                           skip_pct=100.0 * skips / (calls + skips) if calls + skips else 0.0,
                           max_pix_err=mx, mean_pix_err=mean, max_norm_deg=deg)
                rows.append(row)
                print(f"{name},{Cc:.0f},{row['patches']},{eps:.2f},{ms:.3f},{row['hit_pct']:.2f},"
                      f"{row['newton_per_ray']:.4f},{row['skip_pct']:.2f},{mx:.4f},{mean:.4f},{deg:.4f}",
                      file=out, flush=True)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--res", default="1920x1080")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--cs", default="0,1,2,4,8,16")
    ap.add_argument("--eps", default="0.25,0.5,1,2,4,8")
    a = ap.parse_args(argv)
    W, H = (int(x) for x in a.res.lower().split("x"))
    L.load()
    sweep((W, H), a.reps, [float(x) for x in a.cs.split(",")], [float(x) for x in a.eps.split(",")])


if __name__ == "__main__":
    main()
