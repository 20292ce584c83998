#!/usr/bin/env python3
"""Benchmark: Mrays/s and mean Newton iterations per ray of the trace forms on
the 10,018-patch 1920x1080x4-spp scene (64 seeded mixed-degree NURBS), with
hit-t error against the FP64 reference trace, on one MI355X.

One step = one trace of the whole ray batch (8,294,400 rays) with scene and
rays resident in HBM.  The headline value is the best form's Mrays/s; every
form's Mrays/s, iterations per ray and FP64 error statistics are reported in
the same JSON line, with the plain FP32 method (the drop-in) as the baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rt-nurbs_amd"))

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 (AMD published; 78.6 without packed FP32)
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ray_flops(tot, n_rays, exact):
    """Algorithmic FP32 operations of a trace, priced by a fixed table on the
    work the run performed: 40 per ray (reciprocal direction and ray planes),
    33 per box test (66 per binary interior visit, whose two child boxes are
    tested), 32 per binary leaf visit, 180 per patch test (corners and
    bilinear seed), plus each Newton evaluation priced at its degree by the
    reference algorithm's operation count (counted by the kernel).  Work a
    form avoids (culled patches, fewer iterations) is not priced."""
    boxes = 66.0 * tot["interior"] + 32.0 * tot["leaves"] if exact else 33.0 * tot["interior"]
    return 40.0 * n_rays + boxes + 180.0 * tot["patch_tests"] + float(tot["newton_flops"])


def error_stats(g, ref, tie_rel=1e-5):
    """Per-pixel comparison with the FP64 reference trace."""
    h32, h64 = g["hit"], ref["hit"]
    t64 = ref["t"]
    tie = (h64 >= 0) & np.isfinite(ref["t2"]) & ((ref["t2"] - t64) <= tie_rel * np.abs(t64))
    both = (h32 >= 0) & (h64 >= 0)
    same = both & (h32 == h64)
    out = {"rays": int(len(h32)), "hit_frac": float((h32 >= 0).mean()), "ties": int(tie.sum()),
           "patch_mismatch_non_tie": int((both & (h32 != h64) & ~tie).sum()),
           "hit_miss_disagree": int(((h32 >= 0) != (h64 >= 0)).sum())}
    if same.any():
        te = np.abs(g["t"][same].astype(np.float64) - t64[same]) / np.abs(t64[same])
        ue = np.maximum(np.abs(g["u"][same] - ref["u"][same]), np.abs(g["v"][same] - ref["v"][same]))
        n32 = np.cross(g["dpdu"][same].astype(np.float64), g["dpdv"][same].astype(np.float64))
        n64 = np.cross(ref["dpdu"][same], ref["dpdv"][same])
        n32 /= np.linalg.norm(n32, axis=1, keepdims=True) + 1e-300
        n64 /= np.linalg.norm(n64, axis=1, keepdims=True) + 1e-300
        ne = np.abs(n32 - n64).max(axis=1)
        for name, e in (("t_rel", te), ("uv_abs", ue), ("normal_abs", ne)):
            out[name] = {"max": float(e.max()), "p99": float(np.quantile(e, 0.99)), "mean": float(e.mean())}
        out["t_rel_le_1e-5_frac"] = float((te <= 1e-5).mean())
    return out


def time_form(L, scene, rays, prm, steps, warmup, stream):
    hits = L.HitBuffers(rays.n)
    for _ in range(warmup):
        L.trace(scene, rays, prm, hits, totals=False, stream=stream)
    L.synchronize()
    e0, e1 = L.Event(), L.Event()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        L.trace(scene, rays, prm, hits, totals=False, stream=stream)
    e1.record(stream)
    L.synchronize()
    wall = (time.perf_counter() - t0) / steps * 1e3
    kern = e0.elapsed_ms(e1) / steps
    return hits, kern, wall


def cpu_baseline(scene_dl, rays_host, target_s=10.0):
    """The CPU restatement of the reference trace (oracle/) on the host cores,
    over a bounded sample of the same rays and the same scene arrays."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    O.lib()
    cores = max(1, min(16, os.cpu_count() or 1))
    pq = scene_dl["pq"]
    ps = O.PatchSet.__new__(O.PatchSet)
    ps.p = np.ascontiguousarray(pq[:, 0], np.int32)
    ps.q = np.ascontiguousarray(pq[:, 1], np.int32)
    ps.off = np.ascontiguousarray(scene_dl["off"], np.int32)
    ps.cp = np.ascontiguousarray(scene_dl["cp"], np.float32)
    n_all = len(rays_host[0])
    n = 16384
    while True:
        stride = max(1, n_all // n)
        sel = np.arange(0, n_all, stride)[:n]
        sample = [np.ascontiguousarray(a[sel]) for a in rays_host]
        t0 = time.perf_counter()
        O.trace_f32(scene_dl["nodes"], ps, sample, max_iter=8, thr_res=1e-5, thr_det=1e-5,
                    seed_base=0, threads=cores)
        dt = time.perf_counter() - t0
        if dt >= target_s * 0.4 or n >= n_all:
            break
        n = min(n_all, int(n * min(8.0, target_s / max(dt, 1e-3))))
    return {"value": len(sel) / dt / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"{len(sel)} rays (every {stride}th of the batch), plain method FP32 (8 iterations, tol 1e-5), "
                      f"{dt:.1f} s on {cores} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c2", "c1"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 error report")
    ap.add_argument("--forms", default="all", help="comma list of forms to time (plain,wide,...)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(rays shard by image tile, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from rtnurbs import lib as L
    from rtnurbs import scenes
    L.load()

    if args.config == "c2":
        sc = scenes.config2(0)
        spp, jitter = 4, True
        workload = "c2: 64 random NURBS (deg 2-5) refined to 10,018 rational Bezier patches, 1920x1080, 4 jittered spp"
    else:
        sc = scenes.config1(0)
        spp, jitter = 1, False
        workload = "c1: bicubic 8x8 NURBS -> 25 patches, 256x256, 1 spp"
    t0 = time.perf_counter()
    scene = L.Scene(sc.surfaces).build()
    st = scene.stats()
    log(f"scene: {st['n_patches']} patches, {st['n_nodes']} nodes, depth {st['max_depth']}, "
        f"build {time.perf_counter() - t0:.2f}s (decompose {st['build_ms_decompose']:.1f} ms, "
        f"bvh {st['build_ms_bvh']:.1f} ms, wide {st['build_ms_wide']:.1f} ms)")
    rays = L.generate_rays(sc.camera, spp=spp, jitter=jitter, seed=1)
    n = rays.n
    stream = L.stream_create()

    base = dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)
    forms = {
        "plain": dict(mode=L.MODE_EXACT, **base),
        "wide": dict(mode=L.MODE_WIDE, **base),
        "wide_cull": dict(mode=L.MODE_WIDE, hull_cull=1, **base),
        "wide_hull2": dict(mode=L.MODE_WIDE, hull_cull=2, **base),
        "persistent_hull2": dict(mode=L.MODE_PERSISTENT, hull_cull=2, **base),
        "wide_hull2_mixed": dict(mode=L.MODE_WIDE, hull_cull=2, precision=L.PREC_MIXED, cond_thr=1e4, **base),
    }
    if args.forms != "all":
        keep = set(args.forms.split(","))
        forms = {k: v for k, v in forms.items() if k in keep}

    ref = None
    if not args.no_f64:
        t0 = time.perf_counter()
        ref = L.trace_f64(scene, rays, max_iter=32, thr_det=1e-12, thr_res=1e-10).get()
        log(f"fp64 reference: {time.perf_counter() - t0:.2f}s, hits {(ref['hit'] >= 0).mean():.3f}")

    results = {}
    for name, kw in forms.items():
        prm = L.params(**kw)
        # counting run (same kernel, per-ray counters on) for work and iterations
        prm_c = L.params(count=1, **kw)
        hits_c, tot = L.trace(scene, rays, prm_c)
        g = hits_c.get()
        hits, kern_ms, wall_ms = time_form(L, scene, rays, prm, args.steps, args.warmup, stream)
        flops = ray_flops(tot, n, kw["mode"] == L.MODE_EXACT)
        r = {"mrays_s": n / (kern_ms * 1e-3) / 1e6, "kernel_ms": kern_ms, "wall_ms": wall_ms,
             "newton_iters_per_ray": tot["newton_evals"] / n, "newton_calls_per_ray": tot["newton_calls"] / n,
             "patch_tests_per_ray": tot["patch_tests"] / n, "flops_per_ray": flops / n,
             "hull_culled_per_ray": tot["hull_culled"] / n, "fp64_steps": tot["fp64_steps"],
             "tflops": flops / (kern_ms * 1e-3) / 1e12}
        if ref is not None:
            r["fp64"] = error_stats(g, ref)
        results[name] = r
        log(f"{name}: {r['mrays_s']:.1f} Mrays/s ({kern_ms:.2f} ms kernel, {wall_ms:.2f} ms wall), "
            f"iters/ray {r['newton_iters_per_ray']:.2f}, {r['tflops']:.1f} TFLOP/s")

    head = max(results, key=lambda k: results[k]["mrays_s"])
    H = results[head]
    cpu = None
    if not args.no_cpu:
        try:
            dl = scene.download()
            cpu = cpu_baseline(dl, [np.asarray(a) for a in rays.get()])
            log(f"cpu baseline: {cpu['value']:.3f} Mrays/s on {cpu['cores']} threads")
        except Exception as e:  # the baseline is a report, never the measurement
            log(f"cpu baseline failed: {e!r}")
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{head}_{args.config}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out = {
        "metric": "Mrays/s + mean Newton iters/ray at <=1e-5 hit-t err: learned/mixed vs plain",
        "value": round(H["mrays_s"], 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(H["kernel_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded NURBS scene generator)",
        "config": {"workload": workload, "rays": n, "patches": int(st["n_patches"]), "headline_form": head},
        "newton_iters_per_ray": round(H["newton_iters_per_ray"], 4),
        "speedup_vs_plain": round(H["mrays_s"] / results["plain"]["mrays_s"], 3) if "plain" in results else None,
        "forms": results,
        "roofline": {"bound": "valu", "achieved": round(H["tflops"], 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(H["tflops"] / PEAK_FP32_TFLOPS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
