#!/usr/bin/env python3
"""Benchmark of the per-pixel BVH traversal + Newton rational-Bezier
intersection on one MI355X: Mrays/s and mean Newton iterations per ray of
each trace form, with hit-t / (u,v) / normal error against the FP64
reference trace (the plain method in FP64) and patch mismatches against the
FP64 closest-root reference (every tested patch also solved from a 4 x 4 grid
of starts), the plain FP32 method (the drop-in) as the baseline.

Configurations (--config):
  c2  (default) 64 seeded random NURBS (degree 2-5) refined to 10,018
      rational Bezier patches, 1920x1080, 4 jittered spp = 8,294,400 rays.
      One step = one trace of the whole batch with scene and rays resident
      in HBM.  The seed networks of the learned forms are trained at start on
      the GPU (plain-method pairs from training-seed scenes; one of them on
      closest-root targets).
  c1  one bicubic NURBS (25 patches), 256x256, 1 spp.
  c3  config 1 with 20% silhouette rays; FP32 vs the FP64 correction step.
  c5  32 scenes of 1k-200k patches at 256^2..2048^2 x 4 spp with one shadow
      and one mirror ray per hit: one frame = primary launch + secondary
      launch (persistent per-lane or packet form), hipGraph-captured, launch
      shape and forms autotuned.

The JSON line on stdout is short: `value` is the headline form's Mrays/s
(c2: the learned seed + mixed precision + persistent form the target names,
with the same form without the network and the plain method beside it);
every form's full report goes to the detail file (--detail).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c3|c5] [--detail PATH]
       python bench.py --calibrate-copy BYTES   (copy of known size, for PMC scaling)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rt-nurbs_amd"))

PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 peak
PEAK_HBM_GBS = 8000.0
BASE = dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)
METRIC = "Mrays/s + mean Newton iters/ray at ≤1e-5 hit-t err: learned/mixed vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ray_flops(tot, n_rays, exact, box_tot=None):
    """Algorithmic FP32 operations of a trace, priced by a fixed table on the
    work the run performed: 40 per ray (reciprocal direction and ray planes),
    33 per box test (66 per binary interior visit, whose two child boxes are
    tested; 4 per wide node), 32 per binary leaf visit, 180 per patch test
    (corners and bilinear seed), plus each Newton evaluation priced at its
    degree by the reference algorithm's operation count (counted by the
    kernel).  Work a form avoids (rejected patches, fewer iterations) is not
    priced, and neither is work a form adds beyond the per-ray algorithm: a
    packet lane also tests the child boxes of nodes the wave visits for other
    lanes, so the packet forms' box tests are priced from `box_tot`, the
    one-ray form's counts for the same rays (their patch tests and Newton
    work are their own)."""
    bt = box_tot if box_tot is not None else tot
    boxes = 66.0 * bt["interior"] + 32.0 * bt["leaves"] if exact else 33.0 * bt["interior"]
    return 40.0 * n_rays + boxes + 180.0 * tot["patch_tests"] + float(tot["newton_flops"])


def bezier_point64(net, P, Q, u, v):
    """FP64 point of a rational Bezier patch (homogeneous net, (P+1)(Q+1) x 4,
    rows along u) at (u, v): the tensor Bernstein sum, dehomogenised."""
    from math import comb
    bu = np.array([comb(P, i) * u ** i * (1 - u) ** (P - i) for i in range(P + 1)])
    bv = np.array([comb(Q, j) * v ** j * (1 - v) ** (Q - j) for j in range(Q + 1)])
    h = np.einsum("i,j,ijk->k", bu, bv, net.reshape(P + 1, Q + 1, 4).astype(np.float64))
    return h[:3] / h[3]


def hit_distance64(geom, rays_h, idx, prim, u, v):
    """Distance in FP64 from each ray idx[k] to the surface point of its
    recorded hit (patch prim[k] at u[k], v[k]); Newton accepts a hit whose
    two ray-plane residuals sum below thr_res (1e-5 here), so a genuine hit
    lies within that distance of its ray, up to the FP32 rounding of (u, v)."""
    cp, off, pq = geom
    o = np.stack([rays_h[k][idx] for k in range(3)], 1).astype(np.float64)
    d = np.stack([rays_h[k][idx] for k in range(3, 6)], 1).astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out = np.empty(len(idx))
    for k in range(len(idx)):
        P, Q = int(pq[prim[k], 0]), int(pq[prim[k], 1])
        S = bezier_point64(cp[off[prim[k]]:off[prim[k]] + (P + 1) * (Q + 1)], P, Q, float(u[k]), float(v[k]))
        out[k] = np.linalg.norm(np.cross(S - o[k], d[k]))
    return out


def error_stats(g, ref, tie_rel=1e-5, geom=None, rays_h=None):
    """Per-pixel comparison with the FP64 reference trace.  A pixel is a tie
    when the FP64 trace's second-nearest converged hit on another patch lies
    within tie_rel of its nearest in t.  patch_mismatch_t_separated counts the
    non-tie mismatches whose FP32 hit is also more than tie_rel away in t (a
    different surface point, not the same point on a neighbouring patch)."""
    h32, h64 = g["hit"], ref["hit"]
    t64 = ref["t"]
    tie = (h64 >= 0) & np.isfinite(ref["t2"]) & ((ref["t2"] - t64) <= tie_rel * np.abs(t64))
    both = (h32 >= 0) & (h64 >= 0)
    same = both & (h32 == h64)
    mm = both & (h32 != h64) & ~tie
    out = {"rays": int(len(h32)), "hit_frac": float((h32 >= 0).mean()), "ties": int(tie.sum()),
           "patch_mismatch_non_tie": int(mm.sum()),
           "patch_mismatch_t_separated": int((np.abs(g["t"][mm].astype(np.float64) - t64[mm])
                                              > tie_rel * np.abs(t64[mm])).sum()),
           "hit_miss_disagree": int(((h32 >= 0) != (h64 >= 0)).sum())}
    # where the form and the reference disagree: a hit the reference does not
    # have, or a nearer one than the reference's (the form found another root),
    # against the reference's hits the form lost or replaced by farther ones
    extra = (h32 >= 0) & (h64 < 0)
    nearer = mm & (g["t"] < t64)
    out["disagreements"] = {"extra_hit": int(extra.sum()), "nearer_hit": int(nearer.sum()),
                            "lost_hit": int(((h32 < 0) & (h64 >= 0)).sum()), "farther_hit": int((mm & ~nearer).sum())}
    if geom is not None and (extra | nearer).any():
        idx = np.flatnonzero(extra | nearer)[::max(1, int((extra | nearer).sum()) // 2000)]
        dist = hit_distance64(geom, rays_h, idx, h32[idx], g["u"][idx], g["v"][idx])
        out["disagreements"]["extra_or_nearer_checked"] = int(len(idx))
        out["disagreements"]["extra_or_nearer_on_surface_frac"] = float((dist <= 2e-5).mean())
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
        # the same on the pixels where the form reached the reference's root
        # ((u, v) within 1e-3): the precision of the hit, apart from root choice
        root = ue < 1e-3
        out["same_root_pixels"] = int(root.sum())
        out["t_rel_le_1e-5_frac_same_root"] = float((te[root] <= 1e-5).mean()) if root.any() else None
        out["t_rel_max_same_root"] = float(te[root].max()) if root.any() else None
    return out


def time_form(L, scene, rays, prm, steps, warmup, stream, per_launch=None):
    """Kernel time per step from HIP events on the launch stream (mean over the
    timed launches), and wall time.  An event between consecutive launches
    also gives each launch's own time; `per_launch` (a dict) receives their
    median and minimum."""
    hits = L.HitBuffers(rays.n)
    for _ in range(warmup):
        L.trace(scene, rays, prm, hits, totals=False, stream=stream)
    L.synchronize()
    ev = [L.Event() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(steps):
        L.trace(scene, rays, prm, hits, totals=False, stream=stream)
        ev[i + 1].record(stream)
    L.synchronize()
    wall = (time.perf_counter() - t0) / steps * 1e3
    kern = ev[0].elapsed_ms(ev[steps]) / steps
    if per_launch is not None and steps > 0:
        each = np.array([ev[i].elapsed_ms(ev[i + 1]) for i in range(steps)])
        per_launch.update(median=float(np.median(each)), min=float(each.min()))
    return hits, kern, wall


def host_info():
    """CPU model and logical core count of the machine the baseline runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def cpu_baseline(scene_dl, rays_host, target_s=10.0):
    """The CPU restatement of the reference trace (oracle/, plain method,
    FP32) on the host cores, over a bounded sample of the same rays and the
    same scene arrays (the whole batch when it fits the time budget)."""
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
    # one thread over a 1/16 sample of the same rays
    sel1 = sel[::16]
    s1 = [np.ascontiguousarray(a[sel1]) for a in rays_host]
    t0 = time.perf_counter()
    O.trace_f32(scene_dl["nodes"], ps, s1, max_iter=8, thr_res=1e-5, thr_det=1e-5, seed_base=0, threads=1)
    dt1 = time.perf_counter() - t0
    return {"value": len(sel) / dt / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"{len(sel)} of {n_all} rays (every {stride}th), plain method FP32 (8 iterations, tol 1e-5), "
                      f"{dt:.1f} s on {cores} threads",
            "single_thread": {"value": len(sel1) / dt1 / 1e6, "unit": "Mrays/s",
                              "sample": f"{len(sel1)} rays (every {16 * stride}th), {dt1:.2f} s on 1 thread"},
            "host": host_info()}


def cpu_reference(scene_dl, rays_host):
    """The reference's own intersection code (its headers compiled on the host
    by oracle/build_ref.sh, per-pixel loop of intersect.cu:33-193, its own
    constants: 7 iterations, tolerance 1e-4), one thread, over the whole
    batch.  Only for single-degree scenes of degree <= 4, as the reference."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    if not O.ref_available():
        return None
    R = O.ref()
    pq = scene_dl["pq"]
    p, q = int(pq[0, 0]), int(pq[0, 1])
    if not (np.all(pq[:, 0] == p) and np.all(pq[:, 1] == q)) or max(p, q) > 4:
        return None
    n = len(rays_host[0])
    f = [np.ascontiguousarray(a, np.float32) for a in rays_host]
    hit = np.zeros(n, np.int32)
    outs = [np.zeros(n, np.float32) for _ in range(3)] + [np.zeros(3 * n, np.float32) for _ in range(2)]
    newton = np.zeros(n, np.int32)
    cp = np.ascontiguousarray(scene_dl["cp"], np.float32)
    t0 = time.perf_counter()
    R.ref_trace(scene_dl["nodes"].ctypes.data, cp, len(pq), p, q, None, 0.0, 0.0, n, *f, hit, *outs, newton)
    dt = time.perf_counter() - t0
    return {"value": n / dt / 1e6, "unit": "Mrays/s", "cores": 1, "kind": "reference",
            "sample": f"all {n} rays, the reference's trace code compiled on the host (its constants: 7 iterations, "
                      f"tol 1e-4), {dt:.2f} s on one thread"}


def load_traffic(form, config):
    tf = os.path.join(ROOT, "profiles", f"traffic_{form}_{config}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def roofline(tflops, form, config):
    return {"bound": "fp32-valu", "achieved": round(tflops, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / PEAK_FP32_TFLOPS, 4), "traffic": load_traffic(form, config),
            "kernel": form}


def run_forms(L, scene, rays, forms, args, ref, stream, ref_roots=None):
    results = {}
    n = rays.n
    geom, rays_h = None, None
    if ref is not None:
        dl = scene.download()
        geom = (dl["cp"], dl["off"], dl["pq"])
        rays_h = [np.asarray(a) for a in rays.get()]
    for name, kw in forms.items():
        prm = L.params(**kw)
        hits_c, tot = L.trace(scene, rays, L.params(count=1, **kw))   # counting run: work and iterations
        g = hits_c.get()
        each = {}
        hits, kern_ms, wall_ms = time_form(L, scene, rays, prm, args.steps, args.warmup, stream, each)
        overflow = scene.trace_overflow()  # the last timed launch's stack-overflow count
        box = tot
        if kw["mode"] in (L.MODE_PACKET, L.MODE_PERSISTENT_PACKET):
            # box tests priced by the one-ray form's counts of the same rays
            _, box = L.trace(scene, rays, L.params(count=1, **{**kw, "mode": L.MODE_WIDE}))
        flops = ray_flops(tot, n, kw["mode"] == L.MODE_EXACT, box)
        r = {"mrays_s": n / (kern_ms * 1e-3) / 1e6, "kernel_ms": kern_ms, "wall_ms": wall_ms,
             "overflow": overflow + tot["overflow"],
             "box_tests_per_ray": tot["interior"] / n, "box_tests_per_ray_priced": box["interior"] / n,
             "kernel_ms_median": each.get("median"), "kernel_ms_min": each.get("min"),
             "newton_iters_per_ray": tot["newton_evals"] / n, "newton_calls_per_ray": tot["newton_calls"] / n,
             "patch_tests_per_ray": tot["patch_tests"] / n, "flops_per_ray": flops / n,
             "hull_culled_per_ray": tot["hull_culled"] / n, "fp64_steps": tot["fp64_steps"],
             "learned_fallbacks_per_ray": tot["learned_fallbacks"] / n,
             "learned_skips_per_ray": tot["learned_skips"] / n,
             "tflops": flops / (kern_ms * 1e-3) / 1e12}
        if ref is not None:
            r["fp64"] = error_stats(g, ref, geom=geom, rays_h=rays_h)
        if ref_roots is not None:  # against the nearest root the FP64 searches find
            r["fp64_closest_root"] = error_stats(g, ref_roots)
        results[name] = r
        log(f"{name}: {r['mrays_s']:.1f} Mrays/s ({kern_ms:.3f} ms kernel, {wall_ms:.3f} ms wall), "
            f"iters/ray {r['newton_iters_per_ray']:.3f}, {r['tflops']:.1f} TFLOP/s"
            + (f", non-tie patch mismatches {r['fp64']['patch_mismatch_non_tie']}" if ref is not None else ""))
    return results


def seed_ablation(L, scene, rays, net):
    """On the (ray, patch) pairs the plain method solves on this scene (which
    the network never saw): seed error and Newton evaluations from the plain
    bilinear seed and from the network's seed."""
    prm = L.params(mode=L.MODE_WIDE, hull_cull=2, **BASE)
    pr, pp, puv, cnt = L.trace_pairs(scene, rays, prm, 2 * rays.n)
    cnt = min(cnt, 2 * rays.n)
    target = puv.get()[:cnt]
    luv = net.forward(L.pair_features(scene, rays, pr, pp, cnt, kind=net.kind))
    seed_p = L.solve_pairs(scene, rays, pr, pp, cnt, L.params(**{**BASE, "max_iter": 1}))["uv"]
    a = L.solve_pairs(scene, rays, pr, pp, cnt, L.params(**BASE))
    b = L.solve_pairs(scene, rays, pr, pp, cnt, L.params(**BASE), seed_uv=luv)
    both = a["converged"] & b["converged"]
    ep = np.abs(seed_p - target).max(1)
    el = np.abs(luv.get() - target).max(1)
    out = {"pairs": int(cnt), "seed_err_median_plain": float(np.median(ep)), "seed_err_median_learned": float(np.median(el)),
           "iters_per_solve_plain": float(a["iters"][both].mean()), "iters_per_solve_learned": float(b["iters"][both].mean()),
           "converged_plain": float(a["converged"].mean()), "converged_learned": float(b["converged"].mean()),
           "converged_learned_or_fallback": float((a["converged"] | b["converged"]).mean())}
    # the hit output on every solve of the mixed-precision form without the network
    prm_all = L.params(mode=L.MODE_WIDE, hull_cull=2, precision=L.PREC_MIXED, cond_thr=1e4, stall_exit=1, **BASE)
    ar, ap, auv, acnt = L.trace_pairs(scene, rays, prm_all, 4 * rays.n, all=True)
    acnt = min(acnt, 4 * rays.n)
    _, ph = net.forward(L.pair_features(scene, rays, ar, ap, acnt, kind=net.kind), hit=True)
    h, ph = auv.get()[:acnt, 0] >= 0, ph.get()
    order = np.argsort(ph, kind="stable")
    ranks = np.empty(len(ph))
    ranks[order] = np.arange(1, len(ph) + 1)
    npos, nneg = int(h.sum()), int((~h).sum())
    out["hit_output"] = {"solves": int(acnt), "converged": npos,
                         "auc": float((ranks[h].sum() - npos * (npos + 1) / 2) / max(1, npos * nneg))}
    for thr in (0.05, 0.2, 0.5):
        out["hit_output"][f"below_{thr}"] = {"misses": float((ph[~h] < thr).mean()), "hits": float((ph[h] < thr).mean())}
    log(f"seed ablation (kind {net.kind}) on {cnt} solved pairs: seed error {out['seed_err_median_plain']:.3f} -> "
        f"{out['seed_err_median_learned']:.3f}, evaluations per solve {out['iters_per_solve_plain']:.3f} -> "
        f"{out['iters_per_solve_learned']:.3f}; hit output AUC {out['hit_output']['auc']:.4f}, below 0.5: "
        f"{out['hit_output']['below_0.5']['misses']:.3f} of misses, {out['hit_output']['below_0.5']['hits']:.4f} of hits")
    return out


def make_line(args, value, ms, dtype, workload, extra):
    out = {"metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded NURBS scene generators)",
           "config": workload}
    out.update(extra)
    return out


def bench_c2_c1(args, L, scenes, learned):
    if args.config == "c2":
        sc = scenes.config2(0)
        spp, jitter = 4, True
        workload = "10k-patch 1080p: 64 random NURBS (deg 2-5) refined to 10,018 rational Bezier patches, " \
                   "1920x1080, 4 jittered spp"
    else:
        sc = scenes.config1(0)
        spp, jitter = 1, False
        workload = "bicubic 8x8 NURBS -> 25 patches, 256x256, 1 spp"
    t0 = time.perf_counter()
    scene = L.Scene(sc.surfaces).build()
    st = scene.stats()
    log(f"scene: {st['n_patches']} patches, {st['n_nodes']} nodes, depth {st['max_depth']}, "
        f"build {time.perf_counter() - t0:.2f}s (decompose {st['build_ms_decompose']:.1f} ms, "
        f"bvh {st['build_ms_bvh']:.1f} ms, wide {st['build_ms_wide']:.1f} ms)")
    # tiled ray layout (8x8-pixel Morton tiles): the same camera rays, ordered so
    # that a wave's 64 rays cover a compact pixel block; every form gets it
    rays = L.generate_rays(sc.camera, spp=spp, jitter=jitter, seed=1, tiled=True)
    stream = L.stream_create()
    nets, train = {}, None
    # the third network: projected inputs, trained on closest-root targets (each
    # training pair solved from a 4 x 4 grid of starts as well, nearest root kept)
    kinds = {"projected": L.MLP_FEAT_PROJECTED, "box": L.MLP_FEAT_BOX, "projected_root": L.MLP_FEAT_PROJECTED}
    root_grids = {"projected_root": 4}
    if args.load_net:  # parameters saved by an earlier run (profiling passes reuse one training)
        z = np.load(args.load_net)
        nets = {k: L.Mlp(params=z[k], kind=kinds[k]) for k in kinds if k in z}
        train = {"loaded": args.load_net}
    elif args.train_epochs > 0:
        train = {}
        # the same training scenes for both feature kinds: the pairs are
        # collected once per kind (the features differ), then trained
        sl = learned.training_scenes()
        for name, kind in kinds.items():
            t0 = time.perf_counter()
            feats, targets, n, hits = learned.collect_pairs(sl, max_hits=10_000_000, kind=kind,
                                                            root_grid=root_grids.get(name, 0))
            t1 = time.perf_counter()
            L.synchronize()
            net, losses = learned.train_seed_network(feats, targets, n, epochs=args.train_epochs, kind=kind)
            L.synchronize()
            t2 = time.perf_counter()
            del feats, targets
            nets[name] = net
            train[name] = {"pairs": int(n), "converged_pairs": int(hits), "epochs": args.train_epochs,
                           "closest_root_grid": root_grids.get(name, 0),
                           "collect_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 3),
                           "train_mpairs_s": round(n * args.train_epochs / (t2 - t1) / 1e6, 1),
                           "epoch_loss": [round(float(x), 6) for x in losses]}
            log(f"seed network ({name} features): {n} pairs ({hits} converged), {args.train_epochs} epochs, "
                f"loss {losses[0]:.4f} -> {losses[-1]:.4f}, collect {t1 - t0:.2f} s, train {t2 - t1:.2f} s "
                f"({train[name]['train_mpairs_s']} Mpairs/s)")
    if nets and args.save_net:
        np.savez(args.save_net, **{k: v.get_params() for k, v in nets.items()})
    net = nets.get("projected")
    P, W, K, KP = L.MODE_PERSISTENT, L.MODE_WIDE, L.MODE_PACKET, L.MODE_PERSISTENT_PACKET
    mixed = dict(precision=L.PREC_MIXED, cond_thr=1e4)
    forms = {
        "plain": dict(mode=L.MODE_EXACT, **BASE),
        "wide": dict(mode=W, **BASE),
        # one wave traverses the BVH for its 64 rays together (packet form)
        "packet": dict(mode=K, **BASE),
        "wide_hull2": dict(mode=W, hull_cull=2, **BASE),
        "persistent_hull2": dict(mode=P, hull_cull=2, **BASE),
        "wide_hull2_mixed": dict(mode=W, hull_cull=2, **mixed, **BASE),
        "persistent_hull2_mixed": dict(mode=P, hull_cull=2, **mixed, **BASE),
        # abandon solves stuck against the patch boundary (misses) from the 4th evaluation
        "wide_hull2_mixed_stall": dict(mode=W, hull_cull=2, stall_exit=1, **mixed, **BASE),
        "persistent_hull2_mixed_stall": dict(mode=P, hull_cull=2, stall_exit=1, **mixed, **BASE),
        "packet_hull2_mixed": dict(mode=K, hull_cull=2, **mixed, **BASE),
        "packet_hull2_mixed_stall": dict(mode=K, hull_cull=2, stall_exit=1, **mixed, **BASE),
        # the headline form without the network
        "persistent_packet_hull2_mixed_stall": dict(mode=KP, hull_cull=2, stall_exit=1, **mixed, **BASE),
    }
    if net is not None:
        # a 4-evaluation attempt from the network's seed, then the plain seed's full 8; the
        # four lanes of a quad testing the same patch share one network evaluation; pairs the
        # hit output rates below hit_gate (0.7) start from the plain seed directly; in the
        # packet forms the network runs only on patch visits of all 64 lanes.  Measured
        # sweep on this scene: 2 evaluations fall back on a third of the solves, 3 still
        # cost more evaluations per ray than the form without the network, a 6-evaluation
        # fallback loses hits; a lower gate sends more misses through the attempt; the
        # network on sparse visits (silhouettes, partial packets) costs time and saves no
        # evaluations.
        lrn = dict(seed_mode=L.SEED_LEARNED, seed_net=net, learned_iters=4, fallback_iters=8, seed_share=1,
                   hit_gate=args.hit_gate, learned_min_lanes=args.learned_min_lanes)
        # the one-ray forms run the network for any pending lane (a 64-lane threshold
        # would rarely be met by their scattered pairs)
        lrw = {**lrn, "learned_min_lanes": 0}
        forms["wide_learned"] = dict(mode=W, hull_cull=2, **lrw, **BASE)
        forms["persistent_learned_mixed_stall"] = dict(mode=P, hull_cull=2, stall_exit=1, **lrw, **mixed, **BASE)
        forms["packet_learned_mixed_stall"] = dict(mode=K, hull_cull=2, stall_exit=1, **lrn, **mixed, **BASE)
        forms["persistent_packet_learned_mixed"] = dict(mode=KP, hull_cull=2, **lrn, **mixed, **BASE)
        forms["persistent_packet_learned_mixed_stall"] = dict(mode=KP, hull_cull=2, stall_exit=1, **lrn, **mixed, **BASE)
        # the hit output also skips the pairs it rates below hit_skip (no solve at all)
        forms["persistent_packet_learned_skip_mixed_stall"] = dict(mode=KP, hull_cull=2, stall_exit=1,
                                                                   **{**lrn, "hit_skip": args.hit_skip}, **mixed,
                                                                   **BASE)
        forms["persistent_packet_learned_nogate_mixed_stall"] = dict(mode=KP, hull_cull=2, stall_exit=1,
                                                                     **{**lrn, "hit_gate": 0.0}, **mixed, **BASE)
        # the network on every patch visit, however few lanes test the patch
        forms["persistent_packet_learned_alllanes_mixed_stall"] = dict(mode=KP, hull_cull=2, stall_exit=1,
                                                                       **{**lrn, "learned_min_lanes": 0}, **mixed,
                                                                       **BASE)
    if "projected_root" in nets:
        # closest-root targets and no hit gate: every full-packet pair starts from the
        # network's seed, which reaches roots nearer than the bilinear seed's
        lrr = dict(seed_mode=L.SEED_LEARNED, seed_net=nets["projected_root"], learned_iters=4, fallback_iters=8,
                   seed_share=1, hit_gate=0.0, learned_min_lanes=args.learned_min_lanes)
        forms["persistent_packet_learned_root_nogate_mixed_stall"] = dict(mode=KP, hull_cull=2, stall_exit=1, **lrr,
                                                                          **mixed, **BASE)
    if "box" in nets:  # the network on the box-frame inputs (ray, four corners, degrees)
        lrb = dict(seed_mode=L.SEED_LEARNED, seed_net=nets["box"], learned_iters=4, fallback_iters=8, seed_share=1,
                   hit_gate=args.hit_gate, learned_min_lanes=args.learned_min_lanes)
        forms["persistent_packet_learned_box_mixed_stall"] = dict(mode=KP, hull_cull=2, stall_exit=1, **lrb, **mixed,
                                                                  **BASE)
    if args.forms != "all":
        keep = set(args.forms.split(","))
        forms = {k: v for k, v in forms.items() if k in keep}
    ref = None
    if not args.no_f64:
        t0 = time.perf_counter()
        ref = L.trace_f64(scene, rays, max_iter=32, thr_det=1e-12, thr_res=1e-10).get()
        log(f"fp64 reference: {time.perf_counter() - t0:.2f}s, hits {(ref['hit'] >= 0).mean():.3f}")
        # the closest-root reference: every tested patch also solved from a 4 x 4 grid of
        # starts, its nearest converged root kept (roots the plain seed misses count)
        t0 = time.perf_counter()
        ref_roots = L.trace_f64(scene, rays, max_iter=32, thr_det=1e-12, thr_res=1e-10, grid=4).get()
        log(f"fp64 closest-root reference: {time.perf_counter() - t0:.2f}s, hits {(ref_roots['hit'] >= 0).mean():.4f}; "
            f"nearer than the plain-seed reference on {int(((ref_roots['t'] < ref['t'] * (1 - 1e-5)) & (ref['hit'] >= 0)).sum())}"
            f" pixels, hits it lacks on {int(((ref_roots['hit'] >= 0) & (ref['hit'] < 0)).sum())}")
    results = run_forms(L, scene, rays, forms, args, ref, stream, ref_roots=ref_roots if ref is not None else None)
    ablation = {k: seed_ablation(L, scene, rays, v) for k, v in nets.items()} or None
    # headline: the learned seed + mixed precision + persistent form the target names;
    # beside it the same form without the network and the plain method
    north = "persistent_packet_learned_mixed_stall"
    twin = "persistent_packet_hull2_mixed_stall"
    if north in results:
        head = north
    else:
        cand = [k for k in results if "mixed" in k or "learned" in k] or list(results)
        head = max(cand, key=lambda k: results[k]["mrays_s"])
    H = results[head]
    cpu, cpu_port = None, None
    if not args.no_cpu:
        try:
            dl = scene.download()
            rh = [np.asarray(a) for a in rays.get()]
            cpu_port = cpu_baseline(dl, rh)
            log(f"cpu baseline (port): {cpu_port['value']:.3f} Mrays/s on {cpu_port['cores']} threads, "
                f"{cpu_port['single_thread']['value']:.3f} on one ({cpu_port['host']['cpu_model']})")
            cpu = cpu_reference(dl, rh) if args.config == "c1" else None
            if cpu is not None:
                log(f"cpu baseline (reference): {cpu['value']:.3f} Mrays/s on {cpu['cores']} thread")
            else:
                cpu = cpu_port
        except Exception as e:  # the baseline is a report, never the measurement
            log(f"cpu baseline failed: {e!r}")
    plain = results.get("plain", {}).get("mrays_s")
    detail = {"forms": results, "seed_network": train, "seed_ablation": ablation, "cpu_baseline": cpu,
              "cpu_port_baseline": cpu_port if cpu is not cpu_port else None,
              "speedups_vs_plain": {k: round(v["mrays_s"] / plain, 3) for k, v in results.items()} if plain else None}
    write_detail(args, detail)
    extra = {"newton_iters_per_ray": round(H["newton_iters_per_ray"], 4),
             "overflow": max(int(r["overflow"]) for r in results.values()),
             "speedup_vs_plain": round(H["mrays_s"] / plain, 3) if plain else None,
             "summary": {k: brief(results[k]) for k in ("plain", twin, head) if k in results},
             "roofline": roofline(H["tflops"], head, args.config), "cpu_baseline": short_cpu(cpu),
             "kernel_ms_min": round(H["kernel_ms_min"], 4),
             "detail": args.detail}
    wl = {"workload": workload, "rays": rays.n, "patches": int(st["n_patches"]), "headline_form": head,
          "no_network_form": twin if twin in results else None}
    return make_line(args, H["mrays_s"], H["kernel_ms"], "f32", wl, extra)


def brief(r):
    """A form's figures for the stdout line (the full report is in the detail file)."""
    out = {"mrays_s": round(r["mrays_s"], 1), "iters_per_ray": round(r["newton_iters_per_ray"], 3),
           "fp64_steps": r["fp64_steps"]}
    e = r.get("fp64")
    if e:
        out.update(mismatch_non_tie=e["patch_mismatch_non_tie"], hit_miss=e["hit_miss_disagree"],
                   t_le_1e5=round(e.get("t_rel_le_1e-5_frac", float("nan")), 6),
                   t_le_1e5_same_root=round(e.get("t_rel_le_1e-5_frac_same_root") or float("nan"), 7))
        er = r.get("fp64_closest_root")
        if er:  # the same against the FP64 closest-root reference
            out.update(root_mismatch=er["patch_mismatch_non_tie"], root_hit_miss=er["hit_miss_disagree"])
        dg = e.get("disagreements")
        if dg:  # the reference's hits lost (missing, or a farther hit) vs hits it lacks (extra, or nearer)
            out.update(ref_lost=dg["lost_hit"] + dg["farther_hit"], ref_gained=dg["extra_hit"] + dg["nearer_hit"])
    return out


def short_cpu(c):
    if c is None:
        return None
    return {k: c[k] for k in ("value", "unit", "cores", "kind", "sample")} | {"value": round(c["value"], 4)}


def write_detail(args, detail):
    os.makedirs(os.path.dirname(os.path.abspath(args.detail)), exist_ok=True)
    with open(args.detail, "w") as f:
        json.dump(detail, f, indent=1)
    log(f"detail written to {args.detail}")


def bench_c3(args, L, scenes):
    sc = scenes.config1(0)
    scene = L.Scene(sc.surfaces).build()
    rays, mask = scenes.config3_rays(scene, seed=0)
    stream = L.stream_create()
    ref = L.trace_f64(scene, rays).get()
    forms = {"plain": dict(mode=L.MODE_EXACT, **BASE),
             "wide_hull2": dict(mode=L.MODE_WIDE, hull_cull=2, **BASE),
             "wide_hull2_mixed": dict(mode=L.MODE_WIDE, hull_cull=2, precision=L.PREC_MIXED, cond_thr=1e4, **BASE)}
    for thr in (1e2, 1e1, 2.0):  # threshold sweep: how many rays the FP64 step needs for the t bar
        forms[f"wide_hull2_mixed_cond{thr:g}"] = dict(mode=L.MODE_WIDE, hull_cull=2, precision=L.PREC_MIXED,
                                                      cond_thr=thr, **BASE)
    results = run_forms(L, scene, rays, forms, args, ref, stream)
    for name, kw in forms.items():  # error on the silhouette rays alone
        g = L.trace(scene, rays, L.params(**kw))[0].get()
        sub = {k: v[mask] for k, v in g.items()}
        rsub = {k: v[mask] for k, v in ref.items()}
        results[name]["fp64_silhouette"] = error_stats(sub, rsub)
    head = "wide_hull2_mixed"
    H = results[head]
    summ = {}
    for k in ("plain", "wide_hull2", head):
        b = brief(results[k])
        e = results[k].get("fp64_silhouette", {})
        b["silhouette_t_le_1e5"] = round(e.get("t_rel_le_1e-5_frac", float("nan")), 6)
        b["silhouette_t_le_1e5_same_root"] = round(e.get("t_rel_le_1e-5_frac_same_root") or float("nan"), 6)
        b["silhouette_mismatch_non_tie"] = e.get("patch_mismatch_non_tie")
        summ[k] = b
    wl = {"workload": "config 1 scene, 65,536 rays of which 13,107 silhouette rays (grazing 0-5 deg)",
          "rays": rays.n, "headline_form": head}
    cpu = None
    if not args.no_cpu:  # the reference's own trace code on the same rays (a single-degree scene it can run)
        try:
            dl = scene.download()
            rh = [np.asarray(a) for a in rays.get()]
            cpu = cpu_reference(dl, rh) or cpu_baseline(dl, rh)
            log(f"cpu baseline ({cpu['kind']}): {cpu['value']:.3f} Mrays/s on {cpu['cores']} thread(s)")
        except Exception as e:  # the baseline is a report, never the measurement
            log(f"cpu baseline failed: {e!r}")
    write_detail(args, {"forms": results, "cpu_baseline": cpu})
    return make_line(args, H["mrays_s"], H["kernel_ms"], "f32+f64", wl,
                     {"summary": summ, "roofline": roofline(H["tflops"], head, "c3"), "cpu_baseline": short_cpu(cpu),
                      "detail": args.detail})


def bench_c5(args, L, scenes):
    t0 = time.perf_counter()
    items = scenes.config5(0, n_scenes=args.c5_scenes, scale=args.c5_scale)
    built = [L.Scene(s.surfaces).build(bvh=L.BVH_BINNED_SAH) for s in items]
    n_patches = sum(b.stats()["n_patches"] for b in built)
    batch = L.Batch(built)
    log(f"c5: {len(built)} scenes, {n_patches} patches, built in {time.perf_counter() - t0:.1f}s")

    def frame_for(scale):
        cams = [scenes.Camera(s.camera.eye, s.camera.look, s.camera.up, s.camera.fov_deg,
                              max(1, s.camera.W // scale), max(1, s.camera.H // scale)) for s in items]
        counts = [c.W * c.H * 4 for c in cams]
        rays = L.Rays(sum(counts), with_idx=False)
        off = 0
        for cam, nr in zip(cams, counts):
            L.generate_rays(cam, spp=4, jitter=True, seed=1, rays=rays, offset=off, tiled=True)
            off += nr
        ids = L.DeviceArray.from_host(np.repeat(np.arange(len(items), dtype=np.int32), counts))
        fp = L.frame_params(**BASE, hull_cull=2, stall_exit=1, precision=L.PREC_MIXED, cond_thr=1e4)
        return L.Frame(batch, rays, ids, fp), rays, ids

    frame, rays, ids = frame_for(1)
    if args.c5_packet_only:  # no tuning: the default shape, both launches in the packet form (for profiling)
        table = []
        best = {"lanes": 128, "stack": 20, "blocks_per_cu": 8, "refill_min": 64, "packet_primary": 1,
                "packet_secondary": 1}
    else:
        # tuned on the frame itself, not on a lower-resolution proxy: a packet's
        # coherence depends on the resolution (at 1/16 of the pixels the packet
        # launches lose, at full resolution they win)
        t0 = time.perf_counter()
        table = frame.autotune(reps=2)
        best = min((e for e in table if e["ms"] > 0), key=lambda e: e["ms"])
        default = [e for e in table if (e["lanes"], e["stack"], e["blocks_per_cu"], e["refill_min"],
                                        e["packet_primary"], e["packet_secondary"]) == (128, 20, 8, 64, 0, 0)]
        log(f"autotune ({len(table)} candidates on the full frame, {time.perf_counter() - t0:.1f} s): best {best}, "
            f"default {default}")
    frame.set_shape(best["lanes"], best["stack"], best["blocks_per_cu"])
    frame.set_refill(best["refill_min"])
    stream = L.stream_create()
    # the primary and secondary launches in the tuned persistent per-lane form
    # or in the packet form (one wave per 64 rays): each variant timed as a
    # captured graph, the fastest reported
    variants = {}
    for pp, ps in ((1, 1),) if args.c5_packet_only else ((0, 0), (1, 0), (1, 1)):
        frame.set_packet(pp, ps)
        frame.capture(stream)
        for _ in range(args.warmup):
            frame.run(stream)
        L.synchronize()
        e0, e1 = L.Event(), L.Event()
        e0.record(stream)
        for _ in range(args.steps):
            frame.run(stream)
        e1.record(stream)
        L.synchronize()
        variants[(pp, ps)] = e0.elapsed_ms(e1) / args.steps
        st = frame.status()
        total_rays = st["n_primary"] + st["n_secondary"]
        log(f"c5 frame (packet primary {pp}, secondary {ps}): {st['n_primary']} primary + {st['n_secondary']} "
            f"secondary rays, {variants[(pp, ps)]:.2f} ms, {total_rays / variants[(pp, ps)] / 1e3:.1f} Mrays/s")
    log(f"autotuner's choice: packet primary {best['packet_primary']}, secondary {best['packet_secondary']}")
    (pp, ps), ms = min(variants.items(), key=lambda kv: kv[1])
    # algorithmic work of the reported frame: its launches' Newton and patch-test
    # counts, box tests priced by the per-lane form's counts of the same rays
    frame.set_packet(pp, ps)
    c_prim, c_sec = frame.count()
    frame.set_packet(0, 0)
    b_prim, b_sec = frame.count()
    st = frame.status()
    flops = (ray_flops(c_prim, st["n_primary"], False, b_prim) + ray_flops(c_sec, st["n_secondary"], False, b_sec))
    tflops = flops / (ms * 1e-3) / 1e12
    log(f"c5 frame work: {flops / 1e9:.1f} GFLOP priced, {tflops:.1f} TFLOP/s, newton evals "
        f"{c_prim['newton_evals'] + c_sec['newton_evals']}, fp64 steps {c_prim['fp64_steps'] + c_sec['fp64_steps']}")
    names = {0: "persistent per-lane", 1: "packet"}
    wl = {"workload": f"{len(built)} scenes, {n_patches} patches (1k-200k each, degree 2-5), 256^2-2048^2 x 4 spp "
                      "+ shadow and mirror ray per hit, one graph-captured frame",
          "form": f"primary rays {names[pp]}, secondary rays {names[ps]}; projected-hull rejection, mixed precision, "
                  "boundary-stall exit",
          "frame_ms": {f"primary {names[a]}, secondary {names[b]}": round(v, 3) for (a, b), v in variants.items()},
          "primary_rays": st["n_primary"], "secondary_rays": st["n_secondary"], "patches": n_patches,
          "launch_shape": {"lanes": st["shape_lanes"], "stack": st["shape_stack"],
                           "blocks_per_cu": st["shape_blocks_per_cu"], "refill_min": st["refill_min"]}}
    wl["newton_iters_per_ray"] = round((c_prim["newton_evals"] + c_sec["newton_evals"]) / total_rays, 4)
    rf = roofline(tflops, "frame", "c5")
    rf["kernel"] = f"frame: primary {names[pp]} + secondary {names[ps]} launches"
    cpu = None
    if not args.no_cpu:
        # the oracle's plain method on the host cores over a bounded sample of
        # the frame's primary rays, spread over all scenes (strided), each scene's
        # sample traced on its own patches; secondaries are not sampled
        try:
            cpu = cpu_frame_sample(L, built, rays, ids, target_s=12.0)
            log(f"cpu baseline (port): {cpu['value']:.3f} Mrays/s on {cpu['cores']} threads ({cpu['sample']})")
        except Exception as e:  # the baseline is a report, never the measurement
            log(f"cpu baseline failed: {e!r}")
    write_detail(args, {"autotune": table, "frame_ms": wl["frame_ms"], "work_primary": c_prim,
                        "work_secondary": c_sec, "box_work_per_lane": [b_prim, b_sec], "cpu_baseline": cpu})
    return make_line(args, total_rays / ms / 1e3, ms, "f32", wl,
                     {"roofline": rf, "cpu_baseline": short_cpu(cpu), "detail": args.detail})


def cpu_frame_sample(L, built, rays, ids, target_s=12.0):
    """The oracle's plain FP32 method on the host cores over a strided sample of
    a multi-scene frame's primary rays: each sampled ray is traced on its own
    scene's patches.  The sample grows until the traces take about target_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    O.lib()
    cores = max(1, min(16, os.cpu_count() or 1))
    rh = [np.asarray(a) for a in rays.get()]
    sid = ids.get()
    geo = {}
    n_all = len(sid)
    n = 4096
    while True:
        sel = np.unique((np.arange(n, dtype=np.float64) * (n_all / n)).astype(np.int64))  # spread over all scenes
        dt = 0.0
        for k in np.unique(sid[sel]):
            if k not in geo:
                dl = built[int(k)].download()
                ps = O.PatchSet.__new__(O.PatchSet)
                ps.p = np.ascontiguousarray(dl["pq"][:, 0], np.int32)
                ps.q = np.ascontiguousarray(dl["pq"][:, 1], np.int32)
                ps.off = np.ascontiguousarray(dl["off"], np.int32)
                ps.cp = np.ascontiguousarray(dl["cp"], np.float32)
                geo[k] = (dl["nodes"], ps)
            m = sel[sid[sel] == k]
            sample = [np.ascontiguousarray(a[m]) for a in rh]
            t0 = time.perf_counter()
            O.trace_f32(geo[k][0], geo[k][1], sample, max_iter=8, thr_res=1e-5, thr_det=1e-5, seed_base=0,
                        threads=cores)
            dt += time.perf_counter() - t0
        if dt >= target_s * 0.4 or n >= n_all:
            break
        n = min(n_all, int(n * min(8.0, target_s / max(dt, 1e-3))))
    return {"value": len(sel) / dt / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"{len(sel)} of {n_all} primary rays (evenly spaced over all scenes), plain method FP32 "
                      f"(8 iterations, tol 1e-5), {dt:.1f} s on {cores} threads", "host": host_info()}


def calibrate_copy(nbytes):
    """Device-to-device copies of known size (PMC scaling of HBM counters)."""
    from rtnurbs import lib as L
    L.load()
    a = L.DeviceArray((nbytes // 4,), np.float32)
    b = L.DeviceArray((nbytes // 4,), np.float32)
    a.fill_bytes(0)
    L.synchronize()
    for _ in range(3):
        L.check(L.load().rtn_memcpy(b.ptr, a.ptr, a.nbytes, 2))
    L.synchronize()
    print(json.dumps({"copy_bytes": a.nbytes, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c2", "c1", "c3", "c5"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 error report")
    ap.add_argument("--forms", default="all", help="comma list of forms to time (plain,wide,...)")
    ap.add_argument("--train-epochs", type=int, default=16, help="seed network epochs (0: no learned forms)")
    ap.add_argument("--c5-scenes", type=int, default=32)
    ap.add_argument("--c5-scale", type=int, default=1, help="divide config 5 resolutions by this")
    ap.add_argument("--c5-packet-only", action="store_true",
                    help="config 5: skip the autotuner and time only the all-packet frame (profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    ap.add_argument("--save-net", default=None, help="write the trained seed networks' parameters (.npz)")
    ap.add_argument("--load-net", default=None, help="use saved seed-network parameters instead of training")
    ap.add_argument("--hit-gate", type=float, default=0.7, help="learned forms: hit probability below which a "
                    "pair starts from the plain seed")
    ap.add_argument("--learned-min-lanes", type=int, default=64, help="packet learned forms: lanes a patch visit "
                    "needs for the network to run")
    ap.add_argument("--hit-skip", type=float, default=0.002, help="the skip form: hit probability below which a "
                    "pair is not solved")
    ap.add_argument("--detail", default=None, help="per-form report (JSON); default bench_out/<config>_detail.json")
    args = ap.parse_args()
    if args.detail is None:
        args.detail = os.path.join("bench_out", f"{args.config}_detail.json")
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(across GPUs the scene is replicated and rays shard by image tile or sample)",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    from rtnurbs import learned, scenes
    from rtnurbs import lib as L
    L.load()
    if args.config in ("c2", "c1"):
        out = bench_c2_c1(args, L, scenes, learned)
    elif args.config == "c3":
        out = bench_c3(args, L, scenes)
    else:
        out = bench_c5(args, L, scenes)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
