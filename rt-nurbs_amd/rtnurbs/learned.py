"""Training data and training of the learned Newton seed.

Pairs come from the plain method itself: scenes from the config-1 generator
(one bicubic 8x8 NURBS, 25 patches) and, optionally, from the config-2
generator (groups of seeded random NURBS of degree 2..5, curvature-refined),
each seen from random cameras, are traced with the plain seed in the
mixed-precision form without the network (FP64 seed, stall exit); every
Newton solve the trace runs is a training pair: a converged solve with t > 0
gives the target (u, v, 1), any other solve (u, v, 0) (a miss: only the hit
output learns from it).  Generator seeds used here (>= 1000 for config 1,
20000 + k for config 2) are disjoint from the evaluation scenes (config 1
seeds 0 and 2000+, config 2 seed 0).

The network is then trained on the GPU (rtn_mlp_train: BF16 MFMA forward and
backward kernels, Adam) on the pair features of its feature kind.
"""
from __future__ import annotations

import numpy as np

from . import lib as L
from . import scenes

PAIR_PARAMS = dict(mode=L.MODE_WIDE, max_iter=8, thr_res=1e-5, thr_det=1e-5, hull_cull=2, precision=L.PREC_MIXED,
                   cond_thr=1e4, stall_exit=1)
# the closest-root relabelling's solves (plain Newton, each start)
ROOT_PARAMS = dict(max_iter=16, thr_res=1e-5, thr_det=1e-5)


def _orbit_camera(rng, center, radius, W, H):
    """Camera on a random direction around `center` (upper hemisphere for
    height fields), looking at a point near the centre."""
    while True:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if d[2] > 0.25:
            break
    eye = np.asarray(center) + d * radius * rng.uniform(1.0, 2.2)
    look = np.asarray(center) + rng.normal(0.0, 0.1 * radius, size=3)
    up = (0.0, 0.0, 1.0) if abs(d[2]) < 0.95 else (0.0, 1.0, 0.0)
    return scenes.Camera(tuple(eye), tuple(look), up, float(rng.uniform(35.0, 60.0)), W, H)


def training_scenes(n_c1=48, n_c2=24, res=512, seed=0):
    """(surfaces, camera) pairs of the training set."""
    rng = np.random.default_rng(1234 + seed)
    out = []
    for k in range(n_c1):
        s = scenes.bicubic_grid_surface(1000 + k)
        out.append(([s], _orbit_camera(rng, (0.5, 0.5, 0.0), 1.0, res, res)))
    for k in range(n_c2):
        grng = np.random.default_rng(20000 + k + 100 * seed)
        base = [scenes.random_surface(grng, box=4.0) for _ in range(6)]
        surfs = [L.refine(s, scenes.C2_REFINE, scenes.C2_REFINE) for s in base]
        # height-field orientation is random here, so look from any direction
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        eye = np.array([2.0, 2.0, 2.0]) + d * rng.uniform(4.0, 7.0)
        cam = scenes.Camera(tuple(eye), (2.0, 2.0, 2.0), (0.0, 0.0, 1.0) if abs(d[2]) < 0.95 else (0.0, 1.0, 0.0),
                            float(rng.uniform(35.0, 60.0)), res, res)
        out.append((surfs, cam))
    return out


def collect_pairs(scene_list, max_hits=10_000_000, spp=1, kind=L.MLP_FEAT_PROJECTED, log=None, root_grid=0):
    """Traces each scene with the plain seed and gathers (features, targets)
    of every Newton solve into device arrays until max_hits converged solves
    are in; returns (feats (n, 20), targets (n, 3) = (u, v, hit), n, hits).
    root_grid > 0 relabels every pair with its closest root: the pair is
    solved again from the plain seed and from root_grid x root_grid starts
    (16 iterations each) and the nearest converged root with t > 0 becomes the
    target (hit = 1 when any start converges)."""
    cap_all = 3 * max_hits
    feats = L.DeviceArray((cap_all, L.MLP_NIN), np.float32)
    targets = L.DeviceArray((cap_all, 3), np.float32)
    n = hits = 0
    prm = L.params(**PAIR_PARAMS)
    for i, (surfs, cam) in enumerate(scene_list):
        if hits >= max_hits or n >= cap_all:
            break
        sc = L.Scene(surfs).build()
        rays = L.generate_rays(cam, spp=spp, jitter=spp > 1, seed=i + 1)
        cap = max(1, 8 * rays.n)
        pr, pp, puv, cnt = L.trace_pairs(sc, rays, prm, cap, all=True)
        k = min(cnt, cap, cap_all - n)
        if k > 0:
            uv = puv.get()[:k]
            h = uv[:, 0] >= 0.0
            if root_grid > 0:
                sol = L.solve_pairs(sc, rays, pr, pp, k, L.params(**{**ROOT_PARAMS}), grid=root_grid)
                uv = np.where(sol["converged"][:, None], sol["uv"], -1.0).astype(np.float32)
                h = sol["converged"]
            # stop at max_hits converged solves: keep the pairs up to that hit
            if hits + int(h.sum()) > max_hits:
                k = int(np.searchsorted(np.cumsum(h), max_hits - hits, side="left")) + 1
                uv, h = uv[:k], h[:k]
            t = np.concatenate([np.where(h[:, None], uv, 0.0), h[:, None]], 1).astype(np.float32)
            L.pair_features(sc, rays, pr, pp, k, out=feats, out_offset=n, kind=kind)
            L.check(L.load().rtn_memcpy(targets.ptr + n * 12, t.ctypes.data, k * 12, 0))
            n += k
            hits += int(h.sum())
        if log:
            log(f"scene {i}: {cnt} solves (total {n}, {hits} hits)")
    L.synchronize()
    return feats, targets, n, hits


class _View:
    """A prefix of a DeviceArray (first n rows) without copying."""

    def __init__(self, a, n):
        self.ptr, self.shape, self.dtype = a.ptr, (n,) + tuple(a.shape[1:]), a.dtype
        self._keep = a


def train_seed_network(feats, targets, n, epochs=4, batch=32768, lr=6e-3, lr_final=1e-4, seed=0,
                       kind=L.MLP_FEAT_PROJECTED, hit_weight=0.05):
    """Trains a fresh network on the first n pairs; returns (Mlp, epoch losses)."""
    net = L.Mlp(seed=seed, kind=kind)
    losses = net.train(_View(feats, n), _View(targets, n), epochs=epochs, batch=batch, lr=lr, lr_final=lr_final,
                       seed=seed + 1, hit_weight=hit_weight)
    return net, losses


def build_seed_network(max_hits=10_000_000, epochs=4, kind=L.MLP_FEAT_PROJECTED, log=None, **kw):
    """Training set generation and training in one call; returns
    (Mlp, epoch losses, pairs, converged pairs)."""
    sl = training_scenes(**kw)
    feats, targets, n, hits = collect_pairs(sl, max_hits=max_hits, kind=kind, log=log)
    net, losses = train_seed_network(feats, targets, n, epochs=epochs, kind=kind)
    return net, losses, n, hits
