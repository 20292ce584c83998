"""Training data and training of the learned Newton seed.

Pairs come from the plain method itself: scenes from the config-1 generator
(one bicubic 8x8 NURBS, 25 patches) and from the config-2 generator (groups
of seeded random NURBS of degree 2..5, curvature-refined), each seen from
random cameras, are traced with plain seeds; every converged Newton solve
(ray, patch, u, v) with t > 0 is a training pair.  Generator seeds used here
(>= 1000 for config 1, 20000 + k for config 2) are disjoint from the
evaluation scenes (config 1 seed 0, config 2 seed 0).

The network is then trained on the GPU (rtn_mlp_train: BF16 MFMA forward and
backward kernels, Adam) on the pair features.
"""
from __future__ import annotations

import numpy as np

from . import lib as L
from . import scenes

PAIR_PARAMS = dict(mode=L.MODE_WIDE, max_iter=8, thr_res=1e-5, thr_det=1e-5, hull_cull=2)


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


def collect_pairs(scene_list, max_pairs=10_000_000, spp=1, log=None):
    """Traces each scene with plain seeds and gathers (features, targets) of
    the converged solves into device arrays; returns (feats, targets, n)."""
    feats = L.DeviceArray((max_pairs, L.MLP_NIN), np.float32)
    targets = L.DeviceArray((max_pairs, 2), np.float32)
    n = 0
    prm = L.params(**PAIR_PARAMS)
    for i, (surfs, cam) in enumerate(scene_list):
        if n >= max_pairs:
            break
        sc = L.Scene(surfs).build()
        rays = L.generate_rays(cam, spp=spp, jitter=spp > 1, seed=i + 1)
        cap = max(1, 4 * rays.n)
        pr, pp, puv, cnt = L.trace_pairs(sc, rays, prm, cap)
        k = min(cnt, cap, max_pairs - n)
        if k > 0:
            L.pair_features(sc, rays, pr, pp, k, out=feats, out_offset=n)
            L.check(L.load().rtn_memcpy(targets.ptr + n * 8, puv.ptr, k * 8, 2))
            n += k
        if log:
            log(f"scene {i}: {cnt} pairs (total {n})")
    L.synchronize()
    return feats, targets, n


class _View:
    """A prefix of a DeviceArray (first n rows) without copying."""

    def __init__(self, a, n):
        self.ptr, self.shape, self.dtype = a.ptr, (n,) + tuple(a.shape[1:]), a.dtype
        self._keep = a


def train_seed_network(feats, targets, n, epochs=4, batch=32768, lr=6e-3, lr_final=1e-4, seed=0):
    """Trains a fresh network on the first n pairs; returns (Mlp, epoch losses)."""
    net = L.Mlp(seed=seed)
    losses = net.train(_View(feats, n), _View(targets, n), epochs=epochs, batch=batch, lr=lr, lr_final=lr_final,
                       seed=seed + 1)
    return net, losses


def build_seed_network(max_pairs=10_000_000, epochs=4, log=None, **kw):
    """Training set generation and training in one call."""
    sl = training_scenes(**kw)
    feats, targets, n = collect_pairs(sl, max_pairs=max_pairs, log=log)
    net, losses = train_seed_network(feats, targets, n, epochs=epochs)
    return net, losses, n
