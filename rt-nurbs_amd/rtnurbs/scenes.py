"""Seeded scene generators (inputs of the trace path).

The reference has only two hand-built scenes, the rational torus and sphere of
benchmark.cpp:55-79 (also main.cpp:70-98).  Those are restated here, together
with the seeded random surfaces used by the benchmark configurations.  All
arrays are float32; control points are stored homogeneous (w*x, w*y, w*z, w),
u-major with row stride m+1, as nurbsSurf expects (nurbs.h:17-22, main.cpp:96).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Nurbs:
    """One NURBS surface (ParametricSurf::nurbsSurf, nurbs.h:17-22)."""
    n: int            # index of the last control point along u
    m: int            # index of the last control point along v
    p: int
    q: int
    U: np.ndarray     # float32[n+p+2]
    V: np.ndarray     # float32[m+q+2]
    Pw: np.ndarray    # float32[(n+1)*(m+1), 4] homogeneous

    def check(self):
        assert self.U.dtype == np.float32 and self.V.dtype == np.float32
        assert self.Pw.dtype == np.float32
        assert self.U.shape == (self.n + self.p + 2,)
        assert self.V.shape == (self.m + self.q + 2,)
        assert self.Pw.shape == ((self.n + 1) * (self.m + 1), 4)
        return self


@dataclass
class Camera:
    eye: tuple
    look: tuple
    up: tuple
    fov_deg: float
    W: int
    H: int
    znear: float = 0.01
    zfar: float = 1000.0


@dataclass
class Scene:
    surfaces: list = field(default_factory=list)
    camera: Camera | None = None
    name: str = ""

    def save(self, path):
        """Writes the surfaces and camera to a scene file (rtn_scene_save)."""
        from . import lib
        lib.Scene(self.surfaces).save(path, self.camera)


def _homog(x, y, z, w):
    x = np.asarray(x, np.float32); y = np.asarray(y, np.float32)
    z = np.asarray(z, np.float32); w = np.asarray(w, np.float32)
    return np.stack([w * x, w * y, w * z, w], axis=-1).astype(np.float32)


_S = np.float32(0.70710678)
_CX = np.array([1, 1, 0, -1, -1, -1, 0, 1, 1], np.float32)
_CY = np.array([0, 1, 1, 1, 0, -1, -1, -1, 0], np.float32)
_CW = np.array([1, _S, 1, _S, 1, _S, 1, _S, 1], np.float32)
_CIRCLE_KNOTS = np.array([0, 0, 0, .25, .25, .5, .5, .75, .75, 1, 1, 1], np.float32)


def torus(R=1.0, r=0.4, cz=6.0) -> Nurbs:
    """Rational biquadratic torus of benchmark.cpp:55-64 (9x9 net, p=q=2)."""
    R = np.float32(R); r = np.float32(r); cz = np.float32(cz)
    Pw = np.zeros((81, 4), np.float32)
    for i in range(9):
        for j in range(9):
            radial = R + r * _CX[j]
            zz = r * _CY[j]
            w = _CW[i] * _CW[j]
            Pw[i * 9 + j] = [w * radial * _CX[i], w * radial * _CY[i], w * (zz + cz), w]
    return Nurbs(8, 8, 2, 2, _CIRCLE_KNOTS.copy(), _CIRCLE_KNOTS.copy(), Pw).check()


def sphere(Rs=1.0, cz=6.0) -> Nurbs:
    """Rational sphere of benchmark.cpp:66-79 (9x5 net, p=q=2)."""
    Rs = np.float32(Rs); cz = np.float32(cz)
    svr = np.array([0, Rs, Rs, Rs, 0], np.float32)
    svz = np.array([Rs, Rs, 0, -Rs, -Rs], np.float32)
    svw = np.array([1, _S, 1, _S, 1], np.float32)
    Pw = np.zeros((45, 4), np.float32)
    for i in range(9):
        for j in range(5):
            w = _CW[i] * svw[j]
            Pw[i * 5 + j] = [w * svr[j] * _CX[i], w * svr[j] * _CY[i], w * (svz[j] + cz), w]
    V = np.array([0, 0, 0, .5, .5, 1, 1, 1], np.float32)
    return Nurbs(8, 4, 2, 2, _CIRCLE_KNOTS.copy(), V, Pw).check()


def reference_camera(W=1920, H=1080) -> Camera:
    """benchmark.cpp:128-131: lookAt (3,3,0)->(0,0,6), up (0,1,0), 60 deg."""
    return Camera((3, 3, 0), (0, 0, 6), (0, 1, 0), 60.0, W, H)


# ----------------------------------------------------------------- config 1
C1_KNOTS = np.array([0, 0, 0, 0, .2, .4, .6, .8, 1, 1, 1, 1], np.float32)


def bicubic_grid_surface(seed: int) -> Nurbs:
    """One bicubic NURBS: 8x8 net on the unit grid (i/7, j/7), z ~ N(0, 0.2)
    (0.2 read as the standard deviation), w ~ U[0.5, 2], open-uniform knots
    {0^4, .2, .4, .6, .8, 1^4}: 25 rational Bezier patches."""
    rng = np.random.default_rng(seed)
    z = rng.normal(0.0, 0.2, size=(8, 8)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, size=(8, 8)).astype(np.float32)
    ii, jj = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
    x = (ii / np.float32(7)).astype(np.float32)
    y = (jj / np.float32(7)).astype(np.float32)
    Pw = _homog(x, y, z, w).reshape(64, 4)
    return Nurbs(7, 7, 3, 3, C1_KNOTS.copy(), C1_KNOTS.copy(), Pw).check()


def c1_camera(W=256, H=256) -> Camera:
    return Camera((0.5, -0.2, 1.2), (0.5, 0.5, 0.0), (0.0, 0.0, 1.0), 50.0, W, H)


def config1(seed: int = 0, W=256, H=256) -> Scene:
    return Scene([bicubic_grid_surface(seed)], c1_camera(W, H), f"c1-seed{seed}")


# ----------------------------------------------------------------- config 2
def _random_knots(rng, n_cp: int, deg: int) -> np.ndarray:
    """Clamped knot vector with sorted random interior values and random
    interior multiplicities in 1..deg (sum fixed by the control-point count)."""
    slots = n_cp - deg - 1          # interior knot entries
    mults = []
    while sum(mults) < slots:
        mults.append(int(rng.integers(1, deg + 1)))
    if mults:
        mults[-1] -= sum(mults) - slots
    k = len(mults)
    while True:
        vals = np.sort(rng.uniform(0.0, 1.0, size=k))
        gaps = np.diff(np.concatenate([[0.0], vals, [1.0]]))
        if k == 0 or gaps.min() > 0.6 / (k + 1):
            break
    interior = np.repeat(vals.astype(np.float32), mults)
    return np.concatenate([np.zeros(deg + 1, np.float32), interior,
                           np.ones(deg + 1, np.float32)]).astype(np.float32)


def _rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                     [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                     [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])


def random_surface(rng, deg_range=(2, 5), cp_range=(6, 16), spacing=0.36, zsig=0.5,
                   box=10.0) -> Nurbs:
    """Seeded random NURBS: per-direction degree uniform in deg_range, net
    sides uniform in cp_range, grid spacing `spacing`, z ~ N(0, zsig),
    w ~ U[0.5, 2], random interior multiplicities, random rotation and a
    translation that keeps the surface inside [0, box]^3."""
    p = int(rng.integers(deg_range[0], deg_range[1] + 1))
    q = int(rng.integers(deg_range[0], deg_range[1] + 1))
    nu = int(rng.integers(max(cp_range[0], p + 1), cp_range[1] + 1))
    nv = int(rng.integers(max(cp_range[0], q + 1), cp_range[1] + 1))
    U = _random_knots(rng, nu, p)
    V = _random_knots(rng, nv, q)
    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    pts = np.stack([(ii - (nu - 1) / 2) * spacing, (jj - (nv - 1) / 2) * spacing,
                    rng.normal(0.0, zsig, size=(nu, nv))], axis=-1)
    pts = pts @ _rotation(rng).T
    ext = np.abs(pts).max(axis=(0, 1))
    lo = ext + 0.05
    hi = box - ext - 0.05
    t = rng.uniform(np.minimum(lo, hi), np.maximum(lo, hi))
    pts = (pts + t).astype(np.float32)
    w = rng.uniform(0.5, 2.0, size=(nu, nv)).astype(np.float32)
    Pw = _homog(pts[..., 0], pts[..., 1], pts[..., 2], w).reshape(-1, 4)
    return Nurbs(nu - 1, nv - 1, p, q, U, V, Pw).check()


def c2_camera(W=1920, H=1080) -> Camera:
    return Camera((5.0, 5.0, -6.0), (5.0, 5.0, 5.0), (0.0, 1.0, 0.0), 60.0, W, H)


def config2_base(seed: int = 0, n_surf: int = 64) -> list:
    rng = np.random.default_rng(10_000 + seed)
    return [random_surface(rng) for _ in range(n_surf)]


def patch_count(s: Nurbs) -> int:
    """Closed form for the decomposition's patch count: (#distinct interior
    knots + 1) per direction (segs, bezier_patch.h:46-53)."""
    def nseg(K, n, p):
        inner = K[p + 1:n + 1]
        return len(np.unique(inner)) + 1 if len(inner) else 1
    return nseg(s.U, s.n, s.p) * nseg(s.V, s.m, s.q)


C2_REFINE = 0.264   # curvature-refinement constant freezing config 2 at 10,018 patches
C2_PATCHES = 10018


def config2(seed: int = 0, W=1920, H=1080, n_surf: int = 64) -> Scene:
    """64 seeded random NURBS (degree 2..5 per direction, nets 6-16 per side,
    random interior multiplicities, random rigid placement in the 10^3 box),
    refined by the reference's curvatureRefinementU/V (nurbs.h:290-486) at
    C = 0.264 so the scene decomposes into 10,018 rational Bezier patches."""
    from . import lib
    base = config2_base(seed, n_surf)
    surfs = [lib.refine(s, C2_REFINE, C2_REFINE) for s in base]
    return Scene(surfs, c2_camera(W, H), f"c2-seed{seed}")


# ----------------------------------------------------------------- config 3
C3_SILHOUETTE_MOD = 5   # pixels p with p % 5 == 4 (13,107 of 65,536) become silhouette rays


def config3_rays(scene, seed=0, W=256, H=256, max_deg=5.0):
    """Config 1's 256x256 primary rays with 20% of them (pixels p = 4 mod 5,
    13,107 rays) replaced by silhouette rays: grazing angle U(0, 5) degrees
    at random surface points, origin 1..3 units back along the ray.
    Returns (Rays, boolean mask of the silhouette rays)."""
    from . import lib
    cam = c1_camera(W, H)
    prim = lib.generate_rays(cam)
    n = prim.n
    mask = (np.arange(n) % C3_SILHOUETTE_MOD) == C3_SILHOUETTE_MOD - 1
    k = int(mask.sum())
    sil = lib.silhouette_rays(scene, k, seed=seed, max_deg=max_deg)
    host = [a.copy() for a in prim.get()]
    sh = sil.get()
    for a, b in zip(host, sh):
        a[mask] = b
    return lib.Rays.from_host(host), mask


# ----------------------------------------------------------------- config 5
C5_RES = (256, 512, 1024, 2048)


def grid_surface(rng, target_patches, p, q) -> Nurbs:
    """Config-1-style NURBS sized for about `target_patches` Bezier patches:
    net on the unit grid, open-uniform knots with distinct interior values
    (patches = segments_u x segments_v), w ~ U[0.5, 2], z ~ N(0, sigma) with
    sigma = 0.2 * 7 / (net side - 1), i.e. config 1's roughness per net cell."""
    su = max(1, int(round(np.sqrt(target_patches))))
    sv = max(1, int(round(target_patches / su)))
    nu, nv = su + p, sv + q

    def knots(seg, deg):
        inner = (np.arange(1, seg) / seg).astype(np.float32)
        return np.concatenate([np.zeros(deg + 1, np.float32), inner, np.ones(deg + 1, np.float32)])
    sigma = 0.2 * 7.0 / (max(nu, nv) - 1)
    z = rng.normal(0.0, sigma, size=(nu, nv)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, size=(nu, nv)).astype(np.float32)
    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    Pw = _homog((ii / np.float32(nu - 1)).astype(np.float32), (jj / np.float32(nv - 1)).astype(np.float32), z, w)
    return Nurbs(nu - 1, nv - 1, p, q, knots(su, p), knots(sv, q), Pw.reshape(-1, 4)).check()


def config5(seed: int = 0, n_scenes: int = 32, patch_range=(1000, 200000), res=C5_RES, scale=1):
    """32 independent scenes of uneven size: patch counts log-uniform in
    [1k, 200k], degrees uniform in 2..5 per direction, one config-1-style
    surface each, seen by config 1's camera at a resolution drawn from
    256^2..2048^2 (divided by `scale`).  Returns a list of Scene."""
    rng = np.random.default_rng(50_000 + seed)
    out = []
    for k in range(n_scenes):
        target = int(np.exp(rng.uniform(np.log(patch_range[0]), np.log(patch_range[1]))))
        p, q = int(rng.integers(2, 6)), int(rng.integers(2, 6))
        r = int(res[int(rng.integers(0, len(res)))]) // scale
        out.append(Scene([grid_surface(rng, target, p, q)], c1_camera(r, r), f"c5-seed{seed}-{k}"))
    return out
