"""Ray producers and shading on the GPU: silhouette rays (geometry against an
FP64 Bernstein evaluation), secondary rays (shadow towards the light, mirror
by the reflection law), any-hit occlusion against closest-hit, Lambert
shading against numpy, and the mixed-precision form on config 3 (20%
silhouette rays) against the FP64 reference trace."""
import math

import numpy as np
import pytest

from rtnurbs import scenes

pytestmark = pytest.mark.gpu

CFG = dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)


def _bern(n, t):
    return np.array([math.comb(n, i) * t ** i * (1 - t) ** (n - i) for i in range(n + 1)])


def _dbern(n, t):
    if n == 0:
        return np.zeros(1)
    b = _bern(n - 1, t)
    return n * (np.concatenate([[0.0], b]) - np.concatenate([b, [0.0]]))


def _eval64(net, p, q, u, v):
    """FP64 point and partials of a rational Bezier patch (closed-form Bernstein)."""
    P = net.astype(np.float64).reshape(p + 1, q + 1, 4)
    bu, bv, dbu, dbv = _bern(p, u), _bern(q, v), _dbern(p, u), _dbern(q, v)
    pt = np.einsum("i,j,ijk->k", bu, bv, P)
    pu = np.einsum("i,j,ijk->k", dbu, bv, P)
    pv = np.einsum("i,j,ijk->k", bu, dbv, P)
    S = pt[:3] / pt[3]
    return S, (pu[:3] - S * pu[3]) / pt[3], (pv[:3] - S * pv[3]) / pt[3]


def test_silhouette_rays_geometry(rtl):
    sc = scenes.config1(0)
    scene = rtl.Scene(sc.surfaces).build()
    print("silhouette rays n=2000", flush=True)
    rays, ap, au = rtl.silhouette_rays(scene, 2000, seed=3, aim=True)
    o = np.stack(rays.get()[:3], 1).astype(np.float64)
    d = np.stack(rays.get()[3:], 1).astype(np.float64)
    ap, au = ap.get(), au.get()
    dl = scene.download()
    assert np.all(np.abs(np.linalg.norm(d, axis=1) - 1) < 1e-6)
    assert np.all((au[:, 2] >= 0) & (au[:, 2] <= np.radians(5.0) + 1e-6))
    assert np.all((au[:, 3] >= 1.0) & (au[:, 3] <= 3.0))
    for i in range(0, 2000, 7):
        k = ap[i]
        p, q = dl["pq"][k]
        net = dl["cp"][dl["off"][k]:dl["off"][k] + (p + 1) * (q + 1)]
        S, Su, Sv = _eval64(net, p, q, au[i, 0], au[i, 1])
        N = np.cross(Su, Sv)
        N /= np.linalg.norm(N)
        assert np.linalg.norm(o[i] + au[i, 3] * d[i] - S) < 2e-5 * (1 + np.abs(S).max())
        assert abs(math.asin(min(1.0, abs(d[i] @ N))) - au[i, 2]) < 2e-3


def test_secondaries_any_hit_and_shading(rtl):
    sc = scenes.config1(0, 128, 128)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera)
    h, _ = rtl.trace(scene, rays, rtl.params(mode=rtl.MODE_WIDE, hull_cull=2, **CFG))
    g = h.get()
    light = np.array([0.3, -0.4, 1.0])
    L = light / np.linalg.norm(light)
    print("secondaries", flush=True)
    sec, parent, kind, cnt = rtl.secondary_rays(rays, h, light, eps=1e-4)
    nh = int((g["hit"] >= 0).sum())
    assert cnt == 2 * nh
    par, knd = parent.get(), kind.get()
    assert np.array_equal(par, np.tile(np.arange(rays.n), 2))  # fixed kind-major slots r, n + r
    assert np.array_equal(knd >= 0, np.tile(g["hit"] >= 0, 2))
    assert np.all(knd[:rays.n][g["hit"] >= 0] == 0) and np.all(knd[rays.n:][g["hit"] >= 0] == 1)
    valid = knd >= 0
    par, knd = par[valid], knd[valid]
    assert np.array_equal(par[knd == 0], np.flatnonzero(g["hit"] >= 0))
    so = np.stack(sec.get()[:3], 1)[valid].astype(np.float64)
    sd = np.stack(sec.get()[3:], 1)[valid].astype(np.float64)
    ro = np.stack(rays.get()[:3], 1).astype(np.float64)
    rd = np.stack(rays.get()[3:], 1).astype(np.float64)
    S = ro[par] + g["t"][par, None] * rd[par]
    N = np.cross(g["dpdu"][par], g["dpdv"][par]).astype(np.float64)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    N *= -np.sign((N * rd[par]).sum(1))[:, None]
    assert np.abs(so - S).max() < 1e-3
    assert np.all(((so - S) * N).sum(1) > 0)  # offset to the incoming side
    np.testing.assert_allclose(sd[knd == 0], np.tile(L, ((knd == 0).sum(), 1)), atol=1e-6)
    m = knd == 1
    refl = rd[par[m]] - 2 * (rd[par[m]] * N[m]).sum(1)[:, None] * N[m]
    np.testing.assert_allclose(sd[m], refl, atol=1e-5)
    # any-hit occlusion equals closest-hit hit/miss on the same rays
    shadow = rtl.Rays.from_host([a[valid][knd == 0] for a in sec.get()])
    a, ta = rtl.trace(scene, shadow, rtl.params(mode=rtl.MODE_WIDE, hull_cull=2, any_hit=1, count=1, **CFG))
    c, tc = rtl.trace(scene, shadow, rtl.params(mode=rtl.MODE_WIDE, hull_cull=2, count=1, **CFG))
    assert np.array_equal(a.get()["hit"] >= 0, c.get()["hit"] >= 0)
    assert ta["newton_calls"] <= tc["newton_calls"]
    # the same in the packet form (a lane leaves the packet at its first hit)
    ap, tap = rtl.trace(scene, shadow, rtl.params(mode=rtl.MODE_PACKET, hull_cull=2, any_hit=1, count=1, **CFG))
    assert np.array_equal(ap.get()["hit"] >= 0, c.get()["hit"] >= 0)
    assert tap["newton_calls"] <= tc["newton_calls"] * 1.05
    # occlusion queries take no FP64 step under mixed precision (the step cannot
    # change hit or miss); the same hit/miss as the closest-hit trace of the same
    # precision (mixed precision starts from the FP64 seed, so its FP32 Newton
    # follows other trajectories than the plain form's)
    mixed = dict(hull_cull=2, count=1, precision=rtl.PREC_MIXED, cond_thr=1e4, **CFG)
    cm, _ = rtl.trace(scene, shadow, rtl.params(mode=rtl.MODE_WIDE, **mixed))
    for m in (rtl.MODE_WIDE, rtl.MODE_PACKET):
        am, tam = rtl.trace(scene, shadow, rtl.params(mode=m, any_hit=1, **mixed))
        assert tam["fp64_steps"] == 0
        assert np.array_equal(am.get()["hit"] >= 0, cm.get()["hit"] >= 0)
    # Lambert shading (no shadows) against numpy
    img = rtl.shade_lambert(rays, h, 128, 128, 1, light, albedo=0.8).get().ravel()
    hit = g["hit"] >= 0
    Np = np.cross(np.where(hit[:, None], g["dpdu"], 1.0), np.where(hit[:, None], g["dpdv"], 0.5)).astype(np.float64)
    Np /= np.linalg.norm(Np, axis=1, keepdims=True) + 1e-300
    Np *= np.where((Np * rd).sum(1) > 0, -1.0, 1.0)[:, None]
    ref = np.where(hit, 0.8 * np.maximum(0, Np @ L), 0.0)
    np.testing.assert_allclose(img, ref, atol=1e-5)


def test_mixed_precision_on_silhouettes(rtl):
    """Config 3: on the silhouette rays (grazing 0-5 degrees), mixed precision
    at the specified threshold (cond(J) > 1e4, plus the predicted-t-error
    rule) takes FP64 steps and brings hit t within 1e-5 of the FP64 reference
    on at least 99.5% of same-patch pixels (FP32: about 60%), and lowers the
    error of every ray class it corrects."""
    sc = scenes.config1(0)
    scene = rtl.Scene(sc.surfaces).build()
    rays, mask = scenes.config3_rays(scene, seed=0)
    assert rays.n == 65536 and mask.sum() == 13107
    print("config 3 mixed precision", flush=True)
    ref = rtl.trace_f64(scene, rays).get()
    base = dict(mode=rtl.MODE_WIDE, hull_cull=2, **CFG)
    a, _ = rtl.trace(scene, rays, rtl.params(**base))
    b, tb = rtl.trace(scene, rays, rtl.params(precision=rtl.PREC_MIXED, cond_thr=1e4, count=1, **base))
    a, b = a.get(), b.get()
    assert tb["fp64_steps"] > 0

    def t_frac(g, sel):
        same = sel & (g["hit"] >= 0) & (g["hit"] == ref["hit"])
        e = np.abs(g["t"][same].astype(np.float64) - ref["t"][same]) / ref["t"][same]
        return np.mean(e <= 1e-5), e
    fa, _ = t_frac(a, mask)
    fb, _ = t_frac(b, mask)
    print(f"silhouette t within 1e-5: FP32 {fa:.4f}, mixed {fb:.4f}, FP64 steps {tb['fp64_steps']}", flush=True)
    assert fa < 0.8 and fb >= 0.995
    assert t_frac(b, ~mask)[0] >= t_frac(a, ~mask)[0]


def _tiled_source(W, H, spp):
    """Row-major index of each ray of the tiled layout (8x8 tiles, Morton inside)."""
    r = np.arange(W * H * spp)
    s, p = r % spp, r // spp
    tile, w = p // 64, p % 64
    mx = (w & 1) | ((w >> 1) & 2) | ((w >> 2) & 4)
    my = ((w >> 1) & 1) | ((w >> 2) & 2) | ((w >> 3) & 4)
    x = (tile % (W // 8)) * 8 + mx
    y = (tile // (W // 8)) * 8 + my
    return (y * W + x) * spp + s


def test_tiled_ray_layout(rtl):
    """The tiled layout holds the same rays as the reference's row-major one,
    permuted; traced and shaded it gives the same image."""
    sc = scenes.config1(0, 64, 48)
    src = _tiled_source(64, 48, 2)
    assert np.array_equal(np.sort(src), np.arange(64 * 48 * 2))
    a = rtl.generate_rays(sc.camera, spp=2)
    b = rtl.generate_rays(sc.camera, spp=2, tiled=True)
    for ca, cb in zip(a.get(), b.get()):
        assert np.array_equal(cb.view(np.uint32), ca[src].view(np.uint32))
    scene = rtl.Scene(sc.surfaces).build()
    prm = rtl.params(mode=rtl.MODE_WIDE, hull_cull=2, **CFG)
    ha, _ = rtl.trace(scene, a, prm)
    hb, _ = rtl.trace(scene, b, prm)
    light = (0.3, -0.4, 1.0)
    ia = rtl.shade_lambert(a, ha, 64, 48, 2, light).get()
    ib = rtl.shade_lambert(b, hb, 64, 48, 2, light, tiled=True).get()
    # Newton's RNG follows the ray index, so a perturbed solve may differ; shading agrees elsewhere
    assert np.mean(np.abs(ia - ib) < 1e-5) > 0.999
    with pytest.raises(rtl.RtnError):
        rtl.generate_rays(scenes.c1_camera(60, 48), tiled=True)
