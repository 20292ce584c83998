"""Edge cases of the trace entry point, through the C ABI.

- no rays, in every form;
- ray counts that leave the last wave partly empty (1, 63, 65, 4097);
- rays that miss everything, axis-parallel rays (two zero direction
  components, so the slab test divides by zero), rays starting inside the
  scene between the patches;
- a single-patch scene (the BVH root is a leaf, the 4-wide tree one node);
- degree-1 and degree-5 patches.

The exact form is compared with the oracle bit for bit (hit patch, t, u, v,
dpdu, dpdv); the fast forms with one another bit for bit (persistent = wide,
hull-rejected = unrejected, packet = wide but for equal-t ties) and with the
exact form on the hit patches."""
import numpy as np
import pytest

from helpers import c2_small
from rtnurbs import scenes
from test_gpu_exact import _compare_trace, _oracle_scene
from test_gpu_wide import same_but_ties

pytestmark = pytest.mark.gpu

CFG = dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)


def _forms(rtl):
    W, P = rtl.MODE_WIDE, rtl.MODE_PERSISTENT
    return {"wide": dict(mode=W), "persistent": dict(mode=P), "wide_hull2": dict(mode=W, hull_cull=2),
            "persistent_hull2_stall": dict(mode=P, hull_cull=2, stall_exit=1),
            "wide_hull2_stall": dict(mode=W, hull_cull=2, stall_exit=1), "packet": dict(mode=rtl.MODE_PACKET),
            "packet_hull2_stall": dict(mode=rtl.MODE_PACKET, hull_cull=2, stall_exit=1)}


def _check_forms(rtl, scene, rays, ex):
    """Fast forms: persistent = wide and hull-rejected = unrejected bit for bit;
    the stall exit only with its own persistent twin; all agree with exact."""
    out = {}
    for name, kw in _forms(rtl).items():
        h, tot = rtl.trace(scene, rays, rtl.params(count=1, **CFG, **kw))
        assert tot["overflow"] == 0, name
        out[name] = h.get()
    wd = out["wide"]
    m = wd["hit"] >= 0
    for a, b in (("persistent", "wide"), ("wide_hull2", "wide"), ("persistent_hull2_stall", "wide_hull2_stall")):
        x, y = out[a], out[b]
        assert np.array_equal(x["hit"], y["hit"]), (a, b)
        k = y["hit"] >= 0
        for f in ("t", "u", "v"):
            assert np.array_equal(x[f][k].view(np.uint32), y[f][k].view(np.uint32)), (a, b, f)
    same_but_ties(out["packet"], wd)
    same_but_ties(out["packet_hull2_stall"], out["wide_hull2_stall"])
    if len(ex["hit"]):
        assert np.mean(ex["hit"] == wd["hit"]) > 0.99
        same = m & (ex["hit"] == wd["hit"])
        if same.any():
            rel = np.abs(wd["t"][same] - ex["t"][same]) / np.abs(ex["t"][same])
            assert rel.max() < 1e-3
    return out


def _exact_vs_oracle(oracle, rtl, surfaces, rays, what):
    _, _, _, nodes, _, ps = _oracle_scene(oracle, surfaces)
    scene = rtl.Scene(surfaces).build()
    h, tot = rtl.trace(scene, rays, rtl.params(mode=rtl.MODE_EXACT, count=1, **CFG))
    g = h.get()
    assert tot["overflow"] == 0
    r = oracle.trace_f32(nodes, ps, rays.get(), **CFG)
    _compare_trace(g, r, what)
    return scene, g


def test_no_rays(rtl):
    print("case zero rays, every form", flush=True)
    sc = scenes.config1(0, 64, 64)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.Rays(0)
    for name, kw in {"exact": dict(mode=rtl.MODE_EXACT), **_forms(rtl)}.items():
        h, tot = rtl.trace(scene, rays, rtl.params(count=1, **CFG, **kw))
        assert all(v == 0 for v in tot.values()), name
        assert h.get()["hit"].shape == (0,)


@pytest.mark.parametrize("n", [1, 63, 65, 4097])
def test_partial_waves(oracle, rtl, n):
    print(f"case config 1, {n} rays drawn from the camera rays", flush=True)
    sc = scenes.config1(0)
    full = rtl.generate_rays(sc.camera).get()
    sel = np.random.default_rng(n).permutation(len(full[0]))[:n]
    rays = rtl.Rays.from_host([a[sel] for a in full])
    scene, ex = _exact_vs_oracle(oracle, rtl, sc.surfaces, rays, f"{n} rays")
    _check_forms(rtl, scene, rays, ex)


def _special_rays(rng, k):
    """Unit rays in four groups of k: +z axis-parallel from z = -5 over the
    [0,10]^2 footprint, +x axis-parallel from x = -5, away from the scene
    (-z from z = -6), and random directions from random points inside the
    [0,10]^3 box."""
    g = []
    xy = rng.uniform(0.0, 10.0, size=(k, 2))
    g.append((np.c_[xy, np.full(k, -5.0)], np.tile([0.0, 0.0, 1.0], (k, 1))))
    yz = rng.uniform(0.0, 10.0, size=(k, 2))
    g.append((np.c_[np.full(k, -5.0), yz], np.tile([1.0, 0.0, 0.0], (k, 1))))
    g.append((np.c_[rng.uniform(0, 10, size=(k, 2)), np.full(k, -6.0)], np.tile([0.0, 0.0, -1.0], (k, 1))))
    d = rng.normal(size=(k, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    g.append((rng.uniform(0.0, 10.0, size=(k, 3)), d))
    o = np.concatenate([a for a, _ in g]).astype(np.float32)
    d = np.concatenate([b for _, b in g]).astype(np.float32)
    return [o[:, 0], o[:, 1], o[:, 2], d[:, 0], d[:, 1], d[:, 2]]


def test_special_rays(oracle, rtl):
    print("case axis-parallel, outward and inside rays on 8 mixed-degree surfaces", flush=True)
    surfs = c2_small(8, 2)
    k = 1024
    rays = rtl.Rays.from_host(_special_rays(np.random.default_rng(5), k))
    scene, ex = _exact_vs_oracle(oracle, rtl, surfs, rays, "special rays")
    assert (ex["hit"][:k] >= 0).any() and (ex["hit"][k:2 * k] >= 0).any() and (ex["hit"][3 * k:] >= 0).any()
    out = _check_forms(rtl, scene, rays, ex)
    for name, g in [("exact", ex)] + list(out.items()):
        assert (g["hit"][2 * k:3 * k] < 0).all(), name  # rays leaving the scene hit nothing


def _bezier(rng, p, q, at):
    """One rational Bezier surface (clamped knots, no interior knots) near `at`."""
    ii, jj = np.meshgrid(np.arange(p + 1), np.arange(q + 1), indexing="ij")
    pts = np.stack([ii * 2.0 / p, jj * 2.0 / q, rng.normal(0, 0.3, size=ii.shape)], -1) + np.asarray(at)
    w = rng.uniform(0.5, 2.0, size=ii.shape)
    Pw = np.concatenate([pts * w[..., None], w[..., None]], -1).reshape(-1, 4).astype(np.float32)
    U = np.r_[np.zeros(p + 1), np.ones(p + 1)].astype(np.float32)
    V = np.r_[np.zeros(q + 1), np.ones(q + 1)].astype(np.float32)
    return scenes.Nurbs(p, q, p, q, U, V, Pw).check()


@pytest.mark.parametrize("p,q", [(1, 1), (5, 5), (3, 1)])
def test_single_patch_scene(oracle, rtl, p, q):
    print(f"case one Bezier patch of degree ({p},{q})", flush=True)
    s = _bezier(np.random.default_rng(p * 7 + q), p, q, (4.0, 4.0, 5.0))
    rays = rtl.generate_rays(scenes.c2_camera(160, 90))
    scene, ex = _exact_vs_oracle(oracle, rtl, [s], rays, f"single patch ({p},{q})")
    assert scene.stats()["n_patches"] == 1 and (ex["hit"] >= 0).sum() > 100
    _check_forms(rtl, scene, rays, ex)


@pytest.mark.parametrize("deg", [1, 5])
def test_uniform_degree_scenes(oracle, rtl, deg):
    print(f"case six random degree-{deg} surfaces", flush=True)
    rng = np.random.default_rng(40 + deg)
    surfs = [scenes.random_surface(rng, deg_range=(deg, deg)) for _ in range(6)]
    rays = rtl.generate_rays(scenes.c2_camera(240, 135))
    scene, ex = _exact_vs_oracle(oracle, rtl, surfs, rays, f"degree {deg}")
    assert (ex["hit"] >= 0).sum() > 500
    _check_forms(rtl, scene, rays, ex)


def test_closest_root_entry_points_check_their_grid(rtl):
    """rtn_trace_f64_roots and rtn_solve_pairs_roots reject a grid outside
    0..16 with RTN_EINVAL before any launch; grid 0 equals the single-start
    entry points bit for bit; zero pairs and zero rays are no-ops."""
    print("case closest-root entry points: grid checks, grid 0, empty inputs", flush=True)
    sc = scenes.config1(0, 64, 64)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera)
    for g in (-1, 17):
        with pytest.raises(rtl.RtnError) as e:
            rtl.trace_f64(scene, rays, grid=g)
        assert e.value.code == rtl.RTN_EINVAL
    a = rtl.trace_f64(scene, rays).get()
    b = rtl.trace_f64(scene, rays, grid=0).get()
    for k in ("hit", "t", "u", "v", "dpdu", "dpdv", "t2"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k
    assert rtl.trace_f64(scene, rtl.Rays(0), grid=4).get()["hit"].shape == (0,)
    prm = rtl.params(mode=rtl.MODE_WIDE, max_iter=8, thr_res=1e-5, thr_det=1e-5)
    pr, pp, puv, cnt = rtl.trace_pairs(scene, rays, prm, 4 * rays.n, all=True)
    assert cnt > 0
    sp = rtl.params(max_iter=8, thr_res=1e-5, thr_det=1e-5)
    with pytest.raises(rtl.RtnError) as e:
        rtl.solve_pairs(scene, rays, pr, pp, cnt, sp, grid=17)
    assert e.value.code == rtl.RTN_EINVAL
    x = rtl.solve_pairs(scene, rays, pr, pp, cnt, sp)
    y = rtl.solve_pairs(scene, rays, pr, pp, cnt, sp, grid=0)
    for k in ("iters", "converged", "uv", "t"):
        assert np.array_equal(x[k], y[k]), k
    z = rtl.solve_pairs(scene, rays, pr, pp, 0, sp, grid=4)
    assert z["iters"].shape == (0,)
