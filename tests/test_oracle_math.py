"""The CPU oracle against closed forms and invariants (independent numpy FP64
formulas), covering what the reference's fixtures leave open: degree 5, mixed
degrees, exhaustive-SAH optimality, traversal = brute force, analytic
ray-sphere and ray-torus hits on the reference's exact conic surfaces."""
import numpy as np
import pytest

from helpers import c2_small
from rtnurbs import scenes


# ------------------------------------------------------------------ helpers
def bernstein(n, t):
    from math import comb
    return np.array([comb(n, i) * t ** i * (1 - t) ** (n - i) for i in range(n + 1)])


def dbernstein(n, t):
    return n * np.concatenate([[0.0], bernstein(n - 1, t)]) - n * np.concatenate([bernstein(n - 1, t), [0.0]])


def bezier_np(net, p, q, u, v):
    """Homogeneous point and partials of a tensor Bezier patch, FP64 closed form."""
    P = net.reshape(p + 1, q + 1, 4).astype(np.float64)
    Bu, Bv, dBu, dBv = bernstein(p, u), bernstein(q, v), dbernstein(p, u), dbernstein(q, v)
    return (np.einsum("i,j,ijk->k", Bu, Bv, P), np.einsum("i,j,ijk->k", dBu, Bv, P),
            np.einsum("i,j,ijk->k", Bu, dBv, P))


def coxdeboor(K, p, i, x):
    """B-spline basis N_{i,p}(x) by the Cox-de Boor recursion (FP64), with the
    last span closed at the right end."""
    if p == 0:
        if K[i] <= x < K[i + 1]:
            return 1.0
        if x == K[-1] and K[i] < K[i + 1] and K[i + 1] == K[-1]:
            return 1.0
        return 0.0
    a = 0.0 if K[i + p] == K[i] else (x - K[i]) / (K[i + p] - K[i]) * coxdeboor(K, p - 1, i, x)
    b = 0.0 if K[i + p + 1] == K[i + 1] else (K[i + p + 1] - x) / (K[i + p + 1] - K[i + 1]) * coxdeboor(
        K, p - 1, i + 1, x)
    return a + b


def nurbs_point(s, u, v):
    U, V = s.U.astype(np.float64), s.V.astype(np.float64)
    Nu = np.array([coxdeboor(U, s.p, i, u) for i in range(s.n + 1)])
    Nv = np.array([coxdeboor(V, s.q, j, v) for j in range(s.m + 1)])
    P = s.Pw.reshape(s.n + 1, s.m + 1, 4).astype(np.float64)
    h = np.einsum("i,j,ijk->k", Nu, Nv, P)
    return h[:3] / h[3]


# ---------------------------------------------------------------- evaluation
@pytest.mark.parametrize("p,q", [(1, 1), (2, 3), (3, 3), (4, 2), (5, 5), (2, 5)])
def test_patch_eval_matches_bernstein(oracle, p, q):
    rng = np.random.default_rng(p * 10 + q)
    net = np.concatenate([rng.normal(size=((p + 1) * (q + 1), 3)),
                          rng.uniform(0.5, 2, size=((p + 1) * (q + 1), 1))], 1).astype(np.float32)
    for u, v in rng.uniform(0, 1, size=(20, 2)):
        out = np.zeros(12, np.float32)
        oracle.lib().orc_patch_eval_f32(net, p, q, np.float32(u), np.float32(v), out)
        pt, du, dv = bezier_np(net, p, q, float(np.float32(u)), float(np.float32(v)))
        np.testing.assert_allclose(out[0:4], pt, rtol=2e-5, atol=2e-5)
        np.testing.assert_allclose(out[4:8], du, rtol=2e-4, atol=2e-4)
        np.testing.assert_allclose(out[8:12], dv, rtol=2e-4, atol=2e-4)


def test_ray_planes_invariants(oracle):
    rng = np.random.default_rng(3)
    for _ in range(200):
        o = rng.normal(size=3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        pl = np.zeros(8, np.float32)
        oracle.lib().orc_ray_planes_f32(o, d, pl)
        n1, n2 = pl[0:3].astype(np.float64), pl[3:6].astype(np.float64)
        assert abs(np.dot(n1, d)) < 1e-6 and abs(np.dot(n2, d)) < 1e-6
        assert abs(np.linalg.norm(n1) - 1) < 1e-6 and abs(np.dot(n1, n2)) < 1e-6
        assert abs(np.dot(n1, o) + pl[6]) < 1e-5 and abs(np.dot(n2, o) + pl[7]) < 1e-5


def test_slab_matches_brute_force(oracle):
    rng = np.random.default_rng(4)
    for _ in range(300):
        lo = rng.uniform(-1, 1, size=3)
        b = np.concatenate([lo, lo + rng.uniform(0.05, 1, size=3)]).astype(np.float32)
        o = rng.uniform(-2, 2, size=3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        t = oracle.lib().orc_slab_f32(b, o, d)
        ts = np.linspace(0, 6, 60001)
        pts = o[None].astype(np.float64) + ts[:, None] * d[None]
        inside = np.all((pts >= b[:3] - 1e-6) & (pts <= b[3:] + 1e-6), axis=1)
        if inside.any():
            assert np.isfinite(t) and abs(t - ts[inside][0]) < 2e-4
        else:
            assert not np.isfinite(t) or (t > 5.9)


# ------------------------------------------------------------ decomposition
@pytest.mark.parametrize("which", ["c1", "torus", "mixed0", "mixed1", "mixed2"])
def test_decomposition_identity(oracle, which):
    """NURBS surface point (FP64 Cox-de Boor) equals the Bezier patch at local
    coordinates (u-u0)/(u1-u0), for every patch (degree up to 5)."""
    if which == "c1":
        surf = scenes.bicubic_grid_surface(1)
    elif which == "torus":
        surf = scenes.torus()
    else:
        surf = c2_small(3, 7)[int(which[-1])]
    d = oracle.decompose(surf)
    assert d["nbU"] * d["nbV"] == scenes.patch_count(surf)
    rng = np.random.default_rng(9)
    for k in range(0, len(d["cp"]), max(1, len(d["cp"]) // 12)):
        u0, u1, v0, v1 = d["uv"][k].astype(np.float64)
        for a, b in rng.uniform(0, 1, size=(4, 2)):
            pt, _, _ = bezier_np(d["cp"][k], surf.p, surf.q, a, b)
            ref = nurbs_point(surf, u0 + a * (u1 - u0), v0 + b * (v1 - v0))
            np.testing.assert_allclose(pt[:3] / pt[3], ref, rtol=0, atol=3e-5)


def test_decomposition_of_a_bezier_is_identity(oracle):
    rng = np.random.default_rng(2)
    p, q = 4, 3
    Pw = np.concatenate([rng.normal(size=((p + 1) * (q + 1), 3)), np.ones(((p + 1) * (q + 1), 1))],
                        1).astype(np.float32)
    U = np.array([0] * (p + 1) + [1] * (p + 1), np.float32)
    V = np.array([0] * (q + 1) + [1] * (q + 1), np.float32)
    d = oracle.decompose(scenes.Nurbs(p, q, p, q, U, V, Pw))
    assert len(d["cp"]) == 1
    assert np.array_equal(d["cp"][0], Pw)


# ---------------------------------------------------------------------- BVH
def _check_bvh(nodes, used, bounds, n):
    seen = np.zeros(n, int)
    for i in range(used):
        nd = nodes[i]
        if nd["prim_count"] > 0:
            seen[nd["left_first"]: nd["left_first"] + nd["prim_count"]] += 1
        else:
            for c in (nd["left_first"], nd["left_first"] + 1):
                assert np.all(nodes[c]["b"][:3] >= nd["b"][:3]) and np.all(nodes[c]["b"][3:] <= nd["b"][3:])
    assert np.all(seen == 1)


def test_bvh_invariants_and_sah_optimality(oracle):
    rng = np.random.default_rng(5)
    for trial in range(30):
        n = int(rng.integers(2, 9))
        lo = rng.uniform(0, 10, size=(n, 3))
        b = np.concatenate([lo, lo + rng.uniform(0.1, 2, size=(n, 3))], 1).astype(np.float32)
        nodes, perm, used = oracle.build_bvh(b)
        _check_bvh(nodes, used, b, n)
        # root split is the minimum-cost candidate over axes x centroids
        c = 0.5 * (b[:, :3] + b[:, 3:])
        best = np.inf
        for ax in range(3):
            for pos in c[:, ax]:
                L = c[:, ax] < pos
                if L.all() or not L.any():
                    continue
                cost = 0.0
                for side in (L, ~L):
                    e = b[side, 3:].max(0) - b[side, :3].min(0)
                    cost += side.sum() * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0])
                best = min(best, cost)
        if np.isfinite(best):
            l, r = nodes[nodes[0]["left_first"]], nodes[nodes[0]["left_first"] + 1]
            # the cost of the chosen split from the partition counts
            cnt_l = _count_leaves(nodes, nodes[0]["left_first"])
            cnt_r = _count_leaves(nodes, nodes[0]["left_first"] + 1)
            el = l["b"][3:].astype(np.float64) - l["b"][:3]
            er = r["b"][3:].astype(np.float64) - r["b"][:3]
            got = cnt_l * (el[0] * el[1] + el[1] * el[2] + el[2] * el[0]) + \
                cnt_r * (er[0] * er[1] + er[1] * er[2] + er[2] * er[0])
            assert got <= best * (1 + 1e-5)


def _count_leaves(nodes, i):
    nd = nodes[i]
    if nd["prim_count"] > 0:
        return int(nd["prim_count"])
    return _count_leaves(nodes, nd["left_first"]) + _count_leaves(nodes, nd["left_first"] + 1)


# ----------------------------------------------------------------- traversal
def test_traversal_equals_brute_force(oracle):
    """The ordered closest-hit traversal returns the argmin over all patches of
    the converged per-(ray, patch) Newton t, except at near ties."""
    sc = scenes.config1(4, 64, 64)
    P = oracle.c1_pipeline(sc)
    s = sc.surfaces[0]
    r = oracle.trace_f32(P["nodes"], P["ps"], P["rays"], max_iter=8, thr_res=1e-5, thr_det=1e-5)
    L = oracle.lib()
    n = len(P["rays"][0])
    nets = P["nets"]
    mism = 0
    for i in range(0, n, 7):
        o = np.array([P["rays"][k][i] for k in range(3)], np.float32)
        d = np.array([P["rays"][k][i] for k in range(3, 6)], np.float32)
        ts = []
        for k in range(len(nets)):
            seed = np.zeros(2, np.float32)
            L.orc_bilinear_seed_f32(nets[k], s.p, s.q, o, d, seed)
            out = np.zeros(11, np.float32)
            ev = L.orc_newton_f32(nets[k], s.p, s.q, o, d, seed[0], seed[1], 1e-5, 1e-5, i, 8, out)
            if ev > 0:
                t = np.float32(np.float32(out[0] - o[0]) * d[0] + np.float32(out[1] - o[1]) * d[1]
                               + np.float32(out[2] - o[2]) * d[2])
                if t > 0:
                    ts.append((t, k))
        want = min(ts)[1] if ts else -1
        if want != r["hit"][i]:
            srt = sorted(ts)
            near_tie = len(srt) > 1 and (srt[1][0] - srt[0][0]) <= 1e-4 * srt[0][0]
            mism += 0 if near_tie else 1
    assert mism == 0


# --------------------------------------------------------- analytic surfaces
def test_sphere_hits_are_analytic(oracle):
    """The reference's rational sphere is exact: FP64 trace t equals the
    ray-sphere quadratic root."""
    cam = scenes.reference_camera(160, 90)
    P = oracle.c1_pipeline(scenes.Scene([scenes.sphere()], cam))
    r = oracle.trace_f64(P["nodes"], P["ps"], P["rays"], max_iter=40, thr_res=1e-11, thr_det=1e-14)
    o = np.stack(P["rays"][:3], 1).astype(np.float64)
    d = np.stack(P["rays"][3:], 1).astype(np.float64)
    c = np.array([0, 0, 6.0])
    oc = o - c
    b = np.einsum("ij,ij->i", oc, d)
    disc = b * b - (np.einsum("ij,ij->i", oc, oc) - 1.0) * np.einsum("ij,ij->i", d, d)
    hit = r["hit"] >= 0
    assert hit.sum() > 100
    dd = np.einsum("ij,ij->i", d, d)
    near = (-b - np.sqrt(np.maximum(disc, 0))) / dd
    far = (-b + np.sqrt(np.maximum(disc, 0))) / dd
    # one Newton solve per patch from the patch-centre seed: a ray crossing one
    # large patch twice may converge to its far root (reference behaviour).
    # The FP32 circle weight 0.70710678 makes the surface a sphere to ~2e-7.
    err = np.minimum(np.abs(r["t"] - near), np.abs(r["t"] - far)) / near
    assert np.all(err[hit] < 1e-6)
    assert np.mean(np.abs(r["t"][hit] - near[hit]) < 1e-6 * near[hit]) > 0.5
    t_an = np.where(np.abs(r["t"] - near) < np.abs(r["t"] - far), near, far)
    # normals: radial direction
    nrm = np.cross(r["dpdu"][hit], r["dpdv"][hit])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    radial = o[hit] + t_an[hit, None] * d[hit] - c
    radial /= np.linalg.norm(radial, axis=1, keepdims=True)
    assert np.all(np.abs(np.abs(np.einsum("ij,ij->i", nrm, radial)) - 1) < 1e-5)


def test_torus_hits_are_analytic(oracle):
    cam = scenes.reference_camera(120, 68)
    P = oracle.c1_pipeline(scenes.Scene([scenes.torus()], cam))
    r = oracle.trace_f64(P["nodes"], P["ps"], P["rays"], max_iter=40, thr_res=1e-11, thr_det=1e-14)
    hit = np.nonzero(r["hit"] >= 0)[0]
    assert len(hit) > 50
    R0, r0 = 1.0, 0.4
    for i in hit[:: max(1, len(hit) // 60)]:
        o = np.array([P["rays"][k][i] for k in range(3)], np.float64) - [0, 0, 6.0]
        d = np.array([P["rays"][k][i] for k in range(3, 6)], np.float64)
        # (|p|^2 + R^2 - r^2)^2 = 4 R^2 (x^2 + y^2), p = o + t d
        dd, od, oo = d @ d, o @ d, o @ o
        k0 = oo + R0 * R0 - r0 * r0
        coeffs = [dd * dd, 4 * dd * od, 2 * dd * k0 + 4 * od * od - 4 * R0 * R0 * (d[0] ** 2 + d[1] ** 2),
                  4 * od * k0 - 8 * R0 * R0 * (o[0] * d[0] + o[1] * d[1]), k0 * k0 - 4 * R0 * R0 * (o[0] ** 2 + o[1] ** 2)]
        roots = np.roots(coeffs)
        real = np.sort(roots[np.abs(roots.imag) < 1e-7].real)
        real = real[real > 0]
        assert len(real) and np.min(np.abs(real - r["t"][i])) < 1e-6 * real[0]


# --------------------------------------------- closest-root FP64 reference
def test_closest_root_reference_sphere(oracle):
    """With grid starts every hit of the rational sphere is the near root of
    the ray-sphere quadratic, where the plain seed's single solve reaches the
    far root on some rays (the check a closest-root search that kept any
    converged root, or the first, would fail)."""
    cam = scenes.reference_camera(160, 90)
    P = oracle.c1_pipeline(scenes.Scene([scenes.sphere()], cam))
    o = np.stack(P["rays"][:3], 1).astype(np.float64)
    d = np.stack(P["rays"][3:], 1).astype(np.float64)
    oc = o - np.array([0, 0, 6.0])
    b = np.einsum("ij,ij->i", oc, d)
    dd = np.einsum("ij,ij->i", d, d)
    disc = b * b - (np.einsum("ij,ij->i", oc, oc) - 1.0) * dd
    near = (-b - np.sqrt(np.maximum(disc, 0))) / dd
    kw = dict(max_iter=40, thr_res=1e-11, thr_det=1e-14)
    plain = oracle.trace_f64(P["nodes"], P["ps"], P["rays"], **kw)
    roots = oracle.trace_f64(P["nodes"], P["ps"], P["rays"], grid=4, **kw)
    hp, hr = plain["hit"] >= 0, roots["hit"] >= 0
    assert np.all(hr[hp]) and hr.sum() >= hp.sum() > 100
    assert np.all(disc[hr] >= 0)
    assert np.all(np.abs(roots["t"][hr] - near[hr]) < 1e-6 * near[hr])
    assert np.any(np.abs(plain["t"][hp] - near[hp]) > 1e-3 * near[hp])


def test_closest_root_reference_torus(oracle):
    """Every closest-root hit on the rational torus is the smallest positive
    real root of the ray-torus quartic; the plain seed misses it on some."""
    cam = scenes.reference_camera(120, 68)
    P = oracle.c1_pipeline(scenes.Scene([scenes.torus()], cam))
    kw = dict(max_iter=40, thr_res=1e-11, thr_det=1e-14)
    plain = oracle.trace_f64(P["nodes"], P["ps"], P["rays"], **kw)
    roots = oracle.trace_f64(P["nodes"], P["ps"], P["rays"], grid=4, **kw)
    R0, r0 = 1.0, 0.4
    hit = np.nonzero(roots["hit"] >= 0)[0]
    assert len(hit) > 50
    plain_far = 0
    for i in hit:
        o = np.array([P["rays"][k][i] for k in range(3)], np.float64) - [0, 0, 6.0]
        d = np.array([P["rays"][k][i] for k in range(3, 6)], np.float64)
        dd, od, oo = d @ d, o @ d, o @ o
        k0 = oo + R0 * R0 - r0 * r0
        coeffs = [dd * dd, 4 * dd * od, 2 * dd * k0 + 4 * od * od - 4 * R0 * R0 * (d[0] ** 2 + d[1] ** 2),
                  4 * od * k0 - 8 * R0 * R0 * (o[0] * d[0] + o[1] * d[1]), k0 * k0 - 4 * R0 * R0 * (o[0] ** 2 + o[1] ** 2)]
        roots_q = np.roots(coeffs)
        real = np.sort(roots_q[np.abs(roots_q.imag) < 1e-7].real)
        real = real[real > 0]
        assert len(real) and abs(roots["t"][i] - real[0]) < 1e-6 * real[0]
        if plain["hit"][i] < 0 or abs(plain["t"][i] - real[0]) > 1e-6 * real[0]:
            plain_far += 1
    assert plain_far > 0
