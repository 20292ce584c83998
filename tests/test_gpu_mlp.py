"""Learned seed on the GPU against the CPU oracle (oracle/mlp_oracle.py):
BF16 MFMA inference (seed and hit outputs), the training kernels' gradient
(squared seed error on hits + cross-entropy of the hit output), both kinds
of pair features and the training pairs produced by the plain method.

Tolerances: the oracle's quantized mode rounds to BF16 at the same points as
the kernels, leaving FP32 accumulation order, the hardware exp/reciprocal in
GELU and sigmoid, and BF16 rounding-boundary flips: 2e-3 absolute on (u,v),
2e-2 relative L2 per gradient tensor.  Against plain FP64 the BF16 rounding
itself adds about 1e-2: 2e-2 on (u,v), 6e-2 per tensor."""
import os
import sys

import numpy as np
import pytest

from rtnurbs import scenes

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu

TENSORS = [("W1", M.OFF_W1, M.OFF_B1), ("b1", M.OFF_B1, M.OFF_W2), ("W2", M.OFF_W2, M.OFF_B2),
           ("b2", M.OFF_B2, M.OFF_W3), ("W3", M.OFF_W3, M.OFF_B3), ("b3", M.OFF_B3, M.OFF_W4),
           ("W4", M.OFF_W4, M.OFF_B4), ("b4", M.OFF_B4, M.NPARAM)]


def _net(rtl, seed):
    net = rtl.Mlp(seed=seed)
    w = net.get_params()
    rng = np.random.default_rng(seed)
    for _, a, b in TENSORS[1::2]:
        w[a:b] = rng.normal(0, 0.2, b - a)
    net.set_params(w)
    return net, net.get_params().astype(np.float64)


def test_forward_matches_oracle(rtl):
    net, w = _net(rtl, 3)
    rng = np.random.default_rng(0)
    n = 1000  # 15 full waves and a ragged one
    X = rng.normal(0, 1, (n, M.NIN)).astype(np.float32)
    valid = (rng.uniform(size=n) < 0.7).astype(np.int32)  # holes: rows are packed before the products
    valid[:64] = 1
    print("forward n=1000", flush=True)
    Xd = rtl.DeviceArray.from_host(X)
    uv_all, ph = net.forward(Xd, hit=True)
    uv_all, ph = uv_all.get(), ph.get()
    uv_m = net.forward(Xd, rtl.DeviceArray.from_host(valid)).get()
    assert np.array_equal(net.forward(Xd).get(), uv_all)  # the hit output is optional
    # inference uses the polynomial GELU (gelu_infer); the FP64 tanh-GELU
    # network of training is the looser bound
    yq, _ = M.forward(w, X.astype(np.float64), quantize=True, infer=True)
    y, _ = M.forward(w, X.astype(np.float64))
    got = np.concatenate([uv_all, ph[:, None]], 1)
    assert np.abs(got - yq).max() < 2e-3
    assert np.abs(got - y).max() < 2e-2
    m = valid == 1
    assert np.array_equal(uv_m[m], uv_all[m])  # packing does not change a pair's result
    assert np.all(uv_m[~m] == 0)


def test_gradient_matches_oracle(rtl):
    net, w = _net(rtl, 5)
    rng = np.random.default_rng(1)
    n = 300
    X = rng.normal(0, 1, (n, M.NIN)).astype(np.float32)
    T = np.concatenate([rng.uniform(0, 1, (n, 2)), rng.uniform(size=(n, 1)) < 0.6], 1).astype(np.float32)
    lam = 0.2
    print("gradient n=300", flush=True)
    g, loss = net.gradient(rtl.DeviceArray.from_host(X), rtl.DeviceArray.from_host(T), hit_weight=lam)
    lq, gq = M.loss_and_grad(w, X.astype(np.float64), T.astype(np.float64), quantize=True, hit_weight=lam)
    lf, gf = M.loss_and_grad(w, X.astype(np.float64), T.astype(np.float64), hit_weight=lam)
    assert abs(loss - lq) < 1e-3 * lq and abs(loss - lf) < 2e-2 * lf
    for name, a, b in TENSORS:
        rq = np.linalg.norm(g[a:b] - gq[a:b]) / np.linalg.norm(gq[a:b])
        rf = np.linalg.norm(g[a:b] - gf[a:b]) / np.linalg.norm(gf[a:b])
        cos = g[a:b] @ gf[a:b] / (np.linalg.norm(g[a:b]) * np.linalg.norm(gf[a:b]))
        assert rq < 2e-2 and rf < 6e-2 and cos > 0.998, (name, rq, rf, cos)


def _patch_geometry(dl):
    """Bounds (true division by w) and control nets per patch, from a scene download."""
    cp, off, pq = dl["cp"].astype(np.float32), dl["off"], dl["pq"]
    n = len(off)
    bounds = np.zeros((n, 6), np.float32)
    nets = []
    for i in range(n):
        p, q = pq[i]
        net = cp[off[i]:off[i] + (p + 1) * (q + 1)]
        xyz = net[:, :3] / net[:, 3:4]
        bounds[i, :3], bounds[i, 3:] = xyz.min(0), xyz.max(0)
        nets.append(net)
    return bounds, nets


@pytest.mark.parametrize("name,mk", [("c1", lambda: scenes.config1(0, 96, 96)),
                                     ("torus", lambda: scenes.Scene([scenes.torus()],
                                                                     scenes.reference_camera(160, 90)))])
def test_pairs_and_features(rtl, name, mk):
    sc = mk()
    print(f"pairs {name}", flush=True)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera)
    prm = rtl.params(mode=rtl.MODE_WIDE, max_iter=8, thr_res=1e-5, thr_det=1e-5, hull_cull=2)
    pr, pp, puv, cnt = rtl.trace_pairs(scene, rays, prm, 4 * rays.n)
    assert 0 < cnt <= 4 * rays.n
    pr, pp, puv = pr.get()[:cnt], pp.get()[:cnt], puv.get()[:cnt]
    assert np.all((puv >= 0) & (puv <= 1))
    # every solve: the converged ones again, the others with uv = (-1, -1);
    # with the mixed-precision seed and the stall exit as well
    for kw in (dict(), dict(precision=rtl.PREC_MIXED, cond_thr=1e4, stall_exit=1)):
        prm_a = rtl.params(mode=rtl.MODE_WIDE, max_iter=8, thr_res=1e-5, thr_det=1e-5, hull_cull=2, **kw)
        ar, ap, auv, acnt = rtl.trace_pairs(scene, rays, prm_a, 8 * rays.n, all=True)
        ar, ap, auv = ar.get()[:acnt], ap.get()[:acnt], auv.get()[:acnt]
        conv = auv[:, 0] >= 0
        assert np.all((auv[conv] >= 0) & (auv[conv] <= 1)) and np.all(auv[~conv] == -1.0)
        assert conv.sum() > 0 and (~conv).sum() > 0
        _, tot = rtl.trace(scene, rays, rtl.params(count=1, **{k: getattr(prm_a, k) for k in (
            "mode", "max_iter", "thr_res", "thr_det", "hull_cull", "precision", "cond_thr", "stall_exit")}))
        assert acnt == tot["newton_calls"]  # one record per Newton solve of the trace
        if not kw:
            assert sorted(zip(ar[conv].tolist(), ap[conv].tolist())) == sorted(zip(pr.tolist(), pp.tolist()))
    # every closest hit of the same form is among the converged solves
    hits, _ = rtl.trace(scene, rays, prm)
    h = hits.get()
    have = set(zip(pr.tolist(), pp.tolist()))
    for r in np.flatnonzero(h["hit"] >= 0)[::7]:
        assert (int(r), int(h["hit"][r])) in have
    # features against the oracle
    f = rtl.pair_features(scene, rays, rtl.DeviceArray.from_host(pr), rtl.DeviceArray.from_host(pp), cnt).get()
    dl = scene.download()
    bounds, nets = _patch_geometry(dl)
    o = np.stack([a for a in rays.get()[:3]], 1)[pr]
    d = np.stack([a for a in rays.get()[3:]], 1)[pr]
    ref = M.features(o, d, bounds[pp], [nets[k] for k in pp], dl["pq"][pp, 0], dl["pq"][pp, 1])
    # FP32 projections with hardware reciprocals against FP64: far below the
    # BF16 step (2^-8 relative) the features are rounded to
    np.testing.assert_allclose(f, ref, rtol=1e-4, atol=1e-4)
    fb = rtl.pair_features(scene, rays, rtl.DeviceArray.from_host(pr), rtl.DeviceArray.from_host(pp), cnt,
                           kind=rtl.MLP_FEAT_BOX).get()
    refb = M.box_features(o, d, bounds[pp], [nets[k] for k in pp], dl["pq"][pp, 0], dl["pq"][pp, 1])
    np.testing.assert_allclose(fb, refb, rtol=1e-4, atol=1e-4)


def test_solve_pairs_closest_root_on_the_sphere(rtl):
    """rtn_solve_pairs_roots (the closest-root training targets): on the
    rational sphere every converged root is one of the two roots of the
    ray-sphere quadratic (FP32 tolerance), never farther than the single
    solve's, found wherever the single solve converges in front of the
    origin, and on some pairs the near root where the single solve reached
    the far one (a patch on the front of the sphere seen twice)."""
    sc = scenes.Scene([scenes.sphere()], scenes.reference_camera(160, 90))
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera)
    prm = rtl.params(mode=rtl.MODE_WIDE, max_iter=8, thr_res=1e-5, thr_det=1e-5)
    pr, pp, puv, cnt = rtl.trace_pairs(scene, rays, prm, 8 * rays.n, all=True)
    assert cnt > 100
    sp = rtl.params(max_iter=16, thr_res=1e-5, thr_det=1e-5)
    a = rtl.solve_pairs(scene, rays, pr, pp, cnt, sp)
    b = rtl.solve_pairs(scene, rays, pr, pp, cnt, sp, grid=4)
    ha = a["converged"] & (a["t"] > 0)
    hb = b["converged"]
    assert np.all(hb[ha]) and np.all(b["t"][hb] > 0)
    assert np.all(b["t"][ha] <= a["t"][ha] * (1 + 1e-5))
    r = pr.get()[:cnt]
    o = np.stack(rays.get()[:3], 1).astype(np.float64)[r] - [0.0, 0.0, 6.0]
    d = np.stack(rays.get()[3:], 1).astype(np.float64)[r]
    bq = np.einsum("ij,ij->i", o, d)
    dd = np.einsum("ij,ij->i", d, d)
    disc = bq * bq - (np.einsum("ij,ij->i", o, o) - 1.0) * dd
    near = (-bq - np.sqrt(np.maximum(disc, 0))) / dd
    far = (-bq + np.sqrt(np.maximum(disc, 0))) / dd
    assert np.all(disc[hb] >= 0)
    tb = b["t"].astype(np.float64)
    is_near, is_far = np.abs(tb - near) <= 2e-5 * near, np.abs(tb - far) <= 2e-5 * far
    # grazing rays (a chord under 0.1 of the unit sphere) are left out: there
    # the FP32 acceptance (residual 1e-5) allows a larger error in t
    firm = hb & (far - near > 0.1)
    assert firm.sum() > 100 and np.all((is_near | is_far)[firm])
    ta = a["t"].astype(np.float64)
    assert np.any((ha & is_near) & (np.abs(ta - far) <= 2e-5 * far))
