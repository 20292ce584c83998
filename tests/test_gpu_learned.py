"""What the trained seed network gains over the method without it.

The network is trained on the GPU (its own forward/backward kernels) on
plain-method pairs from training seeds of the config-1 generator, then
evaluated on held-out seeds: on the (ray, patch) pairs the plain method
solves, the network's seed is closer to the solution than the plain
bilinear seed, and Newton from it needs fewer evaluations.  In the trace
forms it runs as the first attempt with the plain seed as fallback, and the
trace still finds what the plain method finds (near-identical hit sets)."""
import numpy as np
import pytest

from rtnurbs import learned, scenes

pytestmark = pytest.mark.gpu

CFG = dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)


@pytest.fixture(scope="module")
def net(rtl):
    print("training seed network", flush=True)
    sl = learned.training_scenes(n_c1=16, n_c2=0, res=384)
    feats, targets, n = learned.collect_pairs(sl, max_pairs=3_000_000)
    assert n > 1_000_000
    net, losses = learned.train_seed_network(feats, targets, n, epochs=6)
    assert losses[-1] < 0.7 * losses[0], losses
    return net


def test_learned_seed_beats_plain_seed_on_held_out_scenes(rtl, net):
    prm1 = rtl.params(**{**CFG, "max_iter": 1})
    prm = rtl.params(**CFG)
    err_p, err_l, it_p, it_l, conv_p, conv_l = [], [], [], [], [], []
    for seed in (2000, 2001, 2002):
        s = scenes.bicubic_grid_surface(seed)
        scene = rtl.Scene([s]).build()
        rays = rtl.generate_rays(scenes.c1_camera(160, 160))
        pr, pp, puv, cnt = rtl.trace_pairs(scene, rays, rtl.params(mode=rtl.MODE_WIDE, hull_cull=2, **CFG),
                                           4 * rays.n)
        target = puv.get()[:cnt]
        feats = rtl.pair_features(scene, rays, pr, pp, cnt)
        luv = net.forward(feats)
        plain_seed = rtl.solve_pairs(scene, rays, pr, pp, cnt, prm1)["uv"]
        err_p.append(np.abs(plain_seed - target).max(1))
        err_l.append(np.abs(luv.get() - target).max(1))
        a = rtl.solve_pairs(scene, rays, pr, pp, cnt, prm)
        b = rtl.solve_pairs(scene, rays, pr, pp, cnt, prm, seed_uv=luv)
        both = a["converged"] & b["converged"]
        it_p.append(a["iters"][both])
        it_l.append(b["iters"][both])
        conv_p.append(a["converged"])
        conv_l.append(b["converged"] | a["converged"])  # learned attempt, plain fallback
    ep, el = np.concatenate(err_p), np.concatenate(err_l)
    ip, il = np.concatenate(it_p), np.concatenate(it_l)
    print(f"seed error median plain {np.median(ep):.4f} learned {np.median(el):.4f}; "
          f"iterations plain {ip.mean():.3f} learned {il.mean():.3f}", flush=True)
    assert np.median(el) < 0.6 * np.median(ep)
    assert il.mean() < ip.mean()
    assert np.concatenate(conv_l).mean() >= np.concatenate(conv_p).mean()


def test_learned_trace_keeps_the_hits(rtl, net):
    sc = scenes.config1(2000)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera, spp=4, jitter=True, seed=3, tiled=True)
    a, ta = rtl.trace(scene, rays, rtl.params(mode=rtl.MODE_WIDE, hull_cull=2, count=1, **CFG))
    ga = a.get()
    for mode in (rtl.MODE_WIDE, rtl.MODE_PERSISTENT):
        for share in (0, 1):
            b, tb = rtl.trace(scene, rays, rtl.params(mode=mode, hull_cull=2, seed_mode=rtl.SEED_LEARNED,
                                                      seed_net=net, learned_iters=3, fallback_iters=8,
                                                      seed_share=share, count=1, **CFG))
            gb = b.get()
            assert np.mean(ga["hit"] == gb["hit"]) > 0.99, (mode, share)
            assert (gb["hit"] >= 0).sum() >= 0.995 * (ga["hit"] >= 0).sum(), (mode, share)
            # the learned attempt converges for most hits: fewer fallbacks than Newton calls
            assert 0 < tb["learned_fallbacks"] < 0.8 * tb["newton_calls"], (mode, share)
            # the same patch tests, up to the pruning that different closest hits cause
            assert abs(tb["newton_calls"] - ta["newton_calls"]) < 0.01 * ta["newton_calls"]


def test_learned_packet_forms(rtl, net):
    """The packet form with the learned seed: without seed sharing every pair
    gets its own network evaluation, so it equals the one-ray form but for
    equal-t ties; the persistent packet form traces the same 64-ray chunks
    with the same code, so it equals the packet form bit for bit."""
    from test_gpu_wide import same_but_ties
    sc = scenes.config1(2001)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera, spp=4, jitter=True, seed=5, tiled=True)
    lrn = dict(hull_cull=2, seed_mode=rtl.SEED_LEARNED, seed_net=net, learned_iters=3, fallback_iters=8, **CFG)
    out = {}
    for name, kw in (("wide0", dict(mode=rtl.MODE_WIDE, seed_share=0)), ("packet0", dict(mode=rtl.MODE_PACKET, seed_share=0)),
                     ("packet1", dict(mode=rtl.MODE_PACKET, seed_share=1, stall_exit=1, precision=rtl.PREC_MIXED)),
                     ("ppacket1", dict(mode=rtl.MODE_PERSISTENT_PACKET, seed_share=1, stall_exit=1,
                                       precision=rtl.PREC_MIXED))):
        print(f"case learned {name}", flush=True)
        h, tot = rtl.trace(scene, rays, rtl.params(count=1, **lrn, **kw))
        assert tot["overflow"] == 0 and 0 < tot["learned_fallbacks"] < 0.8 * tot["newton_calls"], name
        out[name] = h.get()
    same_but_ties(out["packet0"], out["wide0"])
    a, b = out["ppacket1"], out["packet1"]
    assert np.array_equal(a["hit"], b["hit"])
    m = b["hit"] >= 0
    for k in ("t", "u", "v"):
        assert np.array_equal(a[k][m].view(np.uint32), b[k][m].view(np.uint32)), k
    # sharing and the stall exit keep the hits of the unshared form
    assert np.mean(b["hit"] == out["packet0"]["hit"]) > 0.99


def test_learned_packet_over_a_batch(rtl, net):
    """The learned seed in the packet form over a multi-scene batch (two
    scenes, rays of both mixed within every wave): without seed sharing it
    equals the wide form over the same batch but for equal-t ties."""
    from test_gpu_frame import _batch_rays
    from test_gpu_wide import same_but_ties
    items = [(rtl.Scene(scenes.config1(k).surfaces).build(), scenes.c1_camera(64, 64)) for k in (2002, 2003)]
    rays, ids, _ = _batch_rays(rtl, items, spp=2)
    perm = np.random.default_rng(1).permutation(rays.n)
    rp = rtl.Rays.from_host([a[perm] for a in rays.get()])
    ip = rtl.DeviceArray.from_host(ids.get()[perm])
    batch = rtl.Batch([sc for sc, _ in items])
    lrn = dict(hull_cull=2, seed_mode=rtl.SEED_LEARNED, seed_net=net, learned_iters=3, fallback_iters=8, seed_share=0,
               precision=rtl.PREC_MIXED, **CFG)
    print("case learned batch", flush=True)
    w, _ = batch.trace(rp, ip, rtl.params(mode=rtl.MODE_WIDE, **lrn))
    p, tp = batch.trace(rp, ip, rtl.params(mode=rtl.MODE_PACKET, count=1, **lrn))
    assert tp["overflow"] == 0 and tp["learned_fallbacks"] > 0
    same_but_ties(p.get(), w.get())
