"""What the trained seed network gains over the method without it.

The network is trained on the GPU (its own forward/backward kernels) at the
stated size, 10M converged pairs (plus the misses the same traces solve) and
16 epochs, on plain-method pairs from training seeds of the config-1
generator, once per feature kind (the box-frame inputs as stated: ray, four
corners, degrees; and the projected control net), then evaluated on held-out
seeds: on the (ray, patch) pairs the plain method solves, the network's seed
is closer to the solution than the plain bilinear seed and Newton from it
needs fewer evaluations; its hit output separates the solves that converge
from those that do not.  In the trace forms it runs as the first attempt
with the plain seed as fallback (and, below the hit gate, the plain seed
alone), and the trace still finds what the plain method finds."""
import numpy as np
import pytest

from rtnurbs import learned, scenes

pytestmark = pytest.mark.gpu

CFG = dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)


def _train(rtl, kind):
    print(f"training seed network kind {kind}", flush=True)
    sl = learned.training_scenes(n_c1=400, n_c2=0, res=384)
    feats, targets, n, hits = learned.collect_pairs(sl, max_hits=10_000_000, kind=kind)
    assert hits == 10_000_000 and n > hits
    net, losses = learned.train_seed_network(feats, targets, n, epochs=16, kind=kind)
    print(f"kind {kind}: {n} pairs, loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.7 * losses[0], losses
    return net


@pytest.fixture(scope="module")
def nets(rtl):
    return {k: _train(rtl, k) for k in (rtl.MLP_FEAT_PROJECTED, rtl.MLP_FEAT_BOX)}


@pytest.fixture(scope="module")
def net(nets, rtl):
    return nets[rtl.MLP_FEAT_PROJECTED]


@pytest.mark.parametrize("kind", [0, 1])
def test_learned_seed_beats_plain_seed_on_held_out_scenes(rtl, nets, kind):
    net = nets[kind]
    prm1 = rtl.params(**{**CFG, "max_iter": 1})
    prm = rtl.params(**CFG)
    err_p, err_l, it_p, it_l, conv_p, conv_l, hp, hl = [], [], [], [], [], [], [], []
    for seed in (2000, 2001, 2002):
        s = scenes.bicubic_grid_surface(seed)
        scene = rtl.Scene([s]).build()
        rays = rtl.generate_rays(scenes.c1_camera(160, 160))
        pr, pp, puv, cnt = rtl.trace_pairs(scene, rays, rtl.params(mode=rtl.MODE_WIDE, hull_cull=2, **CFG),
                                           4 * rays.n)
        target = puv.get()[:cnt]
        feats = rtl.pair_features(scene, rays, pr, pp, cnt, kind=kind)
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
        # the hit output on every solve of the mixed form without the network
        prm_all = rtl.params(**{**learned.PAIR_PARAMS})
        ar, ap, auv, acnt = rtl.trace_pairs(scene, rays, prm_all, 8 * rays.n, all=True)
        _, ph = net.forward(rtl.pair_features(scene, rays, ar, ap, acnt, kind=kind), hit=True)
        hl.append(auv.get()[:acnt, 0] >= 0)
        hp.append(ph.get())
    ep, el = np.concatenate(err_p), np.concatenate(err_l)
    ip, il = np.concatenate(it_p), np.concatenate(it_l)
    h, ph = np.concatenate(hl), np.concatenate(hp)
    # ranking quality of the hit output: the probability that a random
    # converged solve scores above a random non-converged one (ROC AUC)
    order = np.argsort(ph, kind="stable")
    ranks = np.empty(len(ph))
    ranks[order] = np.arange(1, len(ph) + 1)
    npos, nneg = h.sum(), (~h).sum()
    auc = (ranks[h].sum() - npos * (npos + 1) / 2) / (npos * nneg)
    print(f"kind {kind}: seed error median plain {np.median(ep):.4f} learned {np.median(el):.4f}; "
          f"iterations plain {ip.mean():.3f} learned {il.mean():.3f}; hit AUC {auc:.4f} "
          f"({npos} hits, {nneg} misses)", flush=True)
    assert np.median(el) < 0.6 * np.median(ep)
    assert il.mean() < ip.mean()
    assert np.concatenate(conv_l).mean() >= np.concatenate(conv_p).mean()
    assert auc > (0.93 if kind == 0 else 0.85)


def _gated_net(rtl, net, logit):
    """net with its hit output forced to the constant logit (W4 row 2 = 0, b4[2] = logit)."""
    w = net.get_params()
    nh = rtl.MLP_HIDDEN
    off_w4 = rtl.MLP_NPARAM - 3 - 3 * nh  # W4[3][64] b4[3] close the layout
    w[off_w4 + 2 * nh:off_w4 + 3 * nh] = 0.0
    w[-1] = logit
    return rtl.Mlp(params=w, kind=net.kind)


@pytest.mark.parametrize("mode", ["wide", "packet"])
def test_hit_gate_and_skip(rtl, net, mode):
    """A network whose hit probability is 4.5e-5 everywhere: below hit_gate
    every pair starts from the plain seed, so the learned form equals the form
    without the network bit for bit; below hit_skip no pair is solved."""
    low = _gated_net(rtl, net, -10.0)
    sc = scenes.config1(2003)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera, spp=4, jitter=True, seed=7, tiled=True)
    m = rtl.MODE_WIDE if mode == "wide" else rtl.MODE_PACKET
    base = dict(mode=m, hull_cull=2, precision=rtl.PREC_MIXED, cond_thr=1e4, stall_exit=1, count=1, **CFG)
    lrn = dict(seed_mode=rtl.SEED_LEARNED, seed_net=low, learned_iters=3, fallback_iters=8, seed_share=0)
    print(f"case gate {mode}", flush=True)
    a, ta = rtl.trace(scene, rays, rtl.params(**base))
    b, tb = rtl.trace(scene, rays, rtl.params(**base, **lrn, hit_gate=0.5))
    ga, gb = a.get(), b.get()
    assert np.array_equal(ga["hit"], gb["hit"])
    m_ = ga["hit"] >= 0
    for k in ("t", "u", "v"):
        assert np.array_equal(ga[k][m_].view(np.uint32), gb[k][m_].view(np.uint32)), k
    for k in ("newton_calls", "newton_evals", "fp64_steps"):
        assert ta[k] == tb[k], k
    assert tb["learned_fallbacks"] == 0 and tb["learned_skips"] == 0
    c, tc = rtl.trace(scene, rays, rtl.params(**base, **lrn, hit_skip=0.5))
    assert (c.get()["hit"] < 0).all() and tc["newton_calls"] == 0 and tc["learned_skips"] >= ta["newton_calls"]
    # a network that predicts hits everywhere takes the learned attempt on every pair
    high = _gated_net(rtl, net, 10.0)
    d, td = rtl.trace(scene, rays, rtl.params(**base, **{**lrn, "seed_net": high}, hit_gate=0.5, hit_skip=0.4))
    assert td["learned_skips"] == 0 and td["learned_fallbacks"] > 0


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


def _surface_distance64(dl, rays_h, idx, prim, u, v):
    """FP64 distance from ray idx[k] to its recorded hit point (patch prim[k]
    at (u[k], v[k])): the Bernstein tensor sum of the homogeneous net."""
    from math import comb
    out = np.empty(len(idx))
    for k in range(len(idx)):
        P, Q = (int(x) for x in dl["pq"][prim[k]][:2])
        net = dl["cp"][dl["off"][prim[k]]:dl["off"][prim[k]] + (P + 1) * (Q + 1)].astype(np.float64)
        bu = np.array([comb(P, i) * u[k] ** i * (1 - u[k]) ** (P - i) for i in range(P + 1)], np.float64)
        bv = np.array([comb(Q, j) * v[k] ** j * (1 - v[k]) ** (Q - j) for j in range(Q + 1)], np.float64)
        h = np.einsum("i,j,ijk->k", bu, bv, net.reshape(P + 1, Q + 1, 4))
        o = np.array([rays_h[c][idx[k]] for c in range(3)], np.float64)
        d = np.array([rays_h[c][idx[k]] for c in range(3, 6)], np.float64)
        out[k] = np.linalg.norm(np.cross(h[:3] / h[3] - o, d / np.linalg.norm(d)))
    return out


def test_learned_headline_form_on_the_10k_patch_scene(rtl):
    """The learned seed + mixed precision + persistent packet form on the
    10k-patch 1080p scene (8,294,400 rays; network trained on disjoint
    generator seeds): fewer Newton evaluations per ray than the same form
    without the network; against the FP64 reference it loses no more of the
    reference's hits (missing, or replaced by a farther hit) than that form
    does, and where it differs the other way (a hit the reference lacks, or a
    nearer one) its hit is a point of the surface on the ray (FP64 distance
    within 2e-5, the Newton acceptance of 1e-5 plus FP32 rounding)."""
    print("case learned headline c2: training", flush=True)
    feats, targets, n, hits = learned.collect_pairs(learned.training_scenes(), max_hits=10_000_000)
    net, _ = learned.train_seed_network(feats, targets, n, epochs=16)
    del feats, targets
    sc = scenes.config2(0)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera, spp=4, jitter=True, seed=1, tiled=True)
    ref = rtl.trace_f64(scene, rays, max_iter=32, thr_det=1e-12, thr_res=1e-10).get()
    base = dict(mode=rtl.MODE_PERSISTENT_PACKET, hull_cull=2, stall_exit=1, precision=rtl.PREC_MIXED, cond_thr=1e4,
                count=1, **CFG)
    a, ta = rtl.trace(scene, rays, rtl.params(**base))
    b, tb = rtl.trace(scene, rays, rtl.params(**base, seed_mode=rtl.SEED_LEARNED, seed_net=net, learned_iters=4,
                                              fallback_iters=8, seed_share=1, hit_gate=0.7, learned_min_lanes=64))
    assert ta["overflow"] == 0 and tb["overflow"] == 0
    ga, gb = a.get(), b.get()
    t64, h64 = ref["t"], ref["hit"]
    tie = (h64 >= 0) & np.isfinite(ref["t2"]) & ((ref["t2"] - t64) <= 1e-5 * np.abs(t64))

    def classes(g):
        mm = (g["hit"] >= 0) & (h64 >= 0) & (g["hit"] != h64) & ~tie
        nearer = mm & (g["t"] < t64)
        lost = ((g["hit"] < 0) & (h64 >= 0)) | (mm & ~nearer)
        return lost, ((g["hit"] >= 0) & (h64 < 0)) | nearer

    lost_a, _ = classes(ga)
    lost_b, gain_b = classes(gb)
    ia, ib = ta["newton_evals"] / rays.n, tb["newton_evals"] / rays.n
    print(f"evaluations per ray {ia:.4f} without the network, {ib:.4f} with; reference hits lost "
          f"{lost_a.sum()} / {lost_b.sum()}; extra or nearer hits {gain_b.sum()}", flush=True)
    assert ib < ia
    assert lost_b.sum() <= lost_a.sum() * 1.02 + 50
    idx = np.flatnonzero(gain_b)[::max(1, int(gain_b.sum()) // 1500)]
    dist = _surface_distance64(scene.download(), [np.asarray(x) for x in rays.get()], idx, gb["hit"][idx],
                               gb["u"][idx], gb["v"][idx])
    print(f"extra or nearer hits on the surface: {np.mean(dist <= 2e-5):.4f} of {len(idx)}", flush=True)
    assert np.mean(dist <= 2e-5) >= 0.99


def test_learned_min_lanes(rtl, net):
    """The network runs for a wave's pairs on a patch only when at least
    learned_min_lanes lanes have one: a packet of 48 rays with the threshold
    at 49 never runs it and equals the form without the network bit for bit;
    at 1 it runs (a one-evaluation learned attempt falls back)."""
    sc = scenes.config1(2004)
    scene = rtl.Scene(sc.surfaces).build()
    full = rtl.generate_rays(sc.camera, spp=4, jitter=True, seed=9, tiled=True)
    # 48 rays of one 4 x 4-pixel block that sees the surface
    h0 = rtl.trace(scene, full, rtl.params(mode=rtl.MODE_PACKET, hull_cull=2, **CFG))[0].get()["hit"]
    blocks = np.flatnonzero((h0.reshape(-1, 64) >= 0).all(1))
    assert len(blocks) > 0
    sel = slice(int(blocks[len(blocks) // 2]) * 64, int(blocks[len(blocks) // 2]) * 64 + 48)
    rays = rtl.Rays.from_host([np.ascontiguousarray(a[sel]) for a in full.get()])
    base = dict(mode=rtl.MODE_PACKET, hull_cull=2, stall_exit=1, precision=rtl.PREC_MIXED, cond_thr=1e4, count=1,
                **CFG)
    lrn = dict(seed_mode=rtl.SEED_LEARNED, seed_net=net, learned_iters=1, fallback_iters=8, seed_share=0)
    print("case learned_min_lanes", flush=True)
    a, ta = rtl.trace(scene, rays, rtl.params(**base))
    b, tb = rtl.trace(scene, rays, rtl.params(**base, **lrn, learned_min_lanes=49))
    ga, gb = a.get(), b.get()
    assert np.array_equal(ga["hit"], gb["hit"]) and (ga["hit"] >= 0).all()
    for k in ("t", "u", "v"):
        assert np.array_equal(ga[k].view(np.uint32), gb[k].view(np.uint32)), k
    assert tb["learned_fallbacks"] == 0 and ta["newton_evals"] == tb["newton_evals"]
    # one evaluation from the network's seed converges on no pair: every pair falls back
    _, tc = rtl.trace(scene, rays, rtl.params(**base, **lrn, learned_min_lanes=1))
    assert tc["learned_fallbacks"] > 0


def test_learned_parameters_are_checked(rtl, net):
    """Out-of-range learned-seed parameters are rejected with RTN_EINVAL and a
    message, before any launch."""
    sc = scenes.config1(0, 32, 32)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera)
    lrn = dict(mode=rtl.MODE_PACKET, seed_mode=rtl.SEED_LEARNED, seed_net=net, **CFG)
    for bad in (dict(hit_gate=1.5), dict(hit_skip=-0.1), dict(learned_min_lanes=65), dict(learned_min_lanes=-1),
                dict(seed_share=2)):
        with pytest.raises(rtl.RtnError) as ei:
            rtl.trace(scene, rays, rtl.params(**lrn, **bad))
        assert ei.value.code == -1, bad
