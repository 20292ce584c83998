"""Multi-scene batches, the binned-SAH build, frames (primary + shadow +
mirror persistent launches), hipGraph capture and the autotuner.

A batch only relocates each scene's arrays, so tracing rays of several
scenes in one launch must reproduce, bit for bit, tracing each scene on its
own (with Newton's RNG seeded by the same, batch-wide, ray index).  A captured frame must reproduce its eager run on every replay, with
eager work between replays; a tuned launch shape changes scheduling only, so
it must reproduce the default shape's results.  A frame's secondary rays
skip the patch they leave; the same skip through rtn_trace's skip_patch
gives the same hits."""
import numpy as np
import pytest

from rtnurbs import scenes

pytestmark = pytest.mark.gpu

CFG = dict(max_iter=8, thr_res=1e-5, thr_det=1e-5, hull_cull=2)


def _mini_c5(rtl, n=4, res=(64, 96), seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        s = scenes.grid_surface(rng, int(rng.integers(800, 6000)), int(rng.integers(2, 6)), int(rng.integers(2, 6)))
        cam = scenes.c1_camera(res[k % len(res)], res[k % len(res)])
        sc = rtl.Scene([s]).build(bvh=rtl.BVH_BINNED_SAH)
        out.append((sc, cam))
    return out


def _batch_rays(rtl, items, spp=1):
    counts = [c.W * c.H * spp for _, c in items]
    rays = rtl.Rays(sum(counts))
    off = 0
    for (_, cam), n in zip(items, counts):
        rtl.generate_rays(cam, spp=spp, jitter=spp > 1, seed=1, rays=rays, offset=off)
        off += n
    ids = np.repeat(np.arange(len(items), dtype=np.int32), counts)
    return rays, rtl.DeviceArray.from_host(ids), counts


def test_binned_bvh_is_valid(rtl):
    rng = np.random.default_rng(3)
    s = scenes.grid_surface(rng, 5000, 3, 4)
    sc = rtl.Scene([s]).build(bvh=rtl.BVH_BINNED_SAH)
    st = sc.stats()
    assert st["n_patches"] == scenes.patch_count(s)
    nodes = sc.download()["nodes"]
    covered = np.zeros(st["n_patches"], np.int32)
    for nd in nodes:
        if nd["prim_count"] > 0:
            covered[nd["left_first"]:nd["left_first"] + nd["prim_count"]] += 1
        else:
            for c in (nd["left_first"], nd["left_first"] + 1):  # parent box contains the children
                assert np.all(nodes[c]["b"][:3] >= nd["b"][:3]) and np.all(nodes[c]["b"][3:] <= nd["b"][3:])
    assert np.all(covered == 1)
    # the binary (exact-form) and wide traversals agree on this tree
    cam = scenes.c1_camera(96, 96)
    rays = rtl.generate_rays(cam)
    ex, _ = rtl.trace(sc, rays, rtl.params(mode=rtl.MODE_EXACT, max_iter=8, thr_res=1e-5, thr_det=1e-5))
    wd, _ = rtl.trace(sc, rays, rtl.params(mode=rtl.MODE_WIDE, **CFG))
    assert np.mean(ex.get()["hit"] == wd.get()["hit"]) > 0.995


def test_batch_equals_single_scenes(rtl):
    items = [(rtl.Scene(scenes.config1(k).surfaces).build(), scenes.c1_camera(48, 48)) for k in (0, 1)]
    items += _mini_c5(rtl, n=2)
    rays, ids, counts = _batch_rays(rtl, items)
    batch = rtl.Batch([sc for sc, _ in items])
    print("batch trace", flush=True)
    for mode in (rtl.MODE_WIDE, rtl.MODE_PERSISTENT):
        hb, _ = batch.trace(rays, ids, rtl.params(mode=mode, **CFG))
        g = hb.get()
        off = 0
        for k, ((sc, cam), n) in enumerate(zip(items, counts)):
            r1 = rtl.generate_rays(cam)
            # Newton's RNG is seeded by the ray index: global in the batch, so offset here
            h1, _ = rtl.trace(sc, r1, rtl.params(mode=rtl.MODE_WIDE, seed_base=off, **CFG))
            s1 = h1.get()
            hk = g["hit"][off:off + n]
            exp = np.where(s1["hit"] >= 0, s1["hit"] + batch.patch_base[k], -1)
            assert np.array_equal(hk, exp), (mode, k)
            m = s1["hit"] >= 0
            assert np.array_equal(g["t"][off:off + n][m].view(np.uint32), s1["t"][m].view(np.uint32))
            off += n
    # packet forms over a batch: one traversal per scene root present in a wave;
    # the same hits as the wide form but for equal-t ties, persistent = one-shot
    from test_gpu_wide import same_but_ties
    wb, _ = batch.trace(rays, ids, rtl.params(mode=rtl.MODE_WIDE, **CFG))
    pb, _ = batch.trace(rays, ids, rtl.params(mode=rtl.MODE_PACKET, **CFG))
    qb, _ = batch.trace(rays, ids, rtl.params(mode=rtl.MODE_PERSISTENT_PACKET, **CFG))
    same_but_ties(pb.get(), wb.get())
    assert np.array_equal(qb.get()["hit"], pb.get()["hit"])
    # rays of different scenes interleaved within every wave
    perm = np.random.default_rng(0).permutation(rays.n)
    rp = rtl.Rays.from_host([a[perm] for a in rays.get()])
    ip = rtl.DeviceArray.from_host(ids.get()[perm])
    pw, _ = batch.trace(rp, ip, rtl.params(mode=rtl.MODE_WIDE, **CFG))
    pk, tk = batch.trace(rp, ip, rtl.params(mode=rtl.MODE_PACKET, count=1, **CFG))
    assert tk["overflow"] == 0
    same_but_ties(pk.get(), pw.get())
    bad = rtl.DeviceArray.from_host(np.full(rays.n, len(items), np.int32))
    with pytest.raises(rtl.RtnError):
        batch.trace(rays, bad, rtl.params(mode=rtl.MODE_WIDE, **CFG))


def test_frame_graph_replay_and_autotune(rtl):
    items = _mini_c5(rtl, n=4)
    rays, ids, counts = _batch_rays(rtl, items, spp=2)
    batch = rtl.Batch([sc for sc, _ in items])
    fr = rtl.Frame(batch, rays, ids, rtl.frame_params(**{k: v for k, v in CFG.items()}))
    print("frame eager", flush=True)
    fr.run()
    ref = fr.results()
    st = fr.status()
    assert st["overflow"] == 0 and st["n_secondary"] == 2 * int((ref["primary"]["hit"] >= 0).sum())
    # primaries equal a persistent batch trace
    hb, _ = batch.trace(rays, ids, rtl.params(mode=rtl.MODE_PERSISTENT, **CFG))
    assert np.array_equal(hb.get()["hit"], ref["primary"]["hit"])
    # shadow rays end at their first hit: same hit/miss as a closest-hit trace
    # of the same rays (same slots, so the same Newton seeds) that skips the
    # same patches (each ray's origin patch); empty slots miss
    par, kind = ref["parent"], ref["kind"]
    assert np.all(ref["secondary"]["hit"][kind < 0] == -1)
    sr = rtl.Rays.from_host(ref["secondary_rays"])
    sids = rtl.DeviceArray.from_host(ids.get()[par])
    origin = rtl.DeviceArray.from_host(np.where(kind >= 0, ref["primary"]["hit"][par], -1).astype(np.int32))
    hs, _ = batch.trace(sr, sids, rtl.params(mode=rtl.MODE_WIDE, skip_patch=origin, **CFG))
    hs = hs.get()
    sh, mi = kind == 0, kind == 1
    assert np.array_equal(ref["secondary"]["hit"][sh] >= 0, hs["hit"][sh] >= 0)
    assert np.array_equal(ref["secondary"]["hit"][mi], hs["hit"][mi])
    # graph capture and replays, eager work in between
    stream = rtl.stream_create()
    fr.capture(stream)
    assert fr.status()["captured"] == 1
    for rep in range(3):
        print(f"replay {rep}", flush=True)
        fr.run(stream)
        rtl.synchronize()
        got = fr.results()
        assert got["n_secondary"] == ref["n_secondary"]
        for part in ("primary", "secondary"):
            assert np.array_equal(got[part]["hit"], ref[part]["hit"]), (rep, part)
            m = ref[part]["hit"] >= 0
            assert np.array_equal(got[part]["t"][m].view(np.uint32), ref[part]["t"][m].view(np.uint32))
        batch.trace(rays, ids, rtl.params(mode=rtl.MODE_WIDE, **CFG))  # eager work between replays
    # packet form for both launches: the primaries equal the persistent form's
    # but for equal-t ties, and a captured graph replays the eager run
    print("frame packet", flush=True)
    from test_gpu_wide import same_but_ties
    fr.set_packet(1, 1)
    fr.run()
    pk = fr.results()
    same_but_ties(pk["primary"], ref["primary"])
    assert abs(pk["n_secondary"] - ref["n_secondary"]) <= 4
    assert np.mean(pk["secondary"]["hit"] == ref["secondary"]["hit"]) > 0.999
    fr.capture(stream)
    for rep in range(2):
        fr.run(stream)
        rtl.synchronize()
        got = fr.results()
        for part in ("primary", "secondary"):
            assert np.array_equal(got[part]["hit"], pk[part]["hit"]), (rep, part)
        batch.trace(rays, ids, rtl.params(mode=rtl.MODE_WIDE, **CFG))
    fr.set_packet(0, 0)
    # autotuner: tries every built shape, then the packet launches, keeps the
    # fastest; the results are the persistent form's (or, with packets chosen,
    # the packet form's)
    print("autotune", flush=True)
    table = fr.autotune(reps=1)
    assert len(table) >= 9 and any(e["ms"] > 0 for e in table)
    assert {(e["packet_primary"], e["packet_secondary"]) for e in table} >= {(0, 0), (1, 0), (1, 1)}
    st = fr.status()
    assert (st["shape_lanes"], st["shape_stack"]) in {(e["lanes"], e["stack"]) for e in table}
    fr.run()
    got = fr.results()
    want = pk if st["packet"] & 1 else ref
    if st["packet"] in (0, 3):
        for part in ("primary", "secondary"):
            assert np.array_equal(got[part]["hit"], want[part]["hit"]), (st["packet"], part)
    else:
        same_but_ties(got["primary"], want["primary"])
        assert np.mean(got["secondary"]["hit"] == ref["secondary"]["hit"]) > 0.999


def test_secondaries_skip_their_origin_patch(rtl):
    """A secondary ray does not test the patch it leaves.  A solve on that patch
    converges to the ray's own origin (t <= 0, rejected), so skipping it may
    only change the rays whose unskipped trace hits the origin patch again; all
    other rays keep their hit bit for bit, and Newton calls drop."""
    items = _mini_c5(rtl, n=3)
    rays, ids, counts = _batch_rays(rtl, items, spp=2)
    batch = rtl.Batch([sc for sc, _ in items])
    res = {}
    for skip in (0, 1):
        fr = rtl.Frame(batch, rays, ids, rtl.frame_params(skip_origin=skip, **CFG))
        fr.run()
        res[skip] = fr.results()
    par, kind = res[0]["parent"], res[0]["kind"]
    assert np.array_equal(res[1]["secondary_rays"], res[0]["secondary_rays"])
    origin = np.where(kind >= 0, res[0]["primary"]["hit"][par], -1).astype(np.int32)
    a, b = res[0]["secondary"], res[1]["secondary"]
    own = (a["hit"] == origin) & (origin >= 0)
    print(f"secondary rays {int((kind >= 0).sum())}, unskipped hits on the origin patch {int(own.sum())}", flush=True)
    assert np.all(b["hit"][origin >= 0] != origin[origin >= 0])
    same = ~own
    assert np.array_equal(a["hit"][same], b["hit"][same])
    m = same & (a["hit"] >= 0) & (kind == 1)
    assert np.array_equal(a["t"][m].view(np.uint32), b["t"][m].view(np.uint32))
    assert own.mean() < 0.01
    # the same skip through rtn_trace's skip_patch, in every fast form, with counters
    sr = rtl.Rays.from_host(res[0]["secondary_rays"])
    sids = rtl.DeviceArray.from_host(ids.get()[par])
    skip = rtl.DeviceArray.from_host(origin)
    for mode in (rtl.MODE_WIDE, rtl.MODE_PERSISTENT, rtl.MODE_PACKET, rtl.MODE_PERSISTENT_PACKET):
        h0, t0 = batch.trace(sr, sids, rtl.params(mode=mode, count=1, **CFG))
        h1, t1 = batch.trace(sr, sids, rtl.params(mode=mode, count=1, skip_patch=skip, **CFG))
        g0, g1 = h0.get()["hit"], h1.get()["hit"]
        own = (g0 == origin) & (origin >= 0)
        assert np.all(g1[origin >= 0] != origin[origin >= 0])
        assert np.mean(g0[~own] == g1[~own]) > 0.999
        print(f"mode {mode}: newton calls {t0['newton_calls']} -> {t1['newton_calls']}", flush=True)
        assert t1["newton_calls"] < t0["newton_calls"]
