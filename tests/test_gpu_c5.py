"""The batched persistent frame at its stated size: 32 scenes of 1k-200k
patches (1,169,134 in all, degrees 2-5, binned SAH), 256^2-2048^2 x 4
jittered spp (179,044,352 primary rays) plus one shadow and one mirror ray
per hit, in the persistent launches the benchmark times (per-lane persistent
and packet), captured into a hipGraph.  Checked on bounded windows of rays
spread over every scene, so the host never downloads the frame:

- no traversal stack overflows;
- the primaries equal a batch trace of the same rays in the same form, bit
  for bit (hit patch and t);
- a secondary ray never hits the patch it leaves, and empty slots miss;
- the counting run (rtn_frame_count) gives the same results;
- graph replays equal the eager run, bit for bit, with eager work between."""
import numpy as np
import pytest

from rtnurbs import scenes

pytestmark = pytest.mark.gpu

CFG = dict(max_iter=8, thr_res=1e-5, thr_det=1e-5, hull_cull=2, stall_exit=1)


@pytest.fixture(scope="module")
def c5(rtl):
    print("config 5: building 32 scenes", flush=True)
    items = scenes.config5(0)
    built = [rtl.Scene(s.surfaces).build(bvh=rtl.BVH_BINNED_SAH) for s in items]
    n_patches = sum(b.stats()["n_patches"] for b in built)
    batch = rtl.Batch(built)
    counts = [s.camera.W * s.camera.H * 4 for s in items]
    rays = rtl.Rays(sum(counts), with_idx=False)
    off = 0
    for s, n in zip(items, counts):
        rtl.generate_rays(s.camera, spp=4, jitter=True, seed=1, rays=rays, offset=off, tiled=True)
        off += n
    ids = rtl.DeviceArray.from_host(np.repeat(np.arange(len(items), dtype=np.int32), counts))
    starts = np.concatenate([[0], np.cumsum(counts)])
    print(f"config 5: {len(items)} scenes, {n_patches} patches, {rays.n} primary rays", flush=True)
    return batch, rays, ids, starts, n_patches


def _windows(starts, width=2048, per_scene=8):
    """Evenly spaced windows of `width` rays in every scene (the first at its
    start, the last at its end)."""
    wins = []
    for a, b in zip(starts[:-1], starts[1:]):
        for c in np.linspace(a, b - width, per_scene).astype(np.int64):
            wins.append((int(c), width))
    return wins


@pytest.mark.parametrize("packet", [0, 1])
def test_config5_frame_at_full_size(rtl, c5, packet):
    batch, rays, ids, starts, n_patches = c5
    assert len(starts) == 33 and rays.n == 179_044_352 and n_patches == 1_169_134
    fr = rtl.Frame(batch, rays, ids, rtl.frame_params(precision=rtl.PREC_MIXED, cond_thr=1e4, **CFG))
    fr.set_packet(packet, packet)
    print(f"frame eager (packet {packet})", flush=True)
    fr.run()
    rtl.synchronize()
    st = fr.status()
    assert st["overflow"] == 0
    assert st["n_primary"] == rays.n and st["n_secondary_slots"] == 2 * rays.n
    w1 = _windows(starts)
    w2 = w1 + [(rays.n + a, n) for a, n in w1]  # shadow slots, then mirror slots, of the same primaries
    eager = fr.sample(w1, w2)
    n_hit = int((eager["primary"]["hit"] >= 0).sum())
    assert n_hit > 0.05 * len(eager["primary"]["hit"]), n_hit
    # primaries: a batch trace of the same rays in the frame's form
    mode = rtl.MODE_PACKET if packet else rtl.MODE_PERSISTENT
    print("batch trace of the primaries", flush=True)
    hb, tb = batch.trace(rays, ids, rtl.params(mode=mode, precision=rtl.PREC_MIXED, cond_thr=1e4, **CFG),
                         totals=True)
    assert tb["overflow"] == 0
    ref = hb.window(w1)
    del hb
    assert np.array_equal(eager["primary"]["hit"], ref["hit"])
    m = ref["hit"] >= 0
    assert np.array_equal(eager["primary"]["t"][m].view(np.uint32), ref["t"][m].view(np.uint32))
    # secondaries: slot j*n + r belongs to primary r; empty slots miss; no ray hits the patch it leaves
    sec = eager["secondary"]
    par = sec["parent"]
    kind = sec["kind"]
    prim_of_slot = np.concatenate([eager["primary"]["hit"], eager["primary"]["hit"]])
    assert np.all((kind >= 0) == (prim_of_slot >= 0))
    assert np.all(par[kind >= 0] == np.concatenate([np.arange(a, a + n) for a, n in w1] * 2)[kind >= 0])
    assert np.all(sec["hit"][kind < 0] == -1)
    live = kind >= 0
    assert np.all(sec["hit"][live] != prim_of_slot[live])
    assert (sec["hit"][kind == 1] >= 0).any() and (sec["hit"][kind == 0] >= 0).any()
    # the counting run (work totals for pricing) gives the same results
    c_prim, c_sec = fr.count()
    assert c_prim["newton_calls"] > 0 and c_sec["newton_calls"] > 0 and c_prim["overflow"] == 0
    again = fr.sample(w1, w2)
    for part in ("primary", "secondary"):
        assert np.array_equal(again[part]["hit"], eager[part]["hit"]), part
    # graph capture: replays equal the eager run, eager work in between
    stream = rtl.stream_create()
    fr.capture(stream)
    for rep in range(2):
        print(f"replay {rep}", flush=True)
        fr.run(stream)
        rtl.synchronize()
        got = fr.sample(w1, w2)
        for part in ("primary", "secondary"):
            assert np.array_equal(got[part]["hit"], eager[part]["hit"]), (rep, part)
            mm = eager[part]["hit"] >= 0
            assert np.array_equal(got[part]["t"][mm].view(np.uint32), eager[part]["t"][mm].view(np.uint32))
        # eager work between replays: a wide batch trace of the first scene's first rays
        small = rtl.Rays.from_host([rtl.grab_windows(c, [(0, 4096)]) for c in rays.comp])
        batch.trace(small, rtl.DeviceArray.from_host(np.zeros(4096, np.int32)), rtl.params(mode=rtl.MODE_WIDE, **CFG))
    assert fr.status()["overflow"] == 0
