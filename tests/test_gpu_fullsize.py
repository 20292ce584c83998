"""The 10k-patch 1080p scene at its full size (10,018 patches, 1920x1080 x 4
jittered spp = 8,294,400 rays), through size-independent properties:

- the patch count is the frozen closed form;
- the packet form equals the one-ray wide form but for equal-t ties, and the
  persistent packet form equals the packet form bit for bit;
- exact rejection (oriented box + projected hull) changes nothing;
- against the FP64 reference, the FP32 forms keep hit t within 1e-5 on at
  least 99.8% of the same-patch pixels and have at most 0.3% non-tie patch
  mismatches (the plain method's own level); the mixed-precision forms, which
  start Newton from the FP64 reference's seed and take the FP64 step where
  the predicted t error exceeds 2.5e-6, keep t within 1e-5 on at least
  99.999% and have at most 0.05% non-tie mismatches and 0.01% hit/miss
  disagreements (3,146, 474 and 0.9999970 measured)."""
import numpy as np
import pytest

from rtnurbs import scenes
from test_gpu_wide import same_but_ties

pytestmark = pytest.mark.gpu

CFG = dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)


@pytest.fixture(scope="module")
def c2(rtl):
    print("case 10k-patch 1080p x 4 spp", flush=True)
    sc = scenes.config2(0)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera, spp=4, jitter=True, seed=1, tiled=True)
    ref = rtl.trace_f64(scene, rays, max_iter=32, thr_det=1e-12, thr_res=1e-10).get()
    return scene, rays, ref


def _fp64_check(g, ref, what, mixed=False):
    h32, h64 = g["hit"], ref["hit"]
    tie = (h64 >= 0) & np.isfinite(ref["t2"]) & ((ref["t2"] - ref["t"]) <= 1e-5 * np.abs(ref["t"]))
    both = (h32 >= 0) & (h64 >= 0)
    mism = (both & (h32 != h64) & ~tie).sum()
    hm = ((h32 >= 0) != (h64 >= 0)).sum()
    same = both & (h32 == h64)
    rel = np.abs(g["t"][same].astype(np.float64) - ref["t"][same]) / np.abs(ref["t"][same])
    frac = np.mean(rel <= 1e-5)
    print(f"{what}: non-tie mismatches {mism}, hit/miss {hm}, t within 1e-5 {frac:.7f}", flush=True)
    if mixed:
        assert mism <= 0.0005 * len(h32) and hm <= 0.0001 * len(h32), (what, mism, hm)
        assert frac >= 0.99999, (what, frac)
    else:
        assert mism <= 0.003 * len(h32), (what, mism)
        assert frac >= 0.998, (what, frac)


def test_fullsize_forms(rtl, c2):
    scene, rays, ref = c2
    assert rays.n == 8_294_400 and scene.stats()["n_patches"] == scenes.C2_PATCHES
    out = {}
    for name, kw in (("wide", dict(mode=rtl.MODE_WIDE)),
                     ("wide_hull2", dict(mode=rtl.MODE_WIDE, hull_cull=2)),
                     ("packet_hull2_mixed_stall", dict(mode=rtl.MODE_PACKET, hull_cull=2, stall_exit=1,
                                                       precision=rtl.PREC_MIXED, cond_thr=1e4)),
                     ("wide_hull2_mixed_stall", dict(mode=rtl.MODE_WIDE, hull_cull=2, stall_exit=1,
                                                     precision=rtl.PREC_MIXED, cond_thr=1e4)),
                     ("ppacket_hull2_mixed_stall", dict(mode=rtl.MODE_PERSISTENT_PACKET, hull_cull=2, stall_exit=1,
                                                        precision=rtl.PREC_MIXED, cond_thr=1e4))):
        print(f"form {name}", flush=True)
        h, tot = rtl.trace(rays=rays, scene=scene, prm=rtl.params(count=1, **CFG, **kw))
        assert tot["overflow"] == 0, name
        out[name] = h.get()
        _fp64_check(out[name], ref, name, mixed="mixed" in name)
    a, b = out["wide_hull2"], out["wide"]
    assert np.array_equal(a["hit"], b["hit"])
    m = b["hit"] >= 0
    assert np.array_equal(a["t"][m].view(np.uint32), b["t"][m].view(np.uint32))
    same_but_ties(out["packet_hull2_mixed_stall"], out["wide_hull2_mixed_stall"])
    a, b = out["ppacket_hull2_mixed_stall"], out["packet_hull2_mixed_stall"]
    assert np.array_equal(a["hit"], b["hit"])
    m = b["hit"] >= 0
    for k in ("t", "u", "v"):
        assert np.array_equal(a[k][m].view(np.uint32), b[k][m].view(np.uint32)), k


def _same_root_t(g, ref):
    """Fraction of the pixels on the reference's patch and root ((u, v) within
    1e-3) whose hit t is within 1e-5 relative of the reference's."""
    same = (g["hit"] >= 0) & (g["hit"] == ref["hit"])
    ue = np.maximum(np.abs(g["u"][same] - ref["u"][same]), np.abs(g["v"][same] - ref["v"][same]))
    te = np.abs(g["t"][same].astype(np.float64) - ref["t"][same]) / np.abs(ref["t"][same])
    root = ue < 1e-3
    return float(np.mean(te[root] <= 1e-5)), int(root.sum())


def _non_tie_mismatches(g, ref):
    tie = (ref["hit"] >= 0) & np.isfinite(ref["t2"]) & ((ref["t2"] - ref["t"]) <= 1e-5 * np.abs(ref["t"]))
    both = (g["hit"] >= 0) & (ref["hit"] >= 0)
    return int((both & (g["hit"] != ref["hit"]) & ~tie).sum())


def test_fullsize_headline_learned_form(rtl, c2):
    """The headline form (learned seed with hit gate 0.7 on full 64-lane
    visits, 4 + 8 evaluations, mixed precision, stall exit, persistent packet
    chunks) with a network trained as the benchmark trains it (10M converged
    pairs, 16 epochs): no stack overflow; hit t within 1e-5 of the FP64
    trace on >= 99.999% of the pixels where it reached the FP64 trace's root;
    at most 0.3% non-tie patch mismatches and 0.1% hit/miss disagreements
    with the FP64 trace (the network reaches nearer roots the plain seed
    misses); against the FP64 closest-root reference fewer non-tie
    mismatches than the same form without the network, and fewer Newton
    evaluations per ray (measured: 47-51k against 58k; 3.14-3.16 against
    3.25)."""
    from rtnurbs import learned
    scene, rays, ref = c2
    print("training the seed network (10M converged pairs, 16 epochs)", flush=True)
    net, losses, n, hits = learned.build_seed_network(epochs=16)
    assert hits == 10_000_000 and losses[-1] < losses[0]
    base = dict(mode=rtl.MODE_PERSISTENT_PACKET, hull_cull=2, stall_exit=1, precision=rtl.PREC_MIXED, cond_thr=1e4,
                count=1, **CFG)
    lrn = dict(seed_mode=rtl.SEED_LEARNED, seed_net=net, learned_iters=4, fallback_iters=8, seed_share=1,
               hit_gate=0.7, learned_min_lanes=64)
    print("form headline", flush=True)
    hl, tl = rtl.trace(scene, rays, rtl.params(**base, **lrn))
    print("form without the network", flush=True)
    hn, tn = rtl.trace(scene, rays, rtl.params(**base))
    assert tl["overflow"] == 0 and tn["overflow"] == 0
    g, gn = hl.get(), hn.get()
    frac, n_root = _same_root_t(g, ref)
    mism = _non_tie_mismatches(g, ref)
    hm = int(((g["hit"] >= 0) != (ref["hit"] >= 0)).sum())
    print(f"headline: same-root t within 1e-5 {frac:.7f} on {n_root} pixels, non-tie mismatches {mism}, "
          f"hit/miss {hm}", flush=True)
    assert n_root > 2_000_000 and frac >= 0.99999
    assert mism <= 0.003 * rays.n and hm <= 0.001 * rays.n
    print("closest-root reference", flush=True)
    roots = rtl.trace_f64(scene, rays, max_iter=32, thr_det=1e-12, thr_res=1e-10, grid=4).get()
    ml, mn = _non_tie_mismatches(g, roots), _non_tie_mismatches(gn, roots)
    print(f"against the closest root: headline {ml}, without the network {mn}; evaluations per ray "
          f"{tl['newton_evals'] / rays.n:.4f} against {tn['newton_evals'] / rays.n:.4f}", flush=True)
    assert ml < mn
    assert tl["newton_evals"] < tn["newton_evals"]
