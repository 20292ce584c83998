"""GPU fast forms (wide 4-ary BVH, persistent work queue, packet traversal,
hull rejection, mixed precision) against the exact plain method and the FP64
reference.

The fast forms change the arithmetic (FMA, tensor-Bernstein evaluation) and
the traversal order, not the method: per ray the same patches are tested
from the same seeds, so hit patches agree with the exact form except where
two hits are within 1e-4 relative in t (ties) or a Newton solve sits on its
convergence boundary; t agrees to the FP32 Newton tolerance scale.  The
persistent kernel performs the same per-ray arithmetic as the wide kernel,
so the two are compared bit for bit."""
import numpy as np
import pytest

from helpers import c2_small
from rtnurbs import scenes

pytestmark = pytest.mark.gpu

CFG = dict(max_iter=8, thr_res=1e-5, thr_det=1e-5)

CASES = [
    ("c1", lambda: scenes.config1(0)),
    ("mixed", lambda: scenes.Scene(c2_small(8, 2), scenes.c2_camera(240, 135))),
    ("torus", lambda: scenes.Scene([scenes.torus()], scenes.reference_camera(320, 180))),
]


def _run(rtl, scene, rays, **kw):
    h, tot = rtl.trace(scene, rays, rtl.params(**CFG, **kw))
    return h.get(), tot


def same_but_ties(a, b):
    """The packet form against the one-ray form: per (ray, patch) the same
    arithmetic, and a ray tests every patch whose box it enters before its
    closest hit, so the closest hits agree bit for bit except where two
    patches give the same t (the order of the tests then picks one)."""
    d = a["hit"] != b["hit"]
    assert d.sum() <= max(2, 1e-4 * len(d)), d.sum()
    assert np.all((a["hit"][d] >= 0) & (b["hit"][d] >= 0))
    assert np.all(np.abs(a["t"][d] - b["t"][d]) <= 1e-6 * np.abs(b["t"][d]))
    m = ~d & (b["hit"] >= 0)
    for k in ("t", "u", "v"):
        assert np.array_equal(a[k][m].view(np.uint32), b[k][m].view(np.uint32)), k


@pytest.mark.parametrize("name,mk", CASES, ids=[c[0] for c in CASES])
def test_wide_forms_agree(rtl, name, mk):
    sc = mk()
    print(f"case {name}", flush=True)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera)
    ex, _ = _run(rtl, scene, rays, mode=rtl.MODE_EXACT)
    wd, tw = _run(rtl, scene, rays, mode=rtl.MODE_WIDE, count=1)
    assert tw["overflow"] == 0
    ps, tp = _run(rtl, scene, rays, mode=rtl.MODE_PERSISTENT, count=1)
    # persistent == wide, bit for bit (same per-ray arithmetic)
    assert np.array_equal(ps["hit"], wd["hit"])
    m = wd["hit"] >= 0
    for k in ("t", "u", "v"):
        assert np.array_equal(ps[k][m].view(np.uint32), wd[k][m].view(np.uint32)), k
    assert tp["newton_evals"] == tw["newton_evals"] and tp["patch_tests"] == tw["patch_tests"]
    # hull rejection (per-plane sign test, projected hull) drops only pairs whose
    # residual cannot reach the tolerance, and each Newton call restarts its RNG
    # from the ray's seed, so every other test and hence the result is unchanged
    # bit for bit
    culled = {}
    for mode, hull in ((rtl.MODE_WIDE, 1), (rtl.MODE_WIDE, 2), (rtl.MODE_PERSISTENT, 2)):
        h2, t2 = _run(rtl, scene, rays, mode=mode, hull_cull=hull, count=1)
        culled[hull] = t2["hull_culled"]
        assert t2["hull_culled"] > 0 and t2["patch_tests"] == tw["patch_tests"]
        assert t2["newton_calls"] + t2["hull_culled"] == tw["newton_calls"]
        assert np.array_equal(h2["hit"], wd["hit"])
        for k in ("t", "u", "v"):
            assert np.array_equal(h2[k][m].view(np.uint32), wd[k][m].view(np.uint32)), k
    assert culled[2] > culled[1]  # the diagonal directions reject more
    # packet form (one traversal per wave) = one-ray form but for ties, with and without rejection
    pk, tk = _run(rtl, scene, rays, mode=rtl.MODE_PACKET, count=1)
    # the wave's test order (nearest first for its first lane) can test a patch
    # before a nearer hit of this lane would have pruned it: a few % more solves
    assert tk["overflow"] == 0 and abs(tk["newton_evals"] - tw["newton_evals"]) <= 0.05 * tw["newton_evals"]
    same_but_ties(pk, wd)
    pk2, _ = _run(rtl, scene, rays, mode=rtl.MODE_PACKET, hull_cull=2)
    same_but_ties(pk2, wd)
    # wide vs exact: hit patches agree except near ties / boundary solves
    both = (ex["hit"] >= 0) & (wd["hit"] >= 0)
    assert np.mean(ex["hit"] == wd["hit"]) > 0.995
    same = both & (ex["hit"] == wd["hit"])
    rel = np.abs(wd["t"][same] - ex["t"][same]) / np.abs(ex["t"][same])
    assert np.quantile(rel, 0.99) < 1e-5 and rel.max() < 1e-3


def test_mixed_precision_improves_t(rtl):
    """On config 1 the FP64 step lowers the hit-t error of the corrected rays
    against the FP64 reference (threshold 3: primary rays of config 1 have
    cond(J) <= 36, median 2)."""
    sc = scenes.config1(0)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera)
    ref = rtl.trace_f64(scene, rays).get()
    a, _ = _run(rtl, scene, rays, mode=rtl.MODE_WIDE, hull_cull=1)
    b, tb = _run(rtl, scene, rays, mode=rtl.MODE_WIDE, hull_cull=1, precision=rtl.PREC_MIXED, cond_thr=3.0,
                 count=1)
    assert tb["fp64_steps"] > 0
    same = (a["hit"] >= 0) & (a["hit"] == ref["hit"]) & (b["hit"] == ref["hit"])
    ea = np.abs(a["t"][same] - ref["t"][same]) / ref["t"][same]
    eb = np.abs(b["t"][same] - ref["t"][same]) / ref["t"][same]
    assert eb.mean() <= ea.mean()


@pytest.mark.parametrize("name,mk", CASES, ids=[c[0] for c in CASES])
def test_stall_exit(rtl, name, mk):
    """Abandoning solves stuck against the patch boundary saves evaluations
    and keeps the hits: against the FP64 reference the patch mismatches stay
    at the level of the full iteration count."""
    sc = mk()
    print(f"case {name}", flush=True)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera)
    ref = rtl.trace_f64(scene, rays).get()
    full, tf = _run(rtl, scene, rays, mode=rtl.MODE_WIDE, hull_cull=2, count=1)
    st, ts = _run(rtl, scene, rays, mode=rtl.MODE_WIDE, hull_cull=2, stall_exit=1, count=1)
    ps, tp = _run(rtl, scene, rays, mode=rtl.MODE_PERSISTENT, hull_cull=2, stall_exit=1, count=1)
    assert ts["newton_calls"] == tf["newton_calls"]
    assert ts["newton_evals"] < 0.95 * tf["newton_evals"]
    assert np.array_equal(ps["hit"], st["hit"]) and tp["newton_evals"] == ts["newton_evals"]
    pk, _ = _run(rtl, scene, rays, mode=rtl.MODE_PACKET, hull_cull=2, stall_exit=1)
    same_but_ties(pk, st)
    assert np.mean(st["hit"] == full["hit"]) > 0.9995
    # every abandoned solve was a miss of the full count or a hit another patch also provides
    assert (st["hit"] >= 0).sum() >= 0.9995 * (full["hit"] >= 0).sum()
    both = (st["hit"] >= 0) & (st["hit"] == full["hit"])
    assert np.array_equal(st["t"][both].view(np.uint32), full["t"][both].view(np.uint32))
    mis_f = np.sum((full["hit"] != ref["hit"]))
    mis_s = np.sum((st["hit"] != ref["hit"]))
    assert mis_s <= mis_f + max(5, 0.01 * mis_f), (mis_s, mis_f)
