"""GPU FP64 reference trace (rtn_trace_f64) against the CPU oracle's FP64
instantiation of the same traversal (same seed, RNG, caps and pruning slack).

FP64 add/mul/div/sqrt are correctly rounded on both sides and the kernel is
compiled without contraction, so hit patches must agree exactly; t, u, v and
derivatives are compared at 1e-12 relative (FP64 library sqrt on the two
sides is allowed to differ in the last bit)."""
import numpy as np
import pytest

from helpers import c2_small
from rtnurbs import scenes

pytestmark = pytest.mark.gpu


def _oracle_scene(oracle, surfaces):
    nets, degs = [], []
    for s in surfaces:
        d = oracle.decompose(s)
        nets += list(d["cp"])
        degs += [(s.p, s.q)] * len(d["cp"])
    bnd = np.stack([oracle.patch_bounds(n, p, q) for n, (p, q) in zip(nets, degs)])
    nodes, perm, used = oracle.build_bvh(bnd)
    ps = oracle.PatchSet([nets[i] for i in perm], [degs[i] for i in perm])
    return nodes[:used], ps


CASES = [
    ("c1", lambda: scenes.config1(3, W=128, H=128)),
    ("mixed", lambda: scenes.Scene(c2_small(6, 5), scenes.c2_camera(160, 90))),
]


@pytest.mark.parametrize("grid", [0, 4], ids=["plain_seed", "closest_root"])
@pytest.mark.parametrize("name,mk", CASES, ids=[c[0] for c in CASES])
def test_trace_f64_matches_oracle(oracle, rtl, name, mk, grid):
    """grid 0: rtn_trace_f64; grid 4: the closest-root reference
    (rtn_trace_f64_roots), 16 extra Newton starts per tested patch."""
    sc = mk()
    print(f"case {name} grid {grid}", flush=True)
    nodes, ps = _oracle_scene(oracle, sc.surfaces)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera)
    kw = dict(max_iter=32, thr_det=1e-12, thr_res=1e-10, grid=grid)
    g = rtl.trace_f64(scene, rays, **kw).get()
    r = oracle.trace_f64(nodes, ps, rays.get(), prune_slack=1.0 + 4e-5, **kw)
    assert np.array_equal(g["hit"], r["hit"])
    m = r["hit"] >= 0
    assert m.sum() > 0
    for k in ("t", "u", "v", "dpdu", "dpdv"):
        np.testing.assert_allclose(g[k][m], r[k][m], rtol=1e-12, atol=1e-13, err_msg=k)
    fin = np.isfinite(r["t2"])
    assert np.array_equal(fin, np.isfinite(g["t2"]))
    np.testing.assert_allclose(g["t2"][fin], r["t2"][fin], rtol=1e-12)


def test_fp32_plain_error_vs_fp64(rtl):
    """The FP32 plain method against the FP64 reference on config 1: hit-t
    error of converged pixels on the same patch stays within the FP32 Newton
    tolerance scale (1e-5 absolute residual over t ~ 1)."""
    sc = scenes.config1(0)
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera)
    g = rtl.trace(scene, rays, rtl.params(mode=rtl.MODE_EXACT, max_iter=8, thr_res=1e-5, thr_det=1e-5))[0].get()
    r = rtl.trace_f64(scene, rays).get()
    same = (g["hit"] >= 0) & (g["hit"] == r["hit"])
    assert same.mean() > 0.3
    te = np.abs(g["t"][same] - r["t"][same]) / r["t"][same]
    assert np.median(te) < 1e-5


def test_closest_root_reference_is_never_farther(rtl):
    """On a mixed-degree scene the closest-root reference hits wherever the
    plain-seed FP64 reference hits, never farther, and its extra starts find
    nearer roots on some pixels (a size-independent property: the plain
    seed's root is one of the candidates each tested patch compares)."""
    sc = scenes.Scene(c2_small(11, 8), scenes.c2_camera(320, 180))
    scene = rtl.Scene(sc.surfaces).build()
    rays = rtl.generate_rays(sc.camera, spp=2, jitter=True, seed=3)
    a = rtl.trace_f64(scene, rays).get()
    b = rtl.trace_f64(scene, rays, grid=4).get()
    ha, hb = a["hit"] >= 0, b["hit"] >= 0
    assert ha.sum() > 1000
    assert np.all(hb[ha])
    assert np.all(b["t"][ha] <= a["t"][ha] * (1 + 1e-12))
    nearer = ha & (b["t"] < a["t"] * (1 - 1e-5))
    assert (nearer.sum() + (hb & ~ha).sum()) > 0
