"""The CPU oracle against the reference's own outputs (tests/golden/*.npz,
written by tests/golden/make_golden.py from the reference headers compiled in
place).  Every comparison is bit-exact (FP32 bit patterns, integer indices)."""
import os

import numpy as np
import pytest

from helpers import assert_bits_equal
from rtnurbs import scenes

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["c1_seed0", "c1_seed7", "torus", "sphere"]


def load(name):
    return dict(np.load(os.path.join(GOLD, f"{name}.npz")))


def surface_and_camera(g):
    n, m, p, q = (int(x) for x in g["nmpq"])
    s = scenes.Nurbs(n, m, p, q, g["U"], g["V"], g["Pw"])
    c = g["cam"]
    cam = scenes.Camera(tuple(c[0:3]), tuple(c[3:6]), tuple(c[6:9]), float(c[9]), int(c[10]), int(c[11]),
                        float(c[12]), float(c[13]))
    return s, cam


@pytest.mark.parametrize("name", CASES)
def test_camera_and_rays(oracle, name):
    g = load(name)
    s, cam = surface_and_camera(g)
    wfc, cfr = oracle.camera(cam)
    assert_bits_equal(wfc, g["wfc"], "world_from_camera")
    assert_bits_equal(cfr, g["cfr"], "camera_from_raster")
    rays = oracle.gen_rays(wfc, cfr, cam.W, cam.H)
    for k in range(6):
        assert_bits_equal(rays[k][g["sel"]], g["rays"][k], f"ray component {k}")


@pytest.mark.parametrize("name", CASES)
def test_decomposition_bounds_bvh(oracle, name):
    g = load(name)
    s, cam = surface_and_camera(g)
    d = oracle.decompose(s)
    assert (d["nbU"], d["nbV"]) == tuple(g["nbUV"])
    assert d["nbU"] * d["nbV"] == scenes.patch_count(s)
    assert_bits_equal(d["cp"], g["cp"], "patch nets")
    assert np.array_equal(d["uv"], g["uv"])
    b = np.stack([oracle.patch_bounds(n, s.p, s.q) for n in d["cp"]])
    assert np.array_equal(b, g["bounds"])
    nodes, perm, used = oracle.build_bvh(b)
    assert np.array_equal(perm, g["perm"])
    assert np.array_equal(nodes.view(np.uint32).reshape(-1, 8)[:, 6:], g["nodes"][:, 6:])
    assert np.array_equal(nodes["b"], g["nodes"][:, :6].view(np.float32))


@pytest.mark.parametrize("name", CASES)
def test_trace_reference_params(oracle, name):
    """Per-pixel trace with the reference's constants (7 iterations, residual
    1e-4, det 1e-5; seed = pixel index) equals the reference loop."""
    g = load(name)
    s, cam = surface_and_camera(g)
    P = oracle.c1_pipeline(scenes.Scene([s], cam))
    r = oracle.trace_f32(P["nodes"], P["ps"], P["rays"], max_iter=7, thr_res=1e-4, thr_det=1e-5)
    sel = g["sel"]
    assert np.array_equal(r["hit"][sel], g["hit"])
    m = g["hit"] >= 0
    for k in ("t", "u", "v", "dpdu", "dpdv"):
        assert_bits_equal(r[k][sel][m], g[k][m], k)
    assert np.array_equal(r["counters"][sel, 0], g["newton_calls"])


def test_refinement(oracle):
    g = load("refine_torus_C4")
    n, m, p, q = (int(x) for x in g["nmpq"])
    s = scenes.Nurbs(n, m, p, q, g["U"], g["V"], g["Pw"])
    mid = oracle.refine(s, float(g["C"]), None)
    assert mid.n == g["nm_out"][0]
    assert np.array_equal(mid.U, g["U_out"])
    assert_bits_equal(mid.Pw, g["Pw_mid"], "refined U net")
    out = oracle.refine(mid, None, float(g["C"]))
    assert out.m == g["nm_out"][1]
    assert np.array_equal(out.V, g["V_out"])
    assert_bits_equal(out.Pw, g["Pw_out"], "refined V net")


def test_unit_known_answers(oracle):
    g = load("units")
    L = oracle.lib()
    for x, h in zip(g["pcg_in"], g["pcg_out"]):
        assert L.orc_pcg_hash(int(x)) == int(h)
    net = g["net"]
    for i in range(len(g["origins"])):
        o, d = g["origins"][i], g["dirs"][i]
        seed = np.zeros(2, np.float32)
        L.orc_bilinear_seed_f32(net, 3, 3, o, d, seed)
        assert_bits_equal(seed, g["seeds"][i], f"bilinear seed {i}")
        out = np.zeros(11, np.float32)
        ev = L.orc_newton_f32(net, 3, 3, o, d, seed[0], seed[1], 1e-5, 1e-4, i, 7, out)
        assert (ev > 0) == bool(g["newton_hit"][i])
        if ev > 0:
            assert_bits_equal(out, g["newton"][i], f"newton {i}")
        pl = np.zeros(8, np.float32)
        L.orc_ray_planes_f32(o, d, pl)
        assert_bits_equal(pl, g["planes"][i], f"ray planes {i}")
        e = np.zeros(12, np.float32)
        L.orc_patch_eval_f32(net, 3, 3, g["uvs"][i, 0], g["uvs"][i, 1], e)
        assert_bits_equal(e, g["evals"][i], f"patch eval {i}")


@pytest.mark.skipif(not __import__("pyoracle").ref_available(), reason="reference checkout not built here")
def test_oracle_vs_live_reference_more_seeds(oracle):
    """Where the reference headers are compiled (oracle/_ref), more seeds and
    a refined torus against the live reference loop."""
    import ctypes as C
    R = oracle.ref()
    for seed in (11, 12, 13):
        sc = scenes.config1(seed, 128, 128)
        s = sc.surfaces[0]
        P = oracle.c1_pipeline(sc)
        n = sc.camera.W * sc.camera.H
        hit = np.full(n, -1, np.int32)
        t = np.zeros(n, np.float32); u = np.zeros(n, np.float32); v = np.zeros(n, np.float32)
        du = np.zeros((n, 3), np.float32); dv = np.zeros((n, 3), np.float32); nc = np.zeros(n, np.int32)
        R.ref_trace(P["nodes"].ctypes.data, oracle.f32(P["nets"]), len(P["nets"]), s.p, s.q, None, 0.0, 0.0, n,
                    *P["rays"], hit, t, u, v, du, dv, nc)
        r = oracle.trace_f32(P["nodes"], P["ps"], P["rays"])
        assert np.array_equal(r["hit"], hit)
        m = hit >= 0
        assert_bits_equal(r["t"][m], t[m], "t")
        assert np.array_equal(r["counters"][:, 0], nc)
    _ = C
