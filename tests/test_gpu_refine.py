"""Curvature refinement on the GPU (rtn_scene_refine) against the host
implementation (rtn_refine, itself bit-exact with the oracle and the
reference's golden refinement): the refined knot vectors and control nets of
the torus, the sphere and the 64 surfaces of the 10k-patch scene agree bit
for bit, read back through a scene file; the 10k-patch scene then decomposes
into its 10,018 patches and builds the same BVH."""
import numpy as np
import pytest

from rtnurbs import scenes
from test_scene_io import _read_layout

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


@pytest.mark.parametrize("C_u,C_v", [(scenes.C2_REFINE, scenes.C2_REFINE), (4.0, 4.0), (2.0, 0.0), (0.0, 8.0)])
def test_gpu_refinement_matches_host(rtl, tmp_path, C_u, C_v):
    print(f"case refinement C_u={C_u} C_v={C_v} on torus, sphere and 64 random surfaces", flush=True)
    surfs = [scenes.torus(), scenes.sphere()] + scenes.config2_base(0)
    s = rtl.Scene(surfs).refine(C_u, C_v)
    f = tmp_path / "refined.rtns"
    s.save(f)
    _, _, got, _ = _read_layout(f)
    assert len(got) == len(surfs)
    grew = 0
    for (n, p, m, q, U, V, P), src in zip(got, surfs):
        ref = rtl.refine(src, C_u if C_u > 0 else None, C_v if C_v > 0 else None)
        assert (n, p, m, q) == (ref.n, ref.p, ref.m, ref.q)
        assert _same(U, ref.U) and _same(V, ref.V) and _same(P, ref.Pw)
        grew += (n, m) != (src.n, src.m)
    assert grew > len(surfs) // 2


def test_gpu_refined_scene_builds_the_10k_patch_scene(rtl):
    print("case 10k-patch scene refined on the GPU", flush=True)
    host = rtl.Scene(scenes.config2(0).surfaces).build()
    dev = rtl.Scene(scenes.config2_base(0)).refine(scenes.C2_REFINE).build()
    assert dev.stats()["n_patches"] == scenes.C2_PATCHES == host.stats()["n_patches"]
    a, b = host.download(), dev.download()
    for k in ("cp", "off", "pq", "nodes"):
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


def test_refine_rejects_bad_calls(rtl):
    s = rtl.Scene([scenes.torus()])
    with pytest.raises(rtl.RtnError):
        s.refine(-1.0)
    s.build()
    with pytest.raises(rtl.RtnError):
        s.refine(1.0)
