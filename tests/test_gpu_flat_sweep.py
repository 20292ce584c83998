"""The refinement x flat-skip sweep (benchmark.cpp:122-251) through the library:
one CSV row per surface x C x eps, the header the reference prints, flat-skip
taken more often as eps grows, and pixel errors that stay within eps-sized
bounds of the eps = 0 ground truth."""
import io
import math

import numpy as np
import pytest

from rtnurbs import flat_sweep

pytestmark = pytest.mark.gpu


def test_sweep_table(rtl):
    print("case flat sweep 320x180, C in (0, 4), eps in (0.5, 8)", flush=True)
    out = io.StringIO()
    cs, epss = (0, 4), (0.5, 8.0)
    rows = flat_sweep.sweep(res=(320, 180), reps=1, cs=cs, epss=epss, out=out)
    lines = out.getvalue().strip().split("\n")
    assert lines[0] == flat_sweep.HEADER
    assert len(lines) == 1 + len(rows) == 1 + 2 * len(cs) * len(epss)
    assert all(len(l.split(",")) == len(flat_sweep.HEADER.split(",")) for l in lines)
    by = {(r["surface"], r["C"], r["eps_px"]): r for r in rows}
    for surf in ("torus", "sphere"):
        # refinement multiplies the patch count
        assert by[(surf, 4, 0.5)]["patches"] > by[(surf, 0, 0.5)]["patches"]
        for c in cs:
            lo, hi = by[(surf, c, 0.5)], by[(surf, c, 8.0)]
            assert lo["patches"] == hi["patches"]
            assert 0.0 < lo["hit_pct"] < 100.0
            # a larger threshold skips at least as many Newton solves
            assert hi["skip_pct"] >= lo["skip_pct"]
            if c > 0:
                assert hi["skip_pct"] > 0.0
            for r in (lo, hi):
                assert all(math.isfinite(r[k]) for k in ("time_ms", "max_pix_err", "mean_pix_err", "max_norm_deg"))
                assert r["time_ms"] > 0.0 and r["newton_per_ray"] >= 0.0
                assert r["mean_pix_err"] <= r["max_pix_err"]
            # a skipped patch is accepted only when its residual is below eps pixels
            # (intersect.cu:104-123); the mean error therefore stays below eps
            assert lo["mean_pix_err"] < 0.5
            assert hi["mean_pix_err"] < 8.0
            assert abs(hi["hit_pct"] - lo["hit_pct"]) < 5.0
