"""The C ABI's layout twins against the reference's own structs.

A small program compiled with g++ includes the reference's headers where
they lie (Frostbite16/rt-nurbs src/: core/bvh.h, core/camera.h,
math/bezier_patch.h, math/nurbs.h, gpu/soa.h, gpu/interact.h) beside
include/rtnurbs.h and prints sizeof and offsetof of every field of each pair:
BVHNode / rtn_bvh_node, bezierPatch / rtn_bezier_patch, surfaceInteraction /
rtn_surface_interaction, SoAQueue::ray_queue / rtn_ray_queue,
ParametricSurf::nurbsSurf / rtn_nurbs_surf, camera_config / rtn_camera_config.
They must be equal, so reference-side buffers pass to the library unchanged.
(include/rtnurbs.hpp ties its reference-named types to the same rtn_* twins
with static_asserts, so this also pins the C++ layer.)  Runs only where the
reference checkout is present."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("RTN_REFERENCE_SRC", "/root/reference/src")

PAIRS = [  # (reference type, twin, [(reference field, twin field)])
    ("BVHNode", "rtn_bvh_node", [("node_bound.p_min_x", "p_min_x"), ("node_bound.p_max_z", "p_max_z"),
                                 ("left_first", "left_first"), ("prim_count", "prim_count")]),
    ("bezierPatch", "rtn_bezier_patch", [("control_point", "control_point"), ("u0", "u0"), ("u1", "u1"),
                                         ("v0", "v0"), ("v1", "v1"), ("p", "p"), ("q", "q"),
                                         ("heap_pointer", "heap_pointer")]),
    ("surfaceInteraction", "rtn_surface_interaction", [("t_hit", "t_hit"), ("wo", "wo"), ("uv", "uv"),
                                                       ("dpdu", "dpdu"), ("dpdv", "dpdv"), ("dndu", "dndu"),
                                                       ("dndv", "dndv")]),
    ("SoAQueue::ray_queue", "rtn_ray_queue", [(f, f) for f in ("origin_x", "origin_y", "origin_z", "direction_x",
                                                                 "direction_y", "direction_z", "idx_interact")]),
    ("ParametricSurf::nurbsSurf", "rtn_nurbs_surf", [(f, f) for f in ("u_knots", "v_knots", "control_point", "n",
                                                                      "m", "p", "q")]),
    ("camera_config", "rtn_camera_config",
     [(f, f) for f in ("world_from_camera", "screen_from_camera", "NDC_from_screen", "raster_from_NDC",
                       "raster_from_screen", "screen_from_raster", "camera_from_raster", "screen_resolution_x",
                       "screen_resolution_y")] + [("screen_window", "screen_window")]),
]


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference checkout is not present")
@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_layout_twins_match_the_reference(tmp_path):
    lines = ['#include <cstddef>', '#include <cstdio>', '#include "core/bvh.h"', '#include "core/camera.h"',
             '#include "math/bezier_patch.h"', '#include "math/nurbs.h"', '#include "gpu/soa.h"',
             '#include "gpu/interact.h"', '#include "rtnurbs.h"', "int main() {"]
    for ref, twin, fields in PAIRS:
        lines.append(f'  std::printf("{ref} size %zu %zu\\n", sizeof({ref}), sizeof({twin}));')
        for rf, tf in fields:
            lines.append(f'  std::printf("{ref}.{rf} %zu %zu\\n", offsetof({ref}, {rf}), offsetof({twin}, {tf}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["g++", "-std=c++17", "-w", "-I", REF, "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    assert len(out) == sum(1 + len(f) for _, _, f in PAIRS)
    bad = [l for l in out if l.split()[-2] != l.split()[-1]]
    assert not bad, bad
