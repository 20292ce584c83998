"""The ctypes mirrors of the C ABI's structs (rtnurbs.lib) have the C layout:
a small C program compiled with gcc against include/rtnurbs.h prints each
struct's size and every field's offset, and they must equal ctypes' own.
A field added on one side only (a new parameter, a renamed reserved word)
fails here, before any call passes a shifted struct to the library."""
import ctypes
import os
import shutil
import subprocess

import pytest

from rtnurbs import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MIRRORS = {
    "rtn_ray_queue": lib.RayQueue,
    "rtn_trace_params": lib.TraceParams,
    "rtn_mlp_train_params": lib.MlpTrainParams,
    "rtn_build_params": lib.BuildParams,
    "rtn_frame_params": lib.FrameParams,
    "rtn_frame_status_t": lib.FrameStatus,
    "rtn_tune_entry": lib.TuneEntry,
    "rtn_hits": lib.Hits,
    "rtn_trace_totals": lib.TraceTotals,
    "rtn_camera_desc": lib.CameraDesc,
    "rtn_scene_stats": lib.SceneStats,
}


def c_layout(tmp_path):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "rtnurbs.h"', "int main(void) {"]
    for cname, py in MIRRORS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {tuple(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_structs_match_the_c_header(tmp_path):
    c = c_layout(tmp_path)
    bad = []
    for cname, py in MIRRORS.items():
        if c[(cname, "size")] != ctypes.sizeof(py):
            bad.append(f"{cname}: sizeof {c[(cname, 'size')]} vs ctypes {ctypes.sizeof(py)}")
        for f, _ in py._fields_:
            if c[(cname, f)] != getattr(py, f).offset:
                bad.append(f"{cname}.{f}: offset {c[(cname, f)]} vs ctypes {getattr(py, f).offset}")
    assert not bad, bad

