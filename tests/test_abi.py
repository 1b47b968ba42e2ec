"""The C-ABI library loads without a GPU, exports every entry point the
headers declare, and the ctypes mirrors match the C struct layouts that an
embedding solver compiles against."""
import ctypes as C
import os
import re
import subprocess

import pytest

from contour3d_amd import _abi, api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("contour3d.h", "contour3d_amd.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"extern\s+[\w\s\*]+?\b(contour3d_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = api.lib()
    decl = declared_functions()
    assert "contour3d_execute" in decl
    assert set(decl) == set(api.EXPORTS)
    for name in decl:
        assert hasattr(L, name), name


def test_export_table_via_nm():
    out = subprocess.run(["nm", "-D", "--defined-only", api.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    for name in declared_functions():
        assert name in syms, name


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "contour3d_amd.h"
#define P(T, m) printf(#T "." #m " %zu\n", offsetof(T, m))
int main(void) {
  printf("vector %zu\ncolor %zu\ncontour %zu\nline %zu\nimage %zu\n", sizeof(contour3d_vector_t),
         sizeof(contour3d_color_t), sizeof(contour3d_contour_obj_t), sizeof(contour3d_line_obj_t),
         sizeof(contour3d_amd_image_t));
  P(contour3d_amd_image_t, camera_position); P(contour3d_amd_image_t, light_direction);
  P(contour3d_amd_image_t, screen_sizes); P(contour3d_amd_image_t, screen_center);
  P(contour3d_amd_image_t, screen_local); P(contour3d_amd_image_t, bg_color);
  P(contour3d_amd_image_t, num_contours); P(contour3d_amd_image_t, contour_objs);
  P(contour3d_amd_image_t, num_lines); P(contour3d_amd_image_t, line_objs);
  P(contour3d_amd_image_t, rgb); P(contour3d_amd_image_t, depth);
  printf("sopts %zu\nsstats %zu\n", sizeof(contour3d_amd_server_opts_t), sizeof(contour3d_amd_server_stats_t));
  P(contour3d_amd_server_opts_t, hbm_budget); P(contour3d_amd_server_opts_t, max_wait_ms);
  P(contour3d_amd_server_opts_t, snapshot_fields);
  P(contour3d_amd_server_stats_t, mean_batch); P(contour3d_amd_server_stats_t, p99_ms);
  P(contour3d_amd_server_stats_t, frames_per_s); P(contour3d_amd_server_stats_t, peak_admitted_bytes);
  P(contour3d_contour_obj_t, glsizes); P(contour3d_contour_obj_t, grids);
  P(contour3d_contour_obj_t, converter); P(contour3d_contour_obj_t, threshold);
  P(contour3d_contour_obj_t, color); P(contour3d_contour_obj_t, array);
  P(contour3d_line_obj_t, nitems); P(contour3d_line_obj_t, converter);
  P(contour3d_line_obj_t, color); P(contour3d_line_obj_t, width);
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "contour3d_amd", "csrc", "single_rank"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    got = {k: int(v) for k, v in got.items()}
    assert got["vector"] == C.sizeof(_abi.Vector)
    assert got["color"] == C.sizeof(_abi.Color)
    assert got["contour"] == C.sizeof(_abi.ContourObj)
    assert got["line"] == C.sizeof(_abi.LineObj)
    for f in ("glsizes", "grids", "converter", "threshold", "color", "array"):
        assert got[f"contour3d_contour_obj_t.{f}"] == getattr(_abi.ContourObj, f).offset, f
    for f in ("nitems", "converter", "color", "width"):
        assert got[f"contour3d_line_obj_t.{f}"] == getattr(_abi.LineObj, f).offset, f
    assert got["sopts"] == C.sizeof(_abi.ServerOpts)
    assert got["sstats"] == C.sizeof(_abi.ServerStats)
    for f in ("hbm_budget", "max_wait_ms", "snapshot_fields"):
        assert got[f"contour3d_amd_server_opts_t.{f}"] == getattr(_abi.ServerOpts, f).offset, f
    for f in ("mean_batch", "p99_ms", "frames_per_s", "peak_admitted_bytes"):
        assert got[f"contour3d_amd_server_stats_t.{f}"] == getattr(_abi.ServerStats, f).offset, f
    assert got["image"] == C.sizeof(_abi.Image)
    for f, _ in _abi.Image._fields_:
        assert got[f"contour3d_amd_image_t.{f}"] == getattr(_abi.Image, f).offset, f


def test_entry_points_reject_bad_arguments_without_touching_memory():
    L = api.lib()
    # NULL output pointer for a mesh build is refused before any device work
    assert L.contour3d_amd_mesh_build(None, 0, 0, None) == 1
    assert L.contour3d_amd_set_precision(7) == 1
    # the serving entry points refuse work when no server runs
    import ctypes as C
    t = C.c_uint64()
    assert L.contour3d_amd_server_submit(None, 0, C.byref(t)) == 1
    assert L.contour3d_amd_server_wait(1, None) == 1
    assert L.contour3d_amd_server_stop() == 0
