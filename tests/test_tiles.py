"""CPU: the tessellation tiles and work items (contour3d_amd/csrc/c3d_tiles.h,
the header the device kernels use).  tests/tiles/check_tiles.cpp walks every
work item, plane and tile cell of a lattice the way the count, emit and
fused kernels do, and checks that every cell is visited exactly once, that
tiles are consecutive runs of at most 2048 cells numbered in cell order
(so their triangle counts scan into the reference's serial order, k then j
then i, contour/main.c:49-112), that the cell-to-tile lookup inverts them,
and that every corner read stays inside the field.  Lattices cover rows that
fit a tile many times, once, and not at all (split into segments), the
row and plane knobs, and degenerate one-cell axes."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

LATTICES = [
    (1, 1, 1, 0, 0), (1, 4, 6, 0, 0), (63, 63, 63, 0, 0), (511, 511, 511, 0, 0), (2047, 3, 5, 0, 0),
    (2048, 2, 2, 0, 3), (2049, 3, 2, 0, 1), (4100, 2, 3, 0, 2), (5000, 1, 1, 0, 0), (100, 1000, 3, 5, 2),
    (383, 97, 250, 0, 16), (1023, 1023, 4, 0, 8), (255, 127, 383, 3, 100), (6, 6, 6, 2, 1),
]


def test_tiles_cover_every_cell_once_in_serial_order(tmp_path):
    exe = tmp_path / "check_tiles"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "contour3d_amd", "csrc"),
                    os.path.join(ROOT, "tests", "tiles", "check_tiles.cpp"), "-o", str(exe)], check=True)
    args = [str(v) for lat in LATTICES for v in lat]
    r = subprocess.run([str(exe)] + args, capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("ok ") == len(LATTICES) + 1
