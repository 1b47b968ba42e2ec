"""CPU: the raster's narrowed pixel boxes (c3d::tight_box in
contour3d_amd/csrc/c3d_geom.h) hold every pixel the coverage test accepts.

tests/raster_box/check_box.cpp brute-forces the reference's box for random,
thin, collinear, repeated-vertex, pixel-centre-aligned and huge-coordinate
triangles in FP64 and FP32, with the same header the device kernels use.  A
copy of the header with the margin removed must fail the same check, which
shows the check can see a box that is too narrow."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "contour3d_amd", "csrc")
SRC = os.path.join(ROOT, "tests", "raster_box", "check_box.cpp")


def build(tmp_path, include):
    exe = tmp_path / "check_box"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", str(include), SRC, "-o", str(exe)],
                   check=True)
    return exe


def test_narrowed_boxes_hold_every_accepted_pixel(tmp_path):
    exe = build(tmp_path, CSRC)
    for seed in (1, 2):
        r = subprocess.run([str(exe), str(seed), "20000"], capture_output=True, text=True)
        print(r.stdout)
        assert r.returncode == 0, r.stdout
        # pixel-sized triangles: the narrowing must pay (about 16 -> 3.5 candidates)
        m = re.search(r"f64 pixel-sized\s+candidates per triangle\s+([\d.]+) ->\s+([\d.]+)", r.stdout)
        assert float(m.group(2)) < 0.35 * float(m.group(1))


def test_check_sees_a_missing_margin(tmp_path):
    hdr = open(os.path.join(CSRC, "c3d_geom.h")).read()
    line = ("const R M = (R(5.8e-8) * emax + R(12) * g * D * e2 * fabs(P.ainv)) * R(1.00000095367431640625) + "
            "R(16) * eps * qmax;")
    assert line in hdr
    mut = tmp_path / "mut"
    mut.mkdir()
    (mut / "c3d_geom.h").write_text(hdr.replace(line, "const R M = R(0);"))
    exe = build(tmp_path, mut)
    r = subprocess.run([str(exe), "1", "20000"], capture_output=True, text=True)
    assert r.returncode == 1
    assert "MISS f64" in r.stdout
