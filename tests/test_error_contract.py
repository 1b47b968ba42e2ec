"""CPU: parts of the error contract that need no GPU.

contour3d_execute keeps the reference's 0/1 return and its
"[CONTOUR3D ERROR] " stderr line (src/contour3d/logger.c:5-18).  Under an MPI
launch of several ranks each rank's array is a pencil (extend.c:158-181),
which this whole-field library cannot render, so it returns 1 before it
touches the device.  The id limit of a frame (triangles, then line
fragments, in 32 bits) is checked with the same header the device uses."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "contour3d_amd", "csrc")

CALL = r"""
import sys
sys.path.insert(0, %r)
from contour3d_amd import api, fields
import numpy as np
v = fields.example_view(16, 12)
try:
    api.execute(v["camera"], (0, 0, 0), v["light"], v["size"], v["center"], v["local"], v["bg"],
                [dict(field=np.zeros((4, 4, 4)), threshold=0.5, color=(1, 2, 3))], fname="/nonexistent/x.ppm")
except RuntimeError:
    print("RC1")
""" % ROOT


def run_with_env(env_extra):
    env = dict(os.environ)
    for k in ("OMPI_COMM_WORLD_SIZE", "PMI_SIZE", "PMIX_SIZE", "MV2_COMM_WORLD_SIZE", "CONTOUR3D_AMD_WHOLE_FIELD"):
        env.pop(k, None)
    env.update(env_extra)
    env["HIP_VISIBLE_DEVICES"] = "-1"  # the refusal must come before any device work
    return subprocess.run([sys.executable, "-c", CALL], capture_output=True, text=True, env=env, timeout=300)


@pytest.mark.parametrize("var,n", [("OMPI_COMM_WORLD_SIZE", 2), ("PMI_SIZE", 4), ("MV2_COMM_WORLD_SIZE", 8)])
def test_multi_rank_launch_returns_one_with_the_prefix(var, n):
    r = run_with_env({var: str(n)})
    assert "RC1" in r.stdout, r.stdout + r.stderr
    assert f"[CONTOUR3D ERROR] {n} MPI ranks" in r.stderr, r.stderr


def test_single_rank_and_whole_field_declaration_pass_the_guard():
    # one rank, or ranks that declare whole fields, get past the guard (and
    # then fail later here only because this process has no device)
    for env in ({"OMPI_COMM_WORLD_SIZE": "1"}, {"PMI_SIZE": "4", "CONTOUR3D_AMD_WHOLE_FIELD": "1"}):
        r = run_with_env(env)
        assert "MPI ranks" not in r.stderr, r.stderr


def test_frame_id_limit(tmp_path):
    exe = tmp_path / "check_ids"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "frame_ids", "check_ids.cpp"),
                    "-o", str(exe)], check=True)
    two32 = 1 << 32
    cases = [
        (0, 0, 1),
        (two32 - (1 << 24) - 1, 0, 1),            # the largest triangle count
        (two32 - (1 << 24), 0, 0),                # one more
        (two32 - (1 << 24) - 1, (1 << 24), 1),    # last line fragment takes id 2^32 - 1
        (two32 - (1 << 24) - 1, (1 << 24) + 1, 0),  # its successor would wrap to 0
        (10, two32 - 11, 1),
        (10, two32 - 10, 0),
    ]
    args = [str(x) for t, l, _ in cases for x in (t, l)]
    out = subprocess.run([str(exe)] + args, capture_output=True, text=True, check=True).stdout.split()
    assert [int(x) for x in out] == [e for _, _, e in cases]
