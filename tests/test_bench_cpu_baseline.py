"""CPU: bench.py's CPU baseline runs the oracle in one process per z-slab,
as the reference's ranks hold slabs, timed from a shared start.  A field
above every threshold (every cell classified, no triangle) keeps it quick."""
import os
import sys

import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from contour3d_amd import fields  # noqa: E402


def test_cpu_baseline_runs_one_slab_per_rank():
    f = torch.full((2 * 33 + 1, 512, 512), 5.0, dtype=torch.float64)
    view = fields.example_view(64, 48)
    view["bg"] = (0, 0, 0)
    r = bench.cpu_baseline(f, view, ranks=2)
    assert r["cores"] == 2 and r["kind"] == "port" and r["value"] > 0
    assert "2 processes" in r["sample"] and "z-planes 0..66" in r["sample"]
