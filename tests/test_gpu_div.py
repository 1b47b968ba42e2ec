"""GPU: the FP64 division by a view constant (div_shared in
contour3d_amd/csrc/c3d_geom.h) equals the compiler's a / b bit for bit.
tests/div_check/check_div (built by __graft_entry__.build()) compares them
on 805 M (numerator, divisor) pairs covering every exponent and the special
values; the projection of every FP64 frame uses it."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "div_check", "check_div")


def test_shared_division_equals_plain_division():
    assert os.path.exists(EXE), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    print("[gpu-test] shared-denominator division check", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
