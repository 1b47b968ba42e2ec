"""GPU: the failure contract of the reference (src/contour3d/memory.c:14-69,
src/contour3d/logger.c:5-18): a call that cannot allocate returns 1 with a
"[CONTOUR3D ERROR] " line on stderr, and nothing it allocated leaks into
later calls.  Device and pinned allocations are made to fail one at a time
(CONTOUR3D_AMD_TUNE_FAIL_ALLOC = n fails the n-th allocation from now), at
every allocation point a frame, a PPM write and a mesh build pass through;
after each failure the same call must render bit-exactly again."""
import sys

import numpy as np
import pytest

from contour3d_amd import api, fields

from test_gpu_parity import EDGE_LINES, same_bits, view_for

pytestmark = pytest.mark.gpu

PREFIX = "[CONTOUR3D ERROR] "


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def sweep(call, check, capfd, limit=40):
    """For n = 1, 2, ...: start from an empty library, arm the n-th
    allocation to fail and make the call.  It either fails (rc 1, the
    prefixed line on stderr) or, once n exceeds the allocations the call
    makes, succeeds.  After every failure the call is made again, disarmed,
    and checked.  Returns the number of allocation points that failed."""
    failures = 0
    for n in range(1, limit + 1):
        api.finalize()
        capfd.readouterr()
        api.tune(api.TUNE_FAIL_ALLOC, n)
        try:
            out = call()
            ok = True
        except RuntimeError:
            ok = False
        finally:
            api.tune(api.TUNE_FAIL_ALLOC, 0)
        err = capfd.readouterr().err
        if ok:
            check(out)
            return failures
        assert PREFIX in err, err
        failures += 1
        check(call())
    raise AssertionError(f"still failing after {limit} allocation points")


def test_frames_fail_cleanly_at_every_allocation(orc, capfd):
    banner("error contract: execute_rgb under allocation failures")
    f = fields.turbulence((24, 26, 22), kmax=6, seed=3)
    v = view_for(96, 72)
    cont = [dict(field=f, threshold=t, color=c) for t, c in ((-0.4, (250, 60, 30)), (0.4, (30, 60, 250)))]
    ref = orc.execute(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont, EDGE_LINES)

    def call():
        return api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont,
                               EDGE_LINES, want_depth=True)

    def check(out):
        assert np.array_equal(out[0], ref[0])
        assert same_bits(out[1], ref[1])

    n = sweep(call, check, capfd)
    print(f"{n} allocation points failed cleanly")
    assert n >= 8


def test_contour3d_execute_fails_cleanly(orc, capfd, tmp_path):
    banner("error contract: contour3d_execute under allocation failures")
    f = fields.blobs((20, 20, 20), seed=4)
    v = view_for(80, 60)
    cont = [dict(field=f, threshold=0.5, color=(0, 255, 255))]
    ref = tmp_path / "ref.ppm"
    orc.execute(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont, fname=ref)
    out = tmp_path / "out.ppm"

    def call():
        api.execute(v["camera"], (0, 0, 0), v["light"], v["size"], v["center"], v["local"], v["bg"], cont,
                    fname=out)
        return out.read_bytes()

    def check(data):
        assert data == ref.read_bytes()

    assert sweep(call, check, capfd) >= 8


def test_mesh_build_fails_cleanly(orc, capfd):
    banner("error contract: mesh build under allocation failures")
    f = fields.turbulence((18, 20, 16), kmax=5, seed=5)
    cont = [dict(field=f, threshold=0.2, color=(1, 2, 3))]
    ref = orc.mesh(f, 0.2)

    def call():
        m = api.Mesh(cont, precision=api.FP64)
        try:
            return m.copy(0)
        finally:
            m.free()

    def check(got):
        assert np.array_equal(got["cell"].astype(np.int64), ref["cell"])
        assert same_bits(got["pos"], ref["pos"])

    assert sweep(call, check, capfd) >= 5
