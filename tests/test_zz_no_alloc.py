"""GPU: contour3d_amd_execute_rgb_noalloc, the entry point that allocates
nothing (its buffers are regions of one static device array set up by
contour3d_amd_init).  Called through the C ABI, it must render exactly what
contour3d_amd_execute_rgb renders for the same image, in both arithmetic
forms, on two calls in a row (the second smaller, so the array keeps the
first call's larger contents), and refuse what it cannot hold."""
import ctypes as C
import sys

import numpy as np
import pytest

from contour3d_amd import _abi, api, fields

from test_gpu_parity import EDGE_LINES, same_bits, view_for

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def noalloc(v, contours, lines, precision, rgb, depth):
    keep = api._Keep()
    objs = api.contour_objs(contours, keep)
    lobjs = api.line_objs(lines, keep)
    return api.lib().contour3d_amd_execute_rgb_noalloc(
        *api._view_args(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"]), len(contours), objs,
        len(lines), lobjs, api._addr(rgb), None if depth is None else api._addr(depth), int(precision))


def test_noalloc_equals_execute_rgb_on_two_calls(capfd):
    banner("no-alloc entry point against execute_rgb")
    import torch

    assert api.lib().contour3d_amd_init(0) == 0
    big_f = torch.from_numpy(fields.turbulence((64, 60, 56), kmax=12, seed=41)).cuda()
    small_f = torch.from_numpy(fields.blobs((20, 24, 22), seed=42)).cuda()
    cases = [
        ([dict(field=big_f, threshold=t, color=c) for t, c in ((-0.7, (250, 60, 30)), (0.0, (40, 220, 90)),
                                                            (0.7, (60, 120, 255)))], EDGE_LINES, (320, 240)),
        ([dict(field=small_f, threshold=0.5, color=(200, 200, 40))], EDGE_LINES[:1], (96, 72)),
    ]
    for prec in (api.FP64, api.FP32):
        for contours, lines, size in cases:
            v = view_for(*size)
            W, H = size
            ref_rgb, ref_dep = api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"],
                                               contours, lines, precision=prec, want_depth=True)
            rgb = torch.zeros((H, W, 3), dtype=torch.uint8, device="cuda")
            dep = torch.zeros((H, W), dtype=torch.float64, device="cuda")
            assert noalloc(v, contours, lines, prec, rgb, dep) == 0
            torch.cuda.synchronize()
            assert (ref_dep > -1e300).sum() > 100
            assert np.array_equal(rgb.cpu().numpy(), ref_rgb), (prec, size)
            assert same_bits(dep.cpu().numpy(), ref_dep), (prec, size)


def test_noalloc_refuses_what_it_cannot_hold(capfd):
    banner("no-alloc entry point: limits")
    import torch

    assert api.lib().contour3d_amd_init(0) == 0
    f = torch.from_numpy(fields.blobs((16, 16, 16), seed=43)).cuda()
    cont = [dict(field=f, threshold=0.5, color=(1, 2, 3))]
    v = view_for(64, 48)
    rgb = torch.zeros((48, 64, 3), dtype=torch.uint8, device="cuda")
    capfd.readouterr()
    # a host field, a host image and too many pixels are refused with the error line
    assert noalloc(v, [dict(cont[0], field=f.cpu().numpy())], [], api.FP64, rgb, None) == 1
    assert noalloc(v, cont, [], api.FP64, np.zeros((48, 64, 3), np.uint8), None) == 1
    big = view_for(4096, 1024)
    big_rgb = torch.zeros((1024, 4096, 3), dtype=torch.uint8, device="cuda")
    assert noalloc(big, cont, [], api.FP64, big_rgb, None) == 1
    err = capfd.readouterr().err
    assert err.count("[CONTOUR3D ERROR] ") >= 3, err
    # and the entry point still renders afterwards
    ref = api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont)[0]
    assert noalloc(v, cont, [], api.FP64, rgb, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(rgb.cpu().numpy(), ref)


def test_noalloc_fp64_beyond_its_fragment_list(capfd):
    """An FP64 frame with more covered pixel centres than the array's 2^19
    fragment slots takes the z-test passes over the geometry and still
    renders, bit for bit, what execute_rgb renders through its (growing)
    fragment list."""
    banner("no-alloc entry point: FP64 fragment list overflow, 160^3 at 1920x1080")
    import torch

    assert api.lib().contour3d_amd_init(0) == 0
    n = 160
    f = torch.from_numpy(fields.turbulence((n, n, n), kmax=24, seed=44)).cuda()
    g = fields.uniform_grid(n)
    cont = [dict(field=f, grids=(g, g, g), threshold=t, color=c)
            for t, c in ((-0.8, (250, 60, 30)), (0.0, (40, 220, 90)), (0.8, (60, 120, 255)))]
    v = fields.example_view(1920, 1080)
    v["bg"] = (0, 0, 0)
    # zoomed in twice (halved screen vectors) so that the lattice fills the
    # image: about 1.5 M covered pixels where the example view covers 0.46 M
    v["local"] = tuple(tuple(0.5 * x for x in a) for a in v["local"])
    ref_rgb, ref_dep = api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont,
                                       precision=api.FP64, want_depth=True)
    # every covered pixel holds at least one fragment: more than the list holds
    assert int((ref_dep > -1e300).sum()) > (1 << 19)
    rgb = torch.zeros((1080, 1920, 3), dtype=torch.uint8, device="cuda")
    dep = torch.zeros((1080, 1920), dtype=torch.float64, device="cuda")
    assert noalloc(v, cont, [], api.FP64, rgb, dep) == 0
    torch.cuda.synchronize()
    assert np.array_equal(rgb.cpu().numpy(), ref_rgb)
    assert same_bits(dep.cpu().numpy(), ref_dep)
