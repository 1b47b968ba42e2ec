"""GPU: the stored-mesh render (`contour3d_amd_mesh_render`: `k_raster`,
`k_frag_ids`, `k_resolve`) against the CPU oracle and against the frame path.

- FP64: a stored mesh rendered from a camera is bit-identical with the
  oracle's image and z-buffer (render.c:48-163, line.c:26-170), on the image
  cases of the parity tests, lines included.
- FP32: the stored FP32 mesh and the FP32 frame compute the same positions,
  keys and smoothing sums, so their images and depths are identical.
- At the benchmark's size (512^3 turbulence, thresholds -1, 0, 1, 1920x1080)
  the fused frame and the stored-mesh render are two independent device paths
  to the same image: they must agree bit for bit, in both forms.
"""
import sys

import numpy as np
import pytest

from contour3d_amd import api, fields

from test_gpu_parity import IMAGES, same_bits, view_for

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def mesh_render(mesh, v, lines=()):
    return mesh.render(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], lines,
                       want_depth=True)


def frame(v, contours, lines=(), precision=api.FP64):
    return api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], contours, lines,
                           precision=precision, want_depth=True)


@pytest.mark.parametrize("name,contours,lines,size", IMAGES, ids=[c[0] for c in IMAGES])
def test_mesh_render_fp64_bit_exact(orc, name, contours, lines, size):
    banner(f"mesh render fp64 {name} size={size}")
    v = view_for(*size)
    mesh = api.Mesh(contours, precision=api.FP64)
    try:
        rgb, depth = mesh_render(mesh, v, lines)
    finally:
        mesh.free()
    ref_rgb, ref_depth = orc.execute(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"],
                                     contours, lines)
    assert (ref_depth > -1e300).sum() > 50
    assert same_bits(depth, ref_depth)
    assert np.array_equal(rgb, ref_rgb)


@pytest.mark.parametrize("name,contours,lines,size", IMAGES, ids=[c[0] for c in IMAGES])
def test_mesh_render_fp32_equals_fp32_frame(name, contours, lines, size):
    banner(f"mesh render fp32 {name} size={size}")
    v = view_for(*size)
    mesh = api.Mesh(contours, precision=api.FP32)
    try:
        rgb, depth = mesh_render(mesh, v, lines)
    finally:
        mesh.free()
    f_rgb, f_depth = frame(v, contours, lines, precision=api.FP32)
    assert (f_depth > -1e300).sum() > 50
    assert same_bits(depth, f_depth)
    assert np.array_equal(rgb, f_rgb)


def test_fused_frame_equals_stored_mesh_render_at_512():
    import torch

    banner("512^3 turbulence, thresholds -1 0 1, 1920x1080: fused frame against stored-mesh render")
    n = 512
    f = fields.turbulence_device((n, n, n), kmax=64, seed=0)
    torch.cuda.synchronize()
    g = fields.uniform_grid(n)
    cols = ((255, 80, 40), (40, 220, 90), (60, 120, 255))
    cont = [dict(field=f, grids=(g, g, g), threshold=t, color=c) for t, c in zip((-1.0, 0.0, 1.0), cols)]
    v = view_for(1920, 1080)
    for prec, name in ((api.FP64, "fp64"), (api.FP32, "fp32")):
        banner(f"512^3 {name}: frame")
        f_rgb, f_depth = frame(v, cont, precision=prec)
        banner(f"512^3 {name}: stored mesh")
        mesh = api.Mesh(cont, precision=prec)
        try:
            assert sum(mesh.count(o) for o in range(3)) > 100_000_000
            m_rgb, m_depth = mesh_render(mesh, v)
        finally:
            mesh.free()
        covered = (f_depth > -1e300).sum()
        banner(f"512^3 {name}: {covered} covered pixels")
        assert covered > 400_000
        assert same_bits(m_depth, f_depth)
        assert np.array_equal(m_rgb, f_rgb)
    del f
    torch.cuda.empty_cache()
