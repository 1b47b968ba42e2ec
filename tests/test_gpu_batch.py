"""GPU: contour3d_amd_execute_batch.  Every image of a batch must be
bit-identical (pixels and z-buffer) to the same image rendered alone by
contour3d_amd_execute_rgb, in both arithmetic forms, whatever else shares
its frame; FP64 images must also equal the CPU oracle."""
import sys

import numpy as np
import pytest

from contour3d_amd import api, fields

from test_gpu_parity import EDGE_LINES, IMAGES, same_bits, view_for

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def alone(im, precision):
    return api.execute_rgb(im["camera"], im["light"], im["size"], im["center"], im["local"], im["bg"],
                           im["contours"], im.get("lines", ()), precision=precision, want_depth=True)


def uneven_batch():
    out = []
    for name, contours, lines, size in IMAGES:
        im = view_for(*size)
        im.update(contours=contours, lines=lines, want_depth=True)
        out.append(im)
    sv = fields.sphere_view(0.9, 1.8, 72, 64)
    out.append(dict(camera=sv["camera"], light=sv["light"], size=sv["size"], center=sv["center"],
                    local=sv["local"], bg=(9, 9, 9), want_depth=True,
                    contours=[dict(field=fields.blobs((25, 21, 23), seed=7), threshold=0.45, color=(30, 60, 90))],
                    lines=[]))
    # an image with lines only
    im = view_for(64, 48)
    im.update(contours=[], lines=EDGE_LINES, want_depth=True)
    out.append(im)
    return out


@pytest.mark.parametrize("precision", [api.FP64, api.FP32], ids=["fp64", "fp32"])
def test_uneven_batch_equals_images_alone(orc, precision):
    banner(f"batch uneven precision={precision}")
    batch = uneven_batch()
    got = api.execute_batch(batch, precision=precision)
    for q, im in enumerate(batch):
        rgb, dep = alone(im, precision)
        assert np.array_equal(got[q][0], rgb), q
        assert same_bits(got[q][1], dep), q
        if precision == api.FP64:
            ref_rgb, ref_dep = orc.execute(im["camera"], im["light"], im["size"], im["center"], im["local"],
                                           im["bg"], im["contours"], im.get("lines", ()))
            assert np.array_equal(got[q][0], ref_rgb), q
            assert same_bits(got[q][1], ref_dep), q


def small_pass_batch():
    """Field passes of 1 to 26 tiles (1024 cells each): each pass's tiles are
    split over the 8 XCDs, so most of these leave some XCDs nothing of that
    pass.  The last field has six thresholds, two passes over one field."""
    out = []
    shapes = [(3, 3, 3), (9, 7, 5), (12, 11, 10), (20, 18, 16), (33, 30, 29), (6, 40, 7), (24, 22, 20)]
    for q, shape in enumerate(shapes):
        f = fields.blobs(shape, seed=20 + q)
        nthr = 6 if q == len(shapes) - 1 else 1 + q % 4
        im = view_for(48 + 8 * q, 40 + 4 * q)
        im.update(contours=[dict(field=f, threshold=float(t), color=(40 + 30 * k, 200 - 20 * k, 90))
                            for k, t in enumerate(np.linspace(0.25, 0.75, nthr))],
                  lines=[], want_depth=True)
        out.append(im)
    return out


@pytest.mark.parametrize("precision", [api.FP64, api.FP32], ids=["fp64", "fp32"])
def test_many_small_passes_equal_images_alone(orc, precision):
    banner(f"batch of small field passes precision={precision}")
    batch = small_pass_batch()
    got = api.execute_batch(batch, precision=precision)
    drawn = 0
    for q, im in enumerate(batch):
        rgb, dep = alone(im, precision)
        assert np.array_equal(got[q][0], rgb), q
        assert same_bits(got[q][1], dep), q
        drawn += int((dep > -1e300).sum())
        if precision == api.FP64:
            ref_rgb, ref_dep = orc.execute(im["camera"], im["light"], im["size"], im["center"], im["local"],
                                           im["bg"], im["contours"], ())
            assert np.array_equal(got[q][0], ref_rgb), q
            assert same_bits(got[q][1], ref_dep), q
    assert drawn > 500


@pytest.mark.parametrize("precision", [api.FP64, api.FP32], ids=["fp64", "fp32"])
def test_cameras_sharing_one_scene_tessellate_once(orc, precision):
    banner(f"batch shared scene precision={precision}")
    f = fields.turbulence((30, 30, 30), kmax=8, seed=21)
    scene = [dict(field=f, threshold=t, color=c) for t, c in ((-0.8, (255, 40, 40)), (0.6, (40, 40, 255)))]
    batch = []
    for q, (th, ph) in enumerate(((0.3, 0.4), (1.1, 2.0), (2.2, 4.1), (0.7, 5.5))):
        sv = fields.sphere_view(th, ph, 80 + 16 * q, 60 + 8 * q)
        batch.append(dict(camera=sv["camera"], light=sv["light"], size=sv["size"], center=sv["center"],
                          local=sv["local"], bg=(0, 0, 0), contours=scene, want_depth=True))
    api.profile_enable(True)
    api.profile_reset()
    got = api.execute_batch(batch, precision=precision)
    _, _, ntri_batch = api.profile_read()
    api.profile_reset()
    rgb0, dep0 = alone(batch[0], precision)
    _, _, ntri_one = api.profile_read()
    api.profile_enable(False)
    assert ntri_batch == ntri_one > 0
    for q, im in enumerate(batch):
        rgb, dep = (rgb0, dep0) if q == 0 else alone(im, precision)
        assert np.array_equal(got[q][0], rgb), q
        assert same_bits(got[q][1], dep), q


def test_batch_device_outputs_and_capacity_changes():
    banner("batch device outputs, growing and shrinking frames")
    import torch

    big = uneven_batch()
    host = api.execute_batch(big, precision=api.FP64)
    dev = []
    for im in big:
        W, H = im["size"]
        d = dict(im)
        d["rgb"] = torch.zeros((H, W, 3), dtype=torch.uint8, device="cuda")
        d["depth"] = torch.zeros((H, W), dtype=torch.float64, device="cuda")
        dev.append(d)
    api.execute_batch(dev, precision=api.FP64)
    torch.cuda.synchronize()
    for q, d in enumerate(dev):
        assert np.array_equal(d["rgb"].cpu().numpy(), host[q][0]), q
        assert same_bits(d["depth"].cpu().numpy(), host[q][1]), q
    # a smaller frame after a larger one, then the larger one again
    small = api.execute_batch(big[:1], precision=api.FP64)
    assert np.array_equal(small[0][0], host[0][0])
    again = api.execute_batch(big, precision=api.FP64)
    for q in range(len(big)):
        assert np.array_equal(again[q][0], host[q][0]), q
        assert same_bits(again[q][1], host[q][1]), q


def test_batch_errors():
    banner("batch errors")
    v = view_for(16, 16)
    bad = dict(v, contours=[dict(field=np.zeros((1, 4, 4)), threshold=0.0, color=(1, 1, 1))])
    good = dict(v, contours=[dict(field=fields.blobs((10, 10, 10), seed=3), threshold=0.5, color=(1, 1, 1))])
    with pytest.raises(RuntimeError):
        api.execute_batch([good, bad])
    with pytest.raises(RuntimeError):
        api.execute_batch([good, dict(v, size=(0, 16), contours=[])])
    assert api.execute_batch([]) == []
    out = api.execute_batch([good])
    assert out[0][0].shape == (16, 16, 3)


def test_launch_shapes_and_fallback_pass_do_not_change_results(orc):
    """Persistent loops with few workgroups (several tiles and raster
    rounds per workgroup) and a fragment list too small for the frame (the
    FP64 geometry id pass) render the same bits."""
    banner("tuning: tiny grids, fragment-list overflow")
    batch = uneven_batch()
    ref64 = api.execute_batch(batch, precision=api.FP64)
    ref32 = api.execute_batch(batch, precision=api.FP32)
    try:
        for tess, ras, frag in ((8, 8, 0), (24, 3, 16), (0, 0, 16)):
            api.tune(api.TUNE_TESS_GRID, tess)
            api.tune(api.TUNE_RASTER_GRID, ras)
            if frag:
                api.tune(api.TUNE_FRAG_CAPACITY, frag)
            got = api.execute_batch(batch, precision=api.FP64)
            for q in range(len(batch)):
                assert np.array_equal(got[q][0], ref64[q][0]), (tess, ras, frag, q)
                assert same_bits(got[q][1], ref64[q][1]), (tess, ras, frag, q)
            got = api.execute_batch(batch, precision=api.FP32)
            for q in range(len(batch)):
                assert np.array_equal(got[q][0], ref32[q][0]), (tess, ras, frag, q)
                assert same_bits(got[q][1], ref32[q][1]), (tess, ras, frag, q)
        # a single image with the fallback pass, against the oracle
        api.tune(api.TUNE_FRAG_CAPACITY, 1)
        im = batch[2]
        ref_rgb, ref_dep = orc.execute(im["camera"], im["light"], im["size"], im["center"], im["local"], im["bg"],
                                       im["contours"], im["lines"])
        rgb, dep = alone(im, api.FP64)
        assert np.array_equal(rgb, ref_rgb)
        assert same_bits(dep, ref_dep)
    finally:
        api.tune(api.TUNE_TESS_GRID, 0)
        api.tune(api.TUNE_RASTER_GRID, 0)
