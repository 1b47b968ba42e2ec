"""GPU: the serving loop (contour3d_amd_server_*).  Every served image must
be bit-identical to the same image rendered alone by
contour3d_amd_execute_rgb, whatever it was batched with, for any admission
budget and batch size; device fields updated in place between requests are
read at their submit-time values."""
import sys

import numpy as np
import pytest

from contour3d_amd import api, fields

from test_gpu_batch import alone, uneven_batch
from test_gpu_parity import same_bits

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def mixed_requests():
    reqs = []
    for q, im in enumerate(uneven_batch()):
        reqs.append((im, api.FP64 if q % 2 == 0 else api.FP32))
    f = fields.turbulence((28, 30, 26), kmax=7, seed=31)
    scene = [dict(field=f, threshold=t, color=c) for t, c in ((-0.5, (250, 90, 30)), (0.7, (30, 90, 250)))]
    for q, (th, ph) in enumerate(((0.2, 0.1), (0.9, 2.5), (-0.4, 4.0), (1.2, 5.1))):
        sv = fields.sphere_view(th, ph, 72 + 8 * q, 64)
        im = dict(camera=sv["camera"], light=sv["light"], size=sv["size"], center=sv["center"], local=sv["local"],
                  bg=(1, 2, 3), contours=scene, want_depth=True)
        reqs.append((im, api.FP32 if q % 2 else api.FP64))
    return reqs


@pytest.mark.parametrize("max_images,budget", [(16, 200e9), (3, 200e9), (16, 1)], ids=["wide", "narrow", "budget1"])
def test_served_images_equal_images_alone(max_images, budget):
    banner(f"server mix max_images={max_images} budget={budget}")
    reqs = mixed_requests()
    expect = [alone(im, p) for im, p in reqs]
    with api.Server(max_images=max_images, hbm_budget=budget) as srv:
        tickets = [srv.submit(im, p) for im, p in reqs]
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    for q, ((rgb, dep, lat), (ergb, edep)) in enumerate(zip(got, expect)):
        assert np.array_equal(rgb, ergb), q
        assert same_bits(dep, edep), q
        assert lat > 0
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["mean_batch"] <= max_images
    if budget == 1:
        assert st["batches"] == len(reqs)  # nothing fits beside another request
    assert st["peak_admitted_bytes"] > 0


def test_batching_window_starts_lone_requests_and_bursts():
    """With max_wait_ms > 0 an idle worker holds a request until its window
    ends, for company; a lone request and the last of a burst must still
    start then, with nothing else arriving."""
    banner("server: batching window")
    reqs = [(im, api.FP32) for im, _ in mixed_requests()[:5]]
    expect = [alone(im, p) for im, p in reqs]
    with api.Server(max_images=16, max_wait_ms=20.0) as srv:
        rgb, dep, lat = srv.wait(srv.submit(*reqs[0]))
        assert np.array_equal(rgb, expect[0][0]) and same_bits(dep, expect[0][1])
        assert 15.0 <= lat < 5000.0  # held for its window, then rendered
        tickets = [srv.submit(im, p) for im, p in reqs[1:]]
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    for q, ((rgb, dep, _), (ergb, edep)) in enumerate(zip(got, expect[1:])):
        assert np.array_equal(rgb, ergb), q
        assert same_bits(dep, edep), q
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["batches"] <= 3 and st["mean_batch"] > 1.0  # the burst shared frames


def test_ppm_requests_write_the_files_contour3d_execute_writes(tmp_path):
    """submit_ppm: the writer thread writes each image as contour3d_execute
    writes its file (P6 header, then the rows top first, output.c:125-140);
    a file that cannot be written fails that request only."""
    banner("server: PPM output")
    reqs = mixed_requests()[:6]
    expect = [alone(im, p) for im, p in reqs]
    paths = [tmp_path / f"r{q}.ppm" for q in range(len(reqs))]
    with api.Server(max_images=4) as srv:
        tickets = [srv.submit_ppm(im, str(path), p) for (im, p), path in zip(reqs, paths)]
        bad = srv.submit_ppm(reqs[0][0], str(tmp_path / "missing" / "x.ppm"), reqs[0][1])
        with pytest.raises(RuntimeError):  # over the frame's limits: rejected before any pinned image is taken
            srv.submit_ppm(dict(reqs[0][0], size=(70000, 8)), str(tmp_path / "big.ppm"), reqs[0][1])
        got = [srv.wait(t) for t in tickets]
        with pytest.raises(RuntimeError):
            srv.wait(bad)
        st = srv.stats()
    for q, ((im, _), path, (rgb, dep, lat), (ergb, edep)) in enumerate(zip(reqs, paths, got, expect)):
        W, H = im["size"]
        assert path.read_bytes() == f"P6\n{W} {H}\n255\n".encode() + np.ascontiguousarray(ergb).tobytes(), q
        assert rgb is None and same_bits(dep, edep), q
        assert lat > 0
    assert st["completed"] == len(reqs) and st["failed"] == 1


def test_device_fields_updated_in_place_between_requests():
    banner("server: device field updated in place")
    import torch

    f = torch.from_numpy(fields.turbulence((30, 30, 30), kmax=8, seed=40)).cuda()
    v = fields.example_view(96, 72)
    im = dict(camera=v["camera"], light=v["light"], size=(96, 72), center=v["center"], local=v["local"], bg=(0, 0, 0),
              contours=[dict(field=f, threshold=0.2, color=(200, 100, 50))], want_depth=True)
    with api.Server() as srv:
        outs, refs = [], []
        for step in range(3):
            t = srv.submit(im, api.FP64)
            outs.append(srv.wait(t)[:2])
            refs.append(alone(im, api.FP64))
            f.mul_(0.9).add_(0.05)  # the solver advances the field in place
            torch.cuda.synchronize()
    for (a, da), (b, db) in zip(outs, refs):
        assert np.array_equal(a, b)
        assert same_bits(da, db)
    assert not np.array_equal(outs[0][0], outs[2][0])


def test_stop_renders_queued_requests():
    banner("server: stop drains the queue")
    reqs = mixed_requests()[:5]
    expect = [alone(im, p) for im, p in reqs]
    srv = api.Server(max_images=2)
    outs = []
    for im, p in reqs:
        srv.submit(im, p)
    # stop before waiting: every queued request is still rendered
    keep = dict(srv._pending)
    srv.stop()
    for (t, (e, k, out)), (ergb, edep) in zip(sorted(keep.items()), expect):
        assert np.array_equal(out[0], ergb)
        assert same_bits(out[1], edep)
