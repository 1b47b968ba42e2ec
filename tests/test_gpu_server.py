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


def test_poisson_requests_over_fields_a_solver_updates_in_place(orc):
    """C5 in small: seeded Poisson arrivals of requests against device
    fields that a solver advances in place right after every submit.  With
    snapshot_fields each request renders the field as it was at its submit:
    FP64 requests equal the oracle on that version bit for bit, FP32
    requests equal the same version rendered alone."""
    banner("server: Poisson requests, fields updated in place after each submit")
    import time

    import torch

    rng = np.random.default_rng(5)
    pool = [torch.from_numpy(fields.turbulence((n, n + 2, n - 2), kmax=6, seed=60 + q)).cuda()
            for q, n in enumerate((20, 28, 36))]
    torch.cuda.synchronize()
    cols = ((255, 80, 40), (40, 220, 90), (60, 120, 255), (230, 230, 60))
    reqs, seen = [], []
    with api.Server(max_images=16, snapshot_fields=True) as srv:
        t0 = time.perf_counter()
        t_next = 0.0
        for q in range(16):
            t_next += rng.exponential(1.0 / 200.0)
            while time.perf_counter() - t0 < t_next:
                time.sleep(0.0002)
            fi = int(rng.integers(0, len(pool)))
            thr = [float(x) for x in rng.uniform(-1.0, 1.0, size=int(rng.integers(1, 5)))]
            side = int(rng.integers(48, 112))
            sv = fields.sphere_view(float(np.arcsin(rng.uniform(-1, 1))), float(rng.uniform(0, 2 * np.pi)), side,
                                    side)
            prec = api.FP64 if q % 2 == 0 else api.FP32
            version = pool[fi].cpu().numpy().copy()
            im = dict(camera=sv["camera"], light=sv["light"], size=sv["size"], center=sv["center"], local=sv["local"],
                      bg=(0, 0, 0), want_depth=True,
                      contours=[dict(field=pool[fi], threshold=t, color=cols[k]) for k, t in enumerate(thr)])
            ticket = srv.submit(im, prec)
            # the solver's step, at once: the request must not see it
            pool[fi].mul_(-0.97).add_(0.01)
            torch.cuda.synchronize()
            reqs.append((ticket, im, prec))
            seen.append(version)
        got = [srv.wait(t) for t, _, _ in reqs]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for (t, im, prec), version, (rgb, dep, lat) in zip(reqs, seen, got):
        cont = [dict(c, field=version) for c in im["contours"]]
        if prec == api.FP64:
            ref_rgb, ref_dep = orc.execute(im["camera"], im["light"], im["size"], im["center"], im["local"],
                                           im["bg"], cont)
        else:
            ref_rgb, ref_dep = api.execute_rgb(im["camera"], im["light"], im["size"], im["center"], im["local"],
                                               im["bg"], cont, precision=prec, want_depth=True)
        assert np.array_equal(rgb, ref_rgb), t
        assert same_bits(dep, ref_dep), t
        assert lat > 0


def test_snapshot_budget_rejects_at_submit(capfd):
    """Snapshots count against the HBM budget: a request whose field copies
    would exceed it is refused at submit with the error line, and the
    server goes on serving."""
    banner("server: snapshot over the budget")
    import torch

    f = torch.from_numpy(fields.turbulence((24, 24, 24), kmax=6, seed=70)).cuda()
    v = fields.example_view(64, 48)
    im = dict(camera=v["camera"], light=v["light"], size=(64, 48), center=v["center"], local=v["local"],
              bg=(0, 0, 0), want_depth=True, contours=[dict(field=f, threshold=0.1, color=(200, 100, 50))])
    expect = alone(im, api.FP64)
    with api.Server(snapshot_fields=True, hbm_budget=24 ** 3 * 8 + 100) as srv:
        t1 = srv.submit(im, api.FP64)          # fits: its snapshot is held until it completes
        capfd.readouterr()
        big = dict(im, contours=[dict(field=torch.zeros((30, 30, 30), dtype=torch.float64, device="cuda"),
                                      threshold=0.0, color=(1, 1, 1))])
        with pytest.raises(RuntimeError):
            srv.submit(big, api.FP64)
        assert "[CONTOUR3D ERROR] server_submit: no device snapshot" in capfd.readouterr().err
        rgb, dep, _ = srv.wait(t1)
        rgb2, dep2, _ = srv.wait(srv.submit(im, api.FP64))  # the first snapshot went back to the pool
    assert np.array_equal(rgb, expect[0]) and same_bits(dep, expect[1])
    assert np.array_equal(rgb2, expect[0]) and same_bits(dep2, expect[1])


def test_bad_requests_are_refused_at_submit(capfd):
    banner("server: submit-time validation")
    v = fields.example_view(64, 48)
    good = dict(camera=v["camera"], light=v["light"], size=(64, 48), center=v["center"], local=v["local"],
                bg=(0, 0, 0), contours=[dict(field=fields.blobs((10, 10, 10), seed=1), threshold=0.5,
                                              color=(1, 2, 3))])
    with api.Server() as srv:
        for bad in (dict(good, size=(0, 48)),
                    dict(good, contours=[dict(field=np.zeros((1, 5, 5)), threshold=0.0, color=(1, 1, 1))]),
                    dict(good, lines=[dict(edges=((0, 0, 0), (1, 1, 1)), nitems=1, color=(1, 1, 1), width=1.0)])):
            capfd.readouterr()
            with pytest.raises(RuntimeError):
                srv.submit(bad, api.FP64)
            assert "[CONTOUR3D ERROR] " in capfd.readouterr().err
        rgb, _, _ = srv.wait(srv.submit(good, api.FP64))
    assert np.array_equal(rgb, alone(good, api.FP64)[0])


def test_a_request_failing_in_its_frame_fails_alone():
    """One request's frame fails on an injected allocation failure while
    another frame is in flight.  That request alone fails; the frame slot it
    leaves is taken by the next request, never the one still in flight, and
    every other request equals its image rendered alone."""
    banner("server: one frame fails while another is in flight")
    import torch

    big = torch.from_numpy(fields.turbulence((96, 96, 96), kmax=20, seed=80)).cuda()
    v = fields.example_view(512, 512)

    def req(field, thr, size):
        return dict(camera=v["camera"], light=v["light"], size=size, center=v["center"], local=v["local"],
                    bg=(0, 0, 0), want_depth=True,
                    contours=[dict(field=field, threshold=t, color=(200, 100 + 30 * k, 50))
                              for k, t in enumerate(thr)])

    a = req(big, (-1.0, -0.3, 0.3, 1.0), (512, 512))
    others = [req(torch.from_numpy(fields.blobs((20 + 4 * q, 22, 24), seed=90 + q)).cuda(), (0.5,), (96 + 8 * q, 64))
              for q in range(4)]
    expect = [alone(im, api.FP64) for im in [a] + others]
    with api.Server(max_images=1) as srv:
        srv.wait(srv.submit(a, api.FP64))  # frame slot 0 holds buffers for `a`
        api.tune(api.TUNE_FAIL_ALLOC, 1)   # the next allocation fails
        try:
            tickets = [srv.submit(im, api.FP64) for im in [a] + others]
            results = []
            for t in tickets:
                try:
                    results.append(srv.wait(t))
                except RuntimeError:
                    results.append(None)
        finally:
            api.tune(api.TUNE_FAIL_ALLOC, 0)
        st = srv.stats()
    failed = [q for q, r in enumerate(results) if r is None]
    print(f"failed requests: {failed}")
    assert len(failed) <= 1 and st["failed"] == len(failed)
    for q, r in enumerate(results):
        if r is None:
            continue
        assert np.array_equal(r[0], expect[q][0]), q
        assert same_bits(r[1], expect[q][1]), q
