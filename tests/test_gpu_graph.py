"""GPU: frames replayed from captured HIP graphs (CONTOUR3D_AMD_TUNE_GRAPHS).
Every replay must be bit-identical to the same frame rendered eagerly,
with thresholds, colours and cameras changing between replays, unrelated
eager frames and device synchronisations in between, and a replay whose
triangles outgrow the reserved capacity rendered again."""
import sys

import numpy as np
import pytest

from contour3d_amd import api, fields

from test_gpu_batch import uneven_batch
from test_gpu_parity import same_bits, view_for

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def variants():
    """Three frames of one layout: same fields, lattices and image sizes;
    other thresholds, colours and cameras."""
    base = uneven_batch()
    out = []
    for v in range(3):
        frame = []
        for q, im in enumerate(base):
            im = dict(im)
            im["contours"] = [dict(c, threshold=c["threshold"] + 0.03 * v, color=((40 * v + 7 * q) % 256, 90, 200))
                              for c in im["contours"]]
            cam = np.asarray(im["camera"], dtype=np.float64)
            im["camera"] = tuple(cam * (1.0 + 0.02 * v))
            frame.append(im)
        out.append(frame)
    return out


@pytest.mark.parametrize("precision", [api.FP64, api.FP32], ids=["fp64", "fp32"])
def test_graph_replays_equal_eager_frames(precision):
    banner(f"graphs: replays vs eager, precision={precision}")
    import torch

    frames = variants()
    api.tune(api.TUNE_GRAPHS, 0)
    ref = [api.execute_batch(fr, precision=precision) for fr in frames]
    other = view_for(80, 60)
    other_scene = [dict(field=fields.blobs((22, 22, 22), seed=9), threshold=0.5, color=(9, 9, 9))]
    try:
        api.tune(api.TUNE_GRAPHS, 1)
        for rep in range(3):
            for fi, fr in enumerate(frames):
                got = api.execute_batch(fr, precision=precision)
                for q in range(len(fr)):
                    assert np.array_equal(got[q][0], ref[fi][q][0]), (rep, fi, q)
                    assert same_bits(got[q][1], ref[fi][q][1]), (rep, fi, q)
            # unrelated eager work and synchronisations between the replays
            api.execute_rgb(other["camera"], other["light"], other["size"], other["center"], other["local"],
                            other["bg"], other_scene)
            torch.cuda.synchronize()
    finally:
        api.tune(api.TUNE_GRAPHS, 0)


def test_graph_replay_that_outgrows_the_mesh_is_rendered_again():
    banner("graphs: capacity overflow on replay")
    f = fields.turbulence((36, 36, 36), kmax=9, seed=44)
    v = view_for(120, 90)

    def frame(thr):
        return [dict(v, contours=[dict(field=f, threshold=thr, color=(200, 150, 100))], want_depth=True)]

    api.tune(api.TUNE_GRAPHS, 0)
    ref = {t: api.execute_batch(frame(t))[0] for t in (9.0, 0.0, 8.5)}
    try:
        api.tune(api.TUNE_GRAPHS, 1)
        # no triangles: the reserved capacity stays at its minimum
        for t in (9.0, 9.0, 0.0, 0.0, 8.5, 0.0):
            rgb, dep = api.execute_batch(frame(t))[0]
            assert np.array_equal(rgb, ref[t][0]), t
            assert same_bits(dep, ref[t][1]), t
    finally:
        api.tune(api.TUNE_GRAPHS, 0)
