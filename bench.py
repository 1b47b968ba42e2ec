"""Benchmark: one frame of the tessellate -> smooth -> render path.

Workload (the plain method at GPU scale): a 512^3 float64 spectral
turbulence field (modes 1 <= |k| <= 64, k^-5/3 shell spectrum, seeded
phases, mean 0 / variance 1) resident in HBM, thresholds {-1, 0, 1} as three
contour objects, one 1920x1080 image written to HBM.  A step is one
contour3d_amd_execute_rgb call; Gcells/s counts 3 * 511^3 cell-thresholds
per frame.

Prints one JSON line with the throughput of the FP64 form (bit-identical to
the reference), the FP32-geometry form beside it, per-stage times from HIP
events on the library stream, the HBM roofline of the dominant kernel, and
the CPU oracle timed on a bounded slab of the same field.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
N = 512
THRESHOLDS = (-1.0, 0.0, 1.0)
COLORS = ((255, 80, 40), (40, 220, 90), (60, 120, 255))
W, H = 1920, 1080


def algorithmic_bytes(stage, T, npts, npx, fp64):
    """Irreducible HBM bytes per frame of each stage.

    count    the field, read once for all thresholds (8 B per lattice point)
    emit     the field once + per triangle 3 positions and the weighted face
             normal (96 B in FP64, 48 B in FP32): the tessellation bytes
    raster   per triangle its 3 positions, plus the z-buffer keys written
             once and read once per pixel
    resolve  the keys read and the RGB bytes written per pixel (the vertex
             normals of the nearest triangles are gathered here and add
             O(pixels) bytes, not counted)
    """
    r = 8 if fp64 else 4
    if stage == "count":
        return 8 * npts
    if stage == "emit":
        return 8 * npts + 12 * r * T
    if stage == "raster":
        return 9 * r * T + 16 * npx
    if stage == "resolve":
        return 8 * npx + 3 * npx
    return 0


def run_form(api, contours, view, precision, steps, warmup, rgb_dev):
    import torch

    for _ in range(warmup):
        api.execute_rgb(view["camera"], view["light"], view["size"], view["center"], view["local"], view["bg"],
                        contours, precision=precision, rgb=rgb_dev)
    api.profile_reset()
    api.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        api.execute_rgb(view["camera"], view["light"], view["size"], view["center"], view["local"], view["bg"],
                        contours, precision=precision, rgb=rgb_dev)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    api.profile_enable(False)
    ms, calls, ntri = api.profile_read()
    return dt, ms, calls, ntri


def cpu_baseline(field_dev, view):
    """The CPU oracle (single thread) on the first 34 z-planes of the same
    field: 511 x 511 x 33 cells x 3 thresholds into the same 1920x1080 view."""
    from oracle import oracle

    planes = 34
    slab = field_dev[:planes].cpu().numpy()
    g = -0.5 + (np.arange(N) + 0.5) / N
    cont = [dict(field=slab, grids=(g, g, g[:planes]), threshold=t, color=c)
            for t, c in zip(THRESHOLDS, COLORS)]
    t0 = time.perf_counter()
    oracle.execute(view["camera"], view["light"], view["size"], view["center"], view["local"], view["bg"], cont)
    dt = time.perf_counter() - t0
    cells = len(THRESHOLDS) * (N - 1) * (N - 1) * (planes - 1)
    return dict(value=cells / dt / 1e9, unit="Gcells/s", cores=1, kind="port",
                sample=f"oracle, 1 thread, first {planes} z-planes of the 512^3 field "
                       f"({cells / 1e6:.1f} M cell-thresholds) into the 1920x1080 view, {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--n", type=int, default=N, help="lattice points per axis (default 512)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB device-to-device three times (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.calibrate_copy:
        import torch

        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": x.numel() * 8, "copies": 3}))
        return

    import torch

    from contour3d_amd import api, fields

    n = args.n
    print(f"[bench] generating {n}^3 field", file=sys.stderr, flush=True)
    field = fields.turbulence_device((n, n, n), kmax=min(64, n // 2 - 1), seed=0)
    torch.cuda.synchronize()
    view = fields.example_view(W, H)
    view["bg"] = (0, 0, 0)
    g = fields.uniform_grid(n)
    contours = [dict(field=field, grids=(g, g, g), threshold=t, color=c) for t, c in zip(THRESHOLDS, COLORS)]
    rgb_dev = torch.zeros((H, W, 3), dtype=torch.uint8, device="cuda")
    npts = n ** 3
    cells = len(THRESHOLDS) * (n - 1) ** 3
    npx = W * H

    results = {}
    for name, prec in (("fp64", api.FP64), ("fp32", api.FP32)):
        print(f"[bench] {name}: {args.warmup} warmup + {args.steps} timed frames", file=sys.stderr, flush=True)
        dt, ms, calls, ntri = run_form(api, contours, view, prec, args.steps, args.warmup, rgb_dev)
        T = ntri / max(1, args.steps)
        stages = {}
        for s in api.STAGES:
            if s == "smooth":
                continue  # vertex normals are gathered inside the resolve pass
            per = ms[s] / max(1, args.steps)
            b = algorithmic_bytes(s, T, npts, npx, prec == api.FP64)
            stages[s] = dict(ms=round(per, 4), gbs=round(b / (per * 1e-3) / 1e9, 1) if per > 0 and b else None)
        results[name] = dict(dt=dt, T=T, stages=stages, frame_ms=dt / args.steps * 1e3)

    r = results["fp64"]
    dom = max(("count", "emit", "raster", "resolve"), key=lambda s: r["stages"][s]["ms"])
    dom_b = algorithmic_bytes(dom, r["T"], npts, npx, True)
    achieved = dom_b / (r["stages"][dom]["ms"] * 1e-3) / 1e9
    tess_b = algorithmic_bytes("emit", r["T"], npts, npx, True)
    tess_ms = r["stages"]["count"]["ms"] + r["stages"]["scan"]["ms"] + r["stages"]["emit"]["ms"]
    out = {
        "metric": "Gcells/s + frames/s at 512^3 fp64; served p50/p99 latency vs plain GPU path",
        "value": round(cells * args.steps / r["dt"] / 1e9, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["frame_ms"], 3),
        "frames_per_s": round(1e3 / r["frame_ms"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain method at GPU scale: {n}^3 f64 spectral turbulence (|k| 1..64, k^-5/3), "
                               f"thresholds {{-1,0,1}}, {W}x{H} image in HBM",
                   "lattice": [n, n, n], "thresholds": list(THRESHOLDS), "image": [W, H],
                   "triangles_per_frame": int(r["T"])},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None},
        "tessellation": {"ms": round(tess_ms, 4),
                         "gcells_per_s": round(cells / (tess_ms * 1e-3) / 1e9, 2),
                         "hbm_frac": round(tess_b / (tess_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)},
        "stages_fp64": r["stages"],
        "fp32": {"value": round(cells * args.steps / results["fp32"]["dt"] / 1e9, 3), "unit": "Gcells/s",
                 "ms_per_step": round(results["fp32"]["frame_ms"], 3),
                 "frames_per_s": round(1e3 / results["fp32"]["frame_ms"], 2),
                 "stages": results["fp32"]["stages"]},
    }
    if not args.no_cpu:
        print("[bench] cpu oracle baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(field, view)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
