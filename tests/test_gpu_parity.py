"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.  FP64 form: triangle sets, vertex positions, weighted
face normals, vertex normals, z-buffers and pixels are bit-identical.
FP32 form: the triangle set is identical; positional, normal and pixel
errors are bounded as stated in each test."""
import os
import sys

import numpy as np
import pytest

from contour3d_amd import api, fields

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def same_bits(a, b):
    """Bitwise equality with NaN treated as one value (the CPU and the GPU
    produce different NaN payloads)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint64), b[~nb].view(np.uint64))


def cases():
    rng = np.random.default_rng(11)
    out = []
    out.append(("random", rng.standard_normal((12, 13, 14)), [0.0, 0.7], None, None))
    out.append(("blobs", fields.blobs((33, 33, 33), seed=0), [0.5], None, None))
    out.append(("turb3", fields.turbulence((20, 24, 28), kmax=8, seed=1), [-1.0, 0.0, 1.0], None, None))
    f, g = fields.example_field(24)
    out.append(("example_cheb", f, [-0.25, 0.25], g, None))
    n = 14
    r = np.linspace(0.2, 0.5, n)
    th = np.linspace(0.0, 1.5, n + 1)
    z = np.linspace(-0.4, 0.4, n + 2)
    Z, TH, RR = np.meshgrid(z, th, r, indexing="ij")
    cyl = np.sin(6 * TH) * (RR - 0.35) + 0.3 * Z
    out.append(("cylinder_conv", cyl, [0.0, 0.05], (r, th, z), "cylinder"))
    out.append(("shear_conv", rng.standard_normal((9, 10, 11)), [0.1], None, "shear"))
    lin = np.linspace(-0.5, 0.5, 6)
    out.append(("late_conv", rng.standard_normal((6, 6, 6)), [0.0], (lin, lin, lin), "late"))
    out.append(("thin", rng.standard_normal((2, 5, 7)), [0.0], None, None))
    ties = rng.integers(-2, 3, size=(9, 8, 7)).astype(np.float64)
    out.append(("ties", ties, [0.0, 1.0], None, None))
    return out


CASES = cases()


@pytest.mark.parametrize("name,field,thrs,grids,conv", CASES, ids=[c[0] for c in CASES])
def test_mesh_fp64_bit_exact(orc, name, field, thrs, grids, conv):
    banner(f"mesh fp64 {name} shape={field.shape} thr={thrs}")
    cont = []
    for q, t in enumerate(thrs):
        c = dict(field=field, threshold=t, color=(255, 128, 0))
        if grids is not None:
            c["grids"] = grids
        if conv is not None:
            c["converter"] = orc.converter_address(conv)
        cont.append(c)
    mesh = api.Mesh(cont, precision=api.FP64)
    try:
        for q, t in enumerate(thrs):
            if conv is None:
                ref = orc.mesh(field, t, grids=grids)
            else:
                gx, gy, gz = grids if grids is not None else fields.grids_for(field.shape)
                ref = oracle_mesh_converted(orc, field, t, (gx, gy, gz), conv)
            got = mesh.copy(q)
            T = len(ref["area"])
            assert mesh.count(q) == T
            if T == 0:
                continue
            assert np.array_equal(got["cell"].astype(np.int64), ref["cell"])
            assert np.array_equal(got["edge"], ref["edge"])
            assert np.array_equal(got["interior"], ref["interior"])
            assert same_bits(got["pos"], ref["pos"])
            assert same_bits(got["wfn"], ref["area"][:, None] * ref["fn"])
            ins = ref["interior"]
            assert same_bits(got["vn"][ins], ref["vn"][ins])
    finally:
        mesh.free()


def oracle_mesh_converted(orc, field, thr, grids, conv):
    """The oracle's mesh with the lattice mapped through the same C
    converter the product receives."""
    import ctypes as C
    from contour3d_amd import _abi

    nz, ny, nx = field.shape
    fn = C.CFUNCTYPE(_abi.Vector, _abi.Vector)
    L = orc.lib()
    f = getattr(L, "oracle_conv_" + conv)
    f.restype = _abi.Vector
    f.argtypes = [_abi.Vector]
    cx = np.zeros(field.shape)
    cy = np.zeros(field.shape)
    cz = np.zeros(field.shape)
    gx, gy, gz = grids
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = f(_abi.Vector(gx[i], gy[j], gz[k]))
                cx[k, j, i], cy[k, j, i], cz[k, j, i] = r.x, r.y, r.z
    return orc.mesh(field, thr, grids=grids, coords=(cx, cy, cz))


def view_for(W, H):
    v = fields.example_view(W, H)
    return dict(camera=v["camera"], light=v["light"], size=(W, H), center=v["center"], local=v["local"],
                bg=(3, 5, 7))


EDGE_LINES = [dict(edges=((-0.5, -0.5, -0.5), (-0.5, 0.5, -0.5)), nitems=24, color=(255, 255, 255), width=3.0),
              dict(edges=((0.5, -0.5, 0.5), (0.5, 0.5, 0.5)), nitems=24, color=(200, 10, 10), width=2.0),
              dict(edges=((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)), nitems=7, color=(10, 200, 10), width=5.0)]


def image_cases():
    f_ex, g_ex = fields.example_field(32)
    return [
        ("blobs", [dict(field=fields.blobs((40, 40, 40), seed=2), threshold=0.5, color=(0, 255, 255))],
         [], (160, 120)),
        ("turb3", [dict(field=fields.turbulence((32, 32, 32), kmax=10, seed=3), threshold=t, color=c)
                   for t, c in ((-1.0, (255, 0, 0)), (0.0, (0, 255, 0)), (1.0, (0, 0, 255)))], [], (192, 108)),
        ("example", [dict(field=f_ex, grids=g_ex, threshold=-0.25, color=(0, 255, 255)),
                     dict(field=f_ex, grids=g_ex, threshold=0.25, color=(255, 255, 0))], EDGE_LINES, (144, 108)),
        ("two_fields", [dict(field=fields.blobs((20, 26, 30), seed=4), threshold=0.4, color=(250, 250, 250)),
                        dict(field=fields.turbulence((24, 20, 22), kmax=6, seed=5), threshold=0.2,
                             color=(120, 60, 200))], EDGE_LINES[:1], (100, 140)),
    ]


IMAGES = image_cases()


@pytest.mark.parametrize("name,contours,lines,size", IMAGES, ids=[c[0] for c in IMAGES])
def test_image_fp64_bit_exact(orc, name, contours, lines, size):
    banner(f"image fp64 {name} size={size}")
    v = view_for(*size)
    rgb, depth = api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], contours,
                                 lines, precision=api.FP64, want_depth=True)
    ref_rgb, ref_depth = orc.execute(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"],
                                     contours, lines)
    assert (ref_depth > -1e300).sum() > 50
    assert same_bits(depth, ref_depth)
    assert np.array_equal(rgb, ref_rgb)


def test_sphere_camera_and_converter_image(orc):
    banner("image fp64 sphere camera + shear converter")
    v = fields.sphere_view(0.6, 2.1, 96, 96)
    f = fields.turbulence((18, 22, 20), kmax=5, seed=8)
    conv = orc.converter_address("shear")
    cont = [dict(field=f, threshold=0.3, color=(90, 200, 40), converter=conv)]
    lines = [dict(edges=((-0.5, -0.5, 0.0), (0.5, 0.5, 0.0)), nitems=5, color=(1, 2, 3), width=2.5,
                  converter=conv)]
    rgb, depth = api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont,
                                 lines, want_depth=True)
    ref_rgb, ref_depth = orc.execute(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont,
                                     lines)
    assert (ref_depth > -1e300).sum() > 50
    assert same_bits(depth, ref_depth)
    assert np.array_equal(rgb, ref_rgb)


def test_execute_writes_identical_ppm(orc, tmp_path):
    banner("contour3d_execute ppm")
    f = fields.blobs((30, 30, 30), seed=6)
    v = view_for(120, 90)
    cont = [dict(field=f, threshold=0.5, color=(0, 255, 255))]
    out = tmp_path / "gpu.ppm"
    ref = tmp_path / "cpu.ppm"
    api.execute(v["camera"], (0, 0, 0), v["light"], v["size"], v["center"], v["local"], v["bg"], cont,
                EDGE_LINES, fname=out)
    orc.execute(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont, EDGE_LINES, fname=ref)
    assert out.read_bytes() == ref.read_bytes()


def test_device_resident_field_in_place(orc):
    banner("device-resident field")
    import torch

    f = fields.turbulence((26, 26, 26), kmax=7, seed=9)
    v = view_for(128, 96)
    host = [dict(field=f, threshold=0.0, color=(200, 200, 0))]
    dev = [dict(field=torch.from_numpy(f).cuda(), threshold=0.0, color=(200, 200, 0))]
    a, da = api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], host,
                            want_depth=True)
    rgb_dev = torch.zeros((96, 128, 3), dtype=torch.uint8, device="cuda")
    dep_dev = torch.zeros((96, 128), dtype=torch.float64, device="cuda")
    api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], dev, rgb=rgb_dev,
                    depth=dep_dev)
    torch.cuda.synchronize()
    assert np.array_equal(rgb_dev.cpu().numpy(), a)
    assert same_bits(dep_dev.cpu().numpy(), da)


def test_repeated_frames_identical():
    banner("repeat")
    f = fields.turbulence((24, 24, 24), kmax=6, seed=10)
    v = view_for(80, 60)
    cont = [dict(field=f, threshold=t, color=(255, 0, 0)) for t in (-0.5, 0.5)]
    outs = [api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont,
                            want_depth=True) for _ in range(3)]
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0])
        assert same_bits(o[1], outs[0][1])


def test_fp32_form_topology_exact_and_errors_bounded(orc):
    """FP32 form against the FP64 form on a 40^3 turbulence field with three
    thresholds.  Measured on MI355X: max vertex error 2.9e-8, vertex-normal
    angle mean 0.007 deg / max 0.23 deg (slivers), 2 of 4676 pixels off by
    one colour level, depth error 5.6e-6.  The bounds below sit an order of
    magnitude above those and far below what a wrong index or sign gives."""
    banner("fp32 form")
    f = fields.turbulence((40, 40, 40), kmax=10, seed=12)
    thrs = [-1.0, 0.0, 1.0]
    cont = [dict(field=f, threshold=t, color=(255, 255, 255)) for t in thrs]
    m64 = api.Mesh(cont, precision=api.FP64)
    m32 = api.Mesh(cont, precision=api.FP32)
    try:
        for q in range(3):
            a, b = m64.copy(q), m32.copy(q)
            assert np.array_equal(a["cell"], b["cell"])
            assert np.array_equal(a["edges"], b["edges"])
            verr = np.abs(a["pos"] - b["pos"]).max()
            ins = a["interior"] & np.all(np.isfinite(a["vn"]), axis=(1, 2))
            cosang = np.clip(np.sum(a["vn"][ins] * b["vn"][ins], axis=2), -1, 1)
            ang = np.degrees(np.arccos(cosang))
            print(f"thr {thrs[q]}: T={len(a['cell'])} max vertex err {verr:.3g}, "
                  f"normal angle mean {ang.mean():.3g} max {ang.max():.3g} deg")
            assert verr < 1e-6
            assert ang.mean() < 0.05 and ang.max() < 3.0
            # the device comparison tool agrees with the same statistics
            # taken per vertex on the host (vertices with a finite dot product)
            d = np.sum(a["vn"][a["interior"]] * b["vn"][a["interior"]], axis=2)
            angv = np.degrees(np.arccos(np.clip(d[np.isfinite(d)], -1, 1)))
            dev = m64.compare(m32, q)
            assert dev["triangles"] == len(a["cell"]) and dev["topology_mismatches"] == 0
            assert dev["max_vertex_err"] == verr
            assert dev["normals"] == angv.size
            assert abs(dev["max_normal_angle_deg"] - angv.max()) < 1e-6
            assert abs(dev["mean_normal_angle_deg"] - angv.mean()) < 1e-6
            # against itself: no difference beyond acos near 1 (unit vectors
            # are unit to rounding, so the angle floor is ~2e-6 degrees)
            same = m64.compare(m64, q)
            assert same["max_vertex_err"] == 0.0 and same["max_normal_angle_deg"] < 1e-5
            assert same["topology_mismatches"] == 0
    finally:
        m64.free()
        m32.free()
    v = view_for(160, 120)
    r64, d64 = api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont,
                               precision=api.FP64, want_depth=True)
    r32, d32 = api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont,
                               precision=api.FP32, want_depth=True)
    cov = d64 > -1e300
    both = cov & (d32 > -1e300)
    mism = np.any(r64 != r32, axis=2)
    derr = np.abs(d64[both] - d32[both]).max()
    print(f"pixels: covered {cov.sum()}, mismatched {mism.sum()}, depth err {derr:.3g}")
    assert mism.sum() <= 0.01 * cov.sum()
    assert (cov ^ (d32 > -1e300)).sum() <= 0.005 * cov.sum()
    assert derr < 1e-4


def test_errors_return_one():
    banner("error paths")
    f = np.zeros((1, 4, 4))
    v = view_for(16, 16)
    with pytest.raises(RuntimeError):
        api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"],
                        [dict(field=f, threshold=0.0, color=(1, 1, 1))])
    with pytest.raises(RuntimeError):
        api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], [],
                        [dict(edges=((0, 0, 0), (1, 1, 1)), nitems=1, color=(1, 1, 1), width=1.0)])
    with pytest.raises(RuntimeError):
        api.execute(v["camera"], (0, 0, 0), v["light"], v["size"], v["center"], v["local"], v["bg"], [],
                    fname="/nonexistent_dir/x.ppm")
    # the library still works after failures
    g = fields.blobs((12, 12, 12), seed=1)
    api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"],
                    [dict(field=g, threshold=0.5, color=(1, 1, 1))])


def test_frame_outgrowing_the_reserved_cells_renders_again(orc):
    """A waited frame plans against the cell store that earlier frames
    reserved, with no read back of its totals; a frame that needs more is
    flagged by the plan and rendered again with the read back.  After a
    finalize (nothing reserved), a small frame, then one with far more
    triangles: it must equal the oracle, and the profiled triangle count of
    the re-rendered frame must equal that of the next, which fits."""
    banner("frame outgrowing the reserved cell store")
    api.finalize()
    v = view_for(96, 72)
    small = [dict(field=fields.blobs((12, 12, 12), seed=3), threshold=0.5, color=(200, 60, 30))]
    big = [dict(field=fields.turbulence((40, 40, 40), kmax=12, seed=8), threshold=t, color=c)
           for t, c in ((-0.8, (255, 0, 0)), (0.0, (0, 255, 0)), (0.8, (0, 0, 255)))]
    api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], small,
                    precision=api.FP64)
    api.profile_enable(True)
    try:
        counts = []
        for _ in range(2):
            api.profile_reset()
            rgb, dep = api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], big,
                                       precision=api.FP64, want_depth=True)
            counts.append(api.profile_read()[2])
            ref_rgb, ref_dep = orc.execute(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"],
                                           big)
            assert np.array_equal(rgb, ref_rgb)
            assert same_bits(dep, ref_dep)
    finally:
        api.profile_enable(False)
    assert counts[0] == counts[1] > 10000


def test_large_triangles_fill_the_fragment_lists(orc):
    """A 9^3 field at 320x240: each triangle covers hundreds of pixels, so
    the workgroups' LDS fragment lists overflow into the frame's list.  FP64
    stays bit-exact with the oracle, and in both forms the fragment-list
    path renders the same bits as the z-test passes that run after a list
    overflow (forced with a one-entry list)."""
    banner("large triangles: LDS list spill, list vs z-test passes")
    contours = [dict(field=fields.blobs((9, 9, 9), seed=5), threshold=0.45, color=(200, 120, 40)),
                dict(field=fields.blobs((9, 9, 9), seed=6), threshold=0.5, color=(40, 120, 200))]
    v = view_for(320, 240)
    args = (v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], contours)
    ref_rgb, ref_dep = orc.execute(*args)
    assert (ref_dep > -1e300).sum() > 5000
    out = {}
    try:
        for cap in (0, 1):
            api.tune(api.TUNE_FRAG_CAPACITY, cap)
            for prec in (api.FP64, api.FP32):
                out[cap, prec] = api.execute_rgb(*args, precision=prec, want_depth=True)
    finally:
        api.tune(api.TUNE_FRAG_CAPACITY, 0)
    rgb, dep = out[0, api.FP64]
    assert same_bits(dep, ref_dep)
    assert np.array_equal(rgb, ref_rgb)
    for prec in (api.FP64, api.FP32):
        assert np.array_equal(out[0, prec][0], out[1, prec][0]), prec
        assert same_bits(out[0, prec][1], out[1, prec][1]), prec


def test_geometry_cache_follows_grid_values(orc):
    """Lattice geometries are cached by converter and grid values: equal
    grids passed as new arrays reuse the cache, different grids of the same
    size do not.  Alternating the two, each image stays bit-exact."""
    banner("geometry cache keyed on grid values")
    f = fields.blobs((20, 18, 16), seed=9)
    nz, ny, nx = f.shape
    uni = [(-0.5 + (np.arange(m) + 0.5) / m) for m in (nx, ny, nz)]
    cheb = [0.5 * np.sin(np.pi * (np.arange(m) + 0.5) / m - np.pi / 2) for m in (nx, ny, nz)]
    v = view_for(96, 72)
    for grids in (uni, cheb, [g.copy() for g in uni], [g.copy() for g in cheb]):
        cont = [dict(field=f, grids=tuple(grids), threshold=0.45, color=(200, 150, 60))]
        rgb, dep = api.execute_rgb(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont,
                                   precision=api.FP64, want_depth=True)
        ref_rgb, ref_dep = orc.execute(v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont)
        assert (ref_dep > -1e300).sum() > 100
        assert same_bits(dep, ref_dep)
        assert np.array_equal(rgb, ref_rgb)


def test_dense_cells_and_many_thresholds(orc):
    """A checkerboard field crosses every cell at every threshold near 0, so
    a 256-cell chunk yields thousands of triangles (many queue drains per
    chunk), and six thresholds on one field take two field passes.  FP64
    images are bit-exact with the oracle; the FP32 image equals the same
    image rendered in a batch."""
    banner("dense cells, six thresholds on one field")
    n = 18
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    f = np.where((i + j + k) % 2 == 0, 1.0, -1.0) + 0.01 * np.sin(i + 2.0 * j + 3.0 * k)
    cols = [(250, 40, 40), (40, 250, 40), (40, 40, 250), (250, 250, 40), (250, 40, 250), (40, 250, 250)]
    cont = [dict(field=f, threshold=t, color=c) for t, c in zip((-0.6, -0.3, -0.05, 0.05, 0.3, 0.6), cols)]
    v = view_for(160, 120)
    args = (v["camera"], v["light"], v["size"], v["center"], v["local"], v["bg"], cont)
    rgb, dep = api.execute_rgb(*args, precision=api.FP64, want_depth=True)
    ref_rgb, ref_dep = orc.execute(*args)
    assert (ref_dep > -1e300).sum() > 2000
    assert same_bits(dep, ref_dep)
    assert np.array_equal(rgb, ref_rgb)
    rgb32, dep32 = api.execute_rgb(*args, precision=api.FP32, want_depth=True)
    im = dict(camera=v["camera"], light=v["light"], size=v["size"], center=v["center"], local=v["local"],
              bg=v["bg"], contours=cont, want_depth=True)
    other = dict(im, contours=cont[:2])
    got = api.execute_batch([other, im], precision=api.FP32)
    assert np.array_equal(got[1][0], rgb32)
    assert same_bits(got[1][1], dep32)
