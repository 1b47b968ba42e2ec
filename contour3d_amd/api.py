"""ctypes front end of libcontour3d.so.

`execute` mirrors the reference's single public entry point
contour3d_execute (include/contour3d.h:58-86): same arguments, same
0/1-plus-stderr error behaviour (raised here as RuntimeError), same PPM
output.  `execute_rgb`, `Mesh` and the profiling calls wrap the extensions
of include/contour3d_amd.h.  Fields may be numpy arrays (host) or torch
device tensors (used in place, no copy).
"""
import ctypes as C
import os

import numpy as np

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcontour3d.so")

FP64 = 0
FP32 = 1
STAGES = ("count", "scan", "emit", "smooth", "raster", "resolve", "fused")

# every entry point declared in include/contour3d.h and include/contour3d_amd.h
EXPORTS = (
    "contour3d_execute",
    "contour3d_amd_init",
    "contour3d_amd_finalize",
    "contour3d_amd_set_precision",
    "contour3d_amd_identity",
    "contour3d_amd_execute_rgb",
    "contour3d_amd_execute_rgb_noalloc",
    "contour3d_amd_execute_batch",
    "contour3d_amd_server_start",
    "contour3d_amd_server_submit",
    "contour3d_amd_server_submit_ppm",
    "contour3d_amd_server_wait",
    "contour3d_amd_server_stats",
    "contour3d_amd_server_stop",
    "contour3d_amd_mesh_build",
    "contour3d_amd_mesh_count",
    "contour3d_amd_mesh_copy",
    "contour3d_amd_mesh_render",
    "contour3d_amd_mesh_compare",
    "contour3d_amd_mesh_normals",
    "contour3d_amd_mesh_free",
    "contour3d_amd_tune",
    "contour3d_amd_profile_enable",
    "contour3d_amd_profile_read",
    "contour3d_amd_profile_reset",
)

_lib = None


def lib():
    """Load the native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        # torch bundles its own HIP runtime under the same soname; loading it
        # first makes this library bind to that copy, so a process that also
        # uses torch holds one HIP runtime instead of two
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library missing: {LIB_PATH} (run python contour3d_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        V = C.POINTER(_abi.Vector)
        L.contour3d_execute.restype = C.c_int
        L.contour3d_execute.argtypes = [C.c_void_p, V, V, V, C.POINTER(C.c_size_t), V, V,
                                        C.POINTER(_abi.Color), C.c_size_t, C.POINTER(_abi.ContourObj),
                                        C.c_size_t, C.POINTER(_abi.LineObj), C.c_char_p]
        L.contour3d_amd_execute_rgb.restype = C.c_int
        L.contour3d_amd_execute_rgb.argtypes = [V, V, C.POINTER(C.c_size_t), V, V, C.POINTER(_abi.Color),
                                                C.c_size_t, C.POINTER(_abi.ContourObj), C.c_size_t,
                                                C.POINTER(_abi.LineObj), C.c_void_p, C.c_void_p, C.c_int]
        L.contour3d_amd_execute_rgb_noalloc.restype = C.c_int
        L.contour3d_amd_execute_rgb_noalloc.argtypes = L.contour3d_amd_execute_rgb.argtypes
        L.contour3d_amd_execute_batch.restype = C.c_int
        L.contour3d_amd_execute_batch.argtypes = [C.c_size_t, C.POINTER(_abi.Image), C.c_int]
        L.contour3d_amd_tune.argtypes = [C.c_int, C.c_int64]
        L.contour3d_amd_server_start.argtypes = [C.POINTER(_abi.ServerOpts)]
        L.contour3d_amd_server_submit.argtypes = [C.POINTER(_abi.Image), C.c_int, C.POINTER(C.c_uint64)]
        L.contour3d_amd_server_submit_ppm.argtypes = [C.POINTER(_abi.Image), C.c_int, C.c_char_p,
                                                      C.POINTER(C.c_uint64)]
        L.contour3d_amd_server_wait.argtypes = [C.c_uint64, C.POINTER(C.c_double)]
        L.contour3d_amd_server_stats.argtypes = [C.POINTER(_abi.ServerStats)]
        L.contour3d_amd_init.argtypes = [C.c_int]
        L.contour3d_amd_set_precision.argtypes = [C.c_int]
        L.contour3d_amd_mesh_build.argtypes = [C.POINTER(_abi.ContourObj), C.c_size_t, C.c_int,
                                               C.POINTER(C.c_void_p)]
        L.contour3d_amd_mesh_count.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.contour3d_amd_mesh_copy.argtypes = [C.c_void_p, C.c_size_t] + [C.c_void_p] * 5
        L.contour3d_amd_mesh_render.argtypes = [C.c_void_p, V, V, C.POINTER(C.c_size_t), V, V,
                                                C.POINTER(_abi.Color), C.c_size_t, C.POINTER(_abi.LineObj),
                                                C.c_void_p, C.c_void_p]
        L.contour3d_amd_mesh_free.argtypes = [C.c_void_p]
        L.contour3d_amd_mesh_compare.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        L.contour3d_amd_mesh_normals.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
        L.contour3d_amd_profile_enable.argtypes = [C.c_int]
        L.contour3d_amd_profile_read.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


_IDENTITY = None


def identity_address():
    """Address of contour3d_amd_identity (looked up once per process)."""
    global _IDENTITY
    if _IDENTITY is None:
        _IDENTITY = C.cast(lib().contour3d_amd_identity, C.c_void_p).value
    return _IDENTITY


def _addr(a):
    """Data address of a numpy array or a torch tensor."""
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def _shape(a):
    return tuple(a.shape)  # ints for numpy arrays and torch tensors alike


class _Keep:
    """Holds the buffers that ctypes structs point into."""

    def __init__(self):
        self.items = []


_UNIT_GRIDS = {}


def _unit_grid(m):
    """The cell-centred grid of m points on [-0.5, 0.5] (src/main.c:48-54),
    one read-only array per size."""
    g = _UNIT_GRIDS.get(m)
    if g is None:
        g = -0.5 + (np.arange(m) + 0.5) / m
        g.setflags(write=False)
        _UNIT_GRIDS[m] = g
    return g


def _grid_address(g, keep, seen):
    """Address of grid g as contiguous float64; `seen` maps id(g) to it for
    the grids of one call (contours usually share their grid arrays, and
    `keep` holds every array for the call's duration, so no id is reused)."""
    a = seen.get(id(g))
    if a is None:
        h = np.ascontiguousarray(g, dtype=np.float64)
        keep.items += [g, h]
        a = h.__array_interface__["data"][0]
        seen[id(g)] = a
    return a


def contour_objs(contours, keep):
    """ctypes array of contour3d_contour_obj_t from dicts with keys field
    (nz, ny, nx), threshold, color and optional grids / converter."""
    ident = identity_address()
    objs = (_abi.ContourObj * max(1, len(contours)))()
    seen = {}
    for n, c in enumerate(contours):
        f = c["field"]
        if isinstance(f, np.ndarray):
            f = np.ascontiguousarray(f, dtype=np.float64)
        elif not f.is_contiguous():
            raise ValueError("device fields must be contiguous")
        nz, ny, nx = _shape(f)
        grids = c.get("grids")
        if grids is None:
            grids = tuple(_unit_grid(m) for m in (nx, ny, nz))
        keep.items.append(f)
        conv = c.get("converter", ident)
        if not isinstance(conv, int):
            keep.items.append(conv)
            conv = C.cast(conv, C.c_void_p).value
        o = objs[n]
        o.glsizes = (nx, ny, nz)
        o.grids = tuple(_grid_address(g, keep, seen) for g in grids)
        o.converter = conv
        o.threshold = float(c["threshold"])
        o.color = _abi.color(c["color"])
        o.array = _addr(f)
    return objs


def line_objs(lines, keep):
    ident = identity_address()
    objs = (_abi.LineObj * max(1, len(lines)))()
    for n, ln in enumerate(lines):
        conv = ln.get("converter", ident)
        if not isinstance(conv, int):
            keep.items.append(conv)
            conv = C.cast(conv, C.c_void_p).value
        o = objs[n]
        o.edges[0] = _abi.vec(ln["edges"][0])
        o.edges[1] = _abi.vec(ln["edges"][1])
        o.nitems = int(ln["nitems"])
        o.converter = conv
        o.color = _abi.color(ln["color"])
        o.width = float(ln["width"])
    return objs


def _view_args(camera, light, size, center, local, bg):
    sizes = (C.c_size_t * 2)(int(size[0]), int(size[1]))
    loc = (_abi.Vector * 2)(_abi.vec(local[0]), _abi.vec(local[1]))
    return (C.byref(_abi.vec(camera)), C.byref(_abi.vec(light)), sizes, C.byref(_abi.vec(center)), loc,
            C.byref(_abi.color(bg)))


def execute(camera, look_at, light, size, center, local, bg, contours, lines=(), fname="output.ppm"):
    """contour3d_execute: render and write a binary PPM to fname."""
    keep = _Keep()
    objs = contour_objs(contours, keep)
    lobjs = line_objs(lines, keep)
    sizes = (C.c_size_t * 2)(int(size[0]), int(size[1]))
    loc = (_abi.Vector * 2)(_abi.vec(local[0]), _abi.vec(local[1]))
    rc = lib().contour3d_execute(None, C.byref(_abi.vec(camera)), C.byref(_abi.vec(look_at)),
                                 C.byref(_abi.vec(light)), sizes, C.byref(_abi.vec(center)), loc,
                                 C.byref(_abi.color(bg)), len(contours), objs, len(lines), lobjs,
                                 str(fname).encode())
    if rc != 0:
        raise RuntimeError("contour3d_execute failed")


def execute_rgb(camera, light, size, center, local, bg, contours, lines=(), precision=FP64,
                rgb=None, depth=None, want_depth=False):
    """Render into rgb (numpy (H, W, 3) uint8 or a device tensor) and
    optionally depth ((H, W) float64).  Returns (rgb, depth)."""
    W, H = int(size[0]), int(size[1])
    keep = _Keep()
    objs = contour_objs(contours, keep)
    lobjs = line_objs(lines, keep)
    if rgb is None:
        rgb = np.zeros((H, W, 3), np.uint8)
    if depth is None and want_depth:
        depth = np.zeros((H, W), np.float64)
    rc = lib().contour3d_amd_execute_rgb(*_view_args(camera, light, size, center, local, bg),
                                         len(contours), objs, len(lines), lobjs, _addr(rgb),
                                         None if depth is None else _addr(depth), int(precision))
    if rc != 0:
        raise RuntimeError("contour3d_amd_execute_rgb failed")
    return rgb, depth


def execute_batch(images, precision=FP64):
    """Render several images in one frame (contour3d_amd_execute_batch).

    `images` is a list of dicts with the keyword arguments of execute_rgb
    (camera, light, size, center, local, bg, contours, lines, rgb, depth,
    want_depth).  Images whose `contours` list is the same Python object
    share its contour structs, so the library sees equal lists and
    tessellates them once.  Returns a list of (rgb, depth)."""
    keep = _Keep()
    arr = (_abi.Image * max(1, len(images)))()
    made = {}
    out = [_fill_image(arr[n], im, keep, made) for n, im in enumerate(images)]
    if lib().contour3d_amd_execute_batch(len(images), arr, int(precision)) != 0:
        raise RuntimeError("contour3d_amd_execute_batch failed")
    return out


def _fill_image(e, im, keep, made):
    """Fills contour3d_amd_image_t `e` from an execute_batch dict; returns
    its (rgb, depth) outputs.  `made` maps id(contours list) to the ctypes
    array built for that list; the entry keeps the list (so its id is not
    reused) and the buffers the array points into."""
    W, H = int(im["size"][0]), int(im["size"][1])
    cs = im.get("contours", ())
    key = id(cs)
    if key not in made or made[key][3] is not cs:
        ck = _Keep()
        made[key] = (contour_objs(cs, ck), len(cs), ck, cs)
    objs, nc = made[key][0], made[key][1]
    lines = im.get("lines", ())
    lobjs = line_objs(lines, keep)
    keep.items += [objs, lobjs]
    rgb = im.get("rgb")
    if rgb is None:
        rgb = np.zeros((H, W, 3), np.uint8)
    depth = im.get("depth")
    if depth is None and im.get("want_depth"):
        depth = np.zeros((H, W), np.float64)
    e.camera_position = _abi.vec(im["camera"])
    e.light_direction = _abi.vec(im["light"])
    e.screen_sizes[:] = [W, H]
    e.screen_center = _abi.vec(im["center"])
    e.screen_local[0] = _abi.vec(im["local"][0])
    e.screen_local[1] = _abi.vec(im["local"][1])
    e.bg_color = _abi.color(im["bg"])
    e.num_contours = nc
    e.contour_objs = C.cast(objs, C.POINTER(_abi.ContourObj))
    e.num_lines = len(lines)
    e.line_objs = C.cast(lobjs, C.POINTER(_abi.LineObj))
    e.rgb = _addr(rgb)
    e.depth = None if depth is None else _addr(depth)
    keep.items += [rgb, depth]
    return rgb, depth


class Server:
    """The request loop of contour3d_amd_server_*: submit() queues one image
    (an execute_batch dict) and returns a ticket; wait(ticket) returns its
    (rgb, depth, latency_ms).  The buffers behind a request are held here
    until it has been waited for."""

    def __init__(self, max_images=16, hbm_budget=200e9, max_wait_ms=0.0, snapshot_fields=False):
        o = _abi.ServerOpts(int(max_images), int(hbm_budget), float(max_wait_ms), 1 if snapshot_fields else 0)
        if lib().contour3d_amd_server_start(C.byref(o)) != 0:
            raise RuntimeError("contour3d_amd_server_start failed")
        self._pending = {}

    def submit(self, image, precision=FP64):
        keep = _Keep()
        e = _abi.Image()
        made = {}
        out = _fill_image(e, image, keep, made)
        keep.items.append(made)
        t = C.c_uint64()
        if lib().contour3d_amd_server_submit(C.byref(e), int(precision), C.byref(t)) != 0:
            raise RuntimeError("contour3d_amd_server_submit failed")
        self._pending[t.value] = (e, keep, out)
        return t.value

    def submit_ppm(self, image, fname, precision=FP64):
        """As submit, but the server writes the image to `fname` as a binary
        PPM (the bytes contour3d_execute writes) on its writer thread, while
        it renders the next frames; wait(ticket) returns (None, depth,
        latency_ms) once the file is written."""
        keep = _Keep()
        e = _abi.Image()
        made = {}
        # the server renders into a pinned image of its own
        out = _fill_image(e, dict(image, rgb=np.zeros((1, 1, 3), np.uint8)), keep, made)
        keep.items.append(made)
        path = os.fsencode(fname)
        keep.items.append(path)
        t = C.c_uint64()
        if lib().contour3d_amd_server_submit_ppm(C.byref(e), int(precision), C.c_char_p(path), C.byref(t)) != 0:
            raise RuntimeError("contour3d_amd_server_submit_ppm failed")
        self._pending[t.value] = (e, keep, (None, out[1]))
        return t.value

    def wait(self, ticket):
        lat = C.c_double()
        rc = lib().contour3d_amd_server_wait(int(ticket), C.byref(lat))
        e, keep, out = self._pending.pop(ticket)
        if rc != 0:
            raise RuntimeError(f"request {ticket} failed")
        return out[0], out[1], lat.value

    def stats(self):
        st = _abi.ServerStats()
        if lib().contour3d_amd_server_stats(C.byref(st)) != 0:
            raise RuntimeError("contour3d_amd_server_stats failed")
        return {f: getattr(st, f) for f, _ in _abi.ServerStats._fields_}

    def stop(self):
        lib().contour3d_amd_server_stop()
        self._pending.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()


class Mesh:
    """Tessellated and smoothed contour objects held on the device."""

    def __init__(self, contours, precision=FP64):
        self._keep = _Keep()
        objs = contour_objs(contours, self._keep)
        h = C.c_void_p()
        if lib().contour3d_amd_mesh_build(objs, len(contours), int(precision), C.byref(h)) != 0:
            raise RuntimeError("contour3d_amd_mesh_build failed")
        self.handle = h
        self.n = len(contours)
        self.precision = int(precision)

    def count(self, obj):
        n = C.c_size_t()
        if lib().contour3d_amd_mesh_count(self.handle, obj, C.byref(n)) != 0:
            raise RuntimeError("mesh_count failed")
        return n.value

    def copy(self, obj):
        T = self.count(obj)
        out = dict(pos=np.zeros((T, 3, 3)), wfn=np.zeros((T, 3)), vn=np.zeros((T, 3, 3)),
                   cell=np.zeros(T, np.uint32), edges=np.zeros(T, np.uint32))
        if T and lib().contour3d_amd_mesh_copy(self.handle, obj, *(out[k].ctypes.data for k in
                                                                   ("pos", "wfn", "vn", "cell", "edges"))) != 0:
            raise RuntimeError("mesh_copy failed")
        e = out["edges"]
        out["edge"] = np.stack([e & 31, (e >> 5) & 31, (e >> 10) & 31], axis=1).astype(np.int32)
        out["interior"] = ((e >> 15) & 1).astype(bool)
        return out

    def render(self, camera, light, size, center, local, bg, lines=(), rgb=None, depth=None, want_depth=False):
        W, H = int(size[0]), int(size[1])
        keep = _Keep()
        lobjs = line_objs(lines, keep)
        if rgb is None:
            rgb = np.zeros((H, W, 3), np.uint8)
        if depth is None and want_depth:
            depth = np.zeros((H, W), np.float64)
        rc = lib().contour3d_amd_mesh_render(self.handle, *_view_args(camera, light, size, center, local, bg),
                                             len(lines), lobjs, _addr(rgb),
                                             None if depth is None else _addr(depth))
        if rc != 0:
            raise RuntimeError("mesh_render failed")
        return rgb, depth

    def normals(self, obj, out=None):
        """contour3d_amd_mesh_normals: the unit vertex normals of every
        triangle of object `obj`, shape (T, 3, 3), in the mesh's precision;
        `out` may be a device tensor (filled in place)."""
        T = self.count(obj)
        if out is None:
            out = np.zeros((T, 3, 3), np.float64 if self.precision == FP64 else np.float32)
        if T and lib().contour3d_amd_mesh_normals(self.handle, obj, _addr(out)) != 0:
            raise RuntimeError("mesh_normals failed")
        return out

    def compare(self, other, obj):
        """contour3d_amd_mesh_compare of this FP64 mesh with `other`."""
        st = np.zeros(8, np.float64)
        if lib().contour3d_amd_mesh_compare(self.handle, other.handle, obj, st.ctypes.data) != 0:
            raise RuntimeError("mesh_compare failed")
        return dict(triangles=int(st[0]), topology_mismatches=int(st[1]), max_vertex_err=st[2],
                    max_normal_angle_deg=st[3], mean_normal_angle_deg=st[4], normals=int(st[5]),
                    normals_over_1deg=int(st[6]), normals_over_10deg=int(st[7]))

    def free(self):
        if self.handle:
            lib().contour3d_amd_mesh_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def profile_enable(on=True):
    lib().contour3d_amd_profile_enable(1 if on else 0)


def profile_reset():
    lib().contour3d_amd_profile_reset()


def profile_read():
    ms = np.zeros(len(STAGES), np.float64)
    calls = np.zeros(len(STAGES), np.uint64)
    nt = C.c_uint64()
    lib().contour3d_amd_profile_read(ms.ctypes.data, calls.ctypes.data, C.byref(nt))
    return dict(zip(STAGES, ms)), dict(zip(STAGES, calls)), nt.value


TUNE_TESS_GRID = 0
TUNE_RASTER_GRID = 1
TUNE_FRAG_CAPACITY = 2
TUNE_GRAPHS = 3
TUNE_FAIL_ALLOC = 4
TUNE_TESS_ROWS = 5
TUNE_TESS_ZPLANES = 6


def tune(param, value):
    """contour3d_amd_tune: launch-shape and capacity parameters (results do
    not depend on them)."""
    if lib().contour3d_amd_tune(int(param), int(value)) != 0:
        raise RuntimeError("contour3d_amd_tune failed")


AUTOTUNE_KNOBS = {
    TUNE_TESS_GRID: (0, 16384, 1 << 30),   # persistent workgroups (0: 128 per CU), or one per work item
    TUNE_TESS_ROWS: (0, 2, 1),             # rows of cells per tile (0: as many as fit in 2048 cells)
    TUNE_TESS_ZPLANES: (2, 4, 8, 16),      # z-planes per work item (4 by default)
}
KNOB_NAMES = {TUNE_TESS_GRID: "tess_grid", TUNE_TESS_ROWS: "tess_rows", TUNE_TESS_ZPLANES: "tess_zplanes"}


def autotune(run, knobs=None, reps=2):
    """Times run() under every combination of the launch-shape knobs (a dict
    tune parameter -> values; default AUTOTUNE_KNOBS: the tessellation
    workgroup cap and brick shape, which the frame path reads) and keeps the
    fastest.  Every setting renders the same bits.  Returns (best settings
    by knob name, best ms, the sweep as a list of (settings, ms))."""
    import itertools
    import time

    import torch

    knobs = AUTOTUNE_KNOBS if knobs is None else knobs
    params = list(knobs)
    combos = list(itertools.product(*(knobs[p] for p in params)))
    sweep = []
    for values in combos:
        for p, v in zip(params, values):
            tune(p, v)
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        sweep.append(({KNOB_NAMES.get(p, str(p)): v for p, v in zip(params, values)}, ms))
    best = min(range(len(sweep)), key=lambda q: sweep[q][1])
    for p, v in zip(params, combos[best]):
        tune(p, v)
    return sweep[best][0], sweep[best][1], sweep


def set_precision(p):
    if lib().contour3d_amd_set_precision(int(p)) != 0:
        raise RuntimeError("bad precision")


def finalize():
    lib().contour3d_amd_finalize()
