"""ctypes front end of the CPU oracle (oracle/contour3d_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product package
never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from contour3d_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, -std=c99, no FMA)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} missing (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.oracle_mesh.restype = C.c_long
        L.oracle_mesh.argtypes = [C.c_size_t] * 3 + [C.c_void_p] * 7 + [C.c_double] + \
            [C.c_void_p, C.c_void_p, C.c_long] + [C.c_void_p] * 8
        L.oracle_execute.restype = C.c_int
        L.oracle_execute.argtypes = [
            C.POINTER(_abi.Vector), C.POINTER(_abi.Vector), C.POINTER(C.c_size_t),
            C.POINTER(_abi.Vector), C.POINTER(_abi.Vector), C.POINTER(_abi.Color),
            C.c_size_t, C.POINTER(_abi.ContourObj), C.c_size_t, C.POINTER(_abi.LineObj),
            C.c_void_p, C.c_void_p, C.c_char_p]
        L.oracle_project.restype = C.c_long
        L.oracle_project.argtypes = [dp, C.c_size_t, dp, dp]
        L.oracle_raster.restype = None
        L.oracle_raster.argtypes = [dp, C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t,
                                    dp, dp, C.c_void_p, C.c_void_p, dp]
        L.oracle_neighbour_table.restype = None
        L.oracle_neighbour_table.argtypes = [C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def identity_converter_address():
    return C.cast(lib().oracle_identity_converter, C.c_void_p).value


def converter_address(name):
    """Address of one of the oracle's test coordinate maps (cylinder, shear, late)."""
    return C.cast(getattr(lib(), "oracle_conv_" + name), C.c_void_p).value


def _ptr(a):
    return None if a is None else a.ctypes.data


def mesh(field, thr, grids=None, coords=None):
    """Tessellate + vertex normals over the whole field (one threshold).

    field: (nz, ny, nx) float64, x fastest.  grids: (gx, gy, gz) 1-D arrays
    (default: cell-centred uniform grid on [-0.5, 0.5]).  coords: optional
    (cx, cy, cz) lattice coordinate tables of the field's shape.
    Returns a dict of numpy arrays in serial (cell, slot) order.
    """
    field = np.ascontiguousarray(field, dtype=np.float64)
    nz, ny, nx = field.shape
    if grids is None:
        grids = tuple(-0.5 + (np.arange(n) + 0.5) / n for n in (nx, ny, nz))
    gx, gy, gz = (np.ascontiguousarray(g, dtype=np.float64) for g in grids)
    if coords is not None:
        cx, cy, cz = (np.ascontiguousarray(c, dtype=np.float64) for c in coords)
    else:
        cx = cy = cz = None
    ncell = (nx - 1) * (ny - 1) * (nz - 1)
    codes = np.zeros(ncell, np.uint8)
    counts = np.zeros(ncell, np.uint8)
    L = lib()
    T = L.oracle_mesh(nx, ny, nz, _ptr(gx), _ptr(gy), _ptr(gz), _ptr(cx), _ptr(cy), _ptr(cz),
                      _ptr(field), float(thr), _ptr(codes), _ptr(counts), 0,
                      None, None, None, None, None, None, None, None)
    T = int(counts.sum(dtype=np.int64))
    out = dict(
        pos=np.zeros((T, 3, 3)), fn=np.zeros((T, 3)), area=np.zeros(T), vn=np.zeros((T, 3, 3)),
        edge=np.zeros((T, 3), np.int32), cell=np.zeros(T, np.int64), slot=np.zeros(T, np.int32),
        interior=np.zeros(T, np.uint8))
    r = L.oracle_mesh(nx, ny, nz, _ptr(gx), _ptr(gy), _ptr(gz), _ptr(cx), _ptr(cy), _ptr(cz),
                      _ptr(field), float(thr), _ptr(codes), _ptr(counts), T,
                      _ptr(out["pos"]), _ptr(out["fn"]), _ptr(out["area"]), _ptr(out["vn"]),
                      _ptr(out["edge"]), _ptr(out["cell"]), _ptr(out["slot"]),
                      _ptr(out["interior"]))
    if r != T:
        raise RuntimeError(f"oracle_mesh returned {r}, expected {T}")
    out["codes"] = codes.reshape(nz - 1, ny - 1, nx - 1)
    out["counts"] = counts.reshape(nz - 1, ny - 1, nx - 1)
    out["interior"] = out["interior"].astype(bool)
    return out


def execute(camera, light, size, center, local, bg, contours, lines=(), fname=None):
    """Full contour3d_execute restatement on one rank.

    contours: sequence of dicts with keys field (nz,ny,nx), threshold, color,
    and optionally grids (3 arrays) and converter (a ctypes function pointer
    or an integer address; default the identity).
    lines: sequence of dicts with keys edges ((3,),(3,)), nitems, color,
    width, and optionally converter.
    Returns (rgb (H, W, 3) uint8 in PPM row order, depth (H, W) row j = y j).
    """
    W, H = int(size[0]), int(size[1])
    keep = []
    ident = identity_converter_address()
    objs = (_abi.ContourObj * max(1, len(contours)))()
    for n, c in enumerate(contours):
        f = np.ascontiguousarray(c["field"], dtype=np.float64)
        nz, ny, nx = f.shape
        grids = c.get("grids")
        if grids is None:
            grids = tuple(-0.5 + (np.arange(m) + 0.5) / m for m in (nx, ny, nz))
        grids = [np.ascontiguousarray(g, dtype=np.float64) for g in grids]
        keep += [f] + grids
        conv = c.get("converter", ident)
        if not isinstance(conv, int):
            keep.append(conv)
            conv = C.cast(conv, C.c_void_p).value
        o = objs[n]
        o.pencil = 0
        o.glsizes[:] = [nx, ny, nz]
        for d in range(3):
            o.grids[d] = grids[d].ctypes.data
        o.converter = conv
        o.threshold = float(c["threshold"])
        o.color = _abi.color(c["color"])
        o.array = f.ctypes.data
    lobjs = (_abi.LineObj * max(1, len(lines)))()
    for n, ln in enumerate(lines):
        conv = ln.get("converter", ident)
        if not isinstance(conv, int):
            keep.append(conv)
            conv = C.cast(conv, C.c_void_p).value
        o = lobjs[n]
        o.edges[0] = _abi.vec(ln["edges"][0])
        o.edges[1] = _abi.vec(ln["edges"][1])
        o.nitems = int(ln["nitems"])
        o.converter = conv
        o.color = _abi.color(ln["color"])
        o.width = float(ln["width"])
    rgb = np.zeros((H, W, 3), np.uint8)
    depth = np.zeros((H, W), np.float64)
    sizes = (C.c_size_t * 2)(W, H)
    loc = (_abi.Vector * 2)(_abi.vec(local[0]), _abi.vec(local[1]))
    rc = lib().oracle_execute(
        C.byref(_abi.vec(camera)), C.byref(_abi.vec(light)), sizes, C.byref(_abi.vec(center)),
        loc, C.byref(_abi.color(bg)), len(contours), objs, len(lines), lobjs,
        rgb.ctypes.data, depth.ctypes.data, None if fname is None else str(fname).encode())
    if rc != 0:
        raise RuntimeError("oracle_execute failed")
    return rgb, depth


def project(view, pts):
    """Project (n,3) points; view = cam, center, lx, ly (12 doubles)."""
    v = np.ascontiguousarray(view, dtype=np.float64)
    p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(p)
    dp = C.POINTER(C.c_double)
    lib().oracle_project(v.ctypes.data_as(dp), len(p), p.ctypes.data_as(dp), out.ctypes.data_as(dp))
    return out


def raster(view, W, H, bg, pos, vn, colors):
    """Raster triangles in order; view = cam, center, lx, ly, light (15).

    Returns rgb (H, W, 3) and depth (H, W), row j = y index j (unflipped).
    """
    v = np.ascontiguousarray(view, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 9)
    vn = np.ascontiguousarray(vn, dtype=np.float64).reshape(-1, 9)
    cols = np.ascontiguousarray(colors, dtype=np.uint8).reshape(-1, 3)
    bgc = np.ascontiguousarray(bg, dtype=np.uint8)
    rgb = np.zeros((H, W, 3), np.uint8)
    depth = np.zeros((H, W), np.float64)
    dp = C.POINTER(C.c_double)
    lib().oracle_raster(v.ctypes.data_as(dp), W, H, bgc.ctypes.data, len(pos),
                        pos.ctypes.data_as(dp), vn.ctypes.data_as(dp), cols.ctypes.data,
                        rgb.ctypes.data, depth.ctypes.data_as(dp))
    return rgb, depth


def neighbour_table():
    out = np.zeros((19, 4, 4), np.int32)
    n = np.zeros(19, np.int32)
    lib().oracle_neighbour_table(out.ctypes.data, n.ctypes.data)
    return [[tuple(out[e, q]) for q in range(n[e])] for e in range(19)]
