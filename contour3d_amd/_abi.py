"""ctypes mirrors of the public C structs in include/contour3d.h and
include/contour3d_amd.h.

The layouts of contour3d.h follow the reference interface member for member
(include/contour3d.h:10-56); tests/test_abi.py checks sizes and offsets
against the C compiler's view of both headers.
"""
import ctypes as C


class Vector(C.Structure):
    """contour3d_vector_t, include/contour3d.h:10-14."""

    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]

    def __iter__(self):
        return iter((self.x, self.y, self.z))


class Color(C.Structure):
    """contour3d_color_t, include/contour3d.h:16-20."""

    _fields_ = [("r", C.c_uint8), ("g", C.c_uint8), ("b", C.c_uint8)]


# contour3d_vector_t (*)(const contour3d_vector_t)
Converter = C.CFUNCTYPE(Vector, Vector)


class ContourObj(C.Structure):
    """contour3d_contour_obj_t, include/contour3d.h:22-40."""

    _fields_ = [
        ("pencil", C.c_int),
        ("glsizes", C.c_size_t * 3),
        ("grids", C.c_void_p * 3),  # const double* grids[3], set as addresses
        ("converter", C.c_void_p),
        ("threshold", C.c_double),
        ("color", Color),
        ("array", C.c_void_p),
    ]


class LineObj(C.Structure):
    """contour3d_line_obj_t, include/contour3d.h:42-56."""

    _fields_ = [
        ("edges", Vector * 2),
        ("nitems", C.c_size_t),
        ("converter", C.c_void_p),
        ("color", Color),
        ("width", C.c_double),
    ]


class Image(C.Structure):
    """contour3d_amd_image_t (no reference counterpart): the arguments of
    one contour3d_amd_execute_rgb call, one entry of a batch."""

    _fields_ = [
        ("camera_position", Vector),
        ("light_direction", Vector),
        ("screen_sizes", C.c_size_t * 2),
        ("screen_center", Vector),
        ("screen_local", Vector * 2),
        ("bg_color", Color),
        ("num_contours", C.c_size_t),
        ("contour_objs", C.POINTER(ContourObj)),
        ("num_lines", C.c_size_t),
        ("line_objs", C.POINTER(LineObj)),
        ("rgb", C.c_void_p),
        ("depth", C.c_void_p),
    ]


class ServerOpts(C.Structure):
    """contour3d_amd_server_opts_t (no reference counterpart)."""

    _fields_ = [("max_images", C.c_size_t), ("hbm_budget", C.c_uint64), ("max_wait_ms", C.c_double),
                ("snapshot_fields", C.c_int)]


class ServerStats(C.Structure):
    """contour3d_amd_server_stats_t (no reference counterpart)."""

    _fields_ = [
        ("completed", C.c_uint64),
        ("failed", C.c_uint64),
        ("batches", C.c_uint64),
        ("mean_batch", C.c_double),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("max_ms", C.c_double),
        ("frames_per_s", C.c_double),
        ("peak_admitted_bytes", C.c_uint64),
    ]


def vec(v):
    return Vector(float(v[0]), float(v[1]), float(v[2]))


def color(c):
    return Color(int(c[0]), int(c[1]), int(c[2]))


