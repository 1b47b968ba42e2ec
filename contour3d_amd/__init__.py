"""MI355X-native Contour3D tessellate -> smooth -> render path.

The compute path is the HIP library contour3d_amd/lib/libcontour3d.so behind
the C API of include/contour3d.h (+ include/contour3d_amd.h).  This package
is a thin ctypes front end that mirrors the reference's single entry point
contour3d_execute (include/contour3d.h:58-86); it loads the native library
lazily and raises if it is missing -- there is no CPU fallback.
"""
