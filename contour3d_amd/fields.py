"""Seeded synthetic fields and cameras for the benchmark workloads.

Every generator returns a float64 array of shape (nz, ny, nx), x fastest,
which is the memory order the reference indexes with (k*ny + j)*nx + i
(src/contour3d/contour/extend.c:203).  Grids are cell-centred and uniform
on [-0.5, 0.5] like the example's y and z grids (src/main.c:48-54), so the
example camera (src/main.c:127-166) frames every cubic field unchanged.
"""
import ctypes
import ctypes.util
import math

import numpy as np


def uniform_grid(n):
    """x_m = -0.5 + (m + 0.5)/n, src/main.c:48-54 with unit length."""
    return -0.5 + (np.arange(n, dtype=np.float64) + 0.5) / n


def grids_for(shape):
    nz, ny, nx = shape
    return uniform_grid(nx), uniform_grid(ny), uniform_grid(nz)


def blobs(shape, seed=0, nblobs=8):
    """Sum of isotropic Gaussian blobs.

    Centres U[0.2, 0.8]^3 in unit-cube coordinates (shifted by -0.5),
    sigma U[0.05, 0.15], amplitude U[0.5, 1.5], drawn in that order from
    numpy's PCG64 generator.
    """
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.2, 0.8, size=(nblobs, 3)) - 0.5
    s = rng.uniform(0.05, 0.15, size=nblobs)
    a = rng.uniform(0.5, 1.5, size=nblobs)
    gx, gy, gz = grids_for(shape)
    f = np.zeros(shape, dtype=np.float64)
    for b in range(nblobs):
        ex = np.exp(-((gx - c[b, 0]) ** 2) / (2 * s[b] ** 2))
        ey = np.exp(-((gy - c[b, 1]) ** 2) / (2 * s[b] ** 2))
        ez = np.exp(-((gz - c[b, 2]) ** 2) / (2 * s[b] ** 2))
        f += a[b] * ez[:, None, None] * ey[None, :, None] * ex[None, None, :]
    return f


def _spectrum(shape, kmax, seed, kmin=1):
    """Hermitian-half spectrum for irfftn: modes kmin <= |k| <= kmax on the
    unit-periodic box, amplitude |k|^(-11/6) (shell spectrum E(k) ~ k^(-5/3)),
    uniform phases from PCG64(seed) drawn over the whole half-spectrum."""
    nz, ny, nx = shape
    kz = np.fft.fftfreq(nz) * nz
    ky = np.fft.fftfreq(ny) * ny
    kx = np.fft.rfftfreq(nx) * nx
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0.0, 2.0 * math.pi, size=(nz, ny, nx // 2 + 1))
    k2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    k = np.sqrt(k2)
    amp = np.where((k >= kmin) & (k <= kmax), np.power(np.maximum(k, 1e-30), -11.0 / 6.0), 0.0)
    return amp, phase


def turbulence(shape, kmax=64, seed=0, kmin=1):
    """Spectral synthetic turbulence, normalised to mean 0 and variance 1."""
    amp, phase = _spectrum(shape, kmax, seed, kmin)
    spec = amp * np.exp(1j * phase)
    f = np.fft.irfftn(spec, s=shape, axes=(0, 1, 2))
    f -= f.mean()
    f /= f.std()
    return np.ascontiguousarray(f)


def turbulence_device(shape, kmax=64, seed=0, kmin=1, device="cuda"):
    """Same spectrum and phases as turbulence(), transformed on the GPU.

    Used for the large benchmark fields; the inverse FFT rounds differently
    from numpy's, so parity tests use turbulence() bytes instead.
    """
    import torch

    amp, phase = _spectrum(shape, kmax, seed, kmin)
    a = torch.from_numpy(amp).to(device)
    p = torch.from_numpy(phase).to(device)
    del amp, phase
    spec = torch.polar(a, p)
    del a, p
    f = torch.fft.irfftn(spec, s=shape, dim=(0, 1, 2))
    del spec
    f -= f.mean()
    f /= f.std(unbiased=False)
    return f.contiguous()


def blobs_device(shape, seed=0, nblobs=8, device="cuda"):
    """blobs() evaluated on the GPU (same parameters; exp rounds differently
    from numpy's, so parity tests use blobs() bytes)."""
    import torch

    rng = np.random.default_rng(seed)
    c = rng.uniform(0.2, 0.8, size=(nblobs, 3)) - 0.5
    s = rng.uniform(0.05, 0.15, size=nblobs)
    a = rng.uniform(0.5, 1.5, size=nblobs)
    gx, gy, gz = (torch.from_numpy(g).to(device) for g in grids_for(shape))
    f = torch.zeros(shape, dtype=torch.float64, device=device)
    for b in range(nblobs):
        ex = torch.exp(-((gx - c[b, 0]) ** 2) / (2 * s[b] ** 2))
        ey = torch.exp(-((gy - c[b, 1]) ** 2) / (2 * s[b] ** 2))
        ez = torch.exp(-((gz - c[b, 2]) ** 2) / (2 * s[b] ** 2))
        f += a[b] * ez[:, None, None] * ey[None, :, None] * ex[None, None, :]
    return f


def turbulence_fast_device(shape, kmax=64, seed=0, kmin=1, device="cuda"):
    """Spectral turbulence with the spectrum of turbulence() built on the GPU
    and phases from torch's seeded device generator (so the phases differ
    from turbulence()'s PCG64 draws); for generating many large benchmark
    fields quickly."""
    import torch

    nz, ny, nx = shape
    kz = torch.fft.fftfreq(nz, device=device, dtype=torch.float64) * nz
    ky = torch.fft.fftfreq(ny, device=device, dtype=torch.float64) * ny
    kx = torch.fft.rfftfreq(nx, device=device, dtype=torch.float64) * nx
    k = torch.sqrt(kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2)
    amp = torch.where((k >= kmin) & (k <= kmax), torch.clamp(k, min=1e-30) ** (-11.0 / 6.0),
                      torch.zeros((), dtype=torch.float64, device=device))
    del k
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    phase = torch.rand(amp.shape, generator=g, dtype=torch.float64, device=device) * (2 * math.pi)
    spec = torch.polar(amp, phase)
    del amp, phase
    f = torch.fft.irfftn(spec, s=shape, dim=(0, 1, 2))
    del spec
    f -= f.mean()
    f /= f.std(unbiased=False)
    return f.contiguous()


def example_field(n=128):
    """The reference example's field (src/main.c:23-111) on one rank.

    64 products of sines with phases from the C library's rand() in its
    default (unseeded) state, x on a Chebyshev-Gauss grid, y and z uniform,
    then rescaled to [-1, 1].  Returns (field, (gx, gy, gz)).
    """
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.rand.restype = ctypes.c_int
    libc.srand.argtypes = [ctypes.c_uint]
    libc.srand(1)
    RAND_MAX = 2147483647
    pi = 3.1415926535897932
    gx = np.array([-0.5 + 0.5 * (1.0 + math.cos(pi - pi * (m + 0.5) / n)) for m in range(n)])
    gy = np.array([-0.5 + 1.0 / n * 0.5 * (2 * m + 1) for m in range(n)])
    gz = gy.copy()
    f = np.zeros((n, n, n))
    for kz in range(1, 5):
        for ky in range(1, 5):
            for kx in range(1, 5):
                yph = 2.0 * pi * libc.rand() / RAND_MAX
                zph = 2.0 * pi * libc.rand() / RAND_MAX
                sx = np.sin(2.0 * pi * kx * gx)
                sy = np.sin(2.0 * pi * ky * gy + yph)
                sz = np.sin(2.0 * pi * kz * gz + zph)
                f += sz[:, None, None] * sy[None, :, None] * sx[None, None, :]
    lo, hi = f.min(), f.max()
    f = (f - lo) / (hi - lo) * 2.0 - 1.0
    return f, (gx, gy, gz)


# -------------------------------------------------------------------------
# cameras

# src/main.c:127-166
EXAMPLE_CAMERA = dict(
    camera=(8.660254037844386e-1, -1.5, 1.0),
    look_at=(0.0, 0.0, 0.0),
    center=(4.330127018922194e-1, -7.500000000000003e-1, 5.000000000000003e-1),
    local=((1.154700538379251e0, 6.666666666666665e-1, 0.0),
           (-2.5e-1, 4.330127018922195e-1, 8.660254037844386e-1)),
    light=(0.0, 2.0, -1.0),
    bg=(0, 0, 0),
)


def example_view(width, height):
    """Example camera with the horizontal screen vector rescaled so that
    |lx| / |ly| = width / height (square pixels)."""
    v = dict(EXAMPLE_CAMERA)
    lx = np.array(v["local"][0])
    ly = np.array(v["local"][1])
    lx = lx / np.linalg.norm(lx) * np.linalg.norm(ly) * (width / height)
    v["local"] = (tuple(lx), tuple(ly))
    v["size"] = (width, height)
    return v


def _rotate(v, axis, ang):
    """Rodrigues rotation of v about a unit axis."""
    v = np.asarray(v, dtype=np.float64)
    k = np.asarray(axis, dtype=np.float64)
    return v * math.cos(ang) + np.cross(k, v) * math.sin(ang) + k * np.dot(k, v) * (1 - math.cos(ang))


def sphere_view(elevation, azimuth, width, height, focal_camera=3.0, focal_screen=1.5, roll=0.0):
    """A camera on a sphere around the origin looking at it, parametrised by
    elevation/azimuth/roll like CameraAdjuster (CameraAdjuster/src/config.ts:
    140-192): camera at distance focal_camera, screen plane at focal_screen
    from the camera, screen spanning the unit cube's bounding sphere."""
    d = np.array([math.cos(elevation) * math.cos(azimuth),
                  math.cos(elevation) * math.sin(azimuth),
                  math.sin(elevation)])
    cam = focal_camera * d
    ctr = cam - focal_screen * d
    up = np.array([0.0, 0.0, 1.0])
    if abs(np.dot(up, d)) > 0.999:
        up = np.array([0.0, 1.0, 0.0])
    ex = np.cross(up, d)
    ex /= np.linalg.norm(ex)
    ey = np.cross(d, ex)
    ex = _rotate(ex, d, roll)
    ey = _rotate(ey, d, roll)
    # half-extent at the screen plane that covers radius sqrt(3)/2 around 0
    half = focal_screen * math.tan(math.asin(min(0.999, (math.sqrt(3) / 2) / focal_camera))) * 1.05
    ly = ey * (2 * half)
    lx = ex * (2 * half) * (width / height)
    return dict(camera=tuple(cam), look_at=(0.0, 0.0, 0.0), center=tuple(ctr),
                local=(tuple(lx), tuple(ly)), light=tuple(-d + np.array([0.0, 0.0, 0.5])),
                bg=(0, 0, 0), size=(width, height))
