"""Pins the oracle's vertex-normal smoothing (normal.c) against a brute-force
global formulation: a vertex on lattice edge E gets the normalised sum of
area * face normal over every triangle of the mesh with a vertex on E, in
serial (cell, slot) order.  This checks the 19-case neighbour table, its
cell order and the per-cell matching, which the oracle restates."""
import numpy as np
import pytest

from test_oracle_tess import global_key, sphere_field


def brute_force_vertex_normals(m, shape):
    nz, ny, nx = shape
    ncx, ncy = nx - 1, ny - 1
    T = len(m["area"])
    keys = []
    for t in range(T):
        cid = int(m["cell"][t])
        cell = (cid % ncx, (cid // ncx) % ncy, cid // (ncx * ncy))
        keys.append([global_key(cell, int(e)) for e in m["edge"][t]])
    members = {}
    for t in range(T):          # serial order
        for k in set(keys[t]):
            members.setdefault(k, []).append(t)
    out = np.full((T, 3, 3), np.nan)
    for t in range(T):
        if not m["interior"][t]:
            continue
        for v in range(3):
            sx = sy = sz = 0.0
            for u in members[keys[t][v]]:
                a = float(m["area"][u])
                fx, fy, fz = (float(x) for x in m["fn"][u])
                sx = sx + a * fx
                sy = sy + a * fy
                sz = sz + a * fz
            inv = 1.0 / np.sqrt(sx * sx + sy * sy + sz * sz)
            out[t, v] = (sx * inv, sy * inv, sz * inv)
    return out


@pytest.mark.parametrize("shape,seed,thr", [((7, 8, 9), 0, 0.0), ((9, 6, 7), 1, 0.4),
                                            ((6, 6, 6), 2, -0.3)])
def test_vertex_normals_match_global_edge_sum(orc, shape, seed, thr):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape)
    m = orc.mesh(f, thr)
    assert m["interior"].sum() > 20
    bf = brute_force_vertex_normals(m, shape)
    got = m["vn"][m["interior"]]
    exp = bf[m["interior"]]
    assert np.array_equal(got.view(np.uint64), exp.view(np.uint64))
    assert np.all(np.isnan(m["vn"][~m["interior"]]))


def test_plane_vertex_normals_equal_face_normal(orc):
    rng = np.random.default_rng(7)
    a = rng.standard_normal(3)
    g = -0.5 + (np.arange(10) + 0.5) / 10
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    f = a[0] * x + a[1] * y + a[2] * z
    m = orc.mesh(f, 0.05)
    na = -a / np.linalg.norm(a)
    vn = m["vn"][m["interior"]].reshape(-1, 3)
    assert len(vn) > 0
    assert np.max(np.abs(vn - na)) < 1e-12


def test_sphere_vertex_normals_radial(orc):
    """On a sphere the smoothed normals converge to the radial direction at
    first order in h and beat the raw face normals."""
    c = np.array([0.011, -0.023, 0.017])
    means = []
    for n in (17, 33):
        f, g = sphere_field(n, c, 0.33)
        m = orc.mesh(f, 0.0, grids=(g, g, g))
        ins = m["interior"]
        p = m["pos"][ins].reshape(-1, 3)
        vn = m["vn"][ins].reshape(-1, 3)
        radial = -(p - c) / np.linalg.norm(p - c, axis=1)[:, None]
        ang = np.degrees(np.arccos(np.clip(np.sum(vn * radial, axis=1), -1, 1)))
        fn = np.repeat(m["fn"][ins], 3, axis=0)
        angf = np.degrees(np.arccos(np.clip(np.sum(fn * radial, axis=1), -1, 1)))
        assert ang.mean() < 0.8 * angf.mean()
        assert ang.max() < 10.0
        means.append(ang.mean())
    assert means[0] / means[1] > 1.7
