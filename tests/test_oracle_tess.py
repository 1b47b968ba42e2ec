"""Pins the oracle's tessellation (triangulate.c) with checks that do not
restate its formulas: closed-form triangle counts over all 256 corner codes,
exact-rational vertex positions on straddling edges named by the edge ids,
orientation against the gradient of each tetra's affine interpolant, plane
fields, and watertightness / Euler characteristic of closed surfaces."""
from fractions import Fraction

import numpy as np
import pytest

# tetra corner table, triangulate.c:145-152 (corner c = i + 2j + 4k)
TETS = [(0, 7, 3, 1), (0, 7, 1, 5), (0, 7, 5, 4), (0, 7, 4, 6), (0, 7, 6, 2), (0, 7, 2, 3)]
NTRI_BY_POPCOUNT = {0: 0, 1: 1, 2: 2, 3: 1, 4: 0}


def corner_ijk(c):
    return (c & 1, (c >> 1) & 1, c >> 2)


def edge_segment(e):
    """Lattice edge id -> its two cube corners (i, j, k), per the numbering
    documented at normal.c:71-77: x edges 0-3 (j + 2k), y edges 4-7
    (i + 2k), z edges 8-11 (i + 2j), face diagonals 12-17 from the face's
    minimum corner, body diagonal 18."""
    if e < 4:
        j, k = e & 1, e >> 1
        return (0, j, k), (1, j, k)
    if e < 8:
        i, k = (e - 4) & 1, (e - 4) >> 1
        return (i, 0, k), (i, 1, k)
    if e < 12:
        i, j = (e - 8) & 1, (e - 8) >> 1
        return (i, j, 0), (i, j, 1)
    if e < 14:
        x = e - 12
        return (x, 0, 0), (x, 1, 1)
    if e < 16:
        y = e - 14
        return (0, y, 0), (1, y, 1)
    if e < 18:
        z = e - 16
        return (0, 0, z), (1, 1, z)
    return (0, 0, 0), (1, 1, 1)


def global_key(cell_ijk, e):
    a, b = edge_segment(e)
    A = tuple(c + d for c, d in zip(cell_ijk, a))
    B = tuple(c + d for c, d in zip(cell_ijk, b))
    return (A, B) if A < B else (B, A)


def unit_cell_field(code, rng):
    """2x2x2 field: corner c above 0 iff bit c of code; magnitudes random."""
    f = np.zeros((2, 2, 2))
    for c in range(8):
        i, j, k = corner_ijk(c)
        mag = rng.uniform(0.1, 0.9)
        f[k, j, i] = mag if (code >> c) & 1 else -mag
    return f


def tet_gradient(vals, corners):
    """Gradient of the affine interpolant of vals at the 4 corners."""
    P = np.array([corner_ijk(c) for c in corners], dtype=np.float64)
    v = np.array(vals)
    M = P[1:] - P[0]
    return np.linalg.solve(M, v[1:] - v[0])


def test_exhaustive_corner_codes(orc):
    rng = np.random.default_rng(1)
    grids = (np.array([0.0, 1.0]),) * 3
    seen_cases = set()
    for code in range(256):
        f = unit_cell_field(code, rng)
        m = orc.mesh(f, 0.0, grids=grids)
        assert m["codes"][0, 0, 0] == code
        vals = [f[corner_ijk(c)[2], corner_ijk(c)[1], corner_ijk(c)[0]] for c in range(8)]
        expect = 0
        for t in TETS:
            pc = sum((code >> c) & 1 for c in t)
            expect += NTRI_BY_POPCOUNT[pc]
        assert m["counts"][0, 0, 0] == expect == len(m["area"])
        last_tet = -1
        for n in range(len(m["area"])):
            ends = set()
            for v in range(3):
                a, b = edge_segment(int(m["edge"][n, v]))
                ca = a[0] + 2 * a[1] + 4 * a[2]
                cb = b[0] + 2 * b[1] + 4 * b[2]
                ends |= {ca, cb}
                fa, fb = Fraction(vals[ca]), Fraction(vals[cb])
                assert (fa > 0) != (fb > 0), "vertex on a non-straddling edge"
                s = (0 - fa) / (fb - fa)
                exact = [Fraction(a[d]) + s * (b[d] - a[d]) for d in range(3)]
                got = m["pos"][n, v]
                assert max(abs(float(exact[d]) - got[d]) for d in range(3)) < 1e-15
            tets = [q for q, t in enumerate(TETS) if ends <= set(t)]
            assert len(tets) == 1, "triangle edges not inside one tetrahedron"
            q = tets[0]
            assert q >= last_tet, "slot order must follow tetra order"
            last_tet = q
            seen_cases.add((q, sum(((code >> c) & 1) << r for r, c in enumerate(TETS[q]))))
            g = tet_gradient([vals[c] for c in TETS[q]], TETS[q])
            fn = m["fn"][n]
            gu = g / np.linalg.norm(g)
            # from the above-threshold side towards the below side: -grad f
            assert np.dot(fn, gu) < -1 + 1e-9
            assert abs(np.linalg.norm(fn) - 1) < 1e-14
            assert m["area"][n] > 0
    # every tetra sees all 14 non-trivial masks
    assert len(seen_cases) == 6 * 14


@pytest.mark.parametrize("seed", range(6))
def test_affine_fields_all_tet_masks(orc, seed):
    """For every (tetra, mask) an affine field with that mask on that tetra;
    on a 3x3x3 non-uniform lattice every vertex lies on the level set and
    every face normal equals -grad f / |grad f|."""
    rng = np.random.default_rng(100 + seed)
    h = rng.uniform(0.3, 2.0, size=3)
    o = rng.uniform(-1, 1, size=3)
    grids = tuple(o[d] + h[d] * np.array([0.0, 1.0, 2.0]) for d in range(3))
    for t, corners in enumerate(TETS):
        for mask in range(1, 15):
            vals = [rng.uniform(0.2, 1.0) * (1 if (mask >> r) & 1 else -1) for r in range(4)]
            g_idx = tet_gradient(vals, corners)
            P0 = np.array(corner_ijk(corners[0]), dtype=np.float64)
            b = vals[0] - g_idx @ P0
            a_phys = g_idx / h
            I = np.arange(3, dtype=np.float64)
            f = (b + g_idx[0] * I[None, None, :] + g_idx[1] * I[None, :, None]
                 + g_idx[2] * I[:, None, None])
            thr = 0.0
            m = orc.mesh(f, thr, grids=grids)
            assert len(m["area"]) > 0
            na = a_phys / np.linalg.norm(a_phys)
            for n in range(len(m["area"])):
                assert np.dot(m["fn"][n], na) < -1 + 1e-9
                for v in range(3):
                    p = m["pos"][n, v]
                    idx = (p - o) / h
                    assert abs(b + g_idx @ idx - thr) < 1e-11


def sphere_field(n, c, r):
    g = -0.5 + (np.arange(n) + 0.5) / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    return (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 - r * r, g


@pytest.mark.parametrize("n,r", [(12, 0.31), (20, 0.37), (17, 0.2)])
def test_sphere_watertight_and_oriented(orc, n, r):
    c = (0.0123, -0.0271, 0.0317)
    f, g = sphere_field(n, c, r)
    m = orc.mesh(f, 0.0, grids=(g, g, g))
    ncx = n - 1
    directed = {}
    verts = set()
    for t in range(len(m["area"])):
        cid = int(m["cell"][t])
        cell = (cid % ncx, (cid // ncx) % (n - 1), cid // (ncx * (n - 1)))
        keys = [global_key(cell, int(e)) for e in m["edge"][t]]
        verts |= set(keys)
        assert len(set(keys)) == 3
        for a, b in ((0, 1), (1, 2), (2, 0)):
            e = (keys[a], keys[b])
            directed[e] = directed.get(e, 0) + 1
        # inward normals: f grows outwards, normals point to lower f
        ctr = m["pos"][t].mean(axis=0) - np.array(c)
        assert np.dot(m["fn"][t], ctr) < 0
    assert all(v == 1 for v in directed.values())
    for (a, b) in directed:
        assert (b, a) in directed, "open mesh edge"
    E = len(directed) // 2
    assert len(verts) - E + len(m["area"]) == 2


def test_tiny_difference_snaps_to_an_endpoint(orc):
    """|v_head - v_tail| < 1e-8 puts the vertex exactly on an edge end
    (triangulate.c:62-66)."""
    f = np.full((2, 2, 2), -2.5e-9)
    f[0, 0, 0] = 2.5e-9
    grids = (np.array([0.0, 1.0]),) * 3
    m = orc.mesh(f, 0.0, grids=grids)
    assert len(m["area"]) > 0
    for t in range(len(m["area"])):
        for v in range(3):
            a, b = edge_segment(int(m["edge"][t, v]))
            p = tuple(m["pos"][t, v])
            assert p in (tuple(map(float, a)), tuple(map(float, b)))


def test_value_equal_to_threshold_is_below_and_degenerates(orc):
    """threshold < value is strict (triangulate.c:176-179): a corner equal
    to the threshold counts as below, its vertices collapse onto it and the
    zero-area triangle has a NaN face normal (triangulate.c:89-91)."""
    f = np.ones((2, 2, 2))
    f[0, 0, 0] = 0.5
    grids = (np.array([0.0, 1.0]),) * 3
    m = orc.mesh(f, 0.5, grids=grids)
    assert m["codes"][0, 0, 0] == 0xFE
    assert len(m["area"]) == 6
    assert np.all(m["area"] == 0.0)
    assert np.all(np.isnan(m["fn"]))
    assert np.all(m["pos"] == 0.0)


def test_cells_in_serial_order(orc):
    rng = np.random.default_rng(5)
    f = rng.standard_normal((6, 7, 8))
    m = orc.mesh(f, 0.1)
    key = m["cell"] * 16 + m["slot"]
    assert np.all(np.diff(key) > 0)
    assert np.array_equal(np.bincount(m["cell"], minlength=m["counts"].size),
                          m["counts"].ravel())
