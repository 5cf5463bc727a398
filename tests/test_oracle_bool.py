"""Pins of the oracle's boolean meshing (CPU): Isosurface::computeIsosurface
with classify_copy_Voxel / classifyVoxel / generateTriangles_lattice_kernel
(src/Isosurface.cu:44-134, src/MarchingCubes_kernel.cu:158-463, 869-1079,
1640-1784, 1865-2218).

  * single-cell truth tables: every one of the 256 corner patterns of each
    operand, every mode, against the set algebra each mode names;
  * classify_copy_Voxel: val is -1 exactly inside the boolean, the edge
    parameter is the textbook crossing parameter where one operand crosses,
    the average where both cross, untouched on the last point;
  * make_region with a field's own crossing parameters reproduces the
    textbook marching cubes of that field (same cube indices, positions
    within 1e-5 of a voxel);
  * union of two disjoint solids = the two solids meshed alone; the meshes of
    union / intersection / difference of overlapping solids are closed and
    consistently wound once their vertices are welded;
  * idempotence: A ∩ A and A ∪ A mesh like A.
Comparisons against the GPU path are in tests/test_gpu_dropin_bool.py.
"""
import numpy as np
import pytest

from tests.helpers import edge_use_counts, oriented_consistently

CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
ISO, LO, HI = 0.0, 0.2, 0.3


def cell_values(pattern, inside, outside):
    """2x2x2 point values: corner k gets `inside` when bit k of pattern is set."""
    v = np.empty(8, np.float32)
    for k, (dx, dy, dz) in enumerate(CORNER):
        v[dx + 2 * dy + 4 * dz] = inside if (pattern >> k) & 1 else outside
    return v


def cube_of(orc, flags, fixed, dyn, lat):
    verts, occ = orc.bool_classify(flags, fixed, dyn, lat, 2, 2, 2, ISO, LO, HI)
    nverts = orc.tables()[1]
    return verts[0], nverts


def expected_cube(mode, a_bits, b_bits):
    if mode == "union":
        return a_bits | b_bits
    if mode == "intersect":
        return a_bits & b_bits
    return a_bits & ~b_bits & 0xFF  # difference


@pytest.mark.parametrize("mode,flag", [("union", 1), ("intersect", 4), ("diff", 2)])
def test_truth_tables_primitive_pair(orc, mode, flag):
    """Neither fixed nor dynamic: operands F (grid_points val) and D (float)."""
    nverts = orc.tables()[1]
    rng = np.random.default_rng(7)
    for a in range(256):
        b = int(rng.integers(256))
        fixed = np.zeros(8, orc.GRID_POINT)
        fixed["val"] = np.where(cell_values(a, 1, 0) > 0, -1, 1)
        if mode == "diff":
            # difference keeps F inside and D outside: D < iso marks what is removed
            dyn = cell_values(b, -1.0, 1.0)
            verts, _ = orc.bool_classify(flag, fixed, dyn, None, 2, 2, 2, ISO, LO, HI)
            # reference form: (D >= iso) & (F < iso)
            assert verts[0] == nverts[a & ~b & 0xFF], (a, b)
        else:
            dyn = cell_values(b, -1.0, 1.0)
            verts, _ = orc.bool_classify(flag, fixed, dyn, None, 2, 2, 2, ISO, LO, HI)
            assert verts[0] == nverts[expected_cube(mode, a, b)], (a, b)


@pytest.mark.parametrize("mode,flag", [("union", 1), ("intersect", 4), ("diff", 2)])
@pytest.mark.parametrize("which", ["fixed", "dynamic"])
def test_truth_tables_with_lattice_band(orc, mode, flag, which):
    """fixed: D with the band iso1 < L < iso2; dynamic: F with the band (the
    dynamic difference keeps F outside the closed band)."""
    nverts = orc.tables()[1]
    rng = np.random.default_rng(11)
    for a in range(256):
        b = int(rng.integers(256))
        lat = cell_values(b, 0.25, 0.7)  # in band / out of band
        if which == "fixed":
            fx = None
            dyn = cell_values(a, -1.0, 1.0)
            f = flag | orc.B_FIXED
            if mode == "diff":
                want = ~a & b & 0xFF  # (D >= iso) & band
            else:
                want = expected_cube(mode, a, b)
        else:
            fx = np.zeros(8, orc.GRID_POINT)
            fx["val"] = np.where(cell_values(a, 1, 0) > 0, -1, 1)
            dyn = None
            f = flag | orc.B_DYNAMIC
            want = (a & ~b & 0xFF) if mode == "diff" else expected_cube(mode, a, b)
        verts, _ = orc.bool_classify(f, fx, dyn, lat, 2, 2, 2, ISO, LO, HI)
        assert verts[0] == nverts[want], (which, mode, a, b)


def grid(nx, ny, nz, fn):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.ascontiguousarray(fn(x.astype(np.float64), y, z).astype(np.float32).reshape(-1))


def sphere(c, r):
    return lambda x, y, z: np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r


def box(c, h):
    return lambda x, y, z: np.maximum(np.maximum(abs(x - c[0]) - h[0], abs(y - c[1]) - h[1]),
                                      abs(z - c[2]) - h[2])


def fixed_from(orc, field, nx, ny, nz):
    """grid_points of a field: val -1 inside / 1 outside, crossing parameters
    by classify_copy_Voxel (union with the field itself, starting outside)."""
    g = np.zeros(nx * ny * nz, orc.GRID_POINT)
    g["val"] = 1
    return orc.bool_copy(orc.B_UNION, g, field, None, nx, ny, nz, ISO, LO, HI)


def test_copy_marks_and_parameters(orc):
    nx, ny, nz = 9, 8, 7
    a = grid(nx, ny, nz, sphere((4.2, 3.7, 3.1), 2.6))
    out = fixed_from(orc, a, nx, ny, nz)
    n = nx * ny * nz
    np.testing.assert_array_equal(out["val"][:-1], np.where(a[:-1] < ISO, -1, 1))
    assert out["val"][-1] == 1 and out["t_x"][-1] == 0  # the last point is not visited
    g = a.reshape(nz, ny, nx)
    tx = out["t_x"].reshape(nz, ny, nx)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx - 1):
                if z * nx * ny + y * nx + x == n - 1:
                    continue
                f0, f1 = g[z, y, x], g[z, y, x + 1]
                if (f0 < 0) != (f1 < 0):
                    assert tx[z, y, x] == np.float32((ISO - f0) / (f1 - f0))
                else:
                    assert tx[z, y, x] == 0
    # a second operand crossing the same edge is averaged in
    b = grid(nx, ny, nz, sphere((4.6, 3.9, 3.3), 2.4))
    out2 = orc.bool_copy(orc.B_UNION, out, b, None, nx, ny, nz, ISO, LO, HI)
    gb = b.reshape(nz, ny, nx)
    tx2 = out2["t_x"].reshape(nz, ny, nx)
    hits = 0
    for z in range(nz):
        for y in range(ny):
            for x in range(nx - 1):
                f0, f1 = gb[z, y, x], gb[z, y, x + 1]
                t0 = tx[z, y, x]
                if (f0 < 0) != (f1 < 0) and t0 > 0:
                    t = np.float32((ISO - f0) / (f1 - f0))
                    assert tx2[z, y, x] == np.float32((t0 + t) * np.float32(0.5))
                    hits += 1
    assert hits > 0


def weld(pos, total, tol=1e-4):
    p = pos[:total, :3]
    key = np.round(p / tol).astype(np.int64)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    return inv.reshape(-1, 3)


def test_region_equals_textbook_mc(orc):
    nx, ny, nz = 14, 12, 11
    f = grid(nx, ny, nz, sphere((6.3, 5.4, 5.2), 4.1))
    fx = fixed_from(orc, f, nx, ny, nz)
    r = orc.bool_compute_isosurface(orc.B_REGION, fx, None, None, nx, ny, nz, ISO, LO, HI)
    s = orc.compute_isosurface(orc.MODE_STD, f, nx, ny, nz, ISO)
    assert r["active"] == s["active"] and r["total"] == s["total"] > 0
    np.testing.assert_array_equal(r["verts"], s["verts"])
    # the same triangles in the same order, vertices within rounding
    np.testing.assert_allclose(r["pos"][:r["total"]], s["pos"][:s["total"]], atol=1e-5)


def test_union_of_disjoint_solids(orc):
    nx, ny, nz = 20, 12, 12
    a = grid(nx, ny, nz, sphere((4.5, 5.5, 5.5), 3.2))
    b = grid(nx, ny, nz, box((14.5, 5.5, 5.5), (3.3, 3.1, 2.9)))
    fx = fixed_from(orc, a, nx, ny, nz)
    u = orc.bool_compute_isosurface(orc.B_UNION, fx, b, None, nx, ny, nz, ISO, LO, HI)
    ra = orc.bool_compute_isosurface(orc.B_REGION, fx, None, None, nx, ny, nz, ISO, LO, HI)
    fb = fixed_from(orc, b, nx, ny, nz)
    rb = orc.bool_compute_isosurface(orc.B_REGION, fb, None, None, nx, ny, nz, ISO, LO, HI)
    assert u["total"] == ra["total"] + rb["total"]
    key = lambda p, t: np.sort(np.round(p[:t, :3].reshape(-1, 9), 4), axis=0)
    both = np.concatenate([ra["pos"][:ra["total"]], rb["pos"][:rb["total"]]])
    np.testing.assert_allclose(key(u["pos"], u["total"]), key(both, len(both)), atol=2e-4)


@pytest.mark.parametrize("flag", [1, 4, 2])
def test_boolean_meshes_are_closed(orc, flag):
    nx, ny, nz = 18, 17, 16
    a = grid(nx, ny, nz, sphere((8.1, 8.3, 7.9), 5.2))
    b = grid(nx, ny, nz, box((10.4, 9.1, 8.6), (4.1, 3.6, 4.4)))
    fx = fixed_from(orc, a, nx, ny, nz)
    r = orc.bool_compute_isosurface(flag, fx, b, None, nx, ny, nz, ISO, LO, HI)
    assert r["total"] > 0
    tris = weld(r["pos"], r["total"])
    tris = tris[(tris[:, 0] != tris[:, 1]) & (tris[:, 1] != tris[:, 2]) & (tris[:, 0] != tris[:, 2])]
    assert set(edge_use_counts(tris)) == {2}
    assert oriented_consistently(tris)


def test_idempotence(orc):
    nx, ny, nz = 13, 13, 12
    a = grid(nx, ny, nz, sphere((6.2, 6.1, 5.8), 4.3))
    fx = fixed_from(orc, a, nx, ny, nz)
    ra = orc.bool_compute_isosurface(orc.B_REGION, fx, None, None, nx, ny, nz, ISO, LO, HI)
    for flag in (orc.B_UNION, orc.B_INTERSECT):
        r = orc.bool_compute_isosurface(flag, fx, a, None, nx, ny, nz, ISO, LO, HI)
        assert r["total"] == ra["total"]
        # both operands cross the same edges with the same parameter: the
        # average of t and t is t
        np.testing.assert_array_equal(r["pos"][:r["total"]], ra["pos"][:ra["total"]])
    d = orc.bool_compute_isosurface(orc.B_DIFF, fx, a, None, nx, ny, nz, ISO, LO, HI)
    assert d["active"] == 0 and d["total"] == 0
