"""Pins of the oracle's marching cubes (CPU).

  * classification against a vectorised numpy restatement of the definition
    (cube bit k = [f(corner k) < iso], src/MarchingCubes_kernel.cu:3207-3249);
  * exclusive scan against numpy's cumulative sum modulo 2^32 (thrust over
    uint, src/MarchingCubes_kernel.cu:3198-3203), compaction by definition;
  * the canonical indexed mesh: linear fields interpolate exactly, every
    vertex lies on its own grid edge, closed surfaces are 2-manifold with
    consistent winding and the Euler characteristic of their topology, the
    mesh does not depend on the thread count;
  * the reference-shaped non-indexed generate (textbook vertexInterp4) welds
    to the same triangles as the indexed mesh;
  * Gratings::GPU_buffer_normalise_four semantics on small arrays.
"""
import numpy as np
import pytest

from gpucadforam_amd import configs as CF
from tests.helpers import edge_use_counts, euler_characteristic, oriented_consistently

CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def grid_field(nx, ny, nz, fn):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.ascontiguousarray(fn(x, y, z).astype(np.float32).reshape(-1))


def numpy_cubes(f, nx, ny, nz, iso):
    g = f.reshape(nz, ny, nx)
    cube = np.zeros((nz - 1, ny - 1, nx - 1), np.int64)
    for k, (dx, dy, dz) in enumerate(CORNER):
        cube |= (g[dz:nz - 1 + dz, dy:ny - 1 + dy, dx:nx - 1 + dx] < iso).astype(np.int64) << k
    return cube.reshape(-1)


def test_classify_matches_definition(orc):
    rng = np.random.default_rng(3)
    for (nx, ny, nz) in [(2, 2, 2), (7, 5, 3), (33, 17, 9)]:
        f = rng.standard_normal(nx * ny * nz).astype(np.float32)
        f[::7] = 0.0  # values equal to iso are outside (not < iso)
        cube, verts, occ = orc.classify(f, nx, ny, nz, 0.0)
        ref = numpy_cubes(f, nx, ny, nz, 0.0)
        _, nv = orc.tables()
        assert np.array_equal(cube, ref)
        assert np.array_equal(verts, nv[ref])
        assert np.array_equal(occ, (nv[ref] > 0).astype(np.uint32))


def test_all_256_cases_brute_force(orc):
    """Every case on a single cell: cube index = case, triangle count from the
    table, and the emitted triangles are the table row's edges."""
    tri, nv = orc.tables()
    for c in range(256):
        f = np.array([(-1.0 if (c >> k) & 1 else 1.0) for k in range(8)], np.float32)
        g = np.zeros(8, np.float32)
        for k, (dx, dy, dz) in enumerate(CORNER):
            g[dx + 2 * dy + 4 * dz] = f[k]
        cube, verts, occ = orc.classify(g, 2, 2, 2, 0.0)
        assert cube[0] == c and verts[0] == nv[c]
        v, t, e = orc.extract_indexed(g, 2, 2, 2, edge_ids=True)
        assert len(t) == nv[c] // 3
        # map vertex -> local cell edge through the canonical edge id
        edge_of = {}
        for vid, eid in enumerate(e):
            p, ax = divmod(int(eid), 3)
            x, y, z = p % 2, (p // 2) % 2, p // 4
            a = (x, y, z)
            b = tuple(a[i] + (1 if i == ax else 0) for i in range(3))
            ca, cb = CORNER.index(a), CORNER.index(b)
            for le, pair in enumerate([(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7),
                                       (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]):
                if set(pair) == {ca, cb}:
                    edge_of[vid] = le
        got = [edge_of[int(i)] for i in t.reshape(-1)]
        assert got == [int(x) for x in tri[c][: nv[c]]], c


def test_exclusive_scan_wraps_like_uint(orc):
    rng = np.random.default_rng(4)
    a = rng.integers(0, 2 ** 32, 100_000, dtype=np.uint64).astype(np.uint32)
    got = orc.exclusive_scan_u32(a)
    ref = (np.concatenate([[0], np.cumsum(a.astype(np.uint64))[:-1]]) % (1 << 32)).astype(np.uint32)
    assert np.array_equal(got, ref)
    assert orc.exclusive_scan_u32(np.zeros(0, np.uint32)).size == 0


def test_linear_field_is_interpolated_exactly(orc):
    nx, ny, nz = 21, 17, 13
    h = np.array([1 / (nx - 1), 1 / (ny - 1), 1 / (nz - 1)])
    a = np.array([0.31, -0.72, 0.45])
    f = grid_field(nx, ny, nz, lambda x, y, z: a[0] * x * h[0] + a[1] * y * h[1] + a[2] * z * h[2] + 0.07)
    v, t, e = orc.extract_indexed(f, nx, ny, nz, edge_ids=True)
    assert len(t) > 0
    assert np.abs(v.astype(np.float64) @ a + 0.07).max() < 1e-6
    # every vertex sits on its own canonical edge
    p, ax = np.divmod(e.astype(np.int64), 3)
    x, y, z = p % nx, (p // nx) % ny, p // (nx * ny)
    base = np.stack([x * h[0], y * h[1], z * h[2]], 1)
    off = v - base
    for k in range(3):
        on = ax == k
        assert np.all(np.abs(np.delete(off[on], k, axis=1)) < 1e-6)
        assert np.all((off[on, k] >= -1e-7) & (off[on, k] <= h[k] + 1e-7))
    # canonical order: edge ids strictly increasing
    assert np.all(np.diff(e.astype(np.int64)) > 0)


def test_sphere_closed_genus_zero(orc):
    n = 40
    f = grid_field(n, n, n, lambda x, y, z: np.sqrt((x - 19.3) ** 2 + (y - 20.1) ** 2 + (z - 19.7) ** 2) - 13.0)
    v, t, _ = orc.extract_indexed(f, n, n, n, spacing=(1, 1, 1))
    assert edge_use_counts(t) == {2: 3 * len(t) // 2}
    assert oriented_consistently(t)
    chi, all_used = euler_characteristic(len(v), t)
    assert chi == 2 and all_used
    r = np.linalg.norm(v - np.array([19.3, 20.1, 19.7]), axis=1)
    assert np.abs(r - 13.0).max() < 0.05
    # normals of (v1-v0)x(v2-v0) point into the inside (f < 0): towards the centre
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    nrm = np.cross(b - a, c - a)
    assert np.mean(np.einsum("ij,ij->i", nrm, (a + b + c) / 3 - np.array([19.3, 20.1, 19.7])) < 0) > 0.999


def test_torus_genus_one(orc):
    n = 48
    def fn(x, y, z):
        X, Y, Z = x - 23.5, y - 23.5, z - 23.5
        return (np.sqrt(X ** 2 + Y ** 2) - 14.0) ** 2 + Z ** 2 - 25.0
    v, t, _ = orc.extract_indexed(grid_field(n, n, n, fn), n, n, n)
    assert edge_use_counts(t) == {2: 3 * len(t) // 2}
    chi, _ = euler_characteristic(len(v), t)
    assert chi == 0


def test_random_fields_close_up(orc):
    """Random smooth fields that are positive on the grid boundary give closed,
    consistently oriented 2-manifolds (the case table has no cracks)."""
    rng = np.random.default_rng(5)
    for trial in range(4):
        n = 24
        coarse = rng.standard_normal((6, 6, 6))
        z, y, x = np.meshgrid(*(np.linspace(0, 5, n),) * 3, indexing="ij")
        from scipy.ndimage import map_coordinates
        f = map_coordinates(coarse, [z, y, x], order=3)
        f[0, :, :] = f[-1, :, :] = f[:, 0, :] = f[:, -1, :] = f[:, :, 0] = f[:, :, -1] = 1.0
        f = np.ascontiguousarray(f.astype(np.float32).reshape(-1))
        v, t, _ = orc.extract_indexed(f, n, n, n)
        if len(t) == 0:
            continue
        assert set(edge_use_counts(t)) == {2}
        assert oriented_consistently(t)


def test_mesh_independent_of_thread_count(orc):
    p = CF.varying_schwarz_p(64)
    f = orc.lattice_field_f32(p)
    a = orc.extract_indexed(f, 64, 64, 64, nthreads=1)
    b = orc.extract_indexed(f, 64, 64, 64, nthreads=8)
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    assert np.array_equal(a[1], b[1])
    f1 = orc.lattice_field_f32(p, nthreads=1)
    assert np.array_equal(f.view(np.uint32), f1.view(np.uint32))


def test_nonindexed_generate_welds_to_indexed_mesh(orc):
    """Reference-shaped textbook generate (per-voxel edges, voxel-local
    direction) and the canonical indexed mesh describe the same triangles."""
    p = CF.plain_gyroid(40)
    f = orc.lattice_field_f32(p)
    n = 40
    out = orc.compute_isosurface(orc.MODE_STD, f, n, n, n, 0.0)
    v, t, _ = orc.extract_indexed(f, n, n, n, spacing=(1, 1, 1))
    total = out["total"]
    assert total == 3 * len(t)
    pos = out["pos"][:total, :3].reshape(-1, 3, 3)
    ref = v[t]
    # same triangle order (cell order, table order) and positions within
    # rounding of the opposite interpolation direction
    assert np.abs(pos - ref).max() < 1e-5
    assert np.all(out["pos"][:total, 3] == 1.0)
    # counts and scans of the orchestration
    assert out["active"] == int(np.count_nonzero(out["occ"]))
    assert np.array_equal(out["verts_scan"], orc.exclusive_scan_u32(out["verts"]))
    assert np.array_equal(out["occ_scan"], orc.exclusive_scan_u32(out["occ"]))
    act = np.nonzero(out["occ"])[0]
    assert np.array_equal(out["comp"][: len(act)], act.astype(np.uint32))


def test_normalise_four_semantics(orc):
    rng = np.random.default_rng(6)
    nx, ny, nz = 9, 8, 7
    f = rng.uniform(-0.5, 2.0, nx * ny * nz).astype(np.float32)
    ind, nrm = orc.normalise_four(f, nx, ny, nz, 0.2, 0.3)
    a, b = min(0.0, f.min()), max(0.0, f.max())
    k = ((f - np.float32(a)) / (np.float32(b) - np.float32(a))).astype(np.float32)
    g = k.reshape(nz, ny, nx).copy()
    m = ((g >= np.float32(0.2)) & (g <= np.float32(0.3))).astype(np.float32)
    for arr in (g, m):
        arr[0, :, :] = arr[-1, :, :] = arr[:, 0, :] = arr[:, -1, :] = arr[:, :, 0] = arr[:, :, -1] = 0
    assert np.array_equal(nrm, g.reshape(-1))
    assert np.array_equal(ind, m.reshape(-1))
    # all-positive input: the seed 0 becomes the minimum
    ind2, nrm2 = orc.normalise_four(np.full(nx * ny * nz, 3.0, np.float32), nx, ny, nz, 0.2, 0.3)
    assert np.all(nrm2.reshape(nz, ny, nx)[1:-1, 1:-1, 1:-1] == 1.0)


def test_slab_comparison_reproduces_the_whole_mesh(orc):
    """The z-slab comparison used on full-size GPU meshes, checked on the
    oracle's own whole-grid mesh: every slab (bottom, interior, top) matches,
    and a one-plane shift of the slab's triangles is caught."""
    from gpucadforam_amd import configs as CF
    from tests.helpers import mesh_slab_against_oracle
    p = CF.varying_schwarz_p(48)
    f = orc.lattice_field_f32(p)
    sp = (1 / (p.nx - 1), 1 / (p.ny - 1), 1 / (p.nz - 1))
    v, t, _ = orc.extract_indexed(f, p.nx, p.ny, p.nz, spacing=sp)
    plane = p.nx * p.ny
    planes = lambda a, b: f[a * plane:b * plane]
    for z0, z1 in ((0, 6), (17, 30), (40, 47)):
        for totals in (None, (len(v), len(t))):
            nt, nv = mesh_slab_against_oracle(orc, planes, p.nx, p.ny, p.nz, sp, z0, z1,
                                              lambda i, j: v[i:j], lambda i, j: t[i:j], totals)
            assert nt > 0 and nv > 0
    with pytest.raises(AssertionError):
        mesh_slab_against_oracle(orc, planes, p.nx, p.ny, p.nz, sp, 17, 30,
                                 lambda i, j: v[i:j], lambda i, j: t[i + 1:j + 1])
