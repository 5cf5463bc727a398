"""Shared helpers of the test suite (numpy/torch glue and mesh invariants)."""
import numpy as np


def ulp_diff(a, b):
    """Elementwise distance in units in the last place between float32 arrays."""
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    # map the sign-magnitude encoding onto a monotonic integer line
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


def edge_use_counts(tris):
    """Histogram {uses: number of undirected edges} of a triangle list (T,3)."""
    t = np.asarray(tris, np.int64)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    e.sort(axis=1)
    key = e[:, 0] * (1 << 32) + e[:, 1]
    _, cnt = np.unique(key, return_counts=True)
    vals, num = np.unique(cnt, return_counts=True)
    return dict(zip(vals.tolist(), num.tolist()))


def euler_characteristic(nv, tris):
    t = np.asarray(tris, np.int64)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    e.sort(axis=1)
    ne = np.unique(e[:, 0] * (1 << 32) + e[:, 1]).size
    used = np.unique(t).size
    return used - ne + len(t), used == nv


def oriented_consistently(tris):
    """Every directed edge appears at most once (consistent winding)."""
    t = np.asarray(tris, np.int64)
    d = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    key = d[:, 0] * (1 << 32) + d[:, 1]
    return np.unique(key).size == key.size


def to_np(t):
    return t.detach().cpu().numpy()


def mesh_slab_against_oracle(orc, planes, nx, ny, nz, spacing, z0, z1, verts, tris, totals=None):
    """Compares the part of a canonical indexed mesh that a z-slab fixes with
    the oracle's extraction of that slab: the triangles of cells z0 <= z < z1
    (ids and order) and the vertices of the points in planes z0..z1 (bits).

    planes(a, b): float32 field planes [a, b) flattened (host);
    verts(i, j), tris(i, j): the mesh's vertices / triangles [i, j) (host).
    The slab's first vertex and triangle in the whole mesh follow from the
    oracle's counts of planes 0..z0, or, for a slab that ends at the last
    plane, from the mesh's (vertices, triangles) `totals` minus the slab's;
    vertex positions use global z indices (z_index0), so they match the whole
    grid's bit for bit.  Returns (triangles compared, vertices compared)."""
    top_slab = z1 == nz - 1 and totals is not None
    if z0 > 0 and not top_slab:
        v_pre, t0 = orc.mesh_counts(planes(0, z0 + 1), nx, ny, z0 + 1)
        v_plane, _ = orc.mesh_counts(planes(z0, z0 + 1), nx, ny, 1)
        v0 = v_pre - v_plane  # the x/y edges of plane z0 belong to the slab
    else:
        v0 = t0 = 0
    top = min(z1 + 2, nz)  # one plane more, if there is one, for plane z1's z-edges
    v_e, t_e, _ = orc.extract_indexed(planes(z0, top), nx, ny, top - z0, spacing=spacing,
                                      z_index0=z0)
    n_pts = len(v_e)
    if top > z1 + 1:
        n_pts -= orc.mesh_counts(planes(z1 + 1, z1 + 2), nx, ny, 1)[0]
    _, n_slab = orc.mesh_counts(planes(z0, z1 + 1), nx, ny, z1 + 1 - z0)
    if top_slab:
        v0, t0 = totals[0] - n_pts, totals[1] - n_slab
    got_t = np.asarray(tris(t0, t0 + n_slab)).astype(np.int64) - v0
    assert got_t.shape == (n_slab, 3)
    assert np.array_equal(got_t, t_e[:n_slab].astype(np.int64)), "slab %d..%d triangles" % (z0, z1)
    got_v = np.ascontiguousarray(verts(v0, v0 + n_pts), np.float32)
    assert np.array_equal(got_v.view(np.uint32), v_e[:n_pts].view(np.uint32)), \
        "slab %d..%d vertices" % (z0, z1)
    return n_slab, n_pts
