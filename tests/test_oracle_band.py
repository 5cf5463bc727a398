"""Pins of the oracle's band-lattice interpolation (vertexInterp2_new /
vertexInterp3_new, src/MarchingCubes_kernel.cu:3593-3674, 3269-3416) by what
the mathematics fixes, not by restating the branches:

  * a normalised field linear along x puts every vertex exactly where the
    line reaches the band level the edge straddles (lower level first);
  * a node within the 5e-4 snap distance of a level pulls the vertex onto
    the node itself;
  * the endpoint swap makes the vertex independent of the edge's direction:
    the vertices of the mirrored field are the mirror images, bit for bit
    (the triangles differ: the case table is not mirror symmetric);
  * the second band (vol_two, indicator value 2) follows the same rules.

The reference's caller runs it with isoValue 0.25 between indicator values
0 and 1 and band levels 0.20 / 0.30 (src/main.cu:4006-4010)."""
import numpy as np


def linear_band(nx, ny, nz, lo, hi, kx):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    k = kx(x).astype(np.float32) + 0 * y + 0 * z
    ind = ((k >= np.float32(lo)) & (k <= np.float32(hi))).astype(np.float32)
    return ind.reshape(-1), np.ascontiguousarray(k.astype(np.float32).reshape(-1))


def vertices(out):
    return out["pos"][: out["total"], :3].astype(np.float64)


def test_linear_field_vertices_sit_on_the_band_levels(orc):
    n, lo, hi = 12, 0.2, 0.3
    ind, k = linear_band(n, n, n, lo, hi, lambda x: x / (n - 1.0))
    out = orc.compute_isosurface(orc.MODE_BAND1, ind, n, n, n, 0.25, vol1=k, lo1=lo, hi1=hi)
    v = vertices(out)
    assert len(v) > 0
    # the band is the single plane x = 3 (k = 3/11); its faces cross the
    # edges 2-3 (k from 2/11 to 3/11, level 0.2) and 3-4 (level 0.3)
    x_lo = 2 + (lo - 2 / 11) * 11
    x_hi = 3 + (hi - 3 / 11) * 11
    on_lo = np.abs(v[:, 0] - x_lo) < 1e-5
    on_hi = np.abs(v[:, 0] - x_hi) < 1e-5
    assert np.all(on_lo | on_hi)
    assert on_lo.any() and on_hi.any()


def test_nodes_near_a_level_snap(orc):
    n, lo, hi = 12, 0.2, 0.3
    # k(4) = 0.2003: within 5e-4 of the lower level; k(7) = 0.3 exactly
    kx = lambda x: np.where(x < 4, 0.05 * x, np.where(x < 7, 0.2003 + (x - 4) * 0.03, 0.3 + (x - 7) * 0.05))
    ind, k = linear_band(n, n, n, lo, hi, kx)
    out = orc.compute_isosurface(orc.MODE_BAND1, ind, n, n, n, 0.25, vol1=k, lo1=lo, hi1=hi)
    v = vertices(out)
    assert len(v) > 0
    # vertices on the edges leaving the band snap onto the nodes x = 4 and x = 7
    assert set(np.unique(np.round(v[:, 0], 6))) <= {4.0, 7.0}


def test_vertex_does_not_depend_on_edge_direction(orc):
    rng = np.random.default_rng(11)
    n, lo, hi = 14, 0.2, 0.3
    k = rng.uniform(0.0, 0.5, (n, n, n)).astype(np.float32)
    ind = ((k >= np.float32(lo)) & (k <= np.float32(hi))).astype(np.float32)
    c = ((n - 1) / 2.0,) * 3  # grid centred on the origin: mirroring x negates positions
    a = orc.compute_isosurface(orc.MODE_BAND1, ind.reshape(-1), n, n, n, 0.25, center=c,
                               vol1=k.reshape(-1), lo1=lo, hi1=hi)
    kf, indf = k[:, :, ::-1].copy(), ind[:, :, ::-1].copy()
    b = orc.compute_isosurface(orc.MODE_BAND1, indf.reshape(-1), n, n, n, 0.25, center=c,
                               vol1=kf.reshape(-1), lo1=lo, hi1=hi)
    va, vb = vertices(a), vertices(b)
    assert a["total"] == b["total"] > 0
    vb[:, 0] = -vb[:, 0]
    # the case table's triangulation is not mirror symmetric, the vertices are
    ua, ub = np.unique(va, axis=0), np.unique(vb, axis=0)
    assert len(ua) > 1000 and np.array_equal(ua, ub)
    # interpolating from the other end (no swap) would move vertices in the
    # last bits: float32 t from each end, on this field's edge values
    f0 = k.reshape(-1)[:-1].astype(np.float32)
    f1 = k.reshape(-1)[1:].astype(np.float32)
    sel = (np.minimum(f0, f1) <= lo) & (np.maximum(f0, f1) >= lo)
    fa, fb = f0[sel], f1[sel]
    t_fwd = (np.float32(lo) - fa) / (fb - fa)          # from p0 = 0 towards p1 = 1
    t_bwd = (np.float32(lo) - fb) / (fa - fb)          # from p1 = 1 towards p0 = 0
    assert np.count_nonzero(t_fwd != (np.float32(1) - t_bwd)) > 0


def test_second_band_uses_its_own_field_and_levels(orc):
    n = 12
    lo2, hi2 = 0.6, 0.7
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    k2 = (x / (n - 1.0)).astype(np.float32) + 0 * y + 0 * z
    ind = np.where((k2 >= lo2) & (k2 <= hi2), 2.0, 0.0).astype(np.float32)
    zeros = np.zeros_like(k2)
    out = orc.compute_isosurface(orc.MODE_BAND2, ind.reshape(-1), n, n, n, 0.25,
                                 vol1=zeros.reshape(-1), vol2=k2.reshape(-1),
                                 lo1=0.2, hi1=0.3, lo2=lo2, hi2=hi2)
    v = vertices(out)
    assert len(v) > 0
    # band x = 7 (k = 7/11 = 0.636): levels 0.6 on edge 6-7, 0.7 on edge 7-8
    x_lo = 6 + (lo2 - 6 / 11) * 11
    x_hi = 7 + (hi2 - 7 / 11) * 11
    assert np.all((np.abs(v[:, 0] - x_lo) < 1e-5) | (np.abs(v[:, 0] - x_hi) < 1e-5))
