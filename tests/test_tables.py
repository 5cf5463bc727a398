"""The 256-case marching-cubes tables.

Pinned three ways:
  * against the reference's own src/tables.h when the reference tree is
    present (the triangulation must be identical for bit-exact connectivity);
  * derived tables (vertex counts, crossing-edge masks) against their
    definitions;
  * per case, every triangle uses only crossing edges, every crossing edge is
    used, and the triangles of each case close up on random fields (checked in
    the oracle extraction tests).
"""
import os
import re

import numpy as np
import pytest

REF_TABLES = "/root/reference/src/tables.h"

CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGE = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def crossing_mask(c):
    m = 0
    for e, (a, b) in enumerate(EDGE):
        if ((c >> a) & 1) != ((c >> b) & 1):
            m |= 1 << e
    return m


@pytest.fixture(scope="module")
def tables(gc):
    tri = np.zeros(256 * 16, np.uint8)
    nv = np.zeros(256, np.uint8)
    em = np.zeros(256, np.uint16)
    gc.lib.gpucad_mc_tables(tri.ctypes.data, nv.ctypes.data, em.ctypes.data)
    return tri.reshape(256, 16), nv, em


@pytest.mark.skipif(not os.path.exists(REF_TABLES), reason="reference tree not mounted")
def test_tables_match_reference_source(tables):
    src = open(REF_TABLES).read()
    body = src[src.index("triTable[256][16]"):]
    body = body[:body.index("};")]
    rows = re.findall(r"\{([^{}]*)\}", body)
    assert len(rows) == 256
    ref_tri = np.array([[255 if v.strip() == "X" else int(v) for v in r.split(",")] for r in rows],
                       np.uint8)
    nvs = src[src.index("numVertsTable[256]"):]
    nvs = nvs[nvs.index("{") + 1:nvs.index("}")]
    ref_nv = np.array([int(v) for v in nvs.replace("\n", "").split(",") if v.strip()], np.uint8)
    ets = src[src.index("edgeTable[256]"):]
    ets = ets[ets.index("{") + 1:ets.index("}")]
    ref_em = np.array([int(v, 16) for v in ets.replace("\n", "").split(",") if v.strip()],
                      np.uint16)
    tri, nv, em = tables
    assert np.array_equal(tri, ref_tri)
    assert np.array_equal(nv, ref_nv)
    assert np.array_equal(em, ref_em)


def test_vertex_counts_and_masks_follow_definitions(tables):
    tri, nv, em = tables
    for c in range(256):
        n = int(np.sum(tri[c] != 255))
        assert nv[c] == n and n % 3 == 0
        assert np.all(tri[c][n:] == 255)
        assert em[c] == crossing_mask(c)
    hist = {}
    for c in range(256):
        hist[int(nv[c])] = hist.get(int(nv[c]), 0) + 1
    assert hist == {0: 2, 3: 16, 6: 50, 9: 80, 12: 76, 15: 32}


def test_triangles_use_exactly_the_crossing_edges(tables):
    tri, nv, em = tables
    for c in range(256):
        used = 0
        for e in tri[c][: nv[c]]:
            assert (em[c] >> int(e)) & 1, "case %d uses non-crossing edge %d" % (c, e)
            used |= 1 << int(e)
        assert used == em[c], "case %d leaves a crossing edge unused" % c


def test_complement_symmetry_of_counts(tables):
    # the cases c and 255-c cut the same edges
    tri, nv, em = tables
    for c in range(256):
        assert em[c] == em[255 - c]


def test_inside_corners_lie_on_the_normal_side(tables):
    """Triangle normals (v1-v0)x(v2-v0) of single-corner cases point at the
    inside corner (f < iso), the orientation the reference's normals use."""
    tri, nv, em = tables
    for k in range(8):
        for c in (1 << k, 255 ^ (1 << k)):
            inside = np.array(CORNER[k], float) if c == 1 << k else None
            (e0, e1, e2) = tri[c][:3]
            mid = [0.5 * (np.array(CORNER[EDGE[e][0]], float) + np.array(CORNER[EDGE[e][1]], float))
                   for e in (e0, e1, e2)]
            n = np.cross(mid[1] - mid[0], mid[2] - mid[0])
            centroid = sum(mid) / 3
            corner = np.array(CORNER[k], float)
            side = np.dot(n, corner - centroid)
            if inside is not None:
                assert side > 0, "case %d" % c
            else:
                assert side < 0, "case %d" % c


FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
BODY_DIAGONALS = [(0, 6), (1, 7), (2, 4), (3, 5)]


def count_terms(c):
    """|S|, faces with all corners in S, [S = cube minus a body diagonal's ends]."""
    inside = [(c >> k) & 1 for k in range(8)]
    v = sum(inside)
    f = sum(all(inside[k] for k in face) for face in FACES)
    d = sum(1 for i, j in BODY_DIAGONALS if v == 6 and not inside[i] and not inside[j])
    return v, f, d


def test_triangle_count_identity(tables):
    """The extraction's count stage totals triangles per 64-cell word as
    |S| - 2F - 4D over popcounts (no per-cell table lookup); this holds only
    because every case with an outside corner satisfies it."""
    tri, nv, em = tables
    without_diagonal = 0
    for c in range(255):
        v, f, d = count_terms(c)
        assert nv[c] // 3 == v - 2 * f - 4 * d, "case %d" % c
        # the same identity through crossing edges and Euler characteristic
        edges_in = sum(1 for a, b in EDGE if (c >> a) & 1 and (c >> b) & 1)
        polygons = (v - edges_in + f) + (2 if d else 0)
        assert nv[c] // 3 == bin(int(em[c])).count("1") - 2 * polygons
        without_diagonal += nv[c] // 3 != v - 2 * f
    # the body-diagonal term matters in exactly its four cases
    assert without_diagonal == 4


def test_triangle_count_over_all_cells(tables):
    """The count stage sums |S| - 2F - 4D over every cell of a word, active or
    not, and adds 4 per cell with all eight corners inside: the all-outside
    case contributes 0 and the all-inside case exactly -4, so the word total
    equals the table's."""
    tri, nv, em = tables
    assert count_terms(0) == (0, 0, 0)
    v, f, d = count_terms(255)
    assert (v, f, d) == (8, 6, 0) and v - 2 * f - 4 * d == -4 and nv[255] == 0
    rng = np.random.default_rng(7)
    for _ in range(200):
        cases = rng.integers(0, 256, size=64)
        cases[rng.integers(0, 64, size=8)] = 255  # runs of inside cells
        total = 0
        for c in cases:
            v, f, d = count_terms(int(c))
            total += v - 2 * f - 4 * d + (4 if c == 255 else 0)
        assert total == sum(int(nv[c]) // 3 for c in cases)
