"""Pins of three oracle functions by values worked out by hand, each with
planted mistakes that the same checks must reject.

  * create_lattice unit cells (src/lattice_files/Fft_lattice.cu:26-58): on a
    61-point grid the coordinate (x/60 - 0.5)/0.5 is exactly 0 at x = 30,
    +-0.5 at 45/15 and +-1 at the ends, where the six forms take values fixed
    by cos 0 = 1, sin 0 = 0 and the 3.14 scale (not pi), and have the
    symmetries the formulas imply (the reference gyroid is invariant under the
    cyclic axis permutation but not under the y <-> z mirror).
  * File_output::file_write_obj (src/File_output.cu:5-81): a known-answer
    file, written out literally below, for positions that exercise truncation
    toward zero (negative coordinates, -0.4 thousandths giving 0), welding,
    a degenerate face, a repeated face and a rotated face.
  * the band interpolation's equal-values branch
    (src/MarchingCubes_kernel.cu:3664-3671, 3316-3323 and 3393-3400): an
    edge whose indicator pair is 1/0 (or 2/0) with equal normalised values
    outside both band levels gets t = 1 when p0.z == 0 and t = 0 otherwise.

The planted mistakes are numpy arrays or texts with one change each; the
oracle itself is never modified."""
import os

import numpy as np
import pytest

N = 61  # the reference's unit cell (src/main.cu: Nxu = Nyu = Nzu = 61)
C, P, M, LO, HI = 30, 45, 15, 0, 60  # coordinates 0, +0.5, -0.5, -1, +1


def cell_xyz(a):
    """(x, y, z)-indexed view of a linear x-fastest unit cell."""
    return np.asarray(a, np.float64).reshape(N, N, N).transpose(2, 1, 0)


def planted(type_, scale=3.14, swap=None, unscaled2=True):
    """The unit cell with one deliberate change (float32, numpy)."""
    r = (((np.arange(N) * 1.0) / (N - 1)) - 0.5) / 0.5
    X, Y, Z = np.meshgrid(r, r, r, indexing="ij")
    if swap == "yz":
        Y, Z = Z, Y
    elif swap == "xy":
        X, Y = Y, X
    c = lambda v: np.cos(scale * v)
    s = lambda v: np.sin(scale * v)
    if type_ == 0:
        f = c(X) * s(Y) + c(Y) * s(Z) + c(Z) * s(X)
    elif type_ == 1:
        f = c(X) + c(Y) + c(Z)
    elif type_ == 2:
        k = 1.0 if unscaled2 else scale
        cc = lambda v: np.cos(k * v)
        f = 4 * cc(X) * cc(Y) * cc(Z) - (cc(2 * X) * cc(2 * Y) + cc(2 * Y) * cc(2 * Z) +
                                          cc(2 * Z) * cc(2 * X))
    elif type_ == 3:
        f = 2 * (c(X) * c(Y) + c(Y) * c(Z) + c(Z) * c(X)) - (c(2 * X) + c(2 * Y) + c(2 * Z))
    elif type_ == 4:
        f = np.minimum(np.minimum(X * X + Y * Y, Y * Y + Z * Z), Z * Z + X * X)
    else:
        f = c(X) * c(Y) * c(Z) - s(X) * s(Y) * s(Z)
    return f.astype(np.float32)  # already (x, y, z)-indexed


def lattice_failures(type_, F):
    """Checks of one unit cell (x, y, z)-indexed; returns the failed ones."""
    k = 3.14
    fails = []

    def at(x, y, z, want, tol, what):
        if not abs(F[x, y, z] - want) <= tol:
            fails.append("%s: f%s = %.9g, want %.9g" % (what, (x, y, z), F[x, y, z], want))

    def sym(G, what, expect=True):
        same = np.abs(F - G).max() <= 1e-6
        if same != expect:
            fails.append(what)

    if type_ == 0:
        at(C, C, C, 0.0, 0.0, "centre")
        # c(x)s(y) + c(y)s(z) + c(z)s(x) at (+0.5, -0.5, 0): c(1.57)s(-1.57) + 0 + s(1.57)
        at(P, M, C, np.cos(k / 2) * np.sin(-k / 2) + np.sin(k / 2), 1e-6, "mixed point")
        at(M, P, C, np.cos(-k / 2) * np.sin(k / 2) + np.sin(-k / 2), 1e-6, "mixed point 2")
        sym(F.transpose(1, 2, 0), "cyclic permutation")
        sym(F.transpose(0, 2, 1), "y<->z mirror", expect=False)
    elif type_ == 1:
        at(C, C, C, 3.0, 0.0, "centre")
        at(HI, HI, HI, 3 * np.cos(k), 1e-6, "corner (3.14, not pi)")
        at(P, C, C, np.cos(k / 2) + 2, 1e-6, "face axis point")
        sym(F.transpose(1, 0, 2), "x<->y")
        sym(F[::-1, :, :], "even in x")
    elif type_ == 2:
        at(C, C, C, 1.0, 0.0, "centre")
        at(HI, HI, HI, 4 * np.cos(1) ** 3 - 3 * np.cos(2) ** 2, 1e-6, "corner (unscaled)")
        at(P, C, C, 4 * np.cos(0.5) - (2 * np.cos(1) + 1), 1e-6, "axis point")
        sym(F.transpose(2, 0, 1), "permutation")
        sym(F[:, ::-1, :], "even in y")
    elif type_ == 3:
        at(C, C, C, 3.0, 0.0, "centre")
        at(HI, C, C, 2 * (2 * np.cos(k) + 1) - (np.cos(2 * k) + 2), 1e-5, "axis end")
        at(P, P, C, 2 * (np.cos(k / 2) ** 2 + 2 * np.cos(k / 2)) - (2 * np.cos(k) + 1), 1e-5,
           "edge point")
        sym(F.transpose(1, 0, 2), "x<->y")
    elif type_ == 4:
        at(C, C, C, 0.0, 0.0, "centre")
        at(HI, LO, HI, 2.0, 0.0, "corner")
        at(P, P, C, 0.25, 0.0, "two halves")
        at(P, M, HI, 0.5, 0.0, "min picks x^2 + y^2")
        sym(F.transpose(1, 2, 0), "permutation")
    else:
        at(C, C, C, 1.0, 0.0, "centre")
        at(P, P, P, np.cos(k / 2) ** 3 - np.sin(k / 2) ** 3, 1e-6, "(+0.5)^3")
        at(M, M, M, np.cos(k / 2) ** 3 + np.sin(k / 2) ** 3, 1e-6, "(-0.5)^3")
        at(HI, HI, HI, np.cos(k) ** 3 - np.sin(k) ** 3, 1e-6, "corner (3.14, not pi)")
        sym(F.transpose(2, 0, 1), "permutation")
    return fails


@pytest.mark.parametrize("type_", range(6))
def test_create_lattice_hand_values(orc, type_):
    F = cell_xyz(orc.create_lattice(N, N, N, type_))
    assert lattice_failures(type_, F) == []


@pytest.mark.parametrize("type_,mistake", [
    (0, dict(scale=np.pi)), (0, dict(swap="yz")), (0, dict(swap="xy")),
    (1, dict(scale=np.pi)), (3, dict(scale=np.pi)), (5, dict(scale=np.pi)),
    (2, dict(unscaled2=False)), (5, dict(swap="xy", scale=np.pi)),
])
def test_create_lattice_planted_mistakes_fail(type_, mistake):
    assert lattice_failures(type_, planted(type_)) == []  # the unchanged form passes
    assert lattice_failures(type_, planted(type_, **mistake)) != []


def test_create_lattice_type4_planted_mistake_fails():
    F = planted(4)
    assert lattice_failures(4, F) == []
    r = (((np.arange(N) * 1.0) / (N - 1)) - 0.5) / 0.5
    X, Y, Z = np.meshgrid(r, r, r, indexing="ij")
    wrong = np.maximum(np.maximum(X * X + Y * Y, Y * Y + Z * Z), Z * Z + X * X)  # max for min
    assert lattice_failures(4, wrong.astype(np.float32)) != []


# ---- .obj known answer -----------------------------------------------------

def obj_input():
    tris = [
        # T0: truncation toward zero; -0.0019 -> -1 thousandth
        (0.0012, 0.5, -0.0019), (1.2345, 0.0, 0.0), (0.0, 1.0, 0.0),
        # T1: welds onto v1 and v2; -0.0004 truncates to +0
        (0.0015, 0.5, -0.0011), (1.2349, 0.0, 0.0), (-0.0004, 0.0, 1.0),
        # T2: degenerate after the weld (0.0001 -> 0): not written; -2.5 is a new vertex
        (0.0, 1.0, 0.0), (0.0001, 1.0, 0.0), (-2.5, 5.0, 5.0),
        # T3: the same ordered triple as T0: not written again
        (0.0012, 0.5, -0.0019), (1.2345, 0.0, 0.0), (0.0, 1.0, 0.0),
        # T4: T0 rotated: a different ordered triple, written
        (1.2345, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0012, 0.5, -0.0019),
        # T5: a new vertex with more digits; 0.0999 -> 99 thousandths
        (12.3456, 0.0999, 3.0), (1.2345, 0.0, 0.0), (0.0, 1.0, 0.0),
    ]
    pos = np.zeros((len(tris), 4), np.float32)
    pos[:, :3] = np.asarray(tris, np.float32)
    pos[:, 3] = 1.0
    return pos


# vertex lines in first-seen order; faces written "a c b" (winding flipped)
OBJ_EXPECTED = (
    "##Sample latttice new Obj \n"
    "o Solid \n"
    "v 0.001 0.5 -0.001\n"
    "v 1.234 0 0\n"
    "v 0 1 0\n"
    "v 0 0 1\n"
    "v -2.5 5 5\n"
    "v 12.345 0.099 3\n"
    "\n"
    "\n"
    " f  1 3 2\n"
    " f  1 4 2\n"
    " f  2 1 3\n"
    " f  6 3 2\n"
)


def test_obj_writer_known_answer(orc, tmp_path):
    pos = obj_input()
    path = str(tmp_path / "kat.obj")
    assert orc.file_write_obj(pos, len(pos), path) == 0
    assert open(path).read() == OBJ_EXPECTED


@pytest.mark.parametrize("mistake", ["round", "no_flip", "no_face_dedup", "keep_degenerate"])
def test_obj_known_answer_rejects_planted_mistakes(mistake):
    """A small Python writer with one rule changed gives a different text."""
    pos = obj_input()
    out = ["##Sample latttice new Obj \n", "o Solid \n"]
    ids, faces = {}, []
    for v in pos[:, :3]:
        if mistake == "round":
            key = tuple(float(np.float32(round(float(c) * 1000) * 0.001)) for c in v)
        else:
            key = tuple(float(np.float32(int(np.float32(c) * np.float32(1000)) * 0.001)) for c in v)
        if key not in ids:
            ids[key] = len(ids) + 1
            out.append("v %s %s %s\n" % tuple("%g" % c for c in key))
        faces.append(ids[key])
    out += ["\n", "\n"]
    seen = set()
    for i in range(0, len(faces), 3):
        a, b, c = faces[i:i + 3]
        if mistake != "keep_degenerate" and len({a, b, c}) < 3:
            continue
        if mistake != "no_face_dedup" and (a, b, c) in seen:
            continue
        seen.add((a, b, c))
        out.append(" f  %d %d %d\n" % ((a, b, c) if mistake == "no_flip" else (a, c, b)))
    assert "".join(out) != OBJ_EXPECTED


# ---- band interpolation: the equal-values branch ----------------------------

def band_case(nz, corner_z, second_band=False):
    """A 2 x 2 x nz grid whose only indicator-1 (or -2) point is (0, 0, corner_z);
    normalised values 0.5 (and 0.8 for the second band's field) everywhere:
    equal at both ends of every edge and outside the band levels."""
    n = 2 * 2 * nz
    ind = np.zeros(n, np.float32)
    ind[corner_z * 4] = 2.0 if second_band else 1.0
    k1 = np.full(n, 0.5, np.float32)
    k2 = np.full(n, 0.8, np.float32)
    return ind, k1, k2


# hand-worked vertices (voxel size 1, centre 0), in triTable row order:
# cube 254 = {0, 3, 8}, cube 239 = {4, 8, 7} (src/tables.h)
BAND_ONE_CELL = [  # cell at z = 0: every edge of corner 0 has p0.z == 0 -> t = 1
    [(1, 0, 0), (0, 0, 0), (0, 0, 1)],  # edges 0 (p1 = v1), 3 (p1 = v0), 8 (p1 = v4)
]
BAND_TWO_CELLS = [
    [(0, 0, 1), (0, 0, 1), (0, 1, 1)],  # cell z = 0, cube 239: edge 4 (p0 = v4, z = 1 -> t = 0),
                                        # edge 8 (p0 = v0, z = 0 -> t = 1: v4), edge 7 (p0 = v7, t = 0)
    [(0, 0, 1), (0, 1, 1), (0, 0, 1)],  # cell z = 1, cube 254: edges 0, 3, 8 all p0.z = 1 -> t = 0
]


def expected_band(tris):
    pos = np.asarray([v for t in tris for v in t], np.float32)
    nrm = []
    for t in tris:
        a, b, c = (np.asarray(v, np.float64) for v in t)
        nrm += [np.cross(b - a, c - a)] * 3
    return pos, np.asarray(nrm, np.float32)


@pytest.mark.parametrize("second_band", [False, True])
@pytest.mark.parametrize("nz,corner_z,tris", [(2, 0, BAND_ONE_CELL), (3, 1, BAND_TWO_CELLS)])
def test_band_equal_values_branch(orc, nz, corner_z, tris, second_band):
    ind, k1, k2 = band_case(nz, corner_z, second_band)
    mode = orc.MODE_BAND2 if second_band else orc.MODE_BAND1
    out = orc.compute_isosurface(mode, ind, 2, 2, nz, 0.25, vol1=k1, vol2=k2 if second_band else None,
                                 lo1=0.2, hi1=0.3, lo2=0.6, hi2=0.7)
    pos, nrm = expected_band(tris)
    assert out["total"] == len(pos)
    assert np.array_equal(out["pos"][: len(pos), :3], pos)
    assert np.array_equal(out["pos"][: len(pos), 3], np.ones(len(pos), np.float32))
    assert np.array_equal(out["norm"][: len(pos), :3], nrm)


def test_band_equal_values_planted_mistake_fails():
    """With the branch's two outcomes exchanged (t = 0 when p0.z == 0) the
    hand-worked vertices of the z = 0 cell would be the edges' p0 ends."""
    wrong = [(0, 0, 0), (0, 1, 0), (0, 0, 0)]  # edges 0, 3, 8 at p0
    pos, _ = expected_band(BAND_ONE_CELL)
    assert not np.array_equal(np.asarray(wrong, np.float32), pos)
