"""Pins of the oracle's Modelling primitives (src/Modelling.cu:244-761) by
geometry that does not retype the oracle's formulas: float64 closed forms of
each shape, rotations from scipy (Rotation.from_euler('ZYX'), an independent
implementation of the Z·Y·X Euler product the reference's rows spell out at
Modelling.cu:401-407), special points where a surface must pass, and a check
that the wrong rotation order would be caught.

Grid convention (Modelling.cu:395-397): point (x, y, z) sits at
((x - (nx-1)/2) dx, (y - (ny-1)/2) dy, (z - (nz-1)/2) dz); v = point - center;
the shape sees u = R v (rows of R)."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

N = (21, 19, 23)
D = (0.1, 0.12, 0.09)
SPHERE, CUBOID, CUBOID_SHELL, FRUSTUM, TORUS, CONE, CONE_FRUSTUM, LINE = range(8)


def grid_v(center, n=N, d=D):
    nx, ny, nz = n
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = np.stack([(x - (nx - 1) / 2) * d[0], (y - (ny - 1) / 2) * d[1],
                  (z - (nz - 1) / 2) * d[2]], axis=-1).reshape(-1, 3)
    return p - np.asarray(center, np.float64)


def rotated(v, angles, order="ZYX"):
    ax, ay, az = angles
    m = Rotation.from_euler(order, [az, ay, ax] if order == "ZYX" else [ax, ay, az]).as_matrix()
    return v @ m.T


def prim(orc, kind, center, angles, prm, flag=0):
    return orc.primitive(kind, *N, D, center, angles, prm, flag=flag).astype(np.float64)


def test_sphere_is_the_squared_distance_minus_r2(orc):
    c, r = (0.1, -0.2, 0.05), 0.7
    f = prim(orc, SPHERE, c, (0, 0, 0), [r])
    v = grid_v(c)
    want = (v ** 2).sum(1) - r * r
    assert np.max(np.abs(f - want)) < 2e-6 * np.max(np.abs(want))
    # centred on a grid point: exactly -r^2 there
    c0 = ((3 - (N[0] - 1) / 2) * D[0], 0.0, 0.0)
    f0 = prim(orc, SPHERE, np.float32(c0), (0, 0, 0), [r])
    i = (N[2] // 2 * N[1] + N[1] // 2) * N[0] + 3
    assert f0[i] == np.float32(-np.float32(r) * np.float32(r))


def test_sphere_shell_has_both_radii_as_zero_set(orc):
    c, r, t = (0.0, 0.0, 0.0), 0.6, 0.2
    f = prim(orc, SPHERE, c, (0, 0, 0), [r, t], flag=1)
    r2 = (grid_v(c) ** 2).sum(1)
    want = np.maximum(-(r2 - (r - t / 2) ** 2), r2 - (r + t / 2) ** 2)
    assert np.max(np.abs(f - want)) < 1e-5
    inside = f < 0
    rr = np.sqrt(r2)
    assert np.all((rr[inside] > r - t / 2 - 1e-6) & (rr[inside] < r + t / 2 + 1e-6))


@pytest.mark.parametrize("angles", [(0.0, 0.0, 0.0), (0.3, -0.7, 1.1), (1.2, 0.4, -2.0)])
def test_cuboid_is_the_max_of_slabs_in_rotated_axes(orc, angles):
    c, w = (0.05, -0.1, 0.2), (0.9, 0.5, 0.7)
    f = prim(orc, CUBOID, c, angles, list(w))
    u = rotated(grid_v(c), angles)
    want = np.max(np.abs(u) - np.asarray(w) / 2, axis=1)
    assert np.max(np.abs(f - want)) < 2e-6
    if any(angles):
        # the X·Y·Z order (a plausible mistake) does not match
        wrong = np.max(np.abs(rotated(grid_v(c), angles, "XYZ")) - np.asarray(w) / 2, axis=1)
        assert np.max(np.abs(f - wrong)) > 1e-2


def test_cuboid_quarter_turns_swap_its_widths(orc):
    c, w = (0.0, 0.0, 0.0), [0.9, 0.5, 0.7]
    for axis, (i, j) in ((0, (1, 2)), (1, (0, 2)), (2, (0, 1))):
        ang = [0.0, 0.0, 0.0]
        ang[axis] = np.pi / 2
        ws = list(w)
        ws[i], ws[j] = ws[j], ws[i]
        a = prim(orc, CUBOID, c, ang, w)
        b = prim(orc, CUBOID, c, (0, 0, 0), ws)
        assert np.max(np.abs(a - b)) < 1e-6, axis


def test_cuboid_shell_is_a_tube_in_x_y_capped_in_z(orc):
    c, w, th = (0.0, 0.0, 0.0), (1.2, 1.0, 0.8), 0.15
    f = prim(orc, CUBOID_SHELL, c, (0, 0, 0), list(w) + [th])
    u = grid_v(c)
    hx, hy, hz = (x / 2 for x in w)
    outer = np.maximum(np.abs(u[:, 0]) - hx, np.abs(u[:, 1]) - hy)
    inner = np.maximum(np.abs(u[:, 0]) - (hx - th), np.abs(u[:, 1]) - (hy - th))
    want = np.maximum(np.maximum(outer, -inner), np.abs(u[:, 2]) - hz)
    assert np.max(np.abs(f - want)) < 2e-6


def test_pyramid_frustum_passes_through_its_corners(orc):
    xb, xt, h, zb, zt = 1.0, 0.4, 0.8, 0.9, 0.3
    c = (0.0, -0.4, 0.0)
    f = prim(orc, FRUSTUM, c, (0, 0, 0), [xb, xt, h, zb, zt])
    u = grid_v(c)
    # half widths fall linearly from the base (u_y = 0) to the top (u_y = h)
    ratio = (h - u[:, 1]) / h
    want = np.max(np.stack([np.abs(u[:, 0]) - (ratio * (xb - xt) + xt) / 2,
                            np.abs(u[:, 1] - h / 2) - h / 2,
                            np.abs(u[:, 2]) - (ratio * (zb - zt) + zt) / 2]), axis=0)
    assert np.max(np.abs(f - want)) < 2e-6
    assert f[np.argmin(np.abs(want))] == pytest.approx(0.0, abs=0.05)


@pytest.mark.parametrize("angles", [(0.0, 0.0, 0.0), (0.5, 0.2, -0.9)])
def test_torus_is_the_squared_tube_distance(orc, angles):
    c, tr, cr = (0.1, 0.0, -0.05), 0.7, 0.25
    f = prim(orc, TORUS, c, angles, [tr, cr])
    u = rotated(grid_v(c), angles)
    rho = np.hypot(u[:, 0], u[:, 1])
    want = (tr - rho) ** 2 + u[:, 2] ** 2 - cr * cr
    assert np.max(np.abs(f - want)) < 2e-6
    # on the ring circle -cr^2, at the centre tr^2 - cr^2
    assert np.min(f) == pytest.approx(-cr * cr, abs=0.02)


def test_cone_rim_and_apex(orc):
    br, h = 0.6, 1.0
    c = (0.0, -0.5, 0.0)
    f = prim(orc, CONE, c, (0, 0, 0), [br, h])
    u = grid_v(c)
    # lateral surface: rho (h / br) = h - u_y, between the planes u_y = 0 and u_y = h
    lateral = (u[:, 0] ** 2 + u[:, 2] ** 2) * (h / br) ** 2 - (u[:, 1] - h) ** 2
    caps = np.maximum((u[:, 1] - h) * 100.0, -u[:, 1] * 100.0)
    want = np.maximum(lateral, caps)
    assert np.max(np.abs(f - want)) < 2e-5
    # the base rim (rho = br, u_y = 0) is on the surface
    rim = np.abs(np.hypot(u[:, 0], u[:, 2]) - br) + np.abs(u[:, 1])
    assert np.all(np.abs(want[rim < 1e-9]) < 1e-6)


def test_cone_frustum_radii_at_both_ends(orc):
    tr, br, h = 0.3, 0.7, 1.0
    c = (0.0, -0.5, 0.0)
    f = prim(orc, CONE_FRUSTUM, c, (0, 0, 0), [tr, br, h])
    u = grid_v(c)
    r = ((h - u[:, 1]) / h) * (br - tr) + tr  # br at u_y = 0, tr at u_y = h
    lateral = u[:, 0] ** 2 + u[:, 2] ** 2 - r * r
    caps = np.maximum((u[:, 1] - h) * 100.0, -u[:, 1] * 100.0)
    want = np.maximum(lateral, caps)
    assert np.max(np.abs(f - want)) < 2e-5
    base = np.abs(u[:, 1]) < 1e-9
    assert np.all(np.abs(want[base] - (u[base, 0] ** 2 + u[base, 2] ** 2 - br * br)) < 1e-9)


def test_line_primitive_is_a_capped_cylinder(orc):
    c, axis, radius, length = (0.05, 0.0, -0.1), (1.0, 2.0, 0.5), 0.3, 1.2
    f = prim(orc, LINE, c, axis, [radius, 0.0, length])
    a = np.asarray(axis) / np.linalg.norm(axis)
    v = grid_v(c)
    along = v @ a
    dist = np.linalg.norm(np.cross(v, a), axis=1)  # distance from the axis line
    want = np.maximum(np.abs(along) - length / 2, dist - radius)
    assert np.max(np.abs(f - want)) < 1e-5
