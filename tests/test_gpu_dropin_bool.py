"""GPU parity of the drop-in boolean meshing (Isosurface::computeIsosurface,
src/Isosurface.cu:44-134) against the oracle's restatement, through the C-ABI:
classify_copy_Voxel's grid_points (val and edge parameters), voxel counts,
occupancy, both scans, compaction, active/total/nfacets and the non-indexed
pos/norm are compared bit for bit, for every boolean mode (union, intersect,
difference; with a primitive, with a lattice band as "fixed" or "dynamic";
make_region).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from tests.helpers import to_np

pytestmark = pytest.mark.gpu

ISO, LO, HI = 0.0, 0.2, 0.3


def grid(nx, ny, nz, fn):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.ascontiguousarray(fn(x.astype(np.float64), y, z).astype(np.float32).reshape(-1))


def fields(nx, ny, nz):
    """A sphere, a box and a normalised lattice-like field in [0, 1]."""
    a = grid(nx, ny, nz, lambda x, y, z: np.sqrt((x - 0.45 * nx) ** 2 + (y - 0.5 * ny) ** 2 +
                                                 (z - 0.52 * nz) ** 2) - 0.33 * nx)
    b = grid(nx, ny, nz, lambda x, y, z: np.maximum(np.maximum(abs(x - 0.58 * nx) - 0.27 * nx,
                                                               abs(y - 0.47 * ny) - 0.3 * ny),
                                                    abs(z - 0.5 * nz) - 0.25 * nz))
    lat = grid(nx, ny, nz, lambda x, y, z: 0.5 + 0.5 * np.sin(x * 0.61) * np.cos(y * 0.53) *
               np.sin(z * 0.47 + 0.3))
    return a, b, lat


def gp_tensor(orc, arr):
    return torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).copy()).cuda()


def gp_from_tensor(orc, t):
    return to_np(t).view(orc.GRID_POINT)


def grid_args(gc, nx, ny, nz):
    return (gc.Uint3(nx, ny, nz), gc.Uint3(1, nx - 1, (nx - 1) * (ny - 1)),
            gc.Uint3(nx - 1, ny - 1, nz - 1))


MODES = []
for name, op in (("union", 1), ("intersect", 4), ("diff", 2)):
    MODES.append((name, op))
    MODES.append((name + "_fixed", op | 8))
    MODES.append((name + "_dynamic", op | 16))
MODES.append(("make_region", 32))


@pytest.mark.parametrize("name,flags", MODES, ids=[m[0] for m in MODES])
@pytest.mark.parametrize("n", [23, 48])
def test_boolean_pipeline_matches_oracle(gc, orc, dev, name, flags, n):
    nx, ny, nz = n, n - 3, n - 5
    a, b, lat = fields(nx, ny, nz)
    npts = nx * ny * nz
    # grid_points of operand A; the copy pass merges operand B (or the lattice
    # band for the dynamic modes) into it, as the app does before meshing
    g0 = np.zeros(npts, orc.GRID_POINT)
    g0["val"] = np.where(a < ISO, -1, 1)
    g0["t_x"] = np.where(np.arange(npts) % 7 == 0, 0.25, 0.0)  # some prior parameters
    ref_g = orc.bool_copy(flags, g0, b, lat, nx, ny, nz, ISO, LO, HI)
    d_g = gp_tensor(orc, g0)
    d_b = torch.from_numpy(b).cuda()
    d_lat = torch.from_numpy(lat).cuda()
    gs, sh, mk = grid_args(gc, nx, ny, nz)
    gc.check(gc.lib.gpucad_classify_copy_voxel(gc.ptr(d_g), gc.ptr(d_b), gc.ptr(d_lat), gs, LO, HI,
                                               ISO, flags))
    got_g = gp_from_tensor(orc, d_g)
    assert np.array_equal(got_g.view(np.uint32), ref_g.view(np.uint32)), "classify_copy"
    # mesh: fixed = copied grid_points, dynamic = B, lattice = lat
    ref = orc.bool_compute_isosurface(flags, ref_g, b, lat, nx, ny, nz, ISO, LO, HI,
                                      voxel_size=(0.5, 0.5, 0.5), center=(1.0, 2.0, 3.0))
    ncell = (nx - 1) * (ny - 1) * (nz - 1)
    max_verts = max(npts * 4, 300000)
    i32 = lambda: torch.zeros(ncell, dtype=torch.int32, device="cuda")
    verts, verts_scan, occ, occ_scan, comp = i32(), i32(), i32(), i32(), i32()
    pos = torch.full((max_verts, 4), 7.0, dtype=torch.float32, device="cuda")
    norm = torch.full_like(pos, 7.0)
    active, total, nfacets = C.c_uint32(0), C.c_uint32(0), C.c_uint64(12345)
    gc.check(gc.lib.gpucad_compute_isosurface(
        gc.ptr(pos), gc.ptr(norm), ISO, ncell, gc.ptr(verts), gc.ptr(verts_scan), gc.ptr(occ),
        gc.ptr(occ_scan), gs, sh, mk, gc.Float3(0.5, 0.5, 0.5), gc.Float3(1.0, 2.0, 3.0),
        C.byref(active), C.byref(total), gc.ptr(comp), max_verts, gc.ptr(d_g), gc.ptr(d_b),
        gc.ptr(d_lat), LO, HI, flags, C.byref(nfacets)))
    torch.cuda.synchronize()
    u = lambda t: to_np(t).view(np.uint32)
    print(name, n, active.value, total.value, flush=True)
    assert (active.value, total.value) == (ref["active"], ref["total"])
    assert np.array_equal(u(verts), ref["verts"])
    assert np.array_equal(u(occ), ref["occ"])
    assert np.array_equal(u(occ_scan), ref["occ_scan"])
    if ref["active"] == 0:
        assert nfacets.value == 12345  # left as it was, like the reference
        return
    assert nfacets.value == ref["total"] // 3
    assert np.array_equal(u(verts_scan), ref["verts_scan"])
    assert np.array_equal(u(comp)[:ref["active"]], ref["comp"][:ref["active"]])
    t = ref["total"]
    assert np.array_equal(to_np(pos)[:t].view(np.uint32), ref["pos"][:t].view(np.uint32))
    assert np.array_equal(to_np(norm)[:t].view(np.uint32), ref["norm"][:t].view(np.uint32))
    # the first max_verts bytes are cleared, the rest untouched (reference memset)
    cleared = max_verts // 16
    if t < cleared:
        assert np.all(to_np(pos)[t:cleared] == 0)
    assert np.all(to_np(pos)[max(t, cleared + 1):] == 7.0)


def test_boolean_scans_with_leading_empty_tiles(gc, orc, dev):
    """make_region of a ball in the top of the grid (the first 137 scan tiles
    hold no active cell); verts_scan starts with other contents and must end
    as the oracle's, its leading zeros included."""
    nx, ny, nz = 90, 84, 100
    a = grid(nx, ny, nz, lambda x, y, z: np.sqrt((x - 0.5 * nx) ** 2 + (y - 0.5 * ny) ** 2 +
                                                 (z - 0.85 * nz) ** 2) - 8.5)
    npts = nx * ny * nz
    g0 = np.zeros(npts, orc.GRID_POINT)
    g0["val"] = np.where(a < ISO, -1, 1)
    g0["t_y"] = np.where(np.arange(npts) % 5 == 0, 0.5, 0.0)
    flags = 32
    zeros = np.zeros(npts, np.float32)
    ref = orc.bool_compute_isosurface(flags, g0, zeros, zeros, nx, ny, nz, ISO, LO, HI)
    first = int(np.argmax(ref["occ"] != 0))
    assert first >= 100 * 4096
    ncell = (nx - 1) * (ny - 1) * (nz - 1)
    max_verts = max(npts * 4, 300000)
    i32 = lambda: torch.zeros(ncell, dtype=torch.int32, device="cuda")
    verts, occ, occ_scan, comp = i32(), i32(), i32(), i32()
    verts_scan = torch.full((ncell,), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    pos = torch.zeros((max_verts, 4), dtype=torch.float32, device="cuda")
    norm = torch.zeros_like(pos)
    d_g = gp_tensor(orc, g0)
    gs, sh, mk = grid_args(gc, nx, ny, nz)
    active, total, nfacets = C.c_uint32(0), C.c_uint32(0), C.c_uint64(0)
    gc.check(gc.lib.gpucad_compute_isosurface(
        gc.ptr(pos), gc.ptr(norm), ISO, ncell, gc.ptr(verts), gc.ptr(verts_scan), gc.ptr(occ),
        gc.ptr(occ_scan), gs, sh, mk, gc.Float3(1, 1, 1), gc.Float3(0, 0, 0), C.byref(active),
        C.byref(total), gc.ptr(comp), max_verts, gc.ptr(d_g), None, None, LO, HI, flags,
        C.byref(nfacets)))
    torch.cuda.synchronize()
    u = lambda t: to_np(t).view(np.uint32)
    assert (active.value, total.value) == (ref["active"], ref["total"])
    assert np.array_equal(u(verts_scan), ref["verts_scan"])
    assert np.array_equal(u(occ_scan), ref["occ_scan"])
    assert np.array_equal(u(comp)[:ref["active"]], ref["comp"][:ref["active"]])
    t = ref["total"]
    assert np.array_equal(to_np(pos)[:t].view(np.uint32), ref["pos"][:t].view(np.uint32))


def test_stage_entry_points_boolean(gc, orc, dev):
    nx, ny, nz = 30, 27, 25
    a, b, lat = fields(nx, ny, nz)
    g = np.zeros(nx * ny * nz, orc.GRID_POINT)
    g["val"] = np.where(a < ISO, -1, 1)
    g = orc.bool_copy(1, g, b, None, nx, ny, nz, ISO, LO, HI)
    flags = 4  # intersect of the two primitives
    ncell = (nx - 1) * (ny - 1) * (nz - 1)
    verts = torch.zeros(ncell, dtype=torch.int32, device="cuda")
    occ = torch.zeros_like(verts)
    gs, sh, mk = grid_args(gc, nx, ny, nz)
    d_g, d_b = gp_tensor(orc, g), torch.from_numpy(b).cuda()
    gc.check(gc.lib.gpucad_classify_voxel_boolean(gc.ptr(verts), gc.ptr(occ), gc.ptr(d_g),
                                                  gc.ptr(d_b), None, gs, sh, mk, ncell, LO, HI,
                                                  ISO, flags))
    rv, ro = orc.bool_classify(flags, g, b, None, nx, ny, nz, ISO, LO, HI)
    assert np.array_equal(to_np(verts).view(np.uint32), rv)
    assert np.array_equal(to_np(occ).view(np.uint32), ro)
    # generate from the oracle's scans and compaction
    ref = orc.bool_compute_isosurface(flags, g, b, None, nx, ny, nz, ISO, LO, HI)
    max_verts = max(nx * ny * nz * 4, 300000)
    pos = torch.zeros((max_verts, 4), dtype=torch.float32, device="cuda")
    norm = torch.zeros_like(pos)
    comp = torch.from_numpy(ref["comp"].view(np.int32)).cuda()
    scan = torch.from_numpy(ref["verts_scan"].view(np.int32)).cuda()
    gc.check(gc.lib.gpucad_generate_triangles_boolean(
        gc.ptr(pos), gc.ptr(norm), gc.ptr(comp), gc.ptr(scan), gs, sh, mk, gc.Float3(1, 1, 1),
        gc.Float3(0, 0, 0), ISO, ref["active"], max_verts, gc.ptr(d_g), gc.ptr(d_b), None, LO, HI,
        flags))
    t = ref["total"]
    assert t > 0
    assert np.array_equal(to_np(pos)[:t].view(np.uint32), ref["pos"][:t].view(np.uint32))
    assert np.array_equal(to_np(norm)[:t].view(np.uint32), ref["norm"][:t].view(np.uint32))


def test_boolean_missing_operand_rejected(gc, dev):
    nx = ny = nz = 8
    ncell = 7 * 7 * 7
    verts = torch.zeros(ncell, dtype=torch.int32, device="cuda")
    gs, sh, mk = (gc.Uint3(nx, ny, nz), gc.Uint3(1, nx - 1, (nx - 1) * (ny - 1)),
                  gc.Uint3(nx - 1, ny - 1, nz - 1))
    # union with a lattice band ("fixed") reads the lattice: NULL is an error
    st = gc.lib.gpucad_classify_voxel_boolean(gc.ptr(verts), gc.ptr(verts), None, gc.ptr(verts),
                                              None, gs, sh, mk, ncell, LO, HI, ISO, 1 | 8)
    assert st == gc.E_ARG
