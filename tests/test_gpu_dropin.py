"""GPU parity of the drop-in entry points (the reference's operator API).

Each call is made through the C-ABI exactly as the reference's caller makes
it (src/main.cu:4006-4010 and 4115-4132 for the lattice extraction,
src/main.cu:3334-3381 for the primitives) and compared with the oracle's
restatement on the same inputs:
  * voxel counts, occupancy, both scans, the compacted voxel list, active and
    total vertex counts: bit-exact;
  * non-indexed pos/norm: bit-exact (same arithmetic on both sides);
  * reference quirks kept: writes at or past maxVerts-3 are dropped while
    totalVerts is reported in full; zero active voxels give totalVerts = 0;
  * fields: normalise_four bit-exact, primitives within 1 ulp, unit-cell
    lattices within 2e-6 (device vs host libm cosf/sinf);
  * .obj: byte-identical to the oracle's std::map restatement of
    File_output::file_write_obj.
"""
import ctypes as C
import os

import numpy as np
import pytest
import torch

from gpucadforam_amd import configs as CF
from tests.helpers import to_np, ulp_diff

pytestmark = pytest.mark.gpu


def grid_args(gc, nx, ny, nz):
    return (gc.Uint3(nx, ny, nz), gc.Uint3(1, nx - 1, (nx - 1) * (ny - 1)),
            gc.Uint3(nx - 1, ny - 1, nz - 1))


class DropinBuffers:
    def __init__(self, n_cells, max_verts, dev, scan_fill=0):
        self.verts = torch.zeros(n_cells, dtype=torch.int32, device=dev)
        self.verts_scan = torch.full_like(self.verts, scan_fill)
        self.occ = torch.zeros_like(self.verts)
        self.occ_scan = torch.zeros_like(self.verts)
        self.comp = torch.zeros_like(self.verts)
        self.pos = torch.zeros((max_verts, 4), dtype=torch.float32, device=dev)
        self.norm = torch.zeros_like(self.pos)


def run_dropin(gc, mode, vol, nx, ny, nz, iso, vol1=None, vol2=None, lo1=0.0, hi1=0.0, lo2=0.0,
               hi2=0.0, voxel_size=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), max_verts=None,
               scan_fill=0):
    dev = vol.device
    n = (nx - 1) * (ny - 1) * (nz - 1)
    if max_verts is None:
        max_verts = max(nx * ny * nz * 4, 300000)
    b = DropinBuffers(n, max_verts, dev, scan_fill)
    gs, sh, mk = grid_args(gc, nx, ny, nz)
    active, total = C.c_uint32(0), C.c_uint32(0)
    vs, ce = gc.Float3(*voxel_size), gc.Float3(*center)
    common = (gc.ptr(vol), gc.ptr(b.pos), gc.ptr(b.norm), iso, n, gc.ptr(b.verts),
              gc.ptr(b.verts_scan), gc.ptr(b.occ), gc.ptr(b.occ_scan), gs, sh, mk, vs, ce,
              C.byref(active), C.byref(total), gc.ptr(b.comp), max_verts)
    if mode == "lattice":
        st = gc.lib.gpucad_compute_isosurface_lattice(*common, gc.ptr(vol1), gc.ptr(vol2), lo1,
                                                      hi1, lo2, hi2)
    elif mode == "latticeone":
        st = gc.lib.gpucad_compute_isosurface_latticeone(*common, gc.ptr(vol1), lo1, hi1)
    else:
        st = gc.lib.gpucad_compute_isosurface_standard(*common)
    gc.check(st)
    torch.cuda.synchronize()
    return b, active.value, total.value


def compare_with_oracle(orc, b, active, total, ref, check_pos=True):
    assert (active, total) == (ref["active"], ref["total"])
    u = lambda t: to_np(t).view(np.uint32)
    assert np.array_equal(u(b.verts), ref["verts"])
    assert np.array_equal(u(b.occ), ref["occ"])
    assert np.array_equal(u(b.occ_scan), ref["occ_scan"])
    if active:
        assert np.array_equal(u(b.verts_scan), ref["verts_scan"])
        assert np.array_equal(u(b.comp)[:active], ref["comp"][:active])
    if check_pos and total:
        lim = min(total, ref["pos"].shape[0])
        assert np.array_equal(to_np(b.pos)[:lim].view(np.uint32), ref["pos"][:lim].view(np.uint32))
        assert np.array_equal(to_np(b.norm)[:lim].view(np.uint32), ref["norm"][:lim].view(np.uint32))


@pytest.mark.parametrize("ltype", range(6))
def test_unit_lattice_band_pipeline(gc, orc, dev, ltype):
    """The unit-lattice view (src/main.cu:4106-4137): create_lattice on 61^3,
    GPU_buffer_normalise_four, computeIsosurface_latticeone with voxel size 1."""
    n = 61
    f = torch.empty(n ** 3, dtype=torch.float32, device=dev)
    gc.check(gc.lib.gpucad_create_lattice(gc.ptr(f), n, n, n, n ** 3, ltype))
    torch.cuda.synchronize()
    ref_f = orc.create_lattice(n, n, n, ltype)
    assert np.abs(to_np(f) - ref_f).max() < 2e-6
    assert np.abs(ref_f - orc.create_lattice(n, n, n, ltype, f64=True)).max() < 2e-6
    field = to_np(f)
    ind = torch.empty_like(f)
    gc.check(gc.lib.gpucad_normalise_four(gc.ptr(f), gc.ptr(ind), gc.ptr(f), n ** 3, n, n, n,
                                          0.2, 0.3))  # normalised in place, as the caller does
    torch.cuda.synchronize()
    r_ind, r_nrm = orc.normalise_four(field, n, n, n, 0.2, 0.3)
    assert np.array_equal(to_np(ind), r_ind)
    assert np.array_equal(to_np(f).view(np.uint32), r_nrm.view(np.uint32))
    b, active, total = run_dropin(gc, "latticeone", ind, n, n, n, 0.25, vol1=f, lo1=0.2, hi1=0.3)
    ref = orc.compute_isosurface(orc.MODE_BAND1, r_ind, n, n, n, 0.25, vol1=r_nrm, lo1=0.2,
                                 hi1=0.3, max_verts=b.pos.shape[0])
    compare_with_oracle(orc, b, active, total, ref)


def test_spatial_lattice_two_field_pipeline(gc, orc, dev):
    """computeIsosurface_lattice as called at src/main.cu:4006-4010: indicator
    volume, normalised field, a zero second field, iso 0.25, band 0.20/0.30,
    voxel size 0.5 on the fine grid."""
    n = 64
    p = CF.varying_schwarz_p(n)
    f = gc.lattice_field(p)
    field = to_np(f)
    ind = torch.empty_like(f)
    nrm = torch.empty_like(f)
    gc.check(gc.lib.gpucad_normalise_four(gc.ptr(f), gc.ptr(ind), gc.ptr(nrm), n ** 3, n, n, n,
                                          0.2, 0.3))
    zeros = torch.zeros_like(f)
    b, active, total = run_dropin(gc, "lattice", ind, n, n, n, 0.25, vol1=nrm, vol2=zeros,
                                  lo1=0.2, hi1=0.3, lo2=0.0, hi2=0.0, voxel_size=(0.5, 0.5, 0.5))
    r_ind, r_nrm = orc.normalise_four(field, n, n, n, 0.2, 0.3)
    ref = orc.compute_isosurface(orc.MODE_BAND2, r_ind, n, n, n, 0.25, vol1=r_nrm,
                                 vol2=np.zeros_like(r_nrm), lo1=0.2, hi1=0.3,
                                 voxel_size=(0.5, 0.5, 0.5), max_verts=b.pos.shape[0])
    assert total > 1000
    compare_with_oracle(orc, b, active, total, ref)


@pytest.mark.parametrize("n", [33, 128])
def test_standard_pipeline_on_lattice(gc, orc, dev, n):
    p = CF.plain_gyroid(n)
    f = gc.lattice_field(p)
    b, active, total = run_dropin(gc, "standard", f, n, n, n, 0.0, voxel_size=(0.5, 0.25, 1.0),
                                  center=(3.0, -1.0, 0.5))
    ref = orc.compute_isosurface(orc.MODE_STD, to_np(f), n, n, n, 0.0, voxel_size=(0.5, 0.25, 1.0),
                                 center=(3.0, -1.0, 0.5), max_verts=b.pos.shape[0])
    compare_with_oracle(orc, b, active, total, ref)


def test_max_verts_overflow_keeps_reference_behaviour(gc, orc, dev):
    n = 40
    f = gc.lattice_field(CF.plain_gyroid(n))
    full, _, total = run_dropin(gc, "standard", f, n, n, n, 0.0)
    cap = total // 2
    b, active, total2 = run_dropin(gc, "standard", f, n, n, n, 0.0, max_verts=cap)
    assert total2 == total  # totalVerts is not clamped
    # every triangle whose first index is < cap - 3 is written, the rest dropped
    pos = to_np(b.pos)
    ref = to_np(full.pos)[:cap]
    written = np.zeros(cap, bool)
    scan = to_np(full.verts_scan).view(np.uint32)
    verts = to_np(full.verts).view(np.uint32)
    for s_, v_ in zip(scan[verts > 0], verts[verts > 0]):
        for j in range(0, int(v_), 3):
            if s_ + j < cap - 3:
                written[s_ + j:s_ + j + 3] = True
    assert np.array_equal(pos[written].view(np.uint32), ref[written].view(np.uint32))
    assert np.all(pos[~written] == 0)


def test_zero_active_voxels(gc, dev):
    n = 17
    f = torch.ones(n ** 3, dtype=torch.float32, device=dev)
    b, active, total = run_dropin(gc, "standard", f, n, n, n, 0.0)
    assert active == 0 and total == 0


def test_zero_active_voxels_leaves_later_buffers_alone(gc, dev):
    """src/Isosurface.cu:424-436 returns before compactVoxels and the vertex
    scan when no voxel is active: those buffers keep their contents (and may
    be NULL), although the occupancy scan runs and writes its own."""
    n = 19
    cells = (n - 1) ** 3
    vol = torch.ones(n ** 3, dtype=torch.float32, device=dev)
    nrm = torch.zeros_like(vol)
    gs, sh, mk = grid_args(gc, n, n, n)
    b = DropinBuffers(cells, 300000, dev)
    for t in (b.verts_scan, b.comp, b.occ_scan):
        t.fill_(0x5A5A5A5A)
    active, total = C.c_uint32(7), C.c_uint32(7)
    common = (gs, sh, mk, gc.Float3(1, 1, 1), gc.Float3(0, 0, 0), C.byref(active),
              C.byref(total))
    gc.check(gc.lib.gpucad_compute_isosurface_lattice(
        gc.ptr(vol), gc.ptr(b.pos), gc.ptr(b.norm), 0.25, cells, gc.ptr(b.verts),
        gc.ptr(b.verts_scan), gc.ptr(b.occ), gc.ptr(b.occ_scan), *common, gc.ptr(b.comp), 300000,
        gc.ptr(nrm), gc.ptr(nrm), 0.2, 0.3, 0.0, 0.0))
    torch.cuda.synchronize()
    assert active.value == 0 and total.value == 0
    assert torch.all(b.verts_scan == 0x5A5A5A5A) and torch.all(b.comp == 0x5A5A5A5A)
    assert torch.all(b.occ_scan == 0) and torch.all(b.occ == 0) and torch.all(b.verts == 0)
    # the same call with NULL for the buffers the reference does not reach
    gc.check(gc.lib.gpucad_compute_isosurface_lattice(
        gc.ptr(vol), gc.ptr(b.pos), gc.ptr(b.norm), 0.25, cells, gc.ptr(b.verts), None,
        gc.ptr(b.occ), gc.ptr(b.occ_scan), *common, None, 300000, gc.ptr(nrm), gc.ptr(nrm), 0.2,
        0.3, 0.0, 0.0))
    torch.cuda.synchronize()
    assert active.value == 0 and total.value == 0


def ball_field(nx, ny, nz, cz, r):
    """Signed distance to a ball centred at (nx/2, ny/2, cz) of radius r."""
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d = np.sqrt((x - 0.5 * nx) ** 2 + (y - 0.5 * ny) ** 2 + (z - cz) ** 2) - r
    return np.ascontiguousarray(d.astype(np.float32).reshape(-1))


@pytest.mark.parametrize("mode", ["standard", "latticeone"])
def test_scans_with_leading_empty_tiles(gc, orc, dev, mode):
    """Grids whose first scan tiles (4096 cells each) hold no active cell:
    verts_scan starts with other contents and must end as the oracle's, its
    leading zeros included; calls of several sizes in a row (the scans'
    scratch sums are cleared per call), then a call with no active cell,
    which leaves verts_scan alone."""
    cases = [(96, 90, 80, 0.8, 9.0), (61, 70, 66, 0.35, 12.0), (130, 40, 150, 0.9, 6.5),
             (33, 33, 33, 0.5, 6.0)]
    for nx, ny, nz, cz, r in cases:
        vol_np = ball_field(nx, ny, nz, cz * nz, r)
        print("leading empty tiles", mode, nx, ny, nz, flush=True)
        vol = torch.from_numpy(vol_np).to(dev)
        if mode == "standard":
            b, active, total = run_dropin(gc, mode, vol, nx, ny, nz, 0.0, scan_fill=0x5A5A5A5A)
            ref = orc.compute_isosurface(orc.MODE_STD, vol_np, nx, ny, nz, 0.0,
                                         max_verts=b.pos.shape[0])
        else:
            # a band of the normalised field around the sphere (|distance| < 1.5)
            npts = nx * ny * nz
            a_, b_ = min(0.0, float(vol_np.min())), max(0.0, float(vol_np.max()))
            lo, hi = (-1.5 - a_) / (b_ - a_), (1.5 - a_) / (b_ - a_)
            ind = torch.empty_like(vol)
            nrm = torch.empty_like(vol)
            gc.check(gc.lib.gpucad_normalise_four(gc.ptr(vol), gc.ptr(ind), gc.ptr(nrm), npts, nx,
                                                  ny, nz, lo, hi))
            r_ind, r_nrm = orc.normalise_four(vol_np, nx, ny, nz, lo, hi)
            b, active, total = run_dropin(gc, mode, ind, nx, ny, nz, 0.25, vol1=nrm, lo1=lo,
                                          hi1=hi, scan_fill=0x5A5A5A5A)
            ref = orc.compute_isosurface(orc.MODE_BAND1, r_ind, nx, ny, nz, 0.25, vol1=r_nrm,
                                         lo1=lo, hi1=hi, max_verts=b.pos.shape[0])
        first = int(np.argmax(ref["occ"] != 0))
        assert active > 0 and (nz < 40 or first >= 4 * 4096), first
        compare_with_oracle(orc, b, active, total, ref)
    vol = torch.ones(40 * 40 * 40, dtype=torch.float32, device=dev)
    b, active, total = run_dropin(gc, "standard", vol, 40, 40, 40, 0.0, scan_fill=0x5A5A5A5A)
    assert active == 0 and total == 0
    assert torch.all(b.verts_scan == 0x5A5A5A5A)


def test_scans_above_the_prefix_limit(gc, orc, dev):
    """A grid of more than 4096 scan tiles (299 x 239 x 239 cells, 4170
    tiles), which scans its tile sums in one workgroup before the third pass
    instead of each tile adding up the earlier tiles' sums; the first active
    cell lies in tile 1473 and verts_scan starts with other contents."""
    nx, ny, nz = 300, 240, 240
    vol_np = ball_field(nx, ny, nz, 0.6 * nz, 60.0)
    vol = torch.from_numpy(vol_np).to(dev)
    b, active, total = run_dropin(gc, "standard", vol, nx, ny, nz, 0.0, max_verts=3_000_000,
                                  scan_fill=0x5A5A5A5A)
    ref = orc.compute_isosurface(orc.MODE_STD, vol_np, nx, ny, nz, 0.0, max_verts=3_000_000)
    assert (nx - 1) * (ny - 1) * (nz - 1) > 4096 * 4096 and active > 0
    compare_with_oracle(orc, b, active, total, ref)


@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 1_000_003, 9_000_000])
def test_exclusive_scan_u32(gc, orc, dev, n):
    rng = np.random.default_rng(n)
    hi = 2 ** 32 if n < 100 else 16
    a = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    src = torch.from_numpy(a.view(np.int32)).to(dev)
    out = torch.empty_like(src)
    gc.check(gc.lib.gpucad_exclusive_scan_u32(gc.ptr(out), gc.ptr(src), n))
    assert np.array_equal(to_np(out).view(np.uint32), orc.exclusive_scan_u32(a))


def test_stage_entry_points(gc, orc, dev):
    """classify / scan / compact / generate called one by one, as the
    reference's MarchingCubeCuda methods are."""
    n = 48
    p = CF.varying_schwarz_p(n)
    f = gc.lattice_field(p)
    field = to_np(f)
    cells = (n - 1) ** 3
    gs, sh, mk = grid_args(gc, n, n, n)
    verts = torch.zeros(cells, dtype=torch.int32, device=dev)
    occ = torch.zeros_like(verts)
    gc.check(gc.lib.gpucad_classify_voxel_lattice_new(gc.ptr(verts), gc.ptr(occ), gc.ptr(f), gs, sh,
                                                      mk, cells, gc.Float3(1, 1, 1), 0.0))
    cube, rv, ro = orc.classify(field, n, n, n, 0.0)
    assert np.array_equal(to_np(verts).view(np.uint32), rv)
    assert np.array_equal(to_np(occ).view(np.uint32), ro)
    occ_scan = torch.zeros_like(verts)
    verts_scan = torch.zeros_like(verts)
    gc.check(gc.lib.gpucad_exclusive_scan_u32(gc.ptr(occ_scan), gc.ptr(occ), cells))
    gc.check(gc.lib.gpucad_exclusive_scan_u32(gc.ptr(verts_scan), gc.ptr(verts), cells))
    comp = torch.zeros_like(verts)
    gc.check(gc.lib.gpucad_compact_voxels(gc.ptr(comp), gc.ptr(occ), gc.ptr(occ_scan), cells))
    active = int(ro.sum())
    assert np.array_equal(to_np(comp)[:active].view(np.uint32), np.nonzero(ro)[0].astype(np.uint32))
    total = int(rv.sum())
    pos = torch.zeros((total + 8, 4), dtype=torch.float32, device=dev)
    nrm = torch.zeros_like(pos)
    gc.check(gc.lib.gpucad_generate_triangles_standard(
        gc.ptr(pos), gc.ptr(nrm), gc.ptr(comp), gc.ptr(verts_scan), gc.ptr(f), gs, sh, mk,
        gc.Float3(1, 1, 1), gc.Float3(0, 0, 0), 0.0, active, total + 8))
    ref = orc.compute_isosurface(orc.MODE_STD, field, n, n, n, 0.0, max_verts=total + 8)
    assert np.array_equal(to_np(pos)[:total].view(np.uint32), ref["pos"][:total].view(np.uint32))


def test_dropin_rejects_inconsistent_grid(gc, dev):
    n = 9
    f = torch.zeros(n ** 3, dtype=torch.float32, device=dev)
    buf = torch.zeros(1000, dtype=torch.int32, device=dev)
    st = gc.lib.gpucad_classify_voxel_lattice_new(
        gc.ptr(buf), gc.ptr(buf), gc.ptr(f), gc.Uint3(n, n, n), gc.Uint3(1, n - 1, 7),
        gc.Uint3(n - 1, n - 1, n - 1), (n - 1) ** 3, gc.Float3(1, 1, 1), 0.0)
    assert st == gc.E_ARG


PRIMS = [
    # (oracle kind, entry point, args after (data, center, angles|axis), oracle prm, flag)
    (0, "sphere_shell", None, [0.6, 0.2], 1),
    (0, "sphere", None, [0.7, 0.0], 0),
    (1, "cuboid", None, [0.9, 0.6, 0.4], 0),
    (2, "cuboid_shell", None, [0.9, 0.6, 0.4, 0.1], 0),
    (3, "pyramid_frustum", None, [0.8, 0.3, 0.7, 0.6, 0.2], 0),
    (4, "torus", None, [0.5, 0.15], 0),
    (5, "cone", None, [0.4, 0.8], 0),
    (6, "cone_frustum", None, [0.2, 0.5, 0.7], 0),
    (7, "line_cyl", None, [0.3, 0.1, 0.8], 1),
    (7, "line", None, [0.3, 0.1, 0.8], 0),
]


@pytest.mark.parametrize("kind,name,_u,prm,flag", PRIMS, ids=[p_[1] for p_ in PRIMS])
def test_primitives_match_oracle(gc, orc, dev, kind, name, _u, prm, flag):
    nx, ny, nz = 37, 29, 23
    d = (0.05, 0.06, 0.07)
    center = (0.1, -0.05, 0.02)
    angles = (0.3, -0.2, 0.5) if kind != 7 else (0.2, 0.9, 0.4)
    out = torch.empty(nx * ny * nz, dtype=torch.float32, device=dev)
    c3, a3 = gc.Float3(*center), gc.Float3(*angles)
    L = gc.lib
    if kind == 0:
        st = L.gpucad_sphere_with_center(gc.ptr(out), c3, prm[0], prm[1], nx, ny, nz, *d, flag)
    elif kind == 1:
        st = L.gpucad_cuboid(gc.ptr(out), c3, a3, *prm, nx, ny, nz, *d)
    elif kind == 2:
        st = L.gpucad_cuboid_shell(gc.ptr(out), c3, a3, *prm, nx, ny, nz, *d)
    elif kind == 3:
        st = L.gpucad_pyramid_frustum(gc.ptr(out), c3, a3, *prm, nx, ny, nz, *d)
    elif kind == 4:
        st = L.gpucad_torus_with_center(gc.ptr(out), c3, a3, *prm, nx, ny, nz, *d)
    elif kind == 5:
        st = L.gpucad_cone_with_base_radius_height(gc.ptr(out), c3, a3, *prm, nx, ny, nz, *d)
    elif kind == 6:
        st = L.gpucad_cone_frustum(gc.ptr(out), c3, a3, *prm, nx, ny, nz, *d)
    else:
        st = L.gpucad_distance_from_line(gc.ptr(out), c3, a3, *prm, nx, ny, nz, *d, flag)
    gc.check(st)
    torch.cuda.synchronize()
    ref = orc.primitive(kind, nx, ny, nz, d, center, angles, prm, flag)
    assert ulp_diff(to_np(out), ref).max() <= 1


def test_primitive_field_dynamic(gc, dev):
    n = 20
    act = torch.linspace(-1, 1, n ** 3, device=dev)
    iso = torch.zeros(n ** 3, device=dev)
    gc.check(gc.lib.gpucad_primitive_field_dynamic(gc.ptr(act), gc.ptr(iso), n, n, n))
    torch.cuda.synchronize()
    expect = torch.where(act >= 0, torch.full_like(iso, np.finfo(np.float32).max), iso)
    assert torch.equal(iso, expect)


def test_file_write_obj_byte_identical(gc, orc, dev, tmp_path):
    n = 33
    f = torch.empty(n ** 3, dtype=torch.float32, device=dev)
    gc.check(gc.lib.gpucad_create_lattice(gc.ptr(f), n, n, n, n ** 3, 0))
    ind = torch.empty_like(f)
    nrm = torch.empty_like(f)
    gc.check(gc.lib.gpucad_normalise_four(gc.ptr(f), gc.ptr(ind), gc.ptr(nrm), n ** 3, n, n, n,
                                          0.2, 0.3))
    b, active, total = run_dropin(gc, "latticeone", ind, n, n, n, 0.25, vol1=nrm, lo1=0.2,
                                  hi1=0.3, voxel_size=(0.5, 0.5, 0.5))
    assert total > 0
    ours = str(tmp_path / "ours.obj")
    theirs = str(tmp_path / "oracle.obj")
    gc.check(gc.lib.gpucad_file_write_obj(gc.ptr(b.pos), total, ours.encode()))
    assert orc.file_write_obj(to_np(b.pos)[:total], total, theirs) == 0
    assert open(ours, "rb").read() == open(theirs, "rb").read()


def test_write_obj_indexed_roundtrip(gc, dev, tmp_path):
    p = CF.plain_gyroid(40)
    f = gc.lattice_field(p)
    m = gc.extract(f, 40, 40, 40)
    path = str(tmp_path / "mesh.obj")
    gc.check(gc.lib.gpucad_write_obj_indexed(gc.ptr(m.vertices), m.n_vertices, gc.ptr(m.triangles),
                                             m.n_triangles, path.encode(), 0))
    vs, fs = [], []
    for line in open(path):
        if line.startswith("v "):
            vs.append([float(t) for t in line.split()[1:]])
        elif line.startswith("f "):
            fs.append([int(t) - 1 for t in line.split()[1:]])
    assert np.array_equal(np.array(vs, np.float32), to_np(m.vertices))
    assert np.array_equal(np.array(fs), to_np(m.triangles))


def test_compat_classes_drive_the_library(gc, orc, dev, tmp_path):
    """tests/cpp/compat_harness replays the reference caller's unit-lattice
    sequence through the C++ classes of include/gpucad_compat.hpp; the same
    sequence through the C-ABI must give the same counts and identical .obj
    files."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "cpp", "compat_harness")
    assert os.path.exists(exe), "build() compiles the harness"
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [tuple(int(v) for v in line.split()) for line in r.stdout.strip().splitlines()]
    assert [t for t, _, _ in rows] == list(range(6))
    n = 61
    for ltype, active, total in rows:
        f = torch.empty(n ** 3, dtype=torch.float32, device=dev)
        gc.check(gc.lib.gpucad_create_lattice(gc.ptr(f), n, n, n, n ** 3, ltype))
        ind = torch.empty_like(f)
        gc.check(gc.lib.gpucad_normalise_four(gc.ptr(f), gc.ptr(ind), gc.ptr(f), n ** 3, n, n, n,
                                              0.2, 0.3))
        b, a2, t2 = run_dropin(gc, "latticeone", ind, n, n, n, 0.25, vol1=f, lo1=0.2, hi1=0.3)
        assert (a2, t2) == (active, total)
        path = str(tmp_path / ("capi_%d.obj" % ltype))
        gc.check(gc.lib.gpucad_file_write_obj(gc.ptr(b.pos), t2, path.encode()))
        assert open(path, "rb").read() == \
            open(str(tmp_path / ("unit_lattice_%d.obj" % ltype)), "rb").read()


def test_file_write_obj_known_answer(gc, dev, tmp_path):
    """The product writer on the hand-worked .obj input (the literal expected
    file is in tests/test_oracle_pins.py)."""
    from tests.test_oracle_pins import OBJ_EXPECTED, obj_input
    pos = torch.from_numpy(obj_input()).to(dev)
    path = str(tmp_path / "kat.obj")
    gc.check(gc.lib.gpucad_file_write_obj(gc.ptr(pos), pos.shape[0], path.encode()))
    assert open(path).read() == OBJ_EXPECTED


@pytest.mark.parametrize("second_band", [False, True])
@pytest.mark.parametrize("nz,corner_z", [(2, 0), (3, 1)])
def test_band_equal_values_branch(gc, dev, nz, corner_z, second_band):
    """Equal normalised values outside both band levels on an indicator 1/0
    (2/0) edge: t = 1 when p0.z == 0, else 0; hand-worked vertices and normals."""
    from tests.test_oracle_pins import (BAND_ONE_CELL, BAND_TWO_CELLS, band_case,
                                        expected_band)
    ind, k1, k2 = (torch.from_numpy(a).to(dev) for a in band_case(nz, corner_z, second_band))
    if second_band:
        b, active, total = run_dropin(gc, "lattice", ind, 2, 2, nz, 0.25, vol1=k1, vol2=k2,
                                      lo1=0.2, hi1=0.3, lo2=0.6, hi2=0.7)
    else:
        b, active, total = run_dropin(gc, "latticeone", ind, 2, 2, nz, 0.25, vol1=k1, lo1=0.2,
                                      hi1=0.3)
    pos, nrm = expected_band(BAND_ONE_CELL if nz == 2 else BAND_TWO_CELLS)
    assert total == len(pos)
    assert np.array_equal(to_np(b.pos)[:total, :3], pos)
    assert np.array_equal(to_np(b.norm)[:total, :3], nrm)
