"""ORACLE — test infrastructure only (checker and CPU baseline).

ctypes binding of liboracle.so (oracle.cpp, the host restatement of the
reference's hot-path arithmetic).  Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg import this module, and only to check or to time
the CPU path; the product path never calls it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
V4_FLAGS = ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")


def host_isa():
    """"x86-64-v4" when the host CPU has the AVX-512 extensions of that level
    (from /proc/cpuinfo), else "x86-64-v3"."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    return "x86-64-v4" if all(x in flags for x in V4_FLAGS) else "x86-64-v3"
    except OSError:
        pass
    return "x86-64-v3"


# both builds compute the same bits (oracle/Makefile); the one for the host's
# ISA level is loaded when it is built
ISA = host_isa()
LIB_PATH = os.path.join(HERE, "liboracle_v4.so" if ISA == "x86-64-v4" else "liboracle.so")
if not os.path.exists(LIB_PATH):
    ISA, LIB_PATH = "x86-64-v3", os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("liboracle.so not built; run make -C oracle")
    lib = C.CDLL(LIB_PATH)
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
    u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
    vp = C.c_void_p
    sig = {
        "orc_cos_rev": ([f32p, f32p, C.c_int64], None),
        "orc_sin_rev": ([f32p, f32p, C.c_int64], None),
        "orc_atan2": ([f32p, f32p, f32p, C.c_int64], None),
        "orc_tables": ([u8p, u8p], None),
        "orc_lattice_field_f32": ([vp, f32p, C.c_int, C.c_int, C.c_int], None),
        "orc_lattice_field_f64": ([vp, f64p, C.c_int, C.c_int, C.c_int], None),
        "orc_signs": ([f32p, C.c_int, C.c_int, C.c_int, C.c_float, u64p], None),
        "orc_extract_indexed": ([f32p, C.c_int, C.c_int, C.c_int, C.c_float, f32p, f32p, vp,
                                 C.c_uint64, vp, C.c_uint64, u64p, vp, C.c_int, C.c_int], C.c_int),
        "orc_classify": ([f32p, C.c_int, C.c_int, C.c_int, C.c_float, vp, vp, vp], None),
        "orc_exclusive_scan_u32": ([u32p, u32p, C.c_uint64], None),
        "orc_compact": ([u32p, u32p, C.c_uint64, u32p], None),
        "orc_generate": ([C.c_int, f32p, vp, vp, C.c_int, C.c_int, C.c_int, f32p, f32p, C.c_float,
                          C.c_float, C.c_float, C.c_float, C.c_float, u32p, u32p, C.c_uint32,
                          C.c_uint32, f32p, f32p], None),
        "orc_compute_isosurface": ([C.c_int, f32p, vp, vp, C.c_int, C.c_int, C.c_int, f32p, f32p,
                                    C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_uint32, u32p, u32p, u32p, u32p, u32p, f32p, f32p,
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)], C.c_int),
        "orc_normalise_four": ([f32p, f32p, f32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                C.c_float], None),
        "orc_create_lattice": ([f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int], None),
        "orc_primitive": ([C.c_int, f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                           C.c_float, f32p, f32p, f32p, C.c_int, C.c_int], None),
        "orc_file_write_obj": ([f32p, C.c_uint32, C.c_char_p], C.c_int),
        "orc_max_threads": ([], C.c_int),
        "orc_bool_classify": ([C.c_uint, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_float,
                               C.c_float, C.c_float, u32p, u32p], None),
        "orc_bool_copy": ([C.c_uint, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                           C.c_float], None),
        "orc_bool_generate": ([C.c_uint, vp, vp, vp, C.c_int, C.c_int, C.c_int, f32p, f32p,
                               C.c_float, C.c_float, C.c_float, u32p, u32p, C.c_uint32,
                               C.c_uint32, f32p, f32p], None),
        "orc_svl_period_angle": ([f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                  C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int],
                                 None),
        "orc_svl_rescale": ([f32p, f32p, C.c_uint64, C.c_float, C.c_float, C.c_int], None),
        "orc_svl_finding_phi": ([f32p, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_float,
                                 C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int],
                                None),
        "orc_svl_matvec": ([f32p, f32p, C.c_int, C.c_int, C.c_int], None),
        "orc_svl_cg": ([f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                        C.POINTER(C.c_int), C.POINTER(C.c_float)], None),
        "orc_svl_upsample": ([f32p, C.c_int, C.c_int, C.c_int, f32p, C.c_int, C.c_int, C.c_int,
                              C.c_float, C.c_float, C.c_float, C.c_int], None),
        "orc_svl_accumulate": ([f32p, f32p, C.c_uint64, C.c_int, f32p], None),
        "orc_svl_field": ([C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int,
                           C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                           C.c_float, C.c_int, vp, f32p, C.c_int, C.c_int, C.c_int, C.c_float,
                           C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, f32p,
                           np.ctypeslib.ndpointer(np.int32, flags="C"), f32p], None),
        "orc_unit_spectrum": ([f32p, C.c_int, C.c_int, C.c_int, C.c_int, f64p], None),
        "orc_fft_fill": ([f32p, f32p, C.c_int, C.c_int, C.c_int], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _load()


def threads(n=None):
    return n if n else lib.orc_max_threads()


def cos_rev(r):
    r = np.ascontiguousarray(r, np.float32)
    out = np.empty_like(r)
    lib.orc_cos_rev(r, out, r.size)
    return out


def sin_rev(r):
    r = np.ascontiguousarray(r, np.float32)
    out = np.empty_like(r)
    lib.orc_sin_rev(r, out, r.size)
    return out


def atan2(y, x):
    y = np.ascontiguousarray(y, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(y)
    lib.orc_atan2(y, x, out, y.size)
    return out


def tables():
    tri = np.empty(256 * 16, np.uint8)
    nv = np.empty(256, np.uint8)
    lib.orc_tables(tri, nv)
    return tri.reshape(256, 16), nv


def lattice_field_f32(p, z0=0, z1=None, nthreads=None):
    z1 = p.nz if z1 is None else z1
    out = np.empty((z1 - z0) * p.ny * p.nx, np.float32)
    lib.orc_lattice_field_f32(C.addressof(p), out, z0, z1, threads(nthreads))
    return out


def lattice_field_f64(p, z0=0, z1=None, nthreads=None):
    z1 = p.nz if z1 is None else z1
    out = np.empty((z1 - z0) * p.ny * p.nx, np.float64)
    lib.orc_lattice_field_f64(C.addressof(p), out, z0, z1, threads(nthreads))
    return out


def signs(field, nx, ny, nz, iso=0.0):
    W = (nx + 63) // 64
    out = np.empty(W * ny * nz, np.uint64)
    lib.orc_signs(np.ascontiguousarray(field, np.float32), nx, ny, nz, iso, out)
    return out


def extract_indexed(field, nx, ny, nz, iso=0.0, origin=(0, 0, 0), spacing=None, vcap=None,
                    tcap=None, edge_ids=False, nthreads=None, z_index0=0):
    """Canonical indexed mesh: (vertices (V,3) f32, triangles (T,3) u32, edge ids or None).
    z_index0: global z index of the field's first plane (vertex positions of a
    z-slab then equal those of the whole grid's mesh)."""
    field = np.ascontiguousarray(field, np.float32)
    if spacing is None:
        spacing = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    counts = np.zeros(2, np.uint64)
    o = np.asarray(origin, np.float32)
    s = np.asarray(spacing, np.float32)
    if vcap is None:
        lib.orc_extract_indexed(field, nx, ny, nz, iso, o, s, None, 0, None, 0, counts, None,
                                threads(nthreads), z_index0)
        vcap, tcap = int(counts[0]), int(counts[1])
    verts = np.zeros((max(vcap, 1), 3), np.float32)
    tris = np.zeros((max(tcap, 1), 3), np.uint32)
    eids = np.zeros(max(vcap, 1), np.uint64) if edge_ids else None
    lib.orc_extract_indexed(field, nx, ny, nz, iso, o, s, verts.ctypes.data, vcap,
                            tris.ctypes.data, tcap, counts,
                            None if eids is None else eids.ctypes.data, threads(nthreads),
                            z_index0)
    nv, nt = int(counts[0]), int(counts[1])
    return verts[:nv], tris[:nt], (eids[:nv] if edge_ids else None)


def mesh_counts(field, nx, ny, nz, iso=0.0, nthreads=None):
    """(vertices, triangles) of the canonical mesh of a field (no output)."""
    field = np.ascontiguousarray(field, np.float32)
    counts = np.zeros(2, np.uint64)
    o = np.zeros(3, np.float32)
    s = np.ones(3, np.float32)
    lib.orc_extract_indexed(field, nx, ny, nz, iso, o, s, None, 0, None, 0, counts, None,
                            threads(nthreads), 0)
    return int(counts[0]), int(counts[1])


def classify(vol, nx, ny, nz, iso):
    n = (nx - 1) * (ny - 1) * (nz - 1)
    cube = np.empty(n, np.uint8)
    verts = np.empty(n, np.uint32)
    occ = np.empty(n, np.uint32)
    lib.orc_classify(np.ascontiguousarray(vol, np.float32), nx, ny, nz, iso, cube.ctypes.data,
                     verts.ctypes.data, occ.ctypes.data)
    return cube, verts, occ


def exclusive_scan_u32(a):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty_like(a)
    lib.orc_exclusive_scan_u32(a, out, a.size)
    return out


MODE_BAND2, MODE_BAND1, MODE_STD = 0, 1, 2


def compute_isosurface(mode, vol, nx, ny, nz, iso, voxel_size=(1, 1, 1), center=(0, 0, 0),
                       vol1=None, vol2=None, lo1=0.0, hi1=0.0, lo2=0.0, hi2=0.0, max_verts=None):
    n = (nx - 1) * (ny - 1) * (nz - 1)
    if max_verts is None:
        max_verts = max(nx * ny * nz * 4, 300000)
    out = {k: np.zeros(n, np.uint32) for k in ("verts", "verts_scan", "occ", "occ_scan", "comp")}
    pos = np.zeros(4 * max_verts, np.float32)
    norm = np.zeros(4 * max_verts, np.float32)
    active, total = C.c_uint32(0), C.c_uint32(0)
    keep = [np.ascontiguousarray(v, np.float32) if v is not None else None for v in (vol1, vol2)]
    lib.orc_compute_isosurface(
        mode, np.ascontiguousarray(vol, np.float32),
        None if keep[0] is None else keep[0].ctypes.data,
        None if keep[1] is None else keep[1].ctypes.data, nx, ny, nz,
        np.asarray(voxel_size, np.float32), np.asarray(center, np.float32), iso, lo1, hi1, lo2,
        hi2, max_verts, out["verts"], out["verts_scan"], out["occ"], out["occ_scan"], out["comp"],
        pos, norm, C.byref(active), C.byref(total))
    out["active"] = active.value
    out["total"] = total.value
    out["pos"] = pos.reshape(-1, 4)
    out["norm"] = norm.reshape(-1, 4)
    return out


def normalise_four(vol, nx, ny, nz, iso1, iso2):
    vol = np.ascontiguousarray(vol, np.float32)
    ind = np.empty_like(vol)
    nrm = np.empty_like(vol)
    lib.orc_normalise_four(vol, ind, nrm, nx, ny, nz, iso1, iso2)
    return ind, nrm


def create_lattice(nx, ny, nz, type_, f64=False):
    out = np.empty(nx * ny * nz, np.float32)
    lib.orc_create_lattice(out, nx, ny, nz, type_, 1 if f64 else 0)
    return out


def primitive(kind, nx, ny, nz, d, center, angles, prm, flag=0, nthreads=None):
    out = np.empty(nx * ny * nz, np.float32)
    prm = np.asarray(list(prm) + [0.0] * (8 - len(prm)), np.float32)
    lib.orc_primitive(kind, out, nx, ny, nz, d[0], d[1], d[2], np.asarray(center, np.float32),
                      np.asarray(angles, np.float32), prm, flag, threads(nthreads))
    return out


def file_write_obj(pos4, total, filename):
    pos4 = np.ascontiguousarray(pos4, np.float32).reshape(-1)
    return lib.orc_file_write_obj(pos4, total, filename.encode())


# ---------------------------------------------------------------------------
# boolean meshing (Isosurface::computeIsosurface and its stages)
# ---------------------------------------------------------------------------
B_UNION, B_DIFF, B_INTERSECT, B_FIXED, B_DYNAMIC, B_REGION = 1, 2, 4, 8, 16, 32
GRID_POINT = np.dtype([("val", np.int32), ("t_x", np.float32), ("t_y", np.float32),
                       ("t_z", np.float32)])


def _ptr(a, dtype):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype)
    return a, a.ctypes.data


def bool_classify(flags, fixed, dyn, lat, nx, ny, nz, iso, iso1, iso2):
    n = (nx - 1) * (ny - 1) * (nz - 1)
    verts = np.empty(n, np.uint32)
    occ = np.empty(n, np.uint32)
    kf, pf = _ptr(fixed, GRID_POINT)
    kd, pd = _ptr(dyn, np.float32)
    kl, pl = _ptr(lat, np.float32)
    lib.orc_bool_classify(flags, pf, pd, pl, nx, ny, nz, iso, iso1, iso2, verts, occ)
    return verts, occ


def bool_copy(flags, vol_one, vol_two, vol_lat, nx, ny, nz, iso, iso1, iso2):
    """classify_copy_Voxel on a copy of vol_one (grid_points); returns the copy."""
    out = np.array(vol_one, dtype=GRID_POINT, copy=True)
    k2, p2 = _ptr(vol_two, np.float32)
    kl, pl = _ptr(vol_lat, np.float32)
    lib.orc_bool_copy(flags, out.ctypes.data, p2, pl, nx, ny, nz, iso, iso1, iso2)
    return out


def bool_compute_isosurface(flags, fixed, dyn, lat, nx, ny, nz, iso, iso1, iso2,
                            voxel_size=(1, 1, 1), center=(0, 0, 0), max_verts=None):
    """The whole computeIsosurface sequence (classify, scans, compact, generate)."""
    n = (nx - 1) * (ny - 1) * (nz - 1)
    if max_verts is None:
        max_verts = max(nx * ny * nz * 4, 300000)
    verts, occ = bool_classify(flags, fixed, dyn, lat, nx, ny, nz, iso, iso1, iso2)
    occ_scan = exclusive_scan_u32(occ)
    active = int(occ[-1]) + int(occ_scan[-1])
    out = {"verts": verts, "occ": occ, "occ_scan": occ_scan, "active": active}
    pos = np.zeros(4 * max_verts, np.float32)
    norm = np.zeros(4 * max_verts, np.float32)
    if active == 0:
        out.update(total=0, pos=pos.reshape(-1, 4), norm=norm.reshape(-1, 4))
        return out
    comp = np.zeros(n, np.uint32)
    lib.orc_compact(occ, occ_scan, n, comp)
    verts_scan = exclusive_scan_u32(verts)
    total = int(verts[-1]) + int(verts_scan[-1])
    kf, pf = _ptr(fixed, GRID_POINT)
    kd, pd = _ptr(dyn, np.float32)
    kl, pl = _ptr(lat, np.float32)
    lib.orc_bool_generate(flags, pf, pd, pl, nx, ny, nz, np.asarray(voxel_size, np.float32),
                          np.asarray(center, np.float32), iso, iso1, iso2, comp, verts_scan,
                          active, max_verts, pos, norm)
    out.update(comp=comp, verts_scan=verts_scan, total=total, nfacets=total // 3,
               pos=pos.reshape(-1, 4), norm=norm.reshape(-1, 4))
    return out


# ---------------------------------------------------------------------------
# spatially variant lattice synthesis (src/lattice_files/Gratings.cu,
# src/main.cu:3578-3711, 3904-4037)
# ---------------------------------------------------------------------------
def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def svl_period_angle(nx, ny, nz, dx, dy, dz, mean, axis="z", angle=False):
    out = np.empty(nx * ny * nz, np.float32)
    lib.orc_svl_period_angle(out, nx, ny, nz, dx, dy, dz, mean[0], mean[1], mean[2], ord(axis),
                             1 if angle else 0)
    return out


def svl_rescale(a, a1=0.0, b1=0.0, affine=True):
    a = _f32(a).ravel()
    out = np.empty_like(a)
    lib.orc_svl_rescale(a, out, a.size, a1, b1, 1 if affine else 0)
    return out


def svl_finding_phi(nx, ny, nz, i, j, k, lattice_type="n", uniform_type=0, const_period=30.0,
                    periods=(30.0, 30.0, 30.0), period=None, d=(1.0, 1.0, 1.0), lcon=0.16,
                    lcon_1=0.6, sinewave_zaxis=False):
    out = np.empty(nx * ny * nz, np.float32)
    keep = None if period is None else _f32(period)
    per = None if keep is None else keep.ctypes.data
    lib.orc_svl_finding_phi(out, per, nx, ny, nz, i, j, k, d[0], d[1], d[2], ord(lattice_type),
                            uniform_type, const_period, periods[0], periods[1], periods[2], lcon,
                            lcon_1, 1 if sinewave_zaxis else 0)
    return out


def svl_matvec(d, nx, ny, nz):
    d = _f32(d).ravel()
    q = np.empty_like(d)
    lib.orc_svl_matvec(d, q, nx, ny, nz)
    return q


def svl_cg(b, nx, ny, nz, iters=500, end_res=0.01):
    x = _f32(b).ravel().copy()
    it, res = C.c_int(0), C.c_float(0)
    lib.orc_svl_cg(x, nx, ny, nz, iters, end_res, C.byref(it), C.byref(res))
    return x, it.value, res.value


def svl_upsample(c, nx, ny, nz, fine, h, mode=0):
    c = _f32(c).ravel()
    fx, fy, fz = fine
    out = np.empty(fx * fy * fz * (2 if mode else 1), np.float32)
    lib.orc_svl_upsample(c, nx, ny, nz, out, fx, fy, fz, h[0], h[1], h[2], mode)
    return out


def svl_accumulate(svl, grating, index, coeffs):
    svl = _f32(svl).ravel().copy()
    g = _f32(grating).ravel()
    c = _f32(coeffs).ravel()
    lib.orc_svl_accumulate(svl, g, svl.size, index, c)
    return svl


def svl_field(p, coeffs, period=None):
    """The synthesis of a gc_svl_params-like object p (attributes as in
    gpucadforam_amd.SvlParams) with coefficients as (N, 2) float32."""
    c = _f32(coeffs).ravel()
    out = np.empty(p.fx * p.fy * p.fz, np.float32)
    it = np.zeros(3, np.int32)
    res = np.zeros(3, np.float32)
    keep = None if period is None else _f32(period)
    lt = p.lattice_type if isinstance(p.lattice_type, int) else ord(p.lattice_type)
    lib.orc_svl_field(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, lt, p.uniform_type, p.const_period,
                      p.x_period, p.y_period, p.z_period, p.lcon, p.lcon_1, p.sinewave_zaxis,
                      None if keep is None else keep.ctypes.data, c, p.range, p.harmonics,
                      p.cg_max_iter, p.cg_end_res, p.fx, p.fy, p.fz, p.fine_dx, p.fine_dy,
                      p.fine_dz, out, it, res)
    return out, it, res


def unit_spectrum(cell, nx, ny, nz, range_=2):
    side = 2 * range_ + 1
    out = np.empty(2 * side ** 3, np.float64)
    lib.orc_unit_spectrum(_f32(cell).ravel(), nx, ny, nz, range_, out)
    return out[0::2] + 1j * out[1::2]


def fft_fill(half, nx, ny, nz):
    h = _f32(half).ravel()
    full = np.zeros(2 * nx * ny * nz, np.float32)
    lib.orc_fft_fill(h, full, nx, ny, nz)
    return full


# ---- Fft_lattice's transforms (src/lattice_files/Fft_lattice.cu:107-126) ---
# cuFFT 12.3 (the reference's dependency, Makefile:51 -lcufft) semantics,
# restated as DFT matrices applied axis by axis in float64: cufftPlan3d(n0,
# n1, n2) with n2 the fastest axis; unnormalised; forward e^{-2 pi i jk/n},
# inverse e^{+2 pi i jk/n}.  In-place real data pads each row of n2 reals to
# 2 (n2/2 + 1) floats; the half spectrum holds n2/2 + 1 complex values per
# row.  Complex-to-real: inverse along axes 0 and 1, then along the rows the
# Hermitian sum Re X0 + 2 sum_k Re(Xk e^{+...}) (+ the Nyquist term for even
# n2), i.e. the imaginary parts of X0 (and X[n2/2]) are not used.

def _dft_matrix(n, sign):
    j = np.arange(n)
    m = np.outer(j, j) % n  # exact reduction of jk before the angle
    return np.exp(sign * 2j * np.pi * m / n)


def _along(a, axis, mat):
    return np.moveaxis(np.tensordot(mat, a, axes=([1], [axis])), 0, axis)


def fft_r2c_padded(buf, n0, n1, n2):
    """Forward transform of the in-place real layout -> (n0, n1, n2//2+1) complex128."""
    h = n2 // 2 + 1
    x = np.asarray(buf, np.float32).reshape(n0, n1, 2 * h)[:, :, :n2].astype(np.float64)
    X = _along(x, 2, _dft_matrix(n2, -1)[:h])
    X = _along(X, 1, _dft_matrix(n1, -1))
    return _along(X, 0, _dft_matrix(n0, -1))


def fft_c2r(X, n0, n1, n2):
    """Unnormalised inverse of a (n0, n1, n2//2+1) half spectrum -> (n0, n1, n2) float64."""
    h = n2 // 2 + 1
    Y = _along(np.asarray(X, np.complex128).reshape(n0, n1, h), 0, _dft_matrix(n0, 1))
    Y = _along(Y, 1, _dft_matrix(n1, 1))
    j = np.arange(n2)
    kh = (n2 - 1) // 2
    k = np.arange(1, kh + 1)
    ang = 2 * np.pi * (np.outer(j, k) % n2) / n2
    x = Y[..., :1].real + 2 * (Y[..., 1:kh + 1].real @ np.cos(ang).T -
                               Y[..., 1:kh + 1].imag @ np.sin(ang).T)
    if n2 % 2 == 0:
        x = x + Y[..., n2 // 2:n2 // 2 + 1].real * np.where(j % 2 == 0, 1.0, -1.0)
    return x


def fft_c2c_inverse(X, n0, n1, n2):
    """Unnormalised inverse complex transform of (n0, n1, n2) values."""
    Y = np.asarray(X, np.complex128).reshape(n0, n1, n2)
    for axis, n in ((2, n2), (1, n1), (0, n0)):
        Y = _along(Y, axis, _dft_matrix(n, 1))
    return Y
