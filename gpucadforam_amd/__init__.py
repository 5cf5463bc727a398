"""MI355X implicit-geometry engine: Python binding of libgpucad.so.

The product is the C-ABI library built from csrc/ (HIP kernels for gfx950);
this module only loads it with ctypes and offers thin helpers that pass torch
device tensors (device memory and streams are plumbing here).  There is no
fallback: when libgpucad.so is missing, importing this module raises.

torch is imported before the library is loaded so that the library binds to
the HIP runtime torch already loaded (both carry the soname libamdhip64.so.7);
one runtime per process keeps device pointers interchangeable.
"""
import ctypes as C
import os

import torch  # noqa: F401  (must precede loading libgpucad.so)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpucad.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libgpucad.so not found at %s; build it with `python -c 'import __graft_entry__ as g; "
        "g.build()'` (hipcc --offload-arch=gfx950)" % LIB_PATH)

lib = C.CDLL(LIB_PATH)

# ---------------------------------------------------------------------------
# status
# ---------------------------------------------------------------------------
OK, E_ARG, E_OVERFLOW, E_OOM, E_HIP, E_BUSY, E_NOTFOUND, E_IO = range(8)
TPMS_GYROID, TPMS_SCHWARZ_P, TPMS_DIAMOND = 0, 1, 2
PRIM_NONE, PRIM_SPHERE, PRIM_BOX = 0, 1, 2
OP_NONE, OP_INTERSECT, OP_UNION, OP_SUBTRACT = 0, 1, 2, 3
BOOL_UNION, BOOL_DIFF, BOOL_INTERSECT, BOOL_FIXED, BOOL_DYNAMIC, BOOL_MAKE_REGION = 1, 2, 4, 8, 16, 32
MAX_MODES = 8


class GpucadError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("gpucad status %d (%s): %s" % (status, status_string(status), msg))
        self.status = status


class Uint3(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32)]


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Float4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Float2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class SvlParams(C.Structure):
    """gc_svl_params: one spatially variant lattice synthesis (include/gpucad.h)."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
        ("lattice_type", C.c_char), ("pad", C.c_char * 3),
        ("uniform_type", C.c_int32),
        ("const_period", C.c_float), ("x_period", C.c_float), ("y_period", C.c_float),
        ("z_period", C.c_float),
        ("lcon", C.c_float), ("lcon_1", C.c_float),
        ("sinewave_zaxis", C.c_int32),
        ("range", C.c_int32), ("harmonics", C.c_int32),
        ("cg_max_iter", C.c_int32), ("cg_end_res", C.c_float),
        ("fx", C.c_int32), ("fy", C.c_int32), ("fz", C.c_int32),
        ("fine_dx", C.c_float), ("fine_dy", C.c_float), ("fine_dz", C.c_float),
    ]


class SvlResult(C.Structure):
    _fields_ = [("cg_iters", C.c_int32 * 3), ("cg_res", C.c_float * 3)]


class LatticeParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("tpms", C.c_int32),
        ("period0", C.c_float), ("period1", C.c_float),
        ("rot0", C.c_float), ("rot1", C.c_float),
        ("bend_radius", C.c_float),
        ("sheet", C.c_int32),
        ("t0", C.c_float),
        ("n_modes", C.c_int32),
        ("mode_k", (C.c_int32 * 3) * MAX_MODES),
        ("mode_a", C.c_float * MAX_MODES),
        ("mode_phi", C.c_float * MAX_MODES),
        ("prim", C.c_int32),
        ("op", C.c_int32),
        ("prim_center", C.c_float * 3),
        ("prim_radius", C.c_float),
        ("prim_half", C.c_float * 3),
    ]

    def npoints(self):
        return self.nx * self.ny * self.nz


class ExtractArgs(C.Structure):
    _fields_ = [
        ("field", C.c_void_p),
        ("signs", C.c_void_p),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("iso", C.c_float),
        ("origin", Float3),
        ("spacing", Float3),
        ("vertices", C.c_void_p),
        ("vert_capacity", C.c_uint64),
        ("triangles", C.c_void_p),
        ("tri_capacity", C.c_uint64),
        ("normals", C.c_void_p),
        ("workspace", C.c_void_p),
        ("workspace_bytes", C.c_uint64),
        ("d_counts", C.c_void_p),
        ("emit_event", C.c_void_p),
    ]


class MeshCounts(C.Structure):
    _fields_ = [("n_vertices", C.c_uint64), ("n_triangles", C.c_uint64)]


class BatchOptions(C.Structure):
    _fields_ = [
        ("iso", C.c_float),
        ("brick_x", C.c_int32),
        ("brick_lz", C.c_int32),
        ("waves_per_cu", C.c_int32),
        ("use_graph", C.c_int32),
        ("unit_domain", C.c_int32),
        ("vert_capacity", C.c_uint64),
        ("tri_capacity", C.c_uint64),
    ]


class BatchGridResult(C.Structure):
    _fields_ = [
        ("field_offset", C.c_uint64),
        ("vert_offset", C.c_uint64),
        ("n_vertices", C.c_uint64),
        ("tri_offset", C.c_uint64),
        ("n_triangles", C.c_uint64),
        ("status", C.c_int32),
        ("pad", C.c_int32),
    ]


class BatchBuffers(C.Structure):
    _fields_ = [
        ("fields", C.c_void_p),
        ("field_floats", C.c_uint64),
        ("vertices", C.c_void_p),
        ("vert_capacity", C.c_uint64),
        ("triangles", C.c_void_p),
        ("tri_capacity", C.c_uint64),
    ]


class ServerOptions(C.Structure):
    _fields_ = [
        ("budget_bytes", C.c_uint64),
        ("max_batch", C.c_int32),
        ("slots", C.c_int32),
        ("iso", C.c_float),
        ("brick_x", C.c_int32),
        ("brick_lz", C.c_int32),
        ("waves_per_cu", C.c_int32),
    ]


class RequestResult(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64),
        ("status", C.c_int32),
        ("pad", C.c_int32),
        ("field", C.c_void_p),
        ("vertices", C.c_void_p),
        ("triangles", C.c_void_p),
        ("n_vertices", C.c_uint64),
        ("n_triangles", C.c_uint64),
        ("submit_s", C.c_double),
        ("done_s", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
        ("groups", C.c_uint64), ("bytes_in_use", C.c_uint64),
        ("peak_bytes_in_use", C.c_uint64), ("budget_bytes", C.c_uint64),
        ("mean_group_size", C.c_double),
    ]


_vp, _u32, _i32, _u64, _f = C.c_void_p, C.c_uint32, C.c_int32, C.c_uint64, C.c_float

# signatures of every exported entry point (argtypes, restype)
SIGNATURES = {
    "gpucad_version": ([], C.c_int),
    "gpucad_abi_sizes": ([C.POINTER(C.c_uint32), C.c_int32], C.c_int),
    "gpucad_status_string": ([C.c_int], C.c_char_p),
    "gpucad_last_error": ([], C.c_char_p),
    "gpucad_init": ([C.c_int], C.c_int),
    "gpucad_shutdown": ([], C.c_int),
    "gpucad_mc_tables": ([_vp, _vp, _vp], C.c_int),
    "gpucad_sign_words": ([_i32, _i32, _i32], _u64),
    "gpucad_lattice_field": ([C.POINTER(LatticeParams), _vp, _vp, _f, _vp], C.c_int),
    "gpucad_lattice_field_lz": ([C.POINTER(LatticeParams), _vp, _vp, _f, _i32, _vp], C.c_int),
    "gpucad_extract_workspace_bytes": ([_i32, _i32, _i32], _u64),
    "gpucad_extract": ([C.POINTER(ExtractArgs), C.POINTER(MeshCounts), _vp], C.c_int),
    "gpucad_batch_create": ([C.POINTER(LatticeParams), _i32, C.POINTER(BatchOptions),
                             C.POINTER(_vp)], C.c_int),
    "gpucad_batch_run": ([_vp, _vp], C.c_int),
    "gpucad_batch_results": ([_vp, C.POINTER(BatchGridResult), _i32], C.c_int),
    "gpucad_batch_buffers": ([_vp, C.POINTER(BatchBuffers)], C.c_int),
    "gpucad_batch_tuning": ([_vp, C.POINTER(BatchOptions)], C.c_int),
    "gpucad_batch_destroy": ([_vp], C.c_int),
    "gpucad_server_create": ([C.POINTER(ServerOptions), C.POINTER(_vp)], C.c_int),
    "gpucad_server_submit": ([_vp, C.POINTER(LatticeParams), C.POINTER(_u64)], C.c_int),
    "gpucad_server_poll": ([_vp, C.POINTER(RequestResult), _i32, C.POINTER(_i32)], C.c_int),
    "gpucad_server_release": ([_vp, _u64], C.c_int),
    "gpucad_server_stats": ([_vp, C.POINTER(ServerStats)], C.c_int),
    "gpucad_server_destroy": ([_vp], C.c_int),
    "gpucad_server_run_trace": ([_vp, C.POINTER(LatticeParams), C.POINTER(C.c_double), _i32,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "gpucad_plain_run_trace": ([C.POINTER(LatticeParams), C.POINTER(C.c_double), _i32, _f,
                                C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "gpucad_create_lattice": ([_vp, _u32, _u32, _u32, _u32, _u32], C.c_int),
    "gpucad_sphere_with_center": ([_vp, Float3, _f, _f, _i32, _i32, _i32, _f, _f, _f, _i32], C.c_int),
    "gpucad_cuboid": ([_vp, Float3, Float3, _f, _f, _f, _i32, _i32, _i32, _f, _f, _f], C.c_int),
    "gpucad_cuboid_shell": ([_vp, Float3, Float3, _f, _f, _f, _f, _i32, _i32, _i32, _f, _f, _f],
                            C.c_int),
    "gpucad_pyramid_frustum": ([_vp, Float3, Float3, _f, _f, _f, _f, _f, _i32, _i32, _i32, _f, _f,
                                _f], C.c_int),
    "gpucad_torus_with_center": ([_vp, Float3, Float3, _f, _f, _i32, _i32, _i32, _f, _f, _f],
                                 C.c_int),
    "gpucad_cone_with_base_radius_height": ([_vp, Float3, Float3, _f, _f, _i32, _i32, _i32, _f, _f,
                                             _f], C.c_int),
    "gpucad_cone_frustum": ([_vp, Float3, Float3, _f, _f, _f, _i32, _i32, _i32, _f, _f, _f],
                            C.c_int),
    "gpucad_distance_from_line": ([_vp, Float3, Float3, _f, _f, _f, _i32, _i32, _i32, _f, _f, _f,
                                   _i32], C.c_int),
    "gpucad_normalise_four": ([_vp, _vp, _vp, _u64, _i32, _i32, _i32, _f, _f], C.c_int),
    "gpucad_primitive_field_dynamic": ([_vp, _vp, _i32, _i32, _i32], C.c_int),
    "gpucad_primitive_field_fixed": ([_vp, _vp, _i32, _i32, _i32], C.c_int),
    "gpucad_classify_voxel_lattice_new": ([_vp, _vp, _vp, Uint3, Uint3, Uint3, _u32, Float3, _f],
                                          C.c_int),
    "gpucad_exclusive_scan_u32": ([_vp, _vp, _u32], C.c_int),
    "gpucad_compact_voxels": ([_vp, _vp, _vp, _u32], C.c_int),
    "gpucad_generate_triangles_lattice_new": (
        [_vp, _vp, _vp, _vp, _vp, Uint3, Uint3, Uint3, Float3, Float3, _f, _u32, _u32, _u32, _vp,
         _vp, _f, _f, _f, _f], C.c_int),
    "gpucad_generate_triangles_lattice_newone": (
        [_vp, _vp, _vp, _vp, _vp, Uint3, Uint3, Uint3, Float3, Float3, _f, _u32, _u32, _u32, _vp,
         _f, _f], C.c_int),
    "gpucad_generate_triangles_standard": (
        [_vp, _vp, _vp, _vp, _vp, Uint3, Uint3, Uint3, Float3, Float3, _f, _u32, _u32], C.c_int),
    "gpucad_compute_isosurface_lattice": (
        [_vp, _vp, _vp, _f, _u32, _vp, _vp, _vp, _vp, Uint3, Uint3, Uint3, Float3, Float3,
         C.POINTER(_u32), C.POINTER(_u32), _vp, _u32, _vp, _vp, _f, _f, _f, _f], C.c_int),
    "gpucad_compute_isosurface_lattice_noalloc": (
        [_vp, _vp, _vp, _f, _u32, _vp, _vp, _vp, _vp, Uint3, Uint3, Uint3, Float3, Float3,
         C.POINTER(_u32), C.POINTER(_u32), _vp, _u32, _vp, _vp, _f, _f, _f, _f], C.c_int),
    "gpucad_compute_isosurface_latticeone": (
        [_vp, _vp, _vp, _f, _u32, _vp, _vp, _vp, _vp, Uint3, Uint3, Uint3, Float3, Float3,
         C.POINTER(_u32), C.POINTER(_u32), _vp, _u32, _vp, _f, _f], C.c_int),
    "gpucad_compute_isosurface_standard": (
        [_vp, _vp, _vp, _f, _u32, _vp, _vp, _vp, _vp, Uint3, Uint3, Uint3, Float3, Float3,
         C.POINTER(_u32), C.POINTER(_u32), _vp, _u32], C.c_int),
    "gpucad_classify_voxel_boolean": ([_vp, _vp, _vp, _vp, _vp, Uint3, Uint3, Uint3, _u32, _f, _f,
                                       _f, _u32], C.c_int),
    "gpucad_classify_copy_voxel": ([_vp, _vp, _vp, Uint3, _f, _f, _f, _u32], C.c_int),
    "gpucad_generate_triangles_boolean": ([_vp, _vp, _vp, _vp, Uint3, Uint3, Uint3, Float3, Float3,
                                           _f, _u32, _u32, _vp, _vp, _vp, _f, _f, _u32], C.c_int),
    "gpucad_compute_isosurface": ([_vp, _vp, _f, _u32, _vp, _vp, _vp, _vp, Uint3, Uint3, Uint3,
                                   Float3, Float3, C.POINTER(_u32), C.POINTER(_u32), _vp, _u32,
                                   _vp, _vp, _vp, _f, _f, _u32, C.POINTER(_u64)], C.c_int),
    "gpucad_svl_period_data": ([_vp, _i32, _i32, _i32, _f, _f, _f, _f, _f, _f, C.c_char],
                               C.c_int),
    "gpucad_svl_angle_data": ([_vp, _i32, _i32, _i32, _f, _f, _f, _f, _f, _f, C.c_char], C.c_int),
    "gpucad_normalise_three": ([_vp, _vp, _u64, _f, _f], C.c_int),
    "gpucad_normalise_buffer": ([_vp, _vp, _u64], C.c_int),
    "gpucad_svl_finding_phi": ([_vp, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _f, _f, _f,
                                C.c_char, _i32, _f, _f, _f, _f, _f, _f, _i32], C.c_int),
    "gpucad_svl_matvec": ([_vp, _vp, _i32, _i32, _i32], C.c_int),
    "gpucad_svl_cg": ([_vp, _i32, _i32, _i32, _i32, _f, C.POINTER(_i32), C.POINTER(_f)],
                      C.c_int),
    "gpucad_svl_axpby": ([_vp, _f, _vp, _f, _u64], C.c_int),
    "gpucad_svl_dot": ([_vp, _vp, _vp, _u64], C.c_int),
    "gpucad_svl_fillfrac": ([_vp, _vp, _i32, _i32, _i32], C.c_int),
    "gpucad_svl_refine": ([_vp, _i32, _i32, _i32, _vp, _i32, _i32, _i32, _f, _f, _f], C.c_int),
    "gpucad_svl_grating": ([_vp, _i32, _i32, _i32, _vp, _i32, _i32, _i32, _f, _f, _f], C.c_int),
    "gpucad_svl_accumulate": ([_vp, _vp, _u64, _i32, _vp], C.c_int),
    "gpucad_fft_fill": ([_vp, _vp, _i32, _i32, _i32], C.c_int),
    "gpucad_fft_scale": ([_vp, _f, _u64], C.c_int),
    "gpucad_unit_spectrum": ([_vp, _i32, _i32, _i32, _i32, _vp], C.c_int),
    "gpucad_fft_plan3d": ([_i32, _i32, _i32], C.c_int),
    "gpucad_fft_destroy": ([], C.c_int),
    "gpucad_fft_r2c": ([_vp], C.c_int),
    "gpucad_fft_c2r": ([_vp], C.c_int),
    "gpucad_fft_c2c_inverse": ([_vp], C.c_int),
    "gpucad_add_fft_constant": ([_vp, Float2, _i32], C.c_int),
    "gpucad_grating_sum": ([_vp, _vp, _i32, _i32, _i32], C.c_int),
    "gpucad_svl_field": ([C.POINTER(SvlParams), _vp, _vp, _vp, C.POINTER(SvlResult), _vp],
                         C.c_int),
    "gpucad_import_external_buffer": ([_i32, _u64, _u64, C.POINTER(_vp), C.POINTER(_vp)],
                                      C.c_int),
    "gpucad_release_external_buffer": ([_vp], C.c_int),
    "gpucad_import_external_semaphore": ([_i32, C.POINTER(_vp)], C.c_int),
    "gpucad_signal_external_semaphore": ([_vp, _vp], C.c_int),
    "gpucad_wait_external_semaphore": ([_vp, _vp], C.c_int),
    "gpucad_release_external_semaphore": ([_vp], C.c_int),
    "gpucad_file_write_obj": ([_vp, _u32, C.c_char_p], C.c_int),
    "gpucad_write_obj_indexed": ([_vp, _u64, _vp, _u64, C.c_char_p, _i32], C.c_int),
}

for _name, (_args, _res) in SIGNATURES.items():
    _fn = getattr(lib, _name)
    _fn.argtypes = _args
    _fn.restype = _res


def status_string(status):
    return lib.gpucad_status_string(status).decode()


def last_error():
    return lib.gpucad_last_error().decode()


def check(status):
    if status != OK:
        raise GpucadError(status, last_error())
    return status


# ---------------------------------------------------------------------------
# helpers over torch tensors
# ---------------------------------------------------------------------------
def ptr(t):
    """Device (or host) address of a tensor, or None."""
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def init(device=0):
    check(lib.gpucad_init(device))


def sign_words(nx, ny, nz):
    return int(lib.gpucad_sign_words(nx, ny, nz))


def lattice_field(params, out=None, signs=None, iso=0.0, stream=None, lz=0):
    """Evaluates the lattice field into `out` (float32 tensor, nz*ny*nx) and,
    if given, the sign words into `signs` (int64 tensor of sign_words())."""
    n = params.npoints()
    if out is None:
        out = torch.empty(n, dtype=torch.float32, device="cuda")
    assert out.numel() >= n and out.dtype == torch.float32 and out.is_cuda
    if signs is not None:
        assert signs.dtype == torch.int64 and signs.numel() >= sign_words(params.nx, params.ny,
                                                                            params.nz)
    check(lib.gpucad_lattice_field_lz(C.byref(params), ptr(out), ptr(signs), iso, lz,
                                      stream_handle(stream)))
    return out


def extract_workspace_bytes(nx, ny, nz):
    return int(lib.gpucad_extract_workspace_bytes(nx, ny, nz))


class Mesh:
    """Indexed mesh in device memory: vertices (V, 3) float32, triangles (T, 3) int32."""

    def __init__(self, vertices, triangles, nv, nt):
        self.vertices = vertices
        self.triangles = triangles
        self.n_vertices = nv
        self.n_triangles = nt


def extract(field, nx, ny, nz, iso=0.0, origin=(0.0, 0.0, 0.0), spacing=None, signs=None,
            vert_capacity=None, tri_capacity=None, workspace=None, stream=None):
    """Indexed marching cubes of a device field; returns a Mesh (synchronous)."""
    if spacing is None:
        spacing = (1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    if vert_capacity is None:
        vert_capacity = max(1024, nx * ny * nz // 4)
    if tri_capacity is None:
        tri_capacity = 2 * vert_capacity
    wsb = extract_workspace_bytes(nx, ny, nz)
    if workspace is None or workspace.numel() < wsb:
        workspace = torch.empty(wsb, dtype=torch.uint8, device=field.device)
    for _ in range(2):
        verts = torch.empty((vert_capacity, 3), dtype=torch.float32, device=field.device)
        tris = torch.empty((tri_capacity, 3), dtype=torch.int32, device=field.device)
        a = ExtractArgs()
        a.field = field.data_ptr()
        a.signs = None if signs is None else signs.data_ptr()
        a.nx, a.ny, a.nz = nx, ny, nz
        a.iso = iso
        a.origin = Float3(*origin)
        a.spacing = Float3(*spacing)
        a.vertices = verts.data_ptr()
        a.vert_capacity = vert_capacity
        a.triangles = tris.data_ptr()
        a.tri_capacity = tri_capacity
        a.normals = None
        a.workspace = workspace.data_ptr()
        a.workspace_bytes = workspace.numel()
        a.d_counts = None
        counts = MeshCounts()
        st = lib.gpucad_extract(C.byref(a), C.byref(counts), stream_handle(stream))
        if st == E_OVERFLOW and counts.n_vertices < 2**32:
            vert_capacity = max(vert_capacity, int(counts.n_vertices))
            tri_capacity = max(tri_capacity, int(counts.n_triangles))
            continue
        check(st)
        nv, nt = int(counts.n_vertices), int(counts.n_triangles)
        return Mesh(verts[:nv], tris[:nt], nv, nt)
    raise GpucadError(E_OVERFLOW, "capacity retry failed")


def u3(x, y, z):
    return Uint3(x, y, z)


def f3(x, y, z):
    return Float3(x, y, z)


# ---------------------------------------------------------------------------
# batched form
# ---------------------------------------------------------------------------
def _device_view(addr, count, dtype, device):
    """A torch tensor over `count` elements of library-owned device memory
    (valid while the owner lives)."""
    if count == 0:
        return torch.empty(0, dtype=dtype, device=device)

    class _Holder:  # wraps the memory through the CUDA array interface
        __cuda_array_interface__ = {
            "shape": (count,),
            "typestr": {torch.float32: "<f4", torch.int32: "<i4", torch.int64: "<i8"}[dtype],
            "data": (addr, False),
            "version": 2,
        }
    return torch.as_tensor(_Holder(), device=device)


class Batch:
    """Many lattice grids evaluated and meshed per run (gpucad_batch_*).

    grids: list of LatticeParams.  Options left at 0 are autotuned.
    """

    def __init__(self, grids, iso=0.0, brick_x=0, brick_lz=0, waves_per_cu=0, use_graph=True,
                 unit_domain=True, vert_capacity=0, tri_capacity=0):
        arr = (LatticeParams * len(grids))(*grids)
        o = BatchOptions()
        o.iso, o.brick_x, o.brick_lz, o.waves_per_cu = iso, brick_x, brick_lz, waves_per_cu
        o.use_graph, o.unit_domain = int(bool(use_graph)), int(bool(unit_domain))
        o.vert_capacity, o.tri_capacity = vert_capacity, tri_capacity
        h = C.c_void_p()
        check(lib.gpucad_batch_create(arr, len(grids), C.byref(o), C.byref(h)))
        self._h = h
        self.n = len(grids)
        self.grids = list(grids)
        b = BatchBuffers()
        check(lib.gpucad_batch_buffers(self._h, C.byref(b)))
        self.buffers = b
        dev = torch.device("cuda", torch.cuda.current_device())
        self.fields = _device_view(b.fields, b.field_floats, torch.float32, dev)
        self.vertices = _device_view(b.vertices, 3 * b.vert_capacity, torch.float32,
                                     dev).view(-1, 3)
        self.triangles = _device_view(b.triangles, 3 * b.tri_capacity, torch.int32,
                                      dev).view(-1, 3)

    def tuning(self):
        o = BatchOptions()
        check(lib.gpucad_batch_tuning(self._h, C.byref(o)))
        return dict(brick_x=o.brick_x, brick_lz=o.brick_lz, waves_per_cu=o.waves_per_cu,
                    vert_capacity=o.vert_capacity, tri_capacity=o.tri_capacity)

    def run(self, stream=None):
        check(lib.gpucad_batch_run(self._h, stream_handle(stream)))

    def results(self, allow_overflow=False):
        res = (BatchGridResult * self.n)()
        st = lib.gpucad_batch_results(self._h, res, self.n)
        if st != OK and not (allow_overflow and st == E_OVERFLOW):
            check(st)
        return [dict(field_offset=r.field_offset, vert_offset=r.vert_offset,
                     n_vertices=r.n_vertices, tri_offset=r.tri_offset,
                     n_triangles=r.n_triangles, status=r.status) for r in res]

    def field(self, i, res=None):
        p = self.grids[i]
        off = (res or self.results())[i]["field_offset"]
        return self.fields[off:off + p.npoints()]

    def mesh(self, i, res=None):
        r = (res or self.results())[i]
        v = self.vertices[r["vert_offset"]:r["vert_offset"] + r["n_vertices"]]
        t = self.triangles[r["tri_offset"]:r["tri_offset"] + r["n_triangles"]]
        return Mesh(v, t, r["n_vertices"], r["n_triangles"])

    def close(self):
        if self._h:
            check(lib.gpucad_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# served form
# ---------------------------------------------------------------------------
class Server:
    """Request loop with continuous batching under a device-memory budget
    (gpucad_server_*)."""

    def __init__(self, budget_bytes, max_batch=0, slots=0, iso=0.0, brick_x=0, brick_lz=0,
                 waves_per_cu=0):
        o = ServerOptions()
        o.budget_bytes, o.max_batch, o.slots, o.iso = budget_bytes, max_batch, slots, iso
        o.brick_x, o.brick_lz, o.waves_per_cu = brick_x, brick_lz, waves_per_cu
        h = C.c_void_p()
        check(lib.gpucad_server_create(C.byref(o), C.byref(h)))
        self._h = h
        self.device = torch.device("cuda", torch.cuda.current_device())

    def submit(self, params):
        t = C.c_uint64()
        check(lib.gpucad_server_submit(self._h, C.byref(params), C.byref(t)))
        return t.value

    def poll(self, max_results=64):
        out = (RequestResult * max_results)()
        n = C.c_int32()
        check(lib.gpucad_server_poll(self._h, out, max_results, C.byref(n)))
        return [out[i] for i in range(n.value)]

    def release(self, ticket):
        check(lib.gpucad_server_release(self._h, ticket))

    def stats(self):
        s = ServerStats()
        check(lib.gpucad_server_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def result_arrays(self, r, params):
        """Host copies (field, vertices, triangles) of a finished request."""
        f = _device_view(r.field, params.npoints(), torch.float32, self.device).cpu().numpy()
        v = _device_view(r.vertices, 3 * r.n_vertices, torch.float32,
                         self.device).cpu().numpy().reshape(-1, 3)
        t = _device_view(r.triangles, 3 * r.n_triangles, torch.int32,
                         self.device).cpu().numpy().reshape(-1, 3)
        return f, v, t

    def result_tensors(self, r, params):
        """Device views (field, vertices (V, 3), triangles (T, 3)) of a finished
        request, valid until the request is released."""
        f = _device_view(r.field, params.npoints(), torch.float32, self.device)
        v = _device_view(r.vertices, 3 * r.n_vertices, torch.float32, self.device).view(-1, 3)
        t = _device_view(r.triangles, 3 * r.n_triangles, torch.int32, self.device).view(-1, 3)
        return f, v, t

    def run_trace(self, requests, arrivals):
        n = len(requests)
        reqs = (LatticeParams * n)(*requests)
        arr = (C.c_double * n)(*arrivals)
        lat = (C.c_double * n)()
        wall = C.c_double()
        check(lib.gpucad_server_run_trace(self._h, reqs, arr, n, lat, C.byref(wall)))
        return list(lat), wall.value

    def close(self):
        if self._h:
            check(lib.gpucad_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plain_run_trace(requests, arrivals, iso=0.0):
    """The same requests through plain per-request calls, one at a time."""
    n = len(requests)
    reqs = (LatticeParams * n)(*requests)
    arr = (C.c_double * n)(*arrivals)
    lat = (C.c_double * n)()
    wall = C.c_double()
    check(lib.gpucad_plain_run_trace(reqs, arr, n, iso, lat, C.byref(wall)))
    return list(lat), wall.value


# ---------------------------------------------------------------------------
# spatially variant lattice synthesis
# ---------------------------------------------------------------------------
def svl_params(n, lattice_type="n", uniform_type=0, const_period=30.0, x_period=30.0,
               y_period=30.0, z_period=30.0, lcon=0.16, lcon_1=0.6, sinewave_zaxis=False,
               range_=2, harmonics=None, cg_max_iter=500, cg_end_res=0.01, dx=1.0):
    """Parameters of the reference's synthesis loop on an n^3 coarse grid
    (defaults: src/ImguiApp.cpp:91-123, src/main.cu:582-626, 2832-2848):
    fine grid 2n^3 at half spacing, the first half of the coefficients."""
    p = SvlParams()
    p.nx = p.ny = p.nz = n
    p.dx = p.dy = p.dz = dx
    p.lattice_type = lattice_type.encode()
    p.uniform_type = uniform_type
    p.const_period, p.x_period, p.y_period, p.z_period = const_period, x_period, y_period, z_period
    p.lcon, p.lcon_1 = lcon, lcon_1
    p.sinewave_zaxis = 1 if sinewave_zaxis else 0
    p.range = range_
    side = 2 * range_ + 1
    p.harmonics = side ** 3 // 2 if harmonics is None else harmonics
    p.cg_max_iter, p.cg_end_res = cg_max_iter, cg_end_res
    p.fx = p.fy = p.fz = 2 * n
    p.fine_dx = p.fine_dy = p.fine_dz = 0.5 * dx
    return p


def unit_spectrum(cell, n, range_=2):
    """(2r+1)^3 Fourier coefficients (complex64 tensor) of an n^3 unit cell."""
    side = 2 * range_ + 1
    out = torch.empty(side ** 3, dtype=torch.complex64, device=cell.device)
    check(lib.gpucad_unit_spectrum(ptr(cell), n, n, n, range_, ptr(out)))
    return out


def svl_field(params, coeffs, period=None, stream=None):
    """Synthesised lattice field on the fine grid (float32 tensor) and the
    CG statistics of the three unit solves."""
    out = torch.empty(params.fx * params.fy * params.fz, dtype=torch.float32,
                      device=coeffs.device)
    res = SvlResult()
    check(lib.gpucad_svl_field(C.byref(params), ptr(period), ptr(coeffs), ptr(out), C.byref(res),
                               stream_handle(stream)))
    return out, res

