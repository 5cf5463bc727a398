"""The C-ABI boundary (no GPU needed): the library loads, exports every entry
point include/gpucad.h declares, the Python binding covers them with the
right struct layouts, and host-only calls work."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpucad.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpucad_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) >= 30
    for must in ("gpucad_init", "gpucad_lattice_field", "gpucad_extract",
                 "gpucad_compute_isosurface_lattice", "gpucad_file_write_obj"):
        assert must in names


def test_library_exports_every_declared_symbol(gc):
    missing = [n for n in declared_functions() if not hasattr(gc.lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol(gc):
    assert set(declared_functions()) <= set(gc.SIGNATURES)


def test_struct_layouts_match(gc):
    sizes = (C.c_uint32 * 9)()
    assert gc.lib.gpucad_abi_sizes(sizes, 9) == 9
    expect = [C.sizeof(gc.LatticeParams), C.sizeof(gc.ExtractArgs), C.sizeof(gc.MeshCounts),
              C.sizeof(gc.Uint3), C.sizeof(gc.Float3), C.sizeof(gc.Float4), C.sizeof(gc.Float2),
              C.sizeof(gc.SvlParams), C.sizeof(gc.SvlResult)]
    assert list(sizes) == expect
    assert list(sizes)[3:7] == [12, 12, 16, 8]


def test_version_and_status_strings(gc):
    assert gc.lib.gpucad_version() >= 10000
    assert gc.status_string(gc.OK) == "ok"
    assert gc.status_string(gc.E_OVERFLOW) == "output capacity exceeded"
    assert gc.status_string(12345) == "unknown status"


def test_host_tables_without_gpu(gc, orc):
    tri = np.zeros(256 * 16, np.uint8)
    nv = np.zeros(256, np.uint8)
    em = np.zeros(256, np.uint16)
    assert gc.lib.gpucad_mc_tables(tri.ctypes.data, nv.ctypes.data, em.ctypes.data) == gc.OK
    otri, onv = orc.tables()
    assert np.array_equal(tri.reshape(256, 16), otri)
    assert np.array_equal(nv, onv)


def test_sign_words_and_workspace_sizes(gc):
    assert gc.sign_words(64, 3, 2) == 6
    assert gc.sign_words(65, 3, 2) == 12
    assert gc.sign_words(0, 3, 2) == 0
    assert gc.extract_workspace_bytes(1, 5, 5) == 0
    a = gc.extract_workspace_bytes(512, 512, 512)
    assert a >= gc.sign_words(512, 512, 512) * 8


def test_argument_errors_are_reported_without_gpu(gc):
    # invalid arguments are rejected before any device work
    p = gc.LatticeParams()
    p.nx = p.ny = p.nz = 1
    st = gc.lib.gpucad_lattice_field(C.byref(p), None, None, 0.0, None)
    assert st == gc.E_ARG
    assert "field pointer" in gc.last_error() or "points per axis" in gc.last_error()
    assert gc.lib.gpucad_extract(None, None, None) == gc.E_ARG
