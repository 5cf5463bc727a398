"""Graphics-API interop entry points on the CPU: argument checks (these
return before any HIP call, so no GPU is needed)."""
import ctypes as C


def test_import_buffer_rejects_bad_arguments(gc):
    h, p = C.c_void_p(), C.c_void_p()
    assert gc.lib.gpucad_import_external_buffer(-1, 4096, 0, C.byref(h), C.byref(p)) == gc.E_ARG
    assert gc.lib.gpucad_import_external_buffer(3, 0, 0, C.byref(h), C.byref(p)) == gc.E_ARG
    assert gc.lib.gpucad_import_external_buffer(3, 4096, 4096, C.byref(h), C.byref(p)) == gc.E_ARG
    assert gc.lib.gpucad_import_external_buffer(3, 4096, 0, None, C.byref(p)) == gc.E_ARG
    assert h.value is None and p.value is None
    assert gc.lib.gpucad_release_external_buffer(None) == gc.OK


def test_semaphore_entry_points_reject_bad_arguments(gc):
    s = C.c_void_p()
    assert gc.lib.gpucad_import_external_semaphore(-1, C.byref(s)) == gc.E_ARG
    assert gc.lib.gpucad_import_external_semaphore(3, None) == gc.E_ARG
    assert gc.lib.gpucad_signal_external_semaphore(None, None) == gc.E_ARG
    assert gc.lib.gpucad_wait_external_semaphore(None, None) == gc.E_ARG
    assert gc.lib.gpucad_release_external_semaphore(None) == gc.OK
