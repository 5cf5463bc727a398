"""Interop on the GPU: tests/cpp/interop_harness exports a device allocation
as an opaque POSIX fd (HIP's virtual-memory API stands in for Vulkan's
vkGetMemoryFdKHR, which has no device here), imports it through
gpucad_import_external_buffer, checks that the mapping aliases the
allocation, and has gpucad_extract write a mesh's vertices into it; they
must equal the same extraction into ordinary device memory, byte for byte."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def test_extraction_writes_into_an_imported_buffer():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "cpp", "interop_harness")
    print("interop_harness: 64^3 sphere, vertices into an imported fd", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    out = r.stdout.strip()
    if out.startswith("SKIP"):
        pytest.skip(out)
    assert r.returncode == 0 and out.startswith("PASS"), out + r.stderr
