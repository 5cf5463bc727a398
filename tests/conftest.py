"""pytest configuration: the `gpu` marker and shared fixtures.

Tests marked `gpu` need an MI355X (HIP device) and the built libgpucad.so;
everything else runs on the CPU (oracle, host logic, ABI surface).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and libgpucad.so")


@pytest.fixture(scope="session")
def gc():
    import gpucadforam_amd as g
    return g


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as o
    return o


@pytest.fixture(scope="session")
def dev(gc):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test run without a visible GPU")
    gc.init(0)
    return torch.device("cuda:0")
