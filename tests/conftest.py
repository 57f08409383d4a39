"""Shared test configuration: the `gpu` marker, import paths and device helpers."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libhacs_mi355x.so")


@pytest.fixture(scope="session")
def hx_gpu():
    """The loaded library on a live gfx950 device; a GPU test never skips silently."""
    import hacs_amd
    from hacs_amd import _lib
    if not hacs_amd.available():
        raise RuntimeError("gfx950 device unavailable: " + _lib.lib().hx_unavailable_reason().decode())
    import torch
    torch.cuda.init()
    return _lib.lib()


@pytest.fixture(scope="session")
def O():
    """The CPU oracle (test infrastructure)."""
    from oracle import oracle
    oracle.lib()
    return oracle
