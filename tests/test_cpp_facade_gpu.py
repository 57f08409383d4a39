"""The C++ facade (include/hacs_mi355x_engine.hpp) driven by a C++ program built against the library."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_test")


def test_cpp_facade(hx_gpu):
    assert os.path.exists(BIN), "tests/cpp/facade_test is not built (make -C hacs_amd/csrc)"
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0 and "ALL PASSED" in out.stdout, out.stdout + out.stderr
