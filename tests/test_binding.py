"""The reference-side binding (hacs_amd/binding/EngineHip.cpp) against the reference's own headers:
it must compile as the reference's engine::cuda layer (every definition matching a declaration of
Cuda/CudaCtx.h, Cuda/CudaKernels.h, Cuda/CudaRing2k.h and the hook tables), and link with the
reference's compiled ModArithmetic objects (oracle/_ref) and libhacs_mi355x.so into a program that
installs the hooks and runs the reference's elim::rref and ring2k::echelon_blocked through them
(without a device they decline and the host code runs).  The reference tree is read in place in
this container only; nothing of it is kept."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
INC = ["-I" + REF + "/Cuda", "-I" + REF + "/ModArithmetic", "-I" + REF, "-I" + os.path.join(ROOT, "include")]
FLAGS = ["-std=c++17", "-O1", "-DOC_ENABLE_LEGACY=0", "-Wall", "-Werror=return-type"]

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="the reference tree is not present")


def test_binding_compiles_against_reference_headers(tmp_path):
    out = tmp_path / "EngineHip.o"
    r = subprocess.run(["g++", *FLAGS, *INC, "-c", os.path.join(ROOT, "hacs_amd", "binding", "EngineHip.cpp"),
                        "-o", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "warning" not in r.stderr, r.stderr[-2000:]


def test_binding_links_and_runs_the_reference_through_its_hooks(tmp_path):
    ref_lib = os.path.join(ROOT, "oracle", "_ref", "libhacsref.so")
    if not os.path.exists(ref_lib):
        pytest.skip("oracle/_ref is not built")
    exe = tmp_path / "binding_probe"
    r = subprocess.run(["g++", *FLAGS, *INC, os.path.join(ROOT, "tests", "cpp", "binding_probe.cpp"),
                        os.path.join(ROOT, "hacs_amd", "binding", "EngineHip.cpp"), ref_lib,
                        "-L" + os.path.join(ROOT, "hacs_amd"), "-lhacs_mi355x", "-pthread",
                        "-Wl,-rpath," + os.path.join(ROOT, "hacs_amd") + ":" + os.path.dirname(ref_lib),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "FAIL" not in run.stdout
