"""The C++ front end (include/blast/hip/Blast.hpp): compiles on the CPU; the
built test program (tests/cpp/test_api) runs on the GPU."""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_frontend_header_compiles():
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-std=c++20", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "test_api.cpp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.gpu
def test_frontend_program_on_gpu():
    exe = os.path.join(ROOT, "tests", "cpp", "test_api")
    assert os.path.exists(exe), "tests/cpp/test_api is built by __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
