"""The C++ host layer (include/alsvinn_hip.hpp) on the GPU, through the
test program built by __graft_entry__.build()."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_host_layer_against_oracle():
    exe = os.path.join(ROOT, "tests", "cpp", "test_host_api")
    assert os.path.exists(exe), "build the test program with __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
