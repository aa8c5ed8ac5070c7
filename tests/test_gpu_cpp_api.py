"""The C++ particle / field API (include/psum/) on the GPU: a reference-style
step loop written with particle_group::for_each, add_back and interp_diff,
checked against the fused C ABI kernels (tests/cpp/api_step.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "api_step")


def test_cpp_api_step_matches_fused_kernels():
    assert os.path.exists(EXE), "tests/cpp/api_step is built by psum_amd/build.py"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
