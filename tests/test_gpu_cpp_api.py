"""The C++ particle / field API (include/psum/) on the GPU, checked against
the fused C ABI kernels:

  * tests/cpp/api_step.cpp: a reference-style step loop written with
    particle_group::for_each, add_back and interp_diff;
  * tests/cpp/mesh_router.cpp: the triangle-mesh boundary (mesh_generator,
    triangle_mesh_trigger, basic_material_set, boundary_router) against
    psum_hip_mesh_detect / psum_hip_mesh_boundary, bit for bit.
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "api_step")
ROUTER = os.path.join(HERE, "cpp", "mesh_router")


def test_cpp_api_step_matches_fused_kernels():
    assert os.path.exists(EXE), "tests/cpp/api_step is built by psum_amd/build.py"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout


def test_cpp_mesh_router_matches_fused_kernels():
    assert os.path.exists(ROUTER), "tests/cpp/mesh_router is built by psum_amd/build.py"
    print("cpp mesh_router", flush=True)
    r = subprocess.run([ROUTER, os.path.join(HERE, "golden", "race.stl")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
