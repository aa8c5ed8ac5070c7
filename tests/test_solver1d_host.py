"""Poisson_solver_1d of the C++ API (include/psum/poisson_solver_1d.hpp):
a host-only program (tests/cpp/solver1d.cpp, compiled by hipcc with the
other C++ API programs, no GPU call) on the reference's test_psolver_1d.cpp
cases."""
import os
import subprocess

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "solver1d")


def test_poisson_solver_1d_reference_cases():
    assert os.path.exists(EXE), "tests/cpp/solver1d is built by psum_amd/build.py"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
