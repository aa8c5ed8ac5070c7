"""Build the HIP step engine (libpsum_hip.so) in-tree for gfx950.

    python -m psum_amd.build

Every .hip source under csrc/ is compiled with hipcc and linked into
psum_amd/libpsum_hip.so.  Flags:
  -ffp-contract=off   every a*b+c stays two rounded operations, the
                      arithmetic order the parity tests pin bit-exactly
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpsum_hip.so")
ARCH = "gfx950"

FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
    cmd = [_hipcc(), "-c", src, "-o", obj] + FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


CPP_TESTS = [os.path.join(ROOT, "tests", "cpp", "api_step.cpp"), os.path.join(ROOT, "tests", "cpp", "mesh_router.cpp"),
             os.path.join(ROOT, "tests", "cpp", "solver1d.cpp")]


def build_cpp_tests(verbose=True):
    """Compile the C++ API programs the GPU tests run (tests/cpp/*.cpp ->
    executables beside them, linked against libpsum_hip.so)."""
    for src in CPP_TESTS:
        exe = src[:-4]
        cmd = [_hipcc(), "-O2", "-std=c++20", f"--offload-arch={ARCH}", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I", os.path.join(ROOT, "include"), "-o", exe + ".tmp", src,
               "-L", HERE, "-lpsum_hip", "-Wl,-rpath,$ORIGIN/../../psum_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
        if verbose:
            print(f"built {exe}")


HOST_TESTS = [os.path.join(ROOT, "tests", "cpp", "mas_io.cpp")]


def build_host_tests(verbose=True):
    """Host-only C++ test drivers (no GPU code): g++ against include/."""
    for src in HOST_TESTS:
        exe = src[:-4]
        cmd = ["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", exe + ".tmp", src]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src}:\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
        if verbose:
            print(f"built {exe}")


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_cpp_tests(verbose)
    build_host_tests(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
