"""MI355X-native electrostatic PIC step engine (HIP kernels behind a C ABI).

The compiled engine is psum_amd/libpsum_hip.so (build with
`python -m psum_amd.build`).  `psum_amd.engine` is the Python host layer.
"""
from ._lib import LIB_PATH, PsumError, lib  # noqa: F401

__all__ = ["LIB_PATH", "PsumError", "lib"]
