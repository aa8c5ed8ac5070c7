"""The C-ABI library loads on the CPU and exports every function that
include/*.h declares (no compute call needs a GPU here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("psum_hip.h", "psum_field_solver.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"typedef[^;]*;", "", src)
        for m in re.finditer(r"\b(psum_\w+|register_solver|get_solver_funcs)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from psum_amd import _lib
    L = _lib.lib()
    declared = _declared()
    assert len(declared) > 30
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert declared <= bound | {"psum_hip_last_error"}, declared - bound


def test_registry_is_host_only_and_reference_shaped():
    from psum_amd import _lib
    L = _lib.lib()
    t = L.get_solver_funcs(b"hip_multigrid_gpu")
    assert t.init and t.set_matrix and t.solve and t.set_options and t.set_source_replace \
        and t.set_source_addback and t.free
    assert L.get_solver_funcs(b"hip_multigrid32f_gpu").init
    assert not L.get_solver_funcs(b"sycl_multigrid_gpu").init
    assert C.sizeof(_lib.FuncTable) == 7 * C.sizeof(C.c_void_p)


def test_argument_errors_do_not_touch_the_gpu():
    from psum_amd import _lib
    L = _lib.lib()
    g = _lib.PsumGrid()
    g.dim = 7
    rc = L.psum_hip_bin(C.byref(g), 0, None, 10, None, None)
    assert rc == 1
    assert b"invalid grid" in L.psum_hip_last_error()
    assert L.psum_hip_sort(None, -1, 5, None, None, None, 0, None) == 1
    # the fixed-point deposit needs a power-of-two scale (exact rescaling)
    g2 = _lib.PsumGrid()
    g2.dim = 2
    g2.cells[0] = g2.cells[1] = 4
    for d in range(2):
        g2.delta[d] = g2.delta_r[d] = 1.0
    acc = (C.c_int64 * 25)()
    assert L.psum_hip_deposit_cic(C.byref(g2), 0, None, 0, 1.0, 3.0, acc, None, None) == 1
    assert b"power of two" in L.psum_hip_last_error()


def test_grid_struct_layout_matches_oracle():
    from psum_amd import _lib
    import oracle_lib as O
    assert C.sizeof(_lib.PsumGrid) == C.sizeof(O.Grid)
    for (a, ta), (b, tb) in zip(_lib.PsumGrid._fields_, O.Grid._fields_):
        assert a == b and C.sizeof(ta) == C.sizeof(tb)
        assert getattr(_lib.PsumGrid, a).offset == getattr(O.Grid, b).offset


def test_cpp_api_program_is_built():
    """build() compiles the C++ API test program against the headers in
    include/psum/ and the library (it runs only on a GPU)."""
    exe = os.path.join(ROOT, "tests", "cpp", "api_step")
    assert os.path.exists(exe) and os.access(exe, os.X_OK)


@pytest.mark.parametrize("case", ["periodic2d", "walls2d", "walls3d"])
def test_poisson_assembly_matches_oracle(case):
    """psum_poisson_assemble (host side, no GPU) builds the same matrix as the
    oracle's restatement of the reference frontends."""
    import numpy as np
    from psum_amd import _lib
    import oracle_lib as O
    if case == "periodic2d":
        og = O.periodic_grid([12, 9])
        fixed = np.zeros(O.num_nodes(og), np.uint8)
        fixed[0] = 1
    elif case == "walls2d":
        og = O.wall_grid([10, 7])
        shp = O.node_shape(og)
        f = np.zeros(shp, np.uint8)
        f[0, :] = f[-1, :] = f[:, 0] = f[:, -1] = 1
        fixed = f.ravel()
    else:
        og = O.wall_grid([5, 4, 6])
        shp = O.node_shape(og)
        f = np.zeros(shp, np.uint8)
        f[0], f[-1], f[:, 0], f[:, -1], f[:, :, 0], f[:, :, -1] = 1, 1, 1, 1, 1, 1
        fixed = f.ravel()
    r, c, v = O.poisson_coo(og, 1.5, fixed)
    g = _lib.PsumGrid()
    C.memmove(C.byref(g), C.byref(og), C.sizeof(g))
    L = _lib.lib()
    nnz = C.c_int64(0)
    fx = np.ascontiguousarray(fixed, np.uint8)
    assert L.psum_poisson_assemble(C.byref(g), 1.5, fx.ctypes.data, 0, None, None, None, C.byref(nnz)) == 0
    rr = np.zeros(nnz.value, np.uint64)
    cc = np.zeros(nnz.value, np.uint64)
    vv = np.zeros(nnz.value, np.float64)
    assert L.psum_poisson_assemble(C.byref(g), 1.5, fx.ctypes.data, nnz.value, rr.ctypes.data,
                                   cc.ctypes.data, vv.ctypes.data, C.byref(nnz)) == 0
    keep = vv != 0
    a = sorted(zip(rr[keep].tolist(), cc[keep].tolist(), vv[keep].tolist()))
    kr = np.asarray(v) != 0
    b = sorted(zip(np.asarray(r)[kr].tolist(), np.asarray(c)[kr].tolist(), np.asarray(v)[kr].tolist()))
    assert [x[:2] for x in a] == [x[:2] for x in b]
    assert np.allclose([x[2] for x in a], [x[2] for x in b], rtol=1e-14, atol=0)
