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
    # per-species sort intervals: a null stepper is refused before any GPU work
    every = C.c_int(-7)
    assert L.psum_stepper_set_species_sort(None, 0, 5) == 1
    assert L.psum_stepper_species_sort(None, 0, C.byref(every)) == 1
    assert every.value == -7


def test_bench_window_covers_whole_sort_periods():
    """bench.py times windows that start on a step where every species sorts:
    the period is the least common multiple of the per-species intervals
    (zero intervals never sort), and the untimed steps before the window
    bring the step counter to a multiple of it."""
    import importlib.util, os
    spec = importlib.util.spec_from_file_location(
        "bench_mod", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    assert bench._sort_period([20, 200]) == 200
    assert bench._sort_period([20, 50]) == 100
    assert bench._sort_period([0, 20]) == 20
    assert bench._sort_period([0, 0]) == 0

    class FakeStepper:
        def __init__(self, steps):
            self.steps, self.ran = steps, []

        def stats(self):
            return {"steps": self.steps}

        def run(self, n, history=True):
            self.ran.append(n)
            self.steps += n

    st = FakeStepper(41)
    assert bench._align_to_sort(st, 200) == 159 and st.steps == 200
    assert bench._align_to_sort(st, 200) == 0 and st.ran == [159]
    assert bench._align_to_sort(FakeStepper(7), 0) == 0


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
    for name in ("api_step", "mesh_router"):
        exe = os.path.join(ROOT, "tests", "cpp", name)
        assert os.path.exists(exe) and os.access(exe, os.X_OK), name


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


def _product_assemble_ex(og, eps, fixed, nodes, cylindrical=False, eps_fn=None):
    import numpy as np
    from psum_amd import _lib
    g = _lib.PsumGrid()
    C.memmove(C.byref(g), C.byref(og), C.sizeof(g))
    arr = (_lib.PsumMixedNode * max(1, len(nodes)))()
    for k, (e, d, v, gc) in enumerate(nodes):
        arr[k].element, arr[k].direction, arr[k].value_coef, arr[k].grad_coef = e, d, v, gc
    fx = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
    spec = _lib.PsumPoissonSpec()
    spec.cylindrical = int(cylindrical)
    spec.eps = eps
    cb = _lib.EPS_FN(lambda p, u: eps_fn(p)) if eps_fn else _lib.EPS_FN()
    spec.eps_fn = cb
    spec.fixed = None if fx is None else fx.ctypes.data
    spec.mixed = C.cast(arr, C.c_void_p) if nodes else None
    spec.n_mixed = len(nodes)
    L = _lib.lib()
    nnz, mnnz = C.c_int64(0), C.c_int64(0)
    st = L.psum_poisson_assemble_ex(C.byref(g), C.byref(spec), 0, None, None, None, C.byref(nnz),
                                    0, None, None, None, C.byref(mnnz))
    if st != 0:
        return st, None, None
    out = [np.zeros(nnz.value, t) for t in (np.uint64, np.uint64, np.float64)]
    mout = [np.zeros(mnnz.value, t) for t in (np.uint64, np.uint64, np.float64)]
    st = L.psum_poisson_assemble_ex(C.byref(g), C.byref(spec), nnz.value, *[a.ctypes.data for a in out],
                                    C.byref(nnz), mnnz.value, *[a.ctypes.data for a in mout],
                                    C.byref(mnnz))
    return st, out, mout


def _sorted_triplets(t):
    import numpy as np
    r, c, v = (np.asarray(a) for a in t)
    o = np.lexsort((c, r))
    return r[o], c[o], v[o]


def _edge(og, axis, plus):
    import numpy as np
    import oracle_lib as O
    shp = O.node_shape(og)
    idx = np.arange(O.num_nodes(og)).reshape(shp)
    sl = [slice(None)] * og.dim
    sl[axis] = shp[axis] - 1 if plus else 0
    return idx[tuple(sl)].ravel()


@pytest.mark.parametrize("case", ["robin2d", "cyl", "eps_fn", "robin3d"])
def test_poisson_assembly_ex_bitexact_vs_oracle(case):
    """psum_poisson_assemble_ex (host, no GPU): A' and M equal the oracle's
    restatement of the reference frontend entry for entry, bit for bit."""
    import numpy as np
    import oracle_lib as O
    eps_fn, cyl = None, False
    if case == "robin2d":
        og = O.make_grid([-1.0, -0.5], [1.0, 1.5], [9, 7])
        sides = [(1, True, 0.0, 1.0), (1, False, 2.0, 0.5), (0, True, 1.0, 3.0), (0, False, 1.0, 1.0)]
        fixed_sides = [(0, False)]
    elif case == "cyl":
        og = O.make_grid([0.0, 0.0], [1.0, 1.67], [12, 8])
        sides = [(0, True, 0.0, 1.0), (0, False, 0.0, 1.0), (1, False, 0.0, 1.0)]
        fixed_sides = [(1, True)]
        cyl = True
    elif case == "eps_fn":
        og = O.make_grid([0.0, 0.0], [2.0, 1.0], [10, 6])
        sides = [(1, True, 1.0, 2.0)]
        fixed_sides = [(0, False), (0, True)]
        eps_fn = lambda p: 1.0 + 0.5 * np.sin(p[0]) * np.cos(3 * p[1])
    else:
        og = O.make_grid([0.0, -1.0, 0.5], [1.0, 0.5, 1.5], [5, 6, 4])
        sides = [(0, True, 1.0, 2.0), (1, True, 0.0, 1.0), (1, False, 0.5, 1.0), (2, True, 0.0, 1.0),
                 (2, False, 3.0, 0.25)]
        fixed_sides = [(0, False)]
    n = O.num_nodes(og)
    fixed = np.zeros(n, np.uint8)
    for ax, pl in fixed_sides:
        fixed[_edge(og, ax, pl)] = 1
    nodes = []
    for ax, pl, v, gc in sides:
        code = ((0 if pl else 1) if ax == 1 else (2 if pl else 3)) if og.dim == 2 else 2 * ax + (0 if pl else 1)
        nodes += [(int(e), code, v, gc) for e in _edge(og, ax, pl)]
    A, M = O.poisson_assemble_ex(og, 8.854e-12 if cyl else 1.25, fixed, nodes, cyl, eps_fn)
    st, A2, M2 = _product_assemble_ex(og, 8.854e-12 if cyl else 1.25, fixed, nodes, cyl, eps_fn)
    assert st == 0
    for x, y in zip(_sorted_triplets(A), _sorted_triplets(A2)):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(M, M2):          # both row-major
        np.testing.assert_array_equal(x, y)
    assert M2[0].size > 0


def test_poisson_assembly_ex_rejects_bad_input():
    import oracle_lib as O
    og = O.wall_grid([4, 3])
    st, _, _ = _product_assemble_ex(og, 1.0, None, [(4, 0, 1.0, 0.0)])   # grad_coef 0
    assert st != 0
    st, _, _ = _product_assemble_ex(og, 1.0, None, [(400, 0, 0.0, 1.0)])  # element out of range
    assert st != 0
    og3 = O.wall_grid([3, 3, 3])
    st, _, _ = _product_assemble_ex(og3, 1.0, None, [], cylindrical=True)
    assert st != 0
