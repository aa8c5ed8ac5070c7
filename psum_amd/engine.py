"""Python host layer over the HIP step engine's C ABI.

Mirrors the reference interfaces the hot path exposes:
  * simple_grid construction (src/field/simple_grid.hpp:38-51)
  * particle SoA operations: cell binning, stable sort, permute, compaction
    (particle_group::compress, src/particle_container/particle_group.hpp:223-238)
  * add_back / interp_diff + push (src/field/device_field.hpp:270-310)
  * solver_backend (src/field_solver/backend.hpp:11-107) over the registered
    "hip_multigrid_gpu" / "hip_multigrid32f_gpu" backends
  * a native Stepper running the whole reference step loop on the device.

Tensors are torch CUDA tensors (HIP device memory); torch is used only to
own device memory and provide the stream.  Every call goes through
libpsum_hip.so; there is no CPU path.
"""
import ctypes as C
import os

import numpy as np
import torch

from . import _lib

from ._lib import (KEY_INVALID, PSUM_F32, PSUM_F64, FuncTable, PsumError, PsumGrid,
                   PsumSpecies, PsumStepConfig, check, lib)

__all__ = ["make_grid", "periodic_grid", "wall_grid", "num_nodes", "fx_shift", "cell_keys",
           "stable_sort", "permute", "compact", "deposit", "fixed_to_double", "gather_push",
           "field_energy", "deposit_bspline", "gather_push_bspline", "SolverBackend", "Stepper", "Batch", "step_config", "KEY_INVALID",
           "PsumError",
           "get_solver_funcs"]


# ---------------------------------------------------------------- grids

def make_grid(lb, ub, cells, periodic=None, plo=None, plen=None):
    """simple_grid(lb, ub, cells) (simple_grid.hpp:38-51): delta = (ub-lb)/cells,
    delta_r = 1/delta.  periodic[d] = 1 wraps particles into [plo, plo+plen)."""
    d = len(cells)
    g = PsumGrid()
    g.dim = d
    for i in range(d):
        g.cells[i] = int(cells[i])
        g.lb[i] = float(lb[i])
        g.ub[i] = float(ub[i])
        delta = (float(ub[i]) - float(lb[i])) / int(cells[i])
        g.delta[i] = delta
        g.delta_r[i] = 1.0 / delta
        g.periodic[i] = int(periodic[i]) if periodic else 0
        g.plo[i] = float(plo[i]) if plo else 0.0
        g.plen[i] = float(plen[i]) if plen else 0.0
    return g


def periodic_grid(n, dx=1.0):
    """n[d] nodes per dimension at dx/2 + i*dx with period n*dx, the reference's
    periodic node convention (TSI/main.cpp:94: grid({dx/2}, {L - dx/2}, {N - 1}))."""
    d = len(n)
    return make_grid([dx / 2] * d, [ni * dx - dx / 2 for ni in n], [ni - 1 for ni in n],
                     periodic=[1] * d, plo=[0.0] * d, plen=[ni * dx for ni in n])


def wall_grid(cells, dx=1.0):
    """cells[d] cells on [0, cells*dx), absorbing particle walls."""
    d = len(cells)
    return make_grid([0.0] * d, [c * dx for c in cells], cells, periodic=[0] * d)


def num_nodes(g):
    n = 1
    for i in range(g.dim):
        n *= g.cells[i] + 1
    return n


def node_shape(g):
    return tuple(g.cells[i] + 1 for i in range(g.dim))


def fx_shift(w):
    """Deposit fixed-point exponent S = 40 - e, 2^(e-1) <= |w| < 2^e."""
    _, e = np.frexp(abs(w))
    return 40 - int(e)


# ---------------------------------------------------------------- helpers

def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dtype_code(t):
    if t.dtype == torch.float64:
        return PSUM_F64
    if t.dtype == torch.float32:
        return PSUM_F32
    raise PsumError(f"unsupported particle dtype {t.dtype}")


def _ptrs(tensors):
    arr = (C.c_void_p * max(1, len(tensors)))()
    for i, t in enumerate(tensors):
        assert t.is_cuda and t.is_contiguous()
        arr[i] = t.data_ptr()
    return arr


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _check_soa(g, arrs):
    if len(arrs) != g.dim:
        raise PsumError(f"expected {g.dim} component arrays, got {len(arrs)}")
    n = arrs[0].numel()
    for a in arrs:
        if a.numel() != n or a.dtype != arrs[0].dtype or not a.is_cuda or not a.is_contiguous():
            raise PsumError("SoA arrays must be contiguous CUDA tensors of equal size and dtype")
    return n


# --------------------------------------------------------- particle ops

def cell_keys(g, pos):
    """keys[i] = n2i(wrapped cN(x_i)); KEY_INVALID for invalid particles.
    Returns an int32 tensor holding the uint32 bit patterns."""
    n = _check_soa(g, pos)
    keys = torch.empty(n, dtype=torch.int32, device=pos[0].device)
    check(lib().psum_hip_bin(C.byref(g), _dtype_code(pos[0]), _ptrs(pos), n, _p(keys), _stream()),
          "psum_hip_bin")
    return keys


def stable_sort(keys, nkeys):
    """Stable sort of uint32 keys (int32 tensor bit patterns).  Returns
    (sorted_keys, perm)."""
    n = keys.numel()
    out = torch.empty_like(keys)
    perm = torch.empty_like(keys)
    wb = lib().psum_hip_sort_workspace(n)
    work = torch.empty(max(1, wb), dtype=torch.uint8, device=keys.device)
    check(lib().psum_hip_sort(_p(keys), n, nkeys, _p(out), _p(perm), _p(work), wb, _stream()),
          "psum_hip_sort")
    return out, perm


def permute(arrays, perm):
    """[a[perm] for a in arrays] on the device."""
    dst = [torch.empty_like(a) for a in arrays]
    check(lib().psum_hip_permute(_dtype_code(arrays[0]), len(arrays), _ptrs(arrays), _ptrs(dst),
                                 _p(perm), arrays[0].numel(), _stream()), "psum_hip_permute")
    return dst


def compact(pos, vel):
    """Stable compaction of valid particles (finite pos[0])."""
    dim = len(pos)
    n = pos[0].numel()
    op = [torch.empty_like(a) for a in pos]
    ov = [torch.empty_like(a) for a in vel]
    wb = lib().psum_hip_compact_workspace(n)
    work = torch.empty(max(1, wb), dtype=torch.uint8, device=pos[0].device)
    cnt = C.c_int64(0)
    check(lib().psum_hip_compact(_dtype_code(pos[0]), dim, _ptrs(pos), _ptrs(vel), n, _ptrs(op),
                                 _ptrs(ov), C.byref(cnt), _p(work), wb, _stream()),
          "psum_hip_compact")
    m = cnt.value
    return [a[:m] for a in op], [a[:m] for a in ov]


def deposit(g, pos, w, scale=None, acc=None, counters=None):
    """Fixed-point CIC deposit; returns (acc int64 tensor, scale)."""
    n = _check_soa(g, pos)
    if scale is None:
        scale = float(np.ldexp(1.0, fx_shift(w)))
    if acc is None:
        acc = torch.zeros(num_nodes(g), dtype=torch.int64, device=pos[0].device)
    check(lib().psum_hip_deposit_cic(C.byref(g), _dtype_code(pos[0]), _ptrs(pos), n, w, scale,
                                     _p(acc), _p(counters), _stream()), "psum_hip_deposit_cic")
    return acc, scale


def fixed_to_double(acc, scale, offset=0.0):
    out = torch.empty(acc.numel(), dtype=torch.float64, device=acc.device)
    check(lib().psum_hip_fixed_to_double(_p(acc), acc.numel(), 1.0 / scale, offset, _p(out),
                                         _stream()), "psum_hip_fixed_to_double")
    return out


def gather_push(g, pos, vel, phi, qm, dt, mass=1.0, w_dep=None, scale=None, counters=None):
    """Gather grad(phi), kick, drift, boundaries, in place on pos/vel.
    With w_dep the new positions are deposited.  Returns (acc or None, diag)
    where diag = [sum 1/2 m v^2, sum m vx, sum m vy, sum m vz] of the entering
    velocities."""
    n = _check_soa(g, pos)
    _check_soa(g, vel)
    dev = pos[0].device
    if phi.dtype != pos[0].dtype:
        raise PsumError("phi must have the particle dtype")
    nb = lib().psum_hip_push_blocks(n)
    diag = torch.zeros(max(1, nb) * 4, dtype=torch.float64, device=dev)
    acc = None
    if w_dep is not None:
        if scale is None:
            scale = float(np.ldexp(1.0, fx_shift(w_dep)))
        acc = torch.zeros(num_nodes(g), dtype=torch.int64, device=dev)
    check(lib().psum_hip_gather_push(C.byref(g), _dtype_code(pos[0]), _ptrs(pos), _ptrs(vel), n,
                                     _p(phi), qm, dt, mass, w_dep or 0.0, scale or 1.0, _p(acc),
                                     _p(diag), _p(counters), _stream()), "psum_hip_gather_push")
    out = torch.zeros(4, dtype=torch.float64, device=dev)
    if nb > 0:
        work = torch.empty(lib().psum_hip_reduce_workspace() // 8, dtype=torch.float64, device=dev)
        check(lib().psum_hip_reduce_diag(_p(diag), nb, _p(out), _p(work), _stream()),
              "psum_hip_reduce_diag")
    return acc, out


def deposit_bspline(g, pos, w, order, scale=None, acc=None, counters=None):
    """B-spline deposit (order 2..4; 1 = CIC) as deposit() with the
    higher-order weights; returns (acc int64 tensor, scale)."""
    n = _check_soa(g, pos)
    if scale is None:
        scale = float(np.ldexp(1.0, fx_shift(w)))
    if acc is None:
        acc = torch.zeros(num_nodes(g), dtype=torch.int64, device=pos[0].device)
    check(lib().psum_hip_deposit_bspline(C.byref(g), _dtype_code(pos[0]), int(order), _ptrs(pos), n, w,
                                         scale, _p(acc), _p(counters), _stream()),
          "psum_hip_deposit_bspline")
    return acc, scale


def gather_push_bspline(g, pos, vel, phi, qm, dt, order, mass=1.0, counters=None):
    """B-spline gather + kick + drift + boundaries in place (order 2..4; 1 =
    CIC); returns diag as gather_push does."""
    n = _check_soa(g, pos)
    _check_soa(g, vel)
    dev = pos[0].device
    if phi.dtype != pos[0].dtype:
        raise PsumError("phi must have the particle dtype")
    nb = lib().psum_hip_push_blocks(n)
    diag = torch.zeros(max(1, nb) * 4, dtype=torch.float64, device=dev)
    check(lib().psum_hip_gather_push_bspline(C.byref(g), _dtype_code(pos[0]), int(order), _ptrs(pos),
                                             _ptrs(vel), n, _p(phi), qm, dt, mass, _p(diag), _p(counters),
                                             _stream()), "psum_hip_gather_push_bspline")
    out = torch.zeros(4, dtype=torch.float64, device=dev)
    if nb > 0:
        work = torch.empty(lib().psum_hip_reduce_workspace() // 8, dtype=torch.float64, device=dev)
        check(lib().psum_hip_reduce_diag(_p(diag), nb, _p(out), _p(work), _stream()),
              "psum_hip_reduce_diag")
    return out


def field_energy(g, phi, eps=1.0):
    out = torch.zeros(4, dtype=torch.float64, device=phi.device)
    wb = lib().psum_hip_field_energy_workspace(C.byref(g))
    work = torch.empty(max(1, wb), dtype=torch.uint8, device=phi.device)
    check(lib().psum_hip_field_energy(C.byref(g), _p(phi), eps, _p(out), _p(work), wb, _stream()),
          "psum_hip_field_energy")
    return out[0]


# ---------------------------------------------------------------- solver

def get_solver_funcs(name):
    """The registry lookup of the reference ABI (registry.cpp:16-22)."""
    return lib().get_solver_funcs(name.encode())


class SolverBackend:
    """solver_backend (src/field_solver/backend.hpp:11-107) over the HIP
    multigrid backends.  Errors raise PsumError (the C++ backends throw)."""

    def __init__(self, kind="hip_multigrid_gpu"):
        t = get_solver_funcs(kind)
        if not t.init:
            raise PsumError(f"Error: null pointer (no backend named {kind!r})")
        self.h = lib().psum_solver_create(kind.encode())
        if not self.h:
            raise PsumError("Error: failed to create solver. " +
                            lib().psum_hip_last_error().decode(errors="replace"))

    def set_options(self, options):
        check(lib().psum_solver_set_options(self.h, options.encode()), "set_options")

    def set_matrix(self, n_row, n_col, rows, cols, vals):
        rows = np.ascontiguousarray(rows, np.uint64)
        cols = np.ascontiguousarray(cols, np.uint64)
        vals = np.ascontiguousarray(vals, np.float64)
        check(lib().psum_solver_set_matrix(self.h, n_row, n_col, vals.shape[0],
                                           rows.ctypes.data_as(C.c_void_p),
                                           cols.ctypes.data_as(C.c_void_p),
                                           vals.ctypes.data_as(C.c_void_p)), "set_matrix")

    def set_source_replace(self, idxs, vals):
        idxs = np.ascontiguousarray(idxs, np.uint64)
        vals = np.ascontiguousarray(vals, np.float64)
        check(lib().psum_solver_set_source_replace(self.h, idxs.shape[0],
                                                   idxs.ctypes.data_as(C.c_void_p),
                                                   vals.ctypes.data_as(C.c_void_p)), "replace")

    def set_source_addback(self, idxs, vals):
        idxs = np.ascontiguousarray(idxs, np.uint64)
        vals = np.ascontiguousarray(vals, np.float64)
        check(lib().psum_solver_set_source_addback(self.h, idxs.shape[0],
                                                   idxs.ctypes.data_as(C.c_void_p),
                                                   vals.ctypes.data_as(C.c_void_p)), "addback")

    def solve(self, b, x):
        """b, x: float64 CUDA tensors; b is modified in place, x is the warm
        start and the result."""
        assert b.dtype == torch.float64 and x.dtype == torch.float64
        torch.cuda.synchronize()
        check(lib().psum_solver_solve(self.h, _p(b), _p(x)), "solve")

    def stats(self):
        cyc = C.c_int(0)
        rel = C.c_double(0)
        check(lib().psum_solver_stats(self.h, C.byref(cyc), C.byref(rel)), "stats")
        return cyc.value, rel.value

    def close(self):
        if getattr(self, "h", None):
            lib().psum_solver_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


# ---------------------------------------------------------------- stepper

# mixed-boundary side codes (2D: N S E W; 3D: Xpos Xneg Ypos Yneg Zpos Zneg)
DIR2 = {"N": 0, "S": 1, "E": 2, "W": 3}
DIR3 = {"Xpos": 0, "Xneg": 1, "Ypos": 2, "Yneg": 3, "Zpos": 4, "Zneg": 5}


def poisson_assemble(g, eps=8.854e-12, fixed=None, mixed=(), cylindrical=False, eps_fn=None):
    """Host assembly of the reference frontends' FD Poisson system
    (Poisson_solver_2d.hpp:182-410 with the ghost elimination of
    ghost_element_manager.hpp:56-117), through psum_poisson_assemble_ex.
    mixed: (element, direction, value_coef, grad_coef) tuples; eps_fn(pos)
    gives the permittivity at a face midpoint.  Returns ((rows, cols, vals) of
    A', row-major) and ((rows, cols, vals) of M = -B D^-1, row-major)."""
    nodes = list(mixed)
    arr = (_lib.PsumMixedNode * max(1, len(nodes)))()
    for k, (e, d, v, gc) in enumerate(nodes):
        arr[k].element, arr[k].direction, arr[k].value_coef, arr[k].grad_coef = e, d, v, gc
    n = num_nodes(g)
    fx = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
    if fx is not None and fx.size != n:
        raise ValueError("poisson_assemble: fixed mask needs one entry per node")
    spec = _lib.PsumPoissonSpec()
    spec.cylindrical = int(bool(cylindrical))
    spec.eps = float(eps)
    cb = _lib.EPS_FN(lambda p, u: float(eps_fn(p))) if eps_fn is not None else _lib.EPS_FN()
    spec.eps_fn = cb
    spec.fixed = None if fx is None else fx.ctypes.data
    spec.mixed = C.cast(arr, C.c_void_p) if nodes else None
    spec.n_mixed = len(nodes)
    nnz, mnnz = C.c_int64(0), C.c_int64(0)
    check(lib().psum_poisson_assemble_ex(C.byref(g), C.byref(spec), 0, None, None, None, C.byref(nnz),
                                         0, None, None, None, C.byref(mnnz)), "poisson_assemble")
    A = [np.zeros(nnz.value, t) for t in (np.uint64, np.uint64, np.float64)]
    M = [np.zeros(mnnz.value, t) for t in (np.uint64, np.uint64, np.float64)]
    check(lib().psum_poisson_assemble_ex(C.byref(g), C.byref(spec), nnz.value,
                                         *[a.ctypes.data_as(C.c_void_p) for a in A], C.byref(nnz),
                                         mnnz.value, *[a.ctypes.data_as(C.c_void_p) for a in M],
                                         C.byref(mnnz)), "poisson_assemble")
    return tuple(A), tuple(M)


def poisson_addback(M, values):
    """Per row of M (row-major): 0 + sum_m M(r, m) * values[m] in column order
    (Poisson_solver_2d.hpp:71-85).  Returns (rows, sums) for
    set_source_addback."""
    mr, mc, mv = M
    vals = np.asarray(values, np.float64)
    rows, sums = [], []
    q = 0
    while q < len(mr):
        e, acc = q, 0.0
        while e < len(mr) and mr[e] == mr[q]:
            acc += float(mv[e]) * float(vals[int(mc[e])])
            e += 1
        rows.append(int(mr[q]))
        sums.append(acc)
        q = e
    return np.asarray(rows, np.uint64), np.asarray(sums, np.float64)


def step_config(g, species, dtype="f64", mixed_mg=False, depth=4, pre=2, post=1, max_cycles=20,
                tol=1e-10, eps=1.0, dt=0.1, rho_bg=0.0, sort_every=0, fixed=None, fixed_val=None,
                capacity=None, mg_graph=True, extrapolate=False):
    """psum_step_config for the given species (dicts with q, m, w, pos) and
    the host arrays it points at (keep them alive while it is used)."""
    cfg = PsumStepConfig()
    cfg.grid = g
    cfg.dtype = PSUM_F64 if dtype == "f64" else PSUM_F32
    cfg.mixed_mg = int(bool(mixed_mg))
    cfg.nspecies = len(species)
    cfg.depth, cfg.pre, cfg.post, cfg.max_cycles = depth, pre, post, max_cycles
    cfg.sort_every = sort_every
    cfg.mg_graph = int(bool(mg_graph))
    cfg.extrapolate = int(bool(extrapolate))
    cfg.tol, cfg.eps, cfg.dt, cfg.rho_bg = tol, eps, dt, rho_bg
    keep = []
    if fixed is not None:
        fm = np.ascontiguousarray(fixed, np.uint8)
        keep.append(fm)
        cfg.fixed = fm.ctypes.data
    if fixed_val is not None:
        fv = np.ascontiguousarray(fixed_val, np.float64)
        keep.append(fv)
        cfg.fixed_val = fv.ctypes.data
    for i, sp in enumerate(species):
        n = len(sp["pos"][0])
        cfg.species[i] = PsumSpecies(sp["q"], sp["m"], sp["w"], n, capacity if capacity else n)
    return cfg, keep


class Batch:
    """Independent instances stepped together in one replayed hipGraph
    (psum_batch_*).  instances: list of dicts with the Stepper arguments
    (g, species, dtype, ...)."""

    def __init__(self, instances):
        from ._lib import PsumStepConfig as _Cfg
        cfgs = (_Cfg * len(instances))()
        self._keep = []
        for i, a in enumerate(instances):
            a = dict(a)
            g, species = a.pop("g"), a["species"]
            a.pop("species")
            c, k = step_config(g, species, **a)
            cfgs[i] = c
            self._keep += k
        self.h = lib().psum_batch_create(cfgs, len(instances))
        if not self.h:
            raise PsumError("psum_batch_create: " +
                            lib().psum_hip_last_error().decode(errors="replace"))
        self.instances = []
        for i, a in enumerate(instances):
            hi = lib().psum_batch_instance(self.h, i)
            st = Stepper(a["g"], a["species"], dtype=a.get("dtype", "f64"), _handle=hi)
            for s, sp in enumerate(a["species"]):
                st.load(s, sp["pos"], sp["vel"])
            self.instances.append(st)

    def __len__(self):
        return len(self.instances)

    def configure(self, i, fixed_cycles, sort_every=0, sort_shift=0):
        check(lib().psum_batch_configure(self.h, i, fixed_cycles, sort_every, sort_shift),
              "psum_batch_configure")

    def tune(self, warmup=20, probe=10):
        """Autotune every instance; returns a list of dicts (fixed_cycles,
        sort_every, sort_shift, sort_ms, push_ms, push_sort_cost_ms)."""
        from ._lib import PsumBatchTuning
        out = (PsumBatchTuning * len(self.instances))()
        check(lib().psum_batch_tune(self.h, warmup, probe, out), "psum_batch_tune")
        return [{f: getattr(t, f) for f, _ in PsumBatchTuning._fields_ if f != "pad_"} for t in out]

    def reset_history(self, rows):
        check(lib().psum_batch_reset_history(self.h, rows), "psum_batch_reset_history")

    def set_lanes(self, lanes):
        """Replay the step as `lanes` per-group graphs on concurrent streams."""
        check(lib().psum_batch_set_lanes(self.h, lanes), "psum_batch_set_lanes")

    def set_solver_batching(self, on=True):
        """Run the solver cycles of all instances as batched launches (the
        default; needs one lane and 2D fixed-cycle solves) or per instance."""
        check(lib().psum_batch_set_solver_batching(self.h, int(bool(on))),
              "psum_batch_set_solver_batching")

    def solver_launches(self):
        """(batched solver launches per step, per-instance launches they
        replace); (0, 0) when the step does not batch its solves"""
        a, b = C.c_int64(0), C.c_int64(0)
        check(lib().psum_batch_solver_launches(self.h, C.byref(a), C.byref(b)),
              "psum_batch_solver_launches")
        return a.value, b.value

    def run(self, steps):
        """`steps` graph replays; returns the HIP-event milliseconds."""
        ms = C.c_double(0)
        check(lib().psum_batch_run(self.h, steps, C.byref(ms)), "psum_batch_run")
        return ms.value

    def history(self, i, max_rows=1 << 20):
        rows = C.c_int64(0)
        check(lib().psum_batch_history(self.h, i, None, max_rows, C.byref(rows)),
              "psum_batch_history")
        n = rows.value
        out = np.zeros((n, 5), np.float64)
        check(lib().psum_batch_history(self.h, i, out.ctypes.data_as(C.c_void_p), n, C.byref(rows)),
              "psum_batch_history")
        return out

    def close(self):
        if getattr(self, "h", None):
            for st in self.instances:
                st.h = None
            lib().psum_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Stepper:
    """The reference step loop (deposit -> solve -> gather/push) run natively.

    species: list of dicts with q, m, w, pos (list of numpy arrays), vel.
    """

    def __init__(self, g, species, dtype="f64", mixed_mg=False, depth=4, pre=2, post=1,
                 max_cycles=20, tol=1e-10, eps=1.0, dt=0.1, rho_bg=0.0, sort_every=0,
                 fixed=None, fixed_val=None, capacity=None, mg_graph=True, extrapolate=False,
                 interp_order=1, _handle=None):
        self.g = g
        self.dim = g.dim
        self.np_dtype = np.float64 if dtype == "f64" else np.float32
        self.n = [len(sp["pos"][0]) for sp in species]
        self._owned = _handle is None
        if _handle is not None:          # an instance owned by a Batch
            self.h = _handle
            return
        cfg, self._keep = step_config(g, species, dtype, mixed_mg, depth, pre, post, max_cycles,
                                      tol, eps, dt, rho_bg, sort_every, fixed, fixed_val,
                                      capacity, mg_graph, extrapolate)
        self.h = lib().psum_stepper_create(C.byref(cfg))
        if not self.h:
            raise PsumError("psum_stepper_create: " +
                            lib().psum_hip_last_error().decode(errors="replace"))
        if interp_order != 1:
            self.set_interp_order(interp_order)
        for i, sp in enumerate(species):
            self.load(i, sp["pos"], sp["vel"])

    def sort(self):
        """sort every species now (cell bin refined by velocity signs)"""
        check(lib().psum_stepper_sort(self.h), "sort")

    def set_interp_order(self, order):
        """1 = CIC (fused push + deposit), 2..4 = quadratic / cubic / quartic
        B-splines (TSI's interp_order)"""
        check(lib().psum_stepper_set_interp_order(self.h, int(order)), "set_interp_order")

    def tune(self, warmup=20, probe=10):
        """Autotune (psum_stepper_tune): dict with the calibrated fixed cycle
        count, the chosen bin shift and sort intervals ("sort_every": the
        shortest; "species_sort_every": one per species).  The shift and the
        per-species intervals are applied."""
        from ._lib import PsumBatchTuning
        t = PsumBatchTuning()
        check(lib().psum_stepper_tune(self.h, warmup, probe, C.byref(t)), "psum_stepper_tune")
        out = {f: getattr(t, f) for f, _ in PsumBatchTuning._fields_ if f != "pad_"}
        out["species_sort_every"] = [self.species_sort(q) for q in range(len(self.n))]
        return out

    def mg_tuning(self, options):
        """Multigrid kernel-shape options ("fuse 0|1", "tail 0|1", "line 0|1");
        results are bit-identical either way (psum_stepper_mg_tuning)."""
        check(lib().psum_stepper_mg_tuning(self.h, options.encode()), "mg_tuning")

    def set_smoothing(self, pre, post):
        """Smoothing sweeps per level before / after the coarse correction."""
        check(lib().psum_stepper_set_smoothing(self.h, int(pre), int(post)), "set_smoothing")

    def tune_smoothing(self, candidates=((3, 3), (4, 4), (5, 5), (6, 6)), steps=4):
        """Times `steps` steps with each (pre, post) candidate and keeps the
        fastest (psum_stepper_tune_smoothing); returns ((pre, post), {cand:
        ms per step}).  The state advances."""
        n = len(candidates)
        pre = (C.c_int * n)(*[int(c[0]) for c in candidates])
        post = (C.c_int * n)(*[int(c[1]) for c in candidates])
        ms = (C.c_double * n)()
        bp, bq = C.c_int(0), C.c_int(0)
        check(lib().psum_stepper_tune_smoothing(self.h, pre, post, n, int(steps), C.byref(bp), C.byref(bq), ms),
              "tune_smoothing")
        return (bp.value, bq.value), {tuple(c): ms[i] for i, c in enumerate(candidates)}

    def set_schedule(self, fixed_cycles=0, sort_every=0, sort_shift=0):
        """fixed_cycles > 0: that many V-cycles per solve, no tolerance check;
        sort every `sort_every` steps on 2^sort_shift-cell bins (the species
        without an interval of their own, see set_species_sort)."""
        check(lib().psum_stepper_set_schedule(self.h, fixed_cycles, sort_every, sort_shift),
              "set_schedule")

    def set_species_sort(self, s, every):
        """Species s sorts every `every` steps (> 0), or on the stepper's
        sort_every again (0).  tune() sets every species' interval."""
        check(lib().psum_stepper_set_species_sort(self.h, int(s), int(every)), "set_species_sort")

    def species_sort(self, s):
        """The interval species s sorts on (0: never)."""
        v = C.c_int()
        check(lib().psum_stepper_species_sort(self.h, int(s), C.byref(v)), "species_sort")
        return v.value

    def load(self, s, pos, vel):
        """Replace species s with pos / vel: numpy arrays, or device tensors of
        the stepper's dtype (torch CUDA tensors, copied device to device).
        The copy runs on the stepper's own stream, so torch's current stream
        is synchronised first: work queued there that writes the tensors
        (a displacement, a random draw) is complete before the copy reads
        them.  The call returns after the copy."""
        tensors = [a for a in list(pos) + list(vel) if hasattr(a, "data_ptr")]
        if tensors:
            import torch
            for dev in {t.device for t in tensors if t.is_cuda}:
                torch.cuda.current_stream(dev).synchronize()

        def addr(a):
            if hasattr(a, "data_ptr"):
                if a.dtype != _torch_dtype(self.np_dtype) or not a.is_contiguous():
                    raise PsumError("load: tensors must be contiguous and of the stepper dtype")
                return a.data_ptr(), a
            a = np.ascontiguousarray(a, self.np_dtype)
            return a.ctypes.data, a
        keep = [addr(a) for a in list(pos) + list(vel)]
        pa = (C.c_void_p * 3)(*[k[0] for k in keep[:self.dim]])
        va = (C.c_void_p * 3)(*[k[0] for k in keep[self.dim:]])
        check(lib().psum_stepper_load(self.h, s, pa, va, len(pos[0])), "psum_stepper_load")

    def forward_loss(self, steps):
        """L = field energy of phi solved from the positions after `steps`
        steps (no re-sorting); the state advances by `steps`."""
        out = C.c_double(0)
        check(lib().psum_stepper_forward_loss(self.h, steps, C.byref(out)), "forward_loss")
        return out.value

    def record(self, steps, ckpt=20):
        """Checkpointed forward for the adjoint: (L, HIP-event ms)."""
        loss, ms = C.c_double(0), C.c_double(0)
        check(lib().psum_stepper_record(self.h, steps, ckpt, C.byref(loss), C.byref(ms)), "record")
        return loss.value, ms.value

    def reverse(self, seed=1.0):
        """Reverse sweep after record(): (grad_x, grad_v, ms, adjoint V-cycles),
        grad_x[s][d] / grad_v[s][d] torch CUDA tensors of the particle dtype
        holding seed * dL/dx0, seed * dL/dv0.  Restores the recorded state."""
        import torch
        tdt = _torch_dtype(self.np_dtype)
        ns = len(self.n)
        n = [self.count(s) for s in range(ns)]
        gx = [[torch.empty(n[s], dtype=tdt, device="cuda") for _ in range(self.dim)]
              for s in range(ns)]
        gv = [[torch.empty(n[s], dtype=tdt, device="cuda") for _ in range(self.dim)]
              for s in range(ns)]
        px = (C.c_void_p * (3 * ns))()
        pv = (C.c_void_p * (3 * ns))()
        for s in range(ns):
            for d in range(self.dim):
                px[3 * s + d] = gx[s][d].data_ptr()
                pv[3 * s + d] = gv[s][d].data_ptr()
        torch.cuda.synchronize()
        ms = C.c_double(0)
        cyc = C.c_int64(0)
        check(lib().psum_stepper_reverse(self.h, float(seed), px, pv, C.byref(ms), C.byref(cyc)),
              "reverse")
        return gx, gv, ms.value, cyc.value

    def adjoint(self, steps, ckpt=20):
        """(L, grad_x, grad_v, t_record_ms, t_reverse_ms) with numpy gradients."""
        loss, t_f = self.record(steps, ckpt)
        gx, gv, t_b, _ = self.reverse(1.0)
        return (loss, [[a.cpu().numpy() for a in s] for s in gx],
                [[a.cpu().numpy() for a in s] for s in gv], t_f, t_b)

    def run(self, steps, history=True):
        hist = np.zeros((steps, 5), np.float64) if history else None
        check(lib().psum_stepper_run(self.h, steps,
                                     hist.ctypes.data_as(C.c_void_p) if history else None),
              "psum_stepper_run")
        return hist

    def phi(self):
        out = np.zeros(num_nodes(self.g), np.float64)
        check(lib().psum_stepper_phi(self.h, out.ctypes.data_as(C.c_void_p)), "psum_stepper_phi")
        return out

    def rho(self):
        """(rho of the current positions as a node-shaped FP64 array, the
        deposit's fixed-point scale 2^S)"""
        out = np.zeros(num_nodes(self.g), np.float64)
        sc = C.c_double(0)
        check(lib().psum_stepper_rho(self.h, out.ctypes.data_as(C.c_void_p), C.byref(sc)), "rho")
        return out.reshape(node_shape(self.g)), sc.value

    def count(self, s=0):
        """current particle count of species s"""
        pos = (C.c_void_p * 3)()
        vel = (C.c_void_p * 3)()
        n = C.c_int64(0)
        check(lib().psum_stepper_state(self.h, s, pos, vel, C.byref(n)), "psum_stepper_state")
        return n.value

    def state(self, s=0):
        """(pos, vel) of species s copied to host numpy arrays."""
        pos = (C.c_void_p * 3)()
        vel = (C.c_void_p * 3)()
        n = C.c_int64(0)
        check(lib().psum_stepper_state(self.h, s, pos, vel, C.byref(n)), "psum_stepper_state")
        esz = np.dtype(self.np_dtype).itemsize
        out_p, out_v = [], []
        check(lib().psum_stepper_sync(self.h), "psum_stepper_sync")
        for d in range(self.dim):
            for src, out in ((pos[d], out_p), (vel[d], out_v)):
                a = np.empty(n.value, self.np_dtype)
                _hip_memcpy_d2h(a, src, n.value * esz)
                out.append(a)
        return out_p, out_v

    def stats(self):
        cyc = C.c_int64(0)
        rel = C.c_double(0)
        cnt = (C.c_int64 * 4)()
        check(lib().psum_stepper_stats(self.h, C.byref(cyc), C.byref(rel), cnt), "stats")
        steps = C.c_int64(0)
        check(lib().psum_stepper_step_count(self.h, C.byref(steps)), "step_count")
        return {"cycles": cyc.value, "last_rel_res": rel.value, "oob": cnt[0],
                "tile_fallback_blocks": cnt[1], "absorbed": cnt[2], "acc_wraps": cnt[3], "steps": steps.value}

    def set_timing(self, on=True):
        check(lib().psum_stepper_set_timing(self.h, int(on)), "set_timing")

    def push_timing(self):
        """(total ms, launches, particles) of the fused push kernel, HIP events."""
        ms = C.c_double(0)
        nl = C.c_int64(0)
        npart = C.c_int64(0)
        check(lib().psum_stepper_push_timing(self.h, C.byref(ms), C.byref(nl), C.byref(npart)),
              "push_timing")
        return ms.value, nl.value, npart.value

    def save(self, path):
        """write a MAS checkpoint (particles, phi, step counter and solver
        state); restore() continues the run bit-identically"""
        check(lib().psum_stepper_save(self.h, os.fspath(path).encode()), "save")

    def restore(self, path):
        check(lib().psum_stepper_restore(self.h, os.fspath(path).encode()), "restore")

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                lib().psum_stepper_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _torch_dtype(np_dtype):
    import torch
    return torch.float64 if np_dtype == np.float64 else torch.float32


_hip = None


def _hip_memcpy_d2h(dst_np, src_ptr, nbytes):
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    if nbytes == 0:
        return
    rc = _hip.hipMemcpy(dst_np.ctypes.data_as(C.c_void_p), src_ptr, nbytes, 2)  # DeviceToHost
    if rc != 0:
        raise PsumError(f"hipMemcpy failed: {rc}")
