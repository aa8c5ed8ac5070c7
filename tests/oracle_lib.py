"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  The product (psum_amd) never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

KEY_INVALID = 0xFFFFFFFF


class Grid(C.Structure):
    """Same layout as orc_grid / psum_grid."""
    _fields_ = [
        ("dim", C.c_int32),
        ("cells", C.c_int32 * 3),
        ("lb", C.c_double * 3),
        ("ub", C.c_double * 3),
        ("delta", C.c_double * 3),
        ("delta_r", C.c_double * 3),
        ("periodic", C.c_int32 * 3),
        ("pad_", C.c_int32),
        ("plo", C.c_double * 3),
        ("plen", C.c_double * 3),
    ]


def make_grid(lb, ub, cells, periodic=None, plo=None, plen=None):
    """simple_grid(lb, ub, cells) semantics (simple_grid.hpp:38-51)."""
    d = len(cells)
    g = Grid()
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
    """Reference periodic node convention (TSI/main.cpp:94): nodes at
    dx/2 + i*dx, i < n per dimension, particle domain [0, n*dx)."""
    d = len(n)
    return make_grid([dx / 2] * d, [ni * dx - dx / 2 for ni in n], [ni - 1 for ni in n],
                     periodic=[1] * d, plo=[0.0] * d, plen=[ni * dx for ni in n])


def wall_grid(cells, dx=1.0):
    d = len(cells)
    return make_grid([0.0] * d, [c * dx for c in cells], cells, periodic=[0] * d)


def num_nodes(g):
    n = 1
    for i in range(g.dim):
        n *= g.cells[i] + 1
    return n


def node_shape(g):
    return tuple(g.cells[i] + 1 for i in range(g.dim))


_lib = None


MCC_KINDS = {"elastic": 0, "excitation": 1, "ionization": 2, "ion_isotropic": 3, "ion_backward": 4}


class MccChannel(C.Structure):
    """Same layout as orc_mcc_channel / psum_mcc_channel."""
    _fields_ = [
        ("kind", C.c_int32),
        ("nseg", C.c_int32),
        ("exp_form", C.c_int32),
        ("pad_", C.c_int32),
        ("arg_scale", C.c_double),
        ("threshold_ev", C.c_double),
        ("seg_hi", C.c_double * 12),
        ("coef", (C.c_double * 5) * 12),
    ]


class MccProcess(C.Structure):
    """Same layout as orc_mcc_process / psum_mcc_process."""
    _fields_ = [
        ("nchan", C.c_int32),
        ("pad_", C.c_int32),
        ("density", C.c_double),
        ("unit_charge", C.c_double),
        ("incident_mass", C.c_double),
        ("target_mass", C.c_double),
        ("target_temperature", C.c_double),
        ("boltzmann", C.c_double),
        ("chan", MccChannel * 8),
    ]


def mcc_channel(kind, sigma=None, seg_hi=(), coef=None, exp_form=True, arg_scale=1.0, threshold_ev=0.0):
    ch = MccChannel()
    ch.kind = MCC_KINDS[kind] if isinstance(kind, str) else int(kind)
    ch.arg_scale = arg_scale
    ch.threshold_ev = threshold_ev
    if sigma is not None:
        ch.coef[0][4] = sigma
        return ch
    coef = np.asarray(coef, np.float64).reshape(-1, 5)
    ch.nseg = len(coef)
    ch.exp_form = 1 if exp_form else 0
    for i, b in enumerate(seg_hi):
        ch.seg_hi[i] = b
    for i, row in enumerate(coef):
        for j in range(5):
            ch.coef[i][j] = row[j]
    return ch


def mcc_process(channels, density=0.0, incident_mass=9.10956e-31, target_mass=6.67e-27,
                target_temperature=300.0, unit_charge=1.602e-19, boltzmann=1.38065e-23):
    P = MccProcess()
    P.nchan = len(channels)
    for i, ch in enumerate(channels):
        P.chan[i] = ch
    P.density = density
    P.unit_charge = unit_charge
    P.incident_mass = incident_mass
    P.target_mass = target_mass
    P.target_temperature = target_temperature
    P.boltzmann = boltzmann
    return P


class Trace(C.Structure):
    """orc_trace"""
    _fields_ = [("I", C.c_int), ("J", C.c_int), ("K", C.c_int), ("p", C.c_double * 3), ("d", C.c_double * 3),
                ("ih", C.c_int), ("jh", C.c_int), ("kh", C.c_int), ("t", C.c_double),
                ("kx", C.c_int), ("ky", C.c_int), ("kz", C.c_int), ("bx", C.c_int), ("by", C.c_int),
                ("bz", C.c_int), ("last", C.c_int), ("finished", C.c_int)]


class Mesh(C.Structure):
    """orc_mesh"""
    _fields_ = [("I", C.c_int), ("J", C.c_int), ("K", C.c_int), ("box", C.c_double * 6), ("rd", C.c_double * 3),
                ("tri12", C.c_void_p), ("ids", C.c_void_p), ("cell_start", C.c_void_p), ("cell_tri", C.c_void_p),
                ("udf", C.c_void_p)]


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        i64 = C.c_int64
        L.orc_cell_keys_f64.argtypes = [C.POINTER(Grid), P, i64, P]
        L.orc_cell_keys_f32.argtypes = [C.POINTER(Grid), P, i64, P]
        L.orc_stable_sort.argtypes = [P, i64, C.c_uint32, P]
        L.orc_cic_weights.argtypes = [C.POINTER(Grid), P, P, P]
        L.orc_cic_grad_weights.argtypes = [C.POINTER(Grid), P, P, P]
        L.orc_num_nodes.argtypes = [C.POINTER(Grid)]
        L.orc_num_nodes.restype = i64
        for nm in ("orc_deposit_fx_f64", "orc_deposit_fx_f32"):
            getattr(L, nm).argtypes = [C.POINTER(Grid), P, i64, C.c_double, C.c_double, P]
            getattr(L, nm).restype = i64
        L.orc_deposit_ref_f64.argtypes = [C.POINTER(Grid), P, i64, C.c_double, P]
        L.orc_deposit_ref_f64.restype = i64
        L.orc_gather_push_f64.argtypes = [C.POINTER(Grid), P, P, i64, P, C.c_double, C.c_double,
                                          C.c_double, P]
        L.orc_gather_push_f64.restype = i64
        L.orc_gather_push_f32.argtypes = [C.POINTER(Grid), P, P, i64, P, C.c_float, C.c_float,
                                          C.c_double, P]
        L.orc_gather_push_f32.restype = i64
        for nm in ("orc_compact_f64", "orc_compact_f32"):
            getattr(L, nm).argtypes = [C.c_int, P, P, i64, P, P]
            getattr(L, nm).restype = i64
        L.orc_poisson_coo.argtypes = [C.POINTER(Grid), C.c_double, P, P, P, P]
        L.orc_poisson_coo.restype = i64
        L.orc_poisson_assemble_ex.argtypes = [C.POINTER(Grid), C.c_int, C.c_double, EPS_FN, P, P, P,
                                              i64, i64, P, P, P, P, i64, P, P, P, P]
        L.orc_poisson_assemble_ex.restype = C.c_int
        L.orc_poisson_rhs.argtypes = [i64, P, i64, P, P, P, P, i64, P, P, P]
        L.orc_poisson_plant.argtypes = [C.c_int]
        L.orc_bspline_1d.argtypes = [C.c_int, C.c_double, P, P]
        L.orc_bspline_weights.argtypes = [C.POINTER(Grid), C.c_int, P, P, P, P]
        L.orc_bspline_weights.restype = C.c_int
        for nm in ("orc_deposit_bs_f64", "orc_deposit_bs_f32"):
            getattr(L, nm).argtypes = [C.POINTER(Grid), C.c_int, P, i64, C.c_double, C.c_double, P]
            getattr(L, nm).restype = i64
        L.orc_deposit_bs_ref.argtypes = [C.POINTER(Grid), C.c_int, P, i64, C.c_double, P]
        L.orc_deposit_bs_ref.restype = i64
        L.orc_gather_push_bs_f64.argtypes = [C.POINTER(Grid), C.c_int, P, P, i64, P, C.c_double,
                                             C.c_double, C.c_double, P]
        L.orc_gather_push_bs_f64.restype = i64
        L.orc_gather_push_bs_f32.argtypes = [C.POINTER(Grid), C.c_int, P, P, i64, P, C.c_float,
                                             C.c_float, C.c_double, P]
        L.orc_gather_push_bs_f32.restype = i64
        L.orc_bspline_plant.argtypes = [C.c_int]
        L.orc_lcg_next.argtypes = [P]
        L.orc_lcg_next.restype = C.c_double
        L.orc_rand_spherical.argtypes = [P, P]
        L.orc_rand_maxwell.argtypes = [P, C.c_double, C.c_double, C.c_double, P]
        L.orc_mcc_sigma.argtypes = [C.POINTER(MccChannel), C.c_double]
        L.orc_mcc_sigma.restype = C.c_double
        for nm in ("orc_mcc_f64", "orc_mcc_f32"):
            getattr(L, nm).argtypes = [C.POINTER(Grid), C.POINTER(MccProcess), C.c_int, P, P, P, P, i64,
                                       C.c_double, P, P, P, P, P, P, i64, P]
            getattr(L, nm).restype = i64
        L.orc_mcc_plant.argtypes = [C.c_int]
        L.orc_seg_tri.argtypes = [P, P, P, P]
        L.orc_seg_tri.restype = C.c_int
        L.orc_trace_begin.argtypes = [C.c_int, C.c_int, C.c_int, P, P, C.POINTER(Trace)]
        L.orc_trace_begin.restype = C.c_int
        L.orc_trace_next.argtypes = [C.POINTER(Trace)]
        L.orc_trace_next.restype = C.c_int
        L.orc_box_tri_overlap.argtypes = [P, P]
        L.orc_box_tri_overlap.restype = C.c_int
        L.orc_mesh_auto_box.argtypes = [P, i64, P]
        L.orc_mesh_build.argtypes = [P, P, i64, C.c_int, C.c_int, C.c_int, P, P, P, P, P, i64]
        L.orc_mesh_build.restype = i64
        L.orc_mesh_detect.argtypes = [C.POINTER(Mesh), P, P, C.c_int, P, P, P]
        L.orc_mesh_detect.restype = C.c_int
        L.orc_material_action.argtypes = [C.c_int, P, P, P, P, P, C.c_double, C.c_double]
        for nm in ("orc_mesh_boundary_f64", "orc_mesh_boundary_f32"):
            getattr(L, nm).argtypes = [C.POINTER(Mesh), C.c_int, P, P, P, P, i64, C.c_double, C.c_double, P]
            getattr(L, nm).restype = i64
        L.orc_mesh_plant.argtypes = [C.c_int]
        L.orc_compute_udf.argtypes = [C.c_int, C.c_int, C.c_int, P, P, C.c_int, C.c_int]
        L.orc_pic_set_order.argtypes = [C.c_int]
        L.orc_pic_plant.argtypes = [C.c_int]
        L.orc_pic_set_probe.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        L.orc_mg_create.argtypes = [i64, i64, P, P, P, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int, C.c_int]
        L.orc_mg_is_singular.argtypes = [P, P]
        L.orc_mg_is_singular.restype = C.c_int
        L.orc_mg_create.restype = P
        L.orc_mg_levels.argtypes = [P, P]
        L.orc_mg_levels.restype = C.c_int
        L.orc_mg_vcycles.argtypes = [P, P, P, C.c_int]
        L.orc_mg_solve_mixed.argtypes = [P, P, P, C.c_int, C.c_int]
        L.orc_mg_residual_norm.argtypes = [P, P, P]
        L.orc_mg_residual_norm.restype = C.c_double
        L.orc_mg_free.argtypes = [P]
        L.orc_field_energy.argtypes = [C.POINTER(Grid), P, C.c_double]
        L.orc_field_energy.restype = C.c_double
        L.orc_pic_run_f64.argtypes = [C.POINTER(Grid), P, P, i64, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                      P, P, P]
        L.orc_pic_run_f64.restype = i64
        L.orc_pic_run_multi.argtypes = [C.POINTER(Grid), C.c_int, P, P, P, P, P, P, C.c_double,
                                        C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_double, C.c_int, C.c_int, P, P, P, P]
        L.orc_pic_run_multi.restype = i64
        L.orc_pic_loss.argtypes = [C.POINTER(PicProblem), P, P, P]
        L.orc_pic_loss.restype = C.c_double
        L.orc_pic_adjoint.argtypes = [C.POINTER(PicProblem), P, P, P, P]
        L.orc_pic_adjoint.restype = C.c_double
        L.orc_pic_tangent.argtypes = [C.POINTER(PicProblem), P, P, P, P, P]
        L.orc_pic_tangent.restype = C.c_double
        L.orc_pic_tangent_multi.argtypes = [C.POINTER(PicProblem), P, P, C.c_int, P, P, P, P]
        L.orc_pic_tangent_multi.restype = C.c_double
        L.orc_adjoint_plant.argtypes = [C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


MAX_SPECIES = 4


class PicProblem(C.Structure):
    _fields_ = [("g", C.POINTER(Grid)), ("nspecies", C.c_int),
                ("n", C.c_int64 * MAX_SPECIES), ("q", C.c_double * MAX_SPECIES),
                ("m", C.c_double * MAX_SPECIES), ("w", C.c_double * MAX_SPECIES),
                ("rho_bg", C.c_double), ("eps", C.c_double), ("dt", C.c_double),
                ("steps", C.c_int), ("depth", C.c_int), ("pre", C.c_int), ("post", C.c_int),
                ("max_cycles", C.c_int), ("tol", C.c_double), ("fixed", C.c_void_p),
                ("fixed_val", C.c_void_p)]


class Pic:
    """Derivatives of L = field energy after `steps` steps (orc_pic_loss /
    orc_pic_adjoint / orc_pic_tangent).  species: dicts with q, m, w and the
    FP64 pos / vel arrays of the initial state (not modified)."""

    def __init__(self, g, species, rho_bg, dt, steps, eps=1.0, depth=3, pre=2, post=1,
                 tol=1e-13, max_cycles=80, fixed=None, fixed_val=None):
        self.g = g
        self.species = species
        self.fixed = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
        self.fixed_val = None if fixed_val is None else np.ascontiguousarray(fixed_val, np.float64)
        P = PicProblem()
        P.g = C.pointer(g)
        P.nspecies = len(species)
        for i, sp in enumerate(species):
            P.n[i] = len(sp["pos"][0])
            P.q[i], P.m[i], P.w[i] = sp["q"], sp["m"], sp["w"]
        P.rho_bg, P.eps, P.dt = rho_bg, eps, dt
        P.steps, P.depth, P.pre, P.post, P.max_cycles, P.tol = steps, depth, pre, post, max_cycles, tol
        P.fixed = None if self.fixed is None else self.fixed.ctypes.data
        P.fixed_val = None if self.fixed_val is None else self.fixed_val.ctypes.data
        self.P = P

    def _arrs(self, per_species):
        keep = []
        top = (C.c_void_p * len(per_species))()
        for i, arrs in enumerate(per_species):
            arrs = [np.ascontiguousarray(a, np.float64) for a in arrs]
            pa = _ptr_array(arrs)
            keep += [pa, arrs]
            top[i] = C.cast(pa, C.c_void_p).value
        keep.append(top)
        return C.cast(top, C.c_void_p), keep

    def loss(self, pos=None, vel=None):
        """(L, key_hash); pos / vel default to the species' arrays."""
        pos = pos or [sp["pos"] for sp in self.species]
        vel = vel or [sp["vel"] for sp in self.species]
        pp, k1 = self._arrs(pos)
        vp, k2 = self._arrs(vel)
        h = C.c_uint64(0)
        L = lib().orc_pic_loss(C.byref(self.P), pp, vp, C.cast(C.pointer(h), C.c_void_p))
        return L, h.value

    def adjoint(self):
        """(L, grad_x[s][d], grad_v[s][d])"""
        pp, k1 = self._arrs([sp["pos"] for sp in self.species])
        vp, k2 = self._arrs([sp["vel"] for sp in self.species])
        gx = [[np.zeros_like(a, dtype=np.float64) for a in sp["pos"]] for sp in self.species]
        gv = [[np.zeros_like(a, dtype=np.float64) for a in sp["vel"]] for sp in self.species]
        gxp, k3 = self._arrs(gx)
        gvp, k4 = self._arrs(gv)
        L = lib().orc_pic_adjoint(C.byref(self.P), pp, vp, gxp, gvp)
        return L, gx, gv

    def tangent(self, dpos, dvel):
        """(dL, L) for the direction (dpos[s][d], dvel[s][d])"""
        pp, k1 = self._arrs([sp["pos"] for sp in self.species])
        vp, k2 = self._arrs([sp["vel"] for sp in self.species])
        dp, k3 = self._arrs(dpos)
        dv, k4 = self._arrs(dvel)
        L = C.c_double(0)
        dL = lib().orc_pic_tangent(C.byref(self.P), pp, vp, dp, dv, C.cast(C.pointer(L), C.c_void_p))
        return dL, L.value

    def tangent_multi(self, dirs):
        """(dL[j], L) for the directions dirs[j] = (dpos[s][d], dvel[s][d]),
        all along one forward trajectory"""
        pp, k1 = self._arrs([sp["pos"] for sp in self.species])
        vp, k2 = self._arrs([sp["vel"] for sp in self.species])
        nd = len(dirs)
        dpa = (C.c_void_p * nd)()
        dva = (C.c_void_p * nd)()
        keep = []
        for j, (dpos, dvel) in enumerate(dirs):
            dp, k3 = self._arrs(dpos)
            dv, k4 = self._arrs(dvel)
            keep += [dp, dv, k3, k4]
            dpa[j] = dp.value
            dva[j] = dv.value
        dL = np.zeros(nd, np.float64)
        L = C.c_double(0)
        lib().orc_pic_tangent_multi(C.byref(self.P), pp, vp, nd, C.cast(dpa, C.c_void_p),
                                    C.cast(dva, C.c_void_p), _ptr(dL),
                                    C.cast(C.pointer(L), C.c_void_p))
        return dL, L.value


def _ptr_array(arrs):
    """C array of pointers to the given numpy arrays (kept alive by caller)."""
    arr = (C.c_void_p * 3)()
    for i, a in enumerate(arrs):
        arr[i] = a.ctypes.data
    return arr


def cell_keys(g, pos):
    pos = [np.ascontiguousarray(p) for p in pos]
    n = pos[0].shape[0]
    keys = np.empty(n, np.uint32)
    f = lib().orc_cell_keys_f32 if pos[0].dtype == np.float32 else lib().orc_cell_keys_f64
    f(C.byref(g), C.cast(_ptr_array(pos), C.c_void_p), n, _ptr(keys))
    return keys


def stable_sort(keys, nkeys):
    keys = np.ascontiguousarray(keys, np.uint32)
    perm = np.empty(keys.shape[0], np.uint32)
    lib().orc_stable_sort(_ptr(keys), keys.shape[0], nkeys, _ptr(perm))
    return perm


def cic_weights(g, x):
    x = np.ascontiguousarray(x, np.float64)
    idx = np.zeros(8, np.int64)
    coef = np.zeros(8, np.float64)
    ok = lib().orc_cic_weights(C.byref(g), _ptr(x), _ptr(idx), _ptr(coef))
    s = 1 << g.dim
    return bool(ok), idx[:s], coef[:s]


def cic_grad_weights(g, x):
    x = np.ascontiguousarray(x, np.float64)
    idx = np.zeros(8, np.int64)
    gc = np.zeros(24, np.float64)
    ok = lib().orc_cic_grad_weights(C.byref(g), _ptr(x), _ptr(idx), _ptr(gc))
    s = 1 << g.dim
    return bool(ok), idx[:s], gc.reshape(3, 8)[:g.dim, :s]


def fx_shift(w):
    """Deposit fixed-point exponent: S = 40 - e with 2^(e-1) <= |w| < 2^e."""
    _, e = np.frexp(abs(w))
    return 40 - int(e)


def deposit_fx(g, pos, w, scale=None):
    pos = [np.ascontiguousarray(p) for p in pos]
    if scale is None:
        scale = float(np.ldexp(1.0, fx_shift(w)))
    acc = np.zeros(num_nodes(g), np.int64)
    f = lib().orc_deposit_fx_f32 if pos[0].dtype == np.float32 else lib().orc_deposit_fx_f64
    skipped = f(C.byref(g), C.cast(_ptr_array(pos), C.c_void_p), pos[0].shape[0], w, scale,
                _ptr(acc))
    return acc, scale, skipped


def deposit_ref(g, pos, w):
    pos = [np.ascontiguousarray(p, np.float64) for p in pos]
    rho = np.zeros(num_nodes(g), np.float64)
    lib().orc_deposit_ref_f64(C.byref(g), C.cast(_ptr_array(pos), C.c_void_p), pos[0].shape[0],
                              w, _ptr(rho))
    return rho


def gather_push(g, pos, vel, phi, qm, dt, mass=1.0):
    """In place on pos/vel (lists of contiguous arrays).  Returns (absorbed, diag)."""
    diag = np.zeros(4, np.float64)
    pa, va = _ptr_array(pos), _ptr_array(vel)
    phi = np.ascontiguousarray(phi)
    if pos[0].dtype == np.float32:
        absorbed = lib().orc_gather_push_f32(C.byref(g), C.cast(pa, C.c_void_p),
                                             C.cast(va, C.c_void_p), pos[0].shape[0],
                                             _ptr(phi), qm, dt, mass, _ptr(diag))
    else:
        absorbed = lib().orc_gather_push_f64(C.byref(g), C.cast(pa, C.c_void_p),
                                             C.cast(va, C.c_void_p), pos[0].shape[0],
                                             _ptr(phi), qm, dt, mass, _ptr(diag))
    return absorbed, diag


def compact(pos, vel):
    d = len(pos)
    dt = pos[0].dtype
    n = pos[0].shape[0]
    op = [np.empty(n, dt) for _ in range(d)]
    ov = [np.empty(n, dt) for _ in range(d)]
    f = lib().orc_compact_f32 if dt == np.float32 else lib().orc_compact_f64
    m = f(d, C.cast(_ptr_array(pos), C.c_void_p), C.cast(_ptr_array(vel), C.c_void_p), n,
          C.cast(_ptr_array(op), C.c_void_p), C.cast(_ptr_array(ov), C.c_void_p))
    return [a[:m].copy() for a in op], [a[:m].copy() for a in ov]


def poisson_coo(g, eps, fixed):
    fixed = np.ascontiguousarray(fixed, np.uint8)
    nnz = lib().orc_poisson_coo(C.byref(g), eps, _ptr(fixed), None, None, None)
    r = np.empty(nnz, np.uint64)
    c = np.empty(nnz, np.uint64)
    v = np.empty(nnz, np.float64)
    lib().orc_poisson_coo(C.byref(g), eps, _ptr(fixed), _ptr(r), _ptr(c), _ptr(v))
    return r, c, v


EPS_FN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_void_p)


class MixedNode(C.Structure):
    _fields_ = [("element", C.c_uint64), ("direction", C.c_int32), ("pad_", C.c_int32),
                ("value_coef", C.c_double), ("grad_coef", C.c_double)]


# direction codes (2D: N S E W; 3D: Xpos Xneg Ypos Yneg Zpos Zneg)
DIR2 = {"N": 0, "S": 1, "E": 2, "W": 3}
DIR3 = {"Xpos": 0, "Xneg": 1, "Ypos": 2, "Yneg": 3, "Zpos": 4, "Zneg": 5}


def mixed_array(nodes):
    """nodes: iterable of (element, direction, value_coef, grad_coef)"""
    nodes = list(nodes)
    arr = (MixedNode * max(1, len(nodes)))()
    for k, (e, d, v, gc) in enumerate(nodes):
        arr[k].element, arr[k].direction, arr[k].value_coef, arr[k].grad_coef = e, d, v, gc
    return arr, len(nodes)


def poisson_assemble_ex(g, eps=1.0, fixed=None, mixed=(), cylindrical=False, eps_fn=None):
    """General assembly: returns (rows, cols, vals) of A' in column-major
    order and (mrows, mcols, mvals) of M row-major."""
    n = num_nodes(g)
    fx = np.zeros(n, np.uint8) if fixed is None else np.ascontiguousarray(fixed, np.uint8)
    arr, nm = mixed_array(mixed)
    cb = EPS_FN(lambda p, u: eps_fn(p)) if eps_fn is not None else EPS_FN()
    nnz, mnnz = C.c_int64(0), C.c_int64(0)
    rc = lib().orc_poisson_assemble_ex(C.byref(g), int(cylindrical), eps, cb, None, _ptr(fx), arr, nm,
                                       0, None, None, None, C.byref(nnz), 0, None, None, None,
                                       C.byref(mnnz))
    if rc != 0:
        raise ValueError(f"orc_poisson_assemble_ex: {rc}")
    r, c, v = (np.empty(nnz.value, np.uint64), np.empty(nnz.value, np.uint64),
               np.empty(nnz.value, np.float64))
    mr, mc, mv = (np.empty(mnnz.value, np.uint64), np.empty(mnnz.value, np.uint64),
                  np.empty(mnnz.value, np.float64))
    rc = lib().orc_poisson_assemble_ex(C.byref(g), int(cylindrical), eps, cb, None, _ptr(fx), arr, nm,
                                       nnz.value, _ptr(r), _ptr(c), _ptr(v), C.byref(nnz),
                                       mnnz.value, _ptr(mr), _ptr(mc), _ptr(mv), C.byref(mnnz))
    if rc != 0:
        raise ValueError(f"orc_poisson_assemble_ex: {rc}")
    return (r, c, v), (mr, mc, mv)


def poisson_rhs(source, M, mixed_values, fix_idx=(), fix_val=()):
    mr, mc, mv = M
    src = np.ascontiguousarray(source, np.float64)
    mvv = np.ascontiguousarray(mixed_values, np.float64)
    fi = np.ascontiguousarray(fix_idx, np.uint64)
    fv = np.ascontiguousarray(fix_val, np.float64)
    b = np.empty_like(src)
    lib().orc_poisson_rhs(src.size, _ptr(src), mr.size, _ptr(mr), _ptr(mc), _ptr(mv), _ptr(mvv),
                          fi.size, _ptr(fi), _ptr(fv), _ptr(b))
    return b


def bspline_1d(order, s):
    c = np.zeros(5)
    dc = np.zeros(5)
    lib().orc_bspline_1d(order, s, _ptr(c), _ptr(dc))
    return c[:order + 1], dc[:order + 1]


def bspline_weights(g, order, x):
    """(idx, coef, gcoef[D, S]) of the support of x, or None out of range"""
    xx = np.zeros(3)
    xx[:len(x)] = x
    idx = np.zeros(125, np.int64)
    coef = np.zeros(125)
    gc = np.zeros(375)
    S = lib().orc_bspline_weights(C.byref(g), order, _ptr(xx), _ptr(idx), _ptr(coef), _ptr(gc))
    if S == 0:
        return None
    return idx[:S].copy(), coef[:S].copy(), gc.reshape(3, 125)[:g.dim, :S].copy()


def deposit_bs(g, order, pos, w, scale):
    dt = pos[0].dtype
    pos = [np.ascontiguousarray(a) for a in pos]
    acc = np.zeros(num_nodes(g), np.int64)
    f = lib().orc_deposit_bs_f32 if dt == np.float32 else lib().orc_deposit_bs_f64
    skipped = f(C.byref(g), order, C.cast(_ptr_array(pos), C.c_void_p), pos[0].size, w, scale, _ptr(acc))
    return acc, skipped


def deposit_bs_ref(g, order, pos, w):
    pos = [np.ascontiguousarray(a, np.float64) for a in pos]
    rho = np.zeros(num_nodes(g))
    lib().orc_deposit_bs_ref(C.byref(g), order, C.cast(_ptr_array(pos), C.c_void_p), pos[0].size, w, _ptr(rho))
    return rho


def gather_push_bs(g, order, pos, vel, phi, qm, dt, mass=1.0):
    """in place on copies; returns (pos, vel, absorbed, diag)"""
    pos = [np.array(a) for a in pos]
    vel = [np.array(a) for a in vel]
    t = pos[0].dtype
    diag = np.zeros(4)
    phi = np.ascontiguousarray(phi, t)
    f = lib().orc_gather_push_bs_f32 if t == np.float32 else lib().orc_gather_push_bs_f64
    ab = f(C.byref(g), order, C.cast(_ptr_array(pos), C.c_void_p), C.cast(_ptr_array(vel), C.c_void_p),
           pos[0].size, _ptr(phi), qm, dt, mass, _ptr(diag))
    return pos, vel, ab, diag


def pic_set_order(order):
    lib().orc_pic_set_order(order)


def pic_plant(which):
    """Plant a mistake in the whole step (1 = deposit charge sign, 2 = rho_bg
    sign, 3 = push with the previous step's phi, 4 = no periodic
    neutralisation; 0 = none)."""
    lib().orc_pic_plant(which)


class pic_probe:
    """Context manager: while active, every step of orc_pic_run_* records
    sum_s q_s w_s sum_p exp(i k.x_p) for each wave vector in kvecs (shape
    [nk, D]) over the positions the step's solve sees.  After a run,
    .modes(steps) returns a complex array [steps, nk]."""

    def __init__(self, kvecs, max_steps):
        self.k = np.ascontiguousarray(np.atleast_2d(kvecs), np.float64)
        self.out = np.zeros((max_steps, self.k.shape[0], 2), np.float64)

    def __enter__(self):
        lib().orc_pic_set_probe(self.k.shape[0], self.k.ctypes.data, self.out.ctypes.data)
        return self

    def __exit__(self, *exc):
        lib().orc_pic_set_probe(0, None, None)

    def modes(self, steps):
        o = self.out[:steps]
        return o[..., 0] + 1j * o[..., 1]


def bspline_plant(which):
    lib().orc_bspline_plant(which)


def poisson_plant(which):
    lib().orc_poisson_plant(which)


class MG:
    """Reference-restated Galerkin multigrid (FP64 and the 32f mixed form)."""

    def __init__(self, n, rows, cols, vals, I, J, K=1, depth=3, pre=2, post=1, mode=1):
        self.rows = np.ascontiguousarray(rows, np.uint64)
        self.cols = np.ascontiguousarray(cols, np.uint64)
        self.vals = np.ascontiguousarray(vals, np.float64)
        self.n = n
        self.h = lib().orc_mg_create(n, self.vals.shape[0], _ptr(self.rows), _ptr(self.cols),
                                     _ptr(self.vals), I, J, K, depth, pre, post, mode)
        if not self.h:
            raise ValueError("orc_mg_create failed")

    def singular(self):
        pin = np.zeros(1, np.int64)
        s = lib().orc_mg_is_singular(self.h, _ptr(pin))
        return bool(s), int(pin[0])

    def levels(self):
        sizes = np.zeros(64, np.int64)
        nl = lib().orc_mg_levels(self.h, _ptr(sizes))
        return list(sizes[:nl])

    def vcycles(self, b, x, cycles):
        b = np.ascontiguousarray(b, np.float64)
        assert x.dtype == np.float64 and x.flags.c_contiguous
        lib().orc_mg_vcycles(self.h, _ptr(b), _ptr(x), cycles)
        return x

    def solve_mixed(self, b, x, max_cycles, corr_interval=4):
        b = np.ascontiguousarray(b, np.float64)
        lib().orc_mg_solve_mixed(self.h, _ptr(b), _ptr(x), max_cycles, corr_interval)
        return x

    def residual_norm(self, b, x):
        return lib().orc_mg_residual_norm(self.h, _ptr(np.ascontiguousarray(b, np.float64)),
                                          _ptr(np.ascontiguousarray(x, np.float64)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mg_free(self.h)
            self.h = None


def field_energy(g, phi, eps=1.0):
    phi = np.ascontiguousarray(phi, np.float64)
    return lib().orc_field_energy(C.byref(g), _ptr(phi), eps)


def pic_run(g, pos, vel, q, m, w, rho_bg, eps, dt, steps, depth, pre=2, post=1, tol=1e-10,
            max_cycles=20, dep_mode=0, fixed=None):
    """In place on pos/vel (float64).  Returns (hist[steps,5], phi, total_cycles)."""
    hist = np.zeros((steps, 5), np.float64)
    phi = np.zeros(num_nodes(g), np.float64)
    pa, va = _ptr_array(pos), _ptr_array(vel)
    fx = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
    cyc = lib().orc_pic_run_f64(C.byref(g), C.cast(pa, C.c_void_p), C.cast(va, C.c_void_p),
                                pos[0].shape[0], q, m, w, rho_bg, eps, dt, steps, depth, pre, post,
                                tol, max_cycles, dep_mode, None if fx is None else _ptr(fx),
                                _ptr(hist), _ptr(phi))
    if cyc < 0:
        raise RuntimeError("orc_pic_run_f64 failed")
    return hist, phi, cyc


def pic_run_multi(g, species, rho_bg, eps, dt, steps, depth, pre=2, post=1, tol=1e-10,
                  max_cycles=20, dep_mode=0, fixed=None, fixed_val=None):
    """species: list of dicts (q, m, w, pos, vel) with float64 arrays, updated
    in place.  Returns (hist[steps,5], phi, total_cycles)."""
    ns = len(species)
    keep = []
    parr = (C.c_void_p * ns)()
    varr = (C.c_void_p * ns)()
    for i, sp in enumerate(species):
        pa, va = _ptr_array(sp["pos"]), _ptr_array(sp["vel"])
        keep += [pa, va]
        parr[i] = C.cast(pa, C.c_void_p).value
        varr[i] = C.cast(va, C.c_void_p).value
    n = np.array([len(sp["pos"][0]) for sp in species], np.int64)
    q = np.array([sp["q"] for sp in species], np.float64)
    m = np.array([sp["m"] for sp in species], np.float64)
    w = np.array([sp["w"] for sp in species], np.float64)
    hist = np.zeros((steps, 5), np.float64)
    phi = np.zeros(num_nodes(g), np.float64)
    fx = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
    fv = None if fixed_val is None else np.ascontiguousarray(fixed_val, np.float64)
    cyc = lib().orc_pic_run_multi(C.byref(g), ns, C.cast(parr, C.c_void_p),
                                  C.cast(varr, C.c_void_p), _ptr(n), _ptr(q), _ptr(m), _ptr(w),
                                  rho_bg, eps, dt, steps, depth, pre, post, tol, max_cycles,
                                  dep_mode, None if fx is None else _ptr(fx),
                                  None if fv is None else _ptr(fv), _ptr(hist), _ptr(phi))
    if cyc < 0:
        raise RuntimeError("orc_pic_run_multi failed")
    return hist, phi, cyc


def lcg_draws(seed, k):
    """k successive R() values from seed; returns (values, final state)"""
    x = C.c_uint32(seed)
    out = [lib().orc_lcg_next(C.byref(x)) for _ in range(k)]
    return np.array(out), x.value


def rand_spherical(seed):
    x = C.c_uint32(seed)
    o = np.zeros(3)
    lib().orc_rand_spherical(C.byref(x), _ptr(o))
    return o, x.value


def rand_maxwell(seed, temperature, mass, boltzmann=1.38065e-23):
    x = C.c_uint32(seed)
    o = np.zeros(3)
    lib().orc_rand_maxwell(C.byref(x), temperature, mass, boltzmann, _ptr(o))
    return o, x.value


def mcc_sigma(ch, v2):
    return lib().orc_mcc_sigma(C.byref(ch), v2)


def mcc(g, procs, pos, vel, seed, dt, density=None, cap=None):
    """Runs the processes on copies; returns dict(pos, vel, seed, made,
    e_pos, e_vel, e_seed, i_pos, i_vel, i_seed, outcome[n, nproc])."""
    t = pos[0].dtype
    pos = [np.array(a, t) for a in pos]
    vel = [np.array(a, t) for a in vel]
    seed = np.array(seed, np.uint32)
    n = seed.size
    nproc = len(procs)
    arr = (MccProcess * nproc)(*procs)
    cap = n * nproc if cap is None else cap
    D = g.dim
    ep = [np.zeros(max(cap, 1), t) for _ in range(D)]
    ev = [np.zeros(max(cap, 1), t) for _ in range(3)]
    es = np.zeros(max(cap, 1), np.uint32)
    ip = [np.zeros(max(cap, 1), t) for _ in range(D)]
    iv = [np.zeros(max(cap, 1), t) for _ in range(3)]
    is_ = np.zeros(max(cap, 1), np.uint32)
    outc = np.zeros(max(n * nproc, 1), np.int32)
    dens_arr = None
    keep = []
    if density is not None:
        dens_arr = (C.c_void_p * nproc)()
        for k, f in enumerate(density):
            if f is not None:
                f = np.ascontiguousarray(f, np.float64)
                keep.append(f)
                dens_arr[k] = f.ctypes.data
    f = lib().orc_mcc_f32 if t == np.float32 else lib().orc_mcc_f64
    made = f(C.byref(g), arr, nproc, C.cast(dens_arr, C.c_void_p) if dens_arr is not None else None,
             C.cast(_ptr_array(pos), C.c_void_p), C.cast(_ptr_array(vel), C.c_void_p), _ptr(seed), n, dt,
             C.cast(_ptr_array(ep), C.c_void_p), C.cast(_ptr_array(ev), C.c_void_p), _ptr(es),
             C.cast(_ptr_array(ip), C.c_void_p), C.cast(_ptr_array(iv), C.c_void_p), _ptr(is_), cap,
             _ptr(outc))
    m = min(made, cap)
    return dict(pos=pos, vel=vel, seed=seed, made=made,
                e_pos=[a[:m] for a in ep], e_vel=[a[:m] for a in ev], e_seed=es[:m],
                i_pos=[a[:m] for a in ip], i_vel=[a[:m] for a in iv], i_seed=is_[:m],
                outcome=outc[:n * nproc].reshape(n, nproc))


def mcc_plant(which):
    lib().orc_mcc_plant(which)


MAT = {"diffuse": 0, "specular": 1, "maxwell_reflect": 2, "absorb": 3}


def seg_tri(tri, p1, p2):
    """(hit, lambda)"""
    t = np.ascontiguousarray(tri, np.float64).ravel()
    a = np.ascontiguousarray(p1, np.float64)
    b = np.ascontiguousarray(p2, np.float64)
    lam = np.zeros(1)
    h = lib().orc_seg_tri(_ptr(t), _ptr(a), _ptr(b), _ptr(lam))
    return bool(h), lam[0]


def trace_cells(I, J, K, a, b):
    """cell indices visited by the traversal of a->b (grid coordinates)"""
    s = Trace()
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    out = []
    idx = lib().orc_trace_begin(I, J, K, _ptr(a), _ptr(b), C.byref(s))
    while not s.finished:
        out.append(idx)
        idx = lib().orc_trace_next(C.byref(s))
    return out


def box_tri_overlap(tri, box):
    t = np.ascontiguousarray(tri, np.float64).ravel()
    bx = np.ascontiguousarray(box, np.float64)
    return bool(lib().orc_box_tri_overlap(_ptr(t), _ptr(bx)))


def compute_udf(occupied, udf_k=6, max_loop=200):
    """orc_compute_udf of an (I, J, K) occupancy array"""
    occ = np.ascontiguousarray(occupied, np.uint8)
    I, J, K = occ.shape
    out = np.zeros(occ.size, np.float32)
    lib().orc_compute_udf(I, J, K, _ptr(occ), _ptr(out), udf_k, max_loop)
    return out


class OracleMesh:
    """triangle_mesh_trigger built by the oracle; keeps its arrays alive"""

    def __init__(self, tri, ids, I, J, K, box=None):
        self.tri = np.ascontiguousarray(tri, np.float64).reshape(-1, 9)
        self.ids = np.ascontiguousarray(ids, np.int32)
        n = len(self.tri)
        if box is None:
            box = np.zeros(6)
            lib().orc_mesh_auto_box(_ptr(self.tri), n, _ptr(box))
        self.box = np.ascontiguousarray(box, np.float64)
        self.tri12 = np.zeros((n, 12))
        self.kept = np.zeros(n, np.int32)
        ncell = I * J * K
        self.cell_start = np.zeros(ncell + 1, np.uint32)
        total = lib().orc_mesh_build(_ptr(self.tri), _ptr(self.ids), n, I, J, K, _ptr(self.box), _ptr(self.tri12),
                                     _ptr(self.kept), _ptr(self.cell_start), None, 0)
        self.cell_tri = np.zeros(max(total, 1), np.uint32)
        lib().orc_mesh_build(_ptr(self.tri), _ptr(self.ids), n, I, J, K, _ptr(self.box), _ptr(self.tri12),
                             _ptr(self.kept), _ptr(self.cell_start), _ptr(self.cell_tri), total)
        self.cell_tri = self.cell_tri[:total]
        m = Mesh()
        m.I, m.J, m.K = I, J, K
        for c in range(6):
            m.box[c] = self.box[c]
        for c in range(3):
            m.rd[c] = (I, J, K)[c] / (self.box[3 + c] - self.box[c])
        m.tri12 = self.tri12.ctypes.data
        m.ids = self.ids.ctypes.data
        m.cell_start = self.cell_start.ctypes.data
        m.cell_tri = self.cell_tri.ctypes.data if total else None
        m.udf = None
        self.c = m
        self.shape = (I, J, K)

    def set_udf(self, enabled=True):
        """triangle_mesh_trigger::set_acceleration_UDF: the distance field of
        the occupied cells (reference defaults) skips empty space in the
        full traversal"""
        if enabled:
            occ = (np.diff(self.cell_start.astype(np.int64)) > 0).astype(np.uint8)
            self.udf = compute_udf(occ.reshape(self.shape))
            self.c.udf = self.udf.ctypes.data
        else:
            self.c.udf = None

    def detect(self, p1, p2, fast=False):
        """(hit, k, normal, id)"""
        a = np.ascontiguousarray(p1, np.float64)
        b = np.ascontiguousarray(p2, np.float64)
        k = np.zeros(1)
        nrm = np.zeros(3)
        idv = np.zeros(1, np.int32)
        h = lib().orc_mesh_detect(C.byref(self.c), _ptr(a), _ptr(b), int(fast), _ptr(k), _ptr(nrm), _ptr(idv))
        return bool(h), k[0], nrm, int(idv[0])

    def boundary(self, old, pos, vel, seed, temperature=273.15, mass=9.11e-31, fast=False):
        """router.deal on copies; returns dict(pos, vel, seed, hits, hit_id)"""
        t = pos[0].dtype
        old = [np.ascontiguousarray(a, t) for a in old]
        pos = [np.array(a, t) for a in pos]
        vel = [np.array(a, t) for a in vel]
        seed = np.array(seed, np.uint32)
        n = seed.size
        hid = np.zeros(max(n, 1), np.int32)
        f = lib().orc_mesh_boundary_f32 if t == np.float32 else lib().orc_mesh_boundary_f64
        hits = f(C.byref(self.c), int(fast), C.cast(_ptr_array(old), C.c_void_p), C.cast(_ptr_array(pos), C.c_void_p),
                 C.cast(_ptr_array(vel), C.c_void_p), _ptr(seed), n, temperature, mass, _ptr(hid))
        return dict(pos=pos, vel=vel, seed=seed, hits=hits, hit_id=hid[:n])


def material_action(mat, pos, vel, seed, hit, nrm, temperature=273.15, mass=9.11e-31):
    p = np.array(pos, np.float64)
    v = np.array(vel, np.float64)
    s = C.c_uint32(seed)
    h = np.ascontiguousarray(hit, np.float64)
    nn = np.ascontiguousarray(nrm, np.float64)
    lib().orc_material_action(mat, _ptr(p), _ptr(v), C.byref(s), _ptr(h), _ptr(nn), temperature, mass)
    return p, v, s.value


def mesh_plant(which):
    lib().orc_mesh_plant(which)
