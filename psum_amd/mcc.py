"""Monte Carlo collisions on the GPU (psum_hip_mcc).

Host side of the reference's particle_collision module
(src/particle_collision/foundation.hpp:17-107, mcc_model.hpp:184-218):

  * ``channel(kind, ...)`` is one collision type of a processor: its cross
    section as a function of v^2 and its action (the ``collide`` of the
    reference's collision structs, CCP/custom_mccm.hpp:29-104);
  * ``Process`` is a collision_processor: the channels against one
    background target, with the target density (a node field or a
    constant), the incident and target masses and the target temperature;
  * ``collide(g, processes, incident, electrons, ions, dt)`` runs them in
    sequence on every valid particle of ``incident`` (sequence_executor
    inside execute_mcc_model) and appends ionization products (the
    reference's buffer_acc.push_back + clean_buffer).

Particles are ``Particles``: SoA CUDA tensors with spare capacity,
positions in the grid's dim components, three velocity components and a
uint32 random seed per particle (kept as int32 bit patterns).
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import (MCC_ELASTIC, MCC_EXCITATION, MCC_IONIZATION, MCC_ION_BACKWARD,
                   MCC_ION_ISOTROPIC, MCC_MAX_CHAN, MCC_MAX_PROC, MCC_MAX_SEG, PsumError,
                   PsumMccProcess, PsumMccSpecies, check, lib)

UNIT_CHARGE = 1.602e-19      # the reference's models (custom_mccm.hpp:22)
BOLTZMANN = 1.38065e-23      # RandV_Maxwell, rand_functions.hpp:26

KINDS = {"elastic": MCC_ELASTIC, "excitation": MCC_EXCITATION, "ionization": MCC_IONIZATION,
         "ion_isotropic": MCC_ION_ISOTROPIC, "ion_backward": MCC_ION_BACKWARD}


def channel(kind, sigma=None, *, seg_hi=(), coef=None, exp_form=True, arg_scale=1.0, threshold_ev=0.0):
    """One collision channel.  ``sigma``: a constant cross section (m^2);
    otherwise ``coef`` rows (c4, c3, c2, c1, c0) of a piecewise polynomial
    in x = ln(v2 * arg_scale), segment k while x < seg_hi[k] (len(coef) - 1
    breakpoints), exponentiated when ``exp_form`` -- the form of the
    reference's fitted cross sections (CCP/HeCCS.h)."""
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    ch = _lib.PsumMccChannel()
    ch.kind = k
    ch.threshold_ev = float(threshold_ev)
    ch.arg_scale = float(arg_scale)
    if sigma is not None:
        ch.nseg = 0
        ch.coef[0][4] = float(sigma)
        return ch
    coef = np.asarray(coef, np.float64).reshape(-1, 5)
    if not 1 <= len(coef) <= MCC_MAX_SEG or len(seg_hi) != len(coef) - 1:
        raise PsumError("channel: 1..12 segments and one breakpoint fewer")
    ch.nseg = len(coef)
    ch.exp_form = 1 if exp_form else 0
    for i, b in enumerate(seg_hi):
        ch.seg_hi[i] = float(b)
    for i, row in enumerate(coef):
        for j in range(5):
            ch.coef[i][j] = float(row[j])
    return ch


class Process:
    """collision_processor: channels of one incident species against one
    background target (foundation.hpp:67-98).  ``density``: a float (m^-3)
    or an FP64 node field of the grid (CUDA tensor)."""

    def __init__(self, channels, density, incident_mass, target_mass=0.0, target_temperature=0.0,
                 unit_charge=UNIT_CHARGE, boltzmann=BOLTZMANN):
        if not 1 <= len(channels) <= MCC_MAX_CHAN:
            raise PsumError("Process: 1..8 channels")
        self.channels = list(channels)
        self.density = density
        self.c = PsumMccProcess()
        self.c.nchan = len(channels)
        for i, ch in enumerate(channels):
            self.c.chan[i] = ch
        self.c.unit_charge = unit_charge
        self.c.incident_mass = incident_mass
        self.c.target_mass = target_mass
        self.c.target_temperature = target_temperature
        self.c.boltzmann = boltzmann
        if isinstance(density, torch.Tensor):
            if density.dtype != torch.float64 or not density.is_cuda or not density.is_contiguous():
                raise PsumError("Process: a density field must be a contiguous FP64 CUDA tensor")
            self.c.density = 0.0
        else:
            self.c.density = float(density)

    @property
    def field(self):
        return self.density if isinstance(self.density, torch.Tensor) else None


class Particles:
    """SoA particles with spare capacity for collision products."""

    def __init__(self, pos, vel, seed, n=None):
        self.pos, self.vel = list(pos), list(vel)
        self.seed = seed
        cap = self.seed.numel()
        if len(self.vel) != 3:
            raise PsumError("Particles: three velocity components")
        for t in self.pos + self.vel:
            if t.numel() != cap or t.dtype != self.pos[0].dtype or not t.is_cuda or not t.is_contiguous():
                raise PsumError("Particles: contiguous CUDA tensors of one dtype and length")
        if self.seed.dtype != torch.int32:
            raise PsumError("Particles: seeds are an int32 tensor of uint32 bit patterns")
        self.n = cap if n is None else int(n)
        if not 0 <= self.n <= cap:
            raise PsumError("Particles: n above the capacity")

    @classmethod
    def empty(cls, dim, capacity, dtype=torch.float64, device="cuda"):
        nan = float("nan")
        pos = [torch.full((capacity,), nan, dtype=dtype, device=device) for _ in range(dim)]
        vel = [torch.zeros(capacity, dtype=dtype, device=device) for _ in range(3)]
        return cls(pos, vel, torch.zeros(capacity, dtype=torch.int32, device=device), n=0)

    @property
    def capacity(self):
        return self.seed.numel()

    def reserve(self, capacity):
        """Grow the arrays to at least ``capacity`` (contents kept)."""
        if capacity <= self.capacity:
            return

        def grow(t, fill):
            out = torch.full((capacity,), fill, dtype=t.dtype, device=t.device)
            out[: t.numel()] = t
            return out
        self.pos = [grow(t, float("nan")) for t in self.pos]
        self.vel = [grow(t, 0.0) for t in self.vel]
        self.seed = grow(self.seed, 0)

    def c_struct(self):
        s = PsumMccSpecies()
        for d, t in enumerate(self.pos):
            s.pos[d] = t.data_ptr()
        for c, t in enumerate(self.vel):
            s.vel[c] = t.data_ptr()
        s.seed = self.seed.data_ptr()
        s.n = self.n
        s.cap = self.capacity
        return s


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def collide(g, processes, incident, dt, electrons=None, ions=None, outcome=False, grow=True):
    """Run the processes in order on every valid particle of ``incident``
    (execute_mcc_model + execute_all_clean_buffers).  Ionization products
    are appended to ``electrons`` (may be ``incident``) and ``ions``; with
    ``grow`` the arrays are enlarged first when the products would not fit.
    Returns (product pairs appended, outcome tensor or None)."""
    if not 1 <= len(processes) <= MCC_MAX_PROC:
        raise PsumError("collide: 1..4 processes")
    if len(incident.pos) != g.dim:
        raise PsumError(f"collide: expected {g.dim} position components")
    dtype = _lib.PSUM_F64 if incident.pos[0].dtype == torch.float64 else _lib.PSUM_F32
    if incident.pos[0].dtype not in (torch.float64, torch.float32):
        raise PsumError("collide: particles must be float64 or float32")
    for sp in (electrons, ions):
        if sp is not None and sp.pos[0].dtype != incident.pos[0].dtype:
            raise PsumError("collide: product species must have the incident dtype")
    nproc = len(processes)
    procs = (PsumMccProcess * nproc)()
    for k, p in enumerate(processes):
        procs[k] = p.c
    fields = [p.field for p in processes]
    dens = None
    if any(f is not None for f in fields):
        dens = (C.c_void_p * nproc)()
        for k, f in enumerate(fields):
            if f is not None:
                if f.numel() != _num_nodes(g):
                    raise PsumError("collide: density field size differs from the grid's node count")
                dens[k] = f.data_ptr()
    n = incident.n
    dev = incident.seed.device
    work = torch.empty(max(1, lib().psum_hip_mcc_workspace(n, nproc)), dtype=torch.uint8, device=dev)
    out = torch.empty(max(1, n * nproc), dtype=torch.int32, device=dev) if outcome else None
    made = C.c_int64(0)

    def call():
        inc = incident.c_struct()
        es = electrons.c_struct() if electrons is not None else None
        if electrons is incident:
            es = inc
        is_ = ions.c_struct() if ions is not None else None
        return lib().psum_hip_mcc(C.byref(g), dtype, procs, nproc, dens, C.byref(inc),
                                  C.byref(es) if es is not None else None,
                                  C.byref(is_) if is_ is not None else None, float(dt), C.byref(made),
                                  C.c_void_p(out.data_ptr()) if out is not None else None,
                                  C.c_void_p(work.data_ptr()), work.numel(), _stream())

    st = call()
    if st == 1 and grow and electrons is not None and ions is not None and \
            (electrons.n + made.value > electrons.capacity or ions.n + made.value > ions.capacity):
        need = made.value
        electrons.reserve(max(electrons.n + need, 2 * electrons.capacity))
        ions.reserve(max(ions.n + need, 2 * ions.capacity))
        st = call()
    check(st, "psum_hip_mcc")
    if made.value:
        electrons.n += made.value
        ions.n += made.value
    return made.value, out


def _num_nodes(g):
    n = 1
    for d in range(g.dim):
        n *= g.cells[d] + 1
    return n
