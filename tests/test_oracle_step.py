"""Pins of the oracle's whole PIC step (orc_pic_run_multi, oracle/step.c)
by physics and closed forms that do not reuse the oracle's own code:

- one step against an independent numpy restatement of the reference's
  step (CIC deposit, FD Poisson solved directly or by FFT, gradient of the
  bilinear potential, kick then drift; TSI/main.cpp:59-75, 121-127,
  big_bang/stage_6.cpp:33-65) on periodic non-neutral and wall-bounded
  two-species problems in 2D and 3D;
- a quiet-start neutral plasma between grounded walls has no field;
- the Langmuir oscillation of the plain FP64 128x128, 64 e-/cell, dt 0.1
  workload at its full size: frequency of the k = 2 pi/L density mode vs
  omega_p sqrt(1 + 3 k^2 lambda_D^2) with the leapfrog dispersion
  sin(omega dt/2) = omega_0 dt/2;
- total-energy conservation over those 100 steps (TSI/main.cpp:130-146);
- the linear growth rate of a seeded two-stream mode vs the root of the
  warm two-beam dispersion relation (TSI/README.md: "energy grows at the
  linear-theory rate").

Each planted mistake (deposit charge sign, rho_bg sign, the push using the
previous step's potential, no periodic neutralisation) must fail at least
one of these checks.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy.special import wofz

import oracle_lib as O


# --------------------------------------------------------------------------
# independent numpy restatement of one reference step


def _support(lb, dx, nnode, periodic, x):
    """Base node index and fraction per dimension (simple_grid cN, gen_interp)."""
    s = (x - lb) / dx
    c = np.floor(s).astype(np.int64)
    f = s - c
    return c, f


def _stencil_nodes(c, off, nnode, periodic):
    i = c + off
    if periodic:
        i = np.mod(i, nnode)
    return i


def np_deposit(shape, lb, dx, periodic, species):
    D = len(shape)
    rho = np.zeros(shape)
    dV = float(np.prod(dx))
    for s in species:
        cs, fs = zip(*[_support(lb[d], dx[d], shape[d], periodic, s["pos"][d]) for d in range(D)])
        for corner in range(1 << D):
            offs = [(corner >> (D - 1 - d)) & 1 for d in range(D)]
            wgt = np.full(len(s["pos"][0]), s["q"] * s["w"] / dV)
            idx = []
            for d in range(D):
                wgt = wgt * (fs[d] if offs[d] else 1 - fs[d])
                idx.append(_stencil_nodes(cs[d], offs[d], shape[d], periodic))
            np.add.at(rho, tuple(idx), wgt)
    return rho


def np_solve_periodic(rho, dx):
    """-Lap_h phi = rho - mean(rho) on a periodic node grid (5/7-point),
    gauge phi[0] = 0 (the pinned node)."""
    r = rho - rho.mean()
    lam = np.zeros(rho.shape)
    for d, n in enumerate(rho.shape):
        m = np.arange(n)
        sh = [1] * rho.ndim
        sh[d] = n
        lam = lam + ((2 - 2 * np.cos(2 * np.pi * m / n)) / dx[d] ** 2).reshape(sh)
    rh = np.fft.fftn(r)
    lam.flat[0] = 1.0
    ph = rh / lam
    ph.flat[0] = 0.0
    phi = np.real(np.fft.ifftn(ph))
    return phi - phi.flat[0]


def np_solve_walls(rho, dx):
    """-Lap_h phi = rho at interior nodes, phi = 0 on every boundary node."""
    shape = rho.shape
    N = rho.size
    lin = np.arange(N).reshape(shape)
    interior = np.ones(shape, bool)
    for d in range(rho.ndim):
        sl = [slice(None)] * rho.ndim
        sl[d] = 0
        interior[tuple(sl)] = False
        sl[d] = -1
        interior[tuple(sl)] = False
    rows, cols, vals = [], [], []
    ii = lin[interior]
    diag = np.zeros(ii.size)
    for d in range(rho.ndim):
        for step in (-1, 1):
            nb = np.roll(lin, -step, axis=d)[interior]
            rows.append(ii)
            cols.append(nb)
            vals.append(np.full(ii.size, -1.0 / dx[d] ** 2))
            diag += 1.0 / dx[d] ** 2
    bnd = lin[~interior]
    A = sp.coo_matrix((np.concatenate(vals + [diag, np.ones(bnd.size)]),
                       (np.concatenate(rows + [ii, bnd]), np.concatenate(cols + [ii, bnd]))),
                      shape=(N, N)).tocsr()
    b = np.where(interior, rho, 0.0).ravel()
    return spla.spsolve(A, b).reshape(shape)


def np_gather_grad(phi, lb, dx, periodic, pos):
    """grad of the multilinear interpolant of phi at the particles
    (gen_interp_diff, field_interp.hpp:93-113)."""
    D = phi.ndim
    cs, fs = zip(*[_support(lb[d], dx[d], phi.shape[d], periodic, pos[d]) for d in range(D)])
    g = [np.zeros(len(pos[0])) for _ in range(D)]
    for corner in range(1 << D):
        offs = [(corner >> (D - 1 - d)) & 1 for d in range(D)]
        idx = tuple(_stencil_nodes(cs[d], offs[d], phi.shape[d], periodic) for d in range(D))
        val = phi[idx]
        for e in range(D):
            w = np.ones(len(pos[0]))
            for d in range(D):
                if d == e:
                    w = w * ((1.0 if offs[d] else -1.0) / dx[d])
                else:
                    w = w * (fs[d] if offs[d] else 1 - fs[d])
            g[e] += val * w
    return g


def np_one_step(shape, lb, dx, periodic, species, rho_bg, dt):
    rho = np_deposit(shape, lb, dx, periodic, species) + rho_bg
    phi = np_solve_periodic(rho, dx) if periodic else np_solve_walls(rho, dx)
    out = []
    for s in species:
        g = np_gather_grad(phi, lb, dx, periodic, s["pos"])
        v1 = [s["vel"][d] - s["q"] / s["m"] * dt * g[d] for d in range(len(shape))]
        x1 = [s["pos"][d] + v1[d] * dt for d in range(len(shape))]
        out.append((x1, v1))
    return phi, out


def _copy_species(species):
    return [dict(q=s["q"], m=s["m"], w=s["w"], pos=[a.copy() for a in s["pos"]],
                 vel=[a.copy() for a in s["vel"]]) for s in species]


def _oracle_one_step(g, species, rho_bg, dt, fixed=None, depth=3):
    sp_ = _copy_species(species)
    h, phi, _ = O.pic_run_multi(g, sp_, rho_bg, 1.0, dt, 1, depth, tol=1e-13, max_cycles=200,
                                fixed=fixed)
    return phi, sp_


def _blob_species(rng, D, L, n, q, m, w, vth, centre, width):
    pos = [np.mod(centre[d] + width * rng.standard_normal(n), L[d]) for d in range(D)]
    vel = [vth * rng.standard_normal(n) for _ in range(D)]
    return dict(q=q, m=m, w=w, pos=pos, vel=vel)


def _periodic_case(D):
    rng = np.random.default_rng(41 + D)
    n = [24, 20, 16][:D]
    L = [float(v) for v in n]
    species = [_blob_species(rng, D, L, 6000, -1.0, 1.0, 0.05, 1.0, [0.4 * v for v in L], 3.0),
               _blob_species(rng, D, L, 3000, 1.0, 100.0, 0.07, 0.1, [0.6 * v for v in L], 2.0)]
    g = O.periodic_grid(n)
    lb = [0.5] * D
    return g, n, lb, species


def _walls_case(D):
    rng = np.random.default_rng(7 + D)
    cells = [20, 16, 12][:D]
    L = [float(c) for c in cells]
    species = [_blob_species(rng, D, L, 5000, -1.0, 1.0, 0.05, 0.5, [0.45 * v for v in L], 2.0),
               _blob_species(rng, D, L, 4000, 1.0, 40.0, 0.04, 0.05, [0.55 * v for v in L], 2.5)]
    for s in species:   # keep every particle strictly inside the box
        for d in range(D):
            s["pos"][d] = np.clip(s["pos"][d], 1.0, L[d] - 1.0)
    g = O.wall_grid(cells)
    shape = tuple(c + 1 for c in cells)
    fixed = np.zeros(shape, np.uint8)
    for d in range(D):
        sl = [slice(None)] * D
        sl[d] = 0
        fixed[tuple(sl)] = 1
        sl[d] = -1
        fixed[tuple(sl)] = 1
    return g, shape, [0.0] * D, species, fixed.ravel()


def _compare_one_step(g, shape, lb, periodic, species, rho_bg, dt, fixed=None):
    D = len(shape)
    phi_np, want = np_one_step(tuple(shape), lb, [1.0] * D, periodic, species, rho_bg, dt)
    phi_o, got = _oracle_one_step(g, species, rho_bg, dt, fixed=fixed)
    phi_o = phi_o.reshape(shape)
    scale_phi = np.abs(phi_np).max()
    err_phi = np.abs(phi_o - phi_np).max() / scale_phi
    err_v = 0.0
    for s0, (x1, v1), s1 in zip(species, want, got):
        dv = max(np.abs(v1[d] - s0["vel"][d]).max() for d in range(D))
        for d in range(D):
            err_v = max(err_v, np.abs(s1["vel"][d] - v1[d]).max() / dv)
    return err_phi, err_v


CASES = [("periodic", 2), ("periodic", 3), ("walls", 2), ("walls", 3)]


def _run_case(kind, D):
    if kind == "periodic":
        g, n, lb, species = _periodic_case(D)
        # non-neutral on purpose: the background does not cancel the particles
        return _compare_one_step(g, n, lb, True, species, 0.3, 0.1)
    g, shape, lb, species, fixed = _walls_case(D)
    return _compare_one_step(g, shape, lb, False, species, 0.2, 0.1, fixed=fixed)


@pytest.mark.parametrize("kind,D", CASES)
def test_one_step_matches_numpy_restatement(kind, D):
    err_phi, err_v = _run_case(kind, D)
    # the oracle's solve stops at a relative residual of 1e-13
    assert err_phi < 1e-9, err_phi
    assert err_v < 1e-8, err_v


def test_quiet_neutral_box_has_no_field():
    """Electrons on a regular sub-lattice (4 x 4 per cell) over a uniform
    +1 background between grounded walls: every interior node collects
    exactly -1 + 1, so phi is 0 everywhere and no particle is kicked."""
    cells = [16, 12]
    sub = (np.arange(4) + 0.5) / 4
    xs = (np.arange(cells[0])[:, None] + sub[None, :]).ravel()
    ys = (np.arange(cells[1])[:, None] + sub[None, :]).ravel()
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    rng = np.random.default_rng(3)
    n = X.size
    sp_ = [dict(q=-1.0, m=1.0, w=1.0 / 16, pos=[X.ravel().copy(), Y.ravel().copy()],
                vel=[0.01 * rng.standard_normal(n), 0.01 * rng.standard_normal(n)])]
    v0 = [a.copy() for a in sp_[0]["vel"]]
    shape = (cells[0] + 1, cells[1] + 1)
    fixed = np.zeros(shape, np.uint8)
    fixed[0, :] = fixed[-1, :] = fixed[:, 0] = fixed[:, -1] = 1
    h, phi, _ = O.pic_run_multi(O.wall_grid(cells), sp_, 1.0, 1.0, 0.1, 1, 2, tol=1e-13,
                                max_cycles=50, fixed=fixed.ravel())
    assert np.abs(phi).max() == 0.0
    assert h[0, 0] == 0.0
    for d in range(2):
        assert np.array_equal(sp_[0]["vel"][d], v0[d])


# --------------------------------------------------------------------------
# Langmuir oscillation and energy on the plain FP64 workload at full size

C1_N, C1_PPC, C1_DT, C1_STEPS = 128, 64, 0.1, 100


def _c1_run(steps=C1_STEPS):
    from psum_amd.workloads import langmuir_2d
    wl = langmuir_2d(n=C1_N, ppc=C1_PPC, seed=1)
    s = wl["species"][0]
    k = 2 * np.pi / C1_N
    with O.pic_probe([[k, 0.0]], steps) as pr:
        h, _, _ = O.pic_run(O.periodic_grid([C1_N, C1_N]), [a.copy() for a in s["pos"]],
                            [a.copy() for a in s["vel"]], s["q"], s["m"], s["w"], wl["rho_bg"],
                            1.0, C1_DT, steps, depth=4, tol=1e-10, max_cycles=40)
    return h, pr.modes(steps)[:, 0]


def _fit_frequency(y, t, lo=0.5, hi=1.5):
    """Least-squares fit of y = a cos(w t) + b sin(w t) + c; golden-section
    refinement of a coarse scan over w."""
    def res(w):
        A = np.stack([np.cos(w * t), np.sin(w * t), np.ones_like(t)], 1)
        c, *_ = np.linalg.lstsq(A, y, rcond=None)
        return np.sum((A @ c - y) ** 2)
    ws = np.linspace(lo, hi, 401)
    r = [res(w) for w in ws]
    i = int(np.argmin(r))
    a, b = ws[max(i - 1, 0)], ws[min(i + 1, len(ws) - 1)]
    g = (np.sqrt(5) - 1) / 2
    for _ in range(60):
        c1, c2 = b - g * (b - a), a + g * (b - a)
        if res(c1) < res(c2):
            b = c2
        else:
            a = c1
    return 0.5 * (a + b)


def _langmuir_theory():
    k = 2 * np.pi / C1_N
    w0 = np.sqrt(1 + 3 * k * k)                          # Bohm-Gross, omega_p = lambda_D = 1
    return 2 / C1_DT * np.arcsin(w0 * C1_DT / 2)        # leapfrog (kick-drift) dispersion


def _centred_energy_drift(h):
    """|E_n/E_0 - 1| with E_n = field energy(x_n) + the time-centred kinetic
    energy (KE(v_{n-1/2}) + KE(v_{n+1/2}))/2 (history column 1 is the
    kinetic energy of the velocity entering the step)."""
    E = h[:-1, 0] + 0.5 * (h[:-1, 1] + h[1:, 1])
    return float(np.abs(E / E[0] - 1).max())


def _langmuir_checks(h, mode):
    t = np.arange(len(mode)) * C1_DT
    w = _fit_frequency(mode.real, t)
    return abs(w / _langmuir_theory() - 1), _centred_energy_drift(h)


LANGMUIR_FREQ_TOL = 5e-3      # relative, on omega
ENERGY_DRIFT_TOL = 5e-4       # leapfrog: O((omega_p dt)^2) of the ~1.5% field share


@pytest.fixture(scope="module")
def c1_history():
    return _c1_run()


def test_langmuir_frequency_c1(c1_history):
    h, mode = c1_history
    ferr, _ = _langmuir_checks(h, mode)
    assert ferr < LANGMUIR_FREQ_TOL, ferr


def test_total_energy_drift_c1(c1_history):
    h, _ = c1_history
    drift = _centred_energy_drift(h)
    assert drift < ENERGY_DRIFT_TOL, drift
    assert np.all(np.isfinite(h))


# --------------------------------------------------------------------------
# two-stream linear growth


TS_NX, TS_NY, TS_M, TS_V0, TS_DT = 48, 8, 1, 5.0, 0.05


def _plasma_Z(z):
    return 1j * np.sqrt(np.pi) * wofz(z)


def two_stream_growth_theory(k, v0=TS_V0, vth=1.0):
    """Growing root of 1 + sum_b (w_pb^2/(k vth)^2)(1 + zeta_b Z(zeta_b)) = 0,
    zeta_b = (omega - k v_b)/(sqrt 2 k vth), two beams of half the density
    (w_pb^2 = 1/2) at +-v0; Newton from the cold-beam root."""
    def eps(om):
        s = 1.0 + 0j
        for vb in (v0, -v0):
            z = (om - k * vb) / (np.sqrt(2) * k * vth)
            s += 0.5 / (k * vth) ** 2 * (1 + z * _plasma_Z(z))
        return s
    a = (k * v0) ** 2
    om = np.sqrt(complex(a + 0.5 - np.sqrt(0.25 + 2 * a)))
    if om.imag <= 0:
        om = 0.3j
    for _ in range(60):
        f = eps(om)
        h = 1e-7
        om = om - f / ((eps(om + h) - f) / h)
    assert abs(eps(om)) < 1e-10
    return om.imag


def _two_stream_quiet(amp=1e-4, q_per_cell=8, rows=64):
    """Quiet start: every row y_j carries one velocity of a quantile sampled
    Maxwellian (beams alternate between rows), with positions on a regular
    x-lattice of q_per_cell per cell, so the density stays uniform except for
    the seeded displacement x += amp sin(k x)/k of mode TS_M."""
    from scipy.stats import norm
    nb = rows // 2
    vq = norm.ppf((np.arange(nb) + 0.5) / nb)
    xs = (np.arange(TS_NX * q_per_cell) + 0.5) / q_per_cell
    k = 2 * np.pi * TS_M / TS_NX
    xs = np.mod(xs + amp * np.sin(k * xs) / k, TS_NX)
    X, Y, VX = [], [], []
    for j in range(rows):
        v = (TS_V0 if j % 2 == 0 else -TS_V0) + vq[j // 2]
        X.append(xs)
        Y.append(np.full(xs.size, (j + 0.5) * TS_NY / rows))
        VX.append(np.full(xs.size, v))
    X, Y, VX = map(np.concatenate, (X, Y, VX))
    w = TS_NX * TS_NY / X.size
    return dict(q=-1.0, m=1.0, w=w, pos=[X, Y], vel=[VX, np.zeros_like(VX)]), k


def _two_stream_growth(steps=480):
    s, k = _two_stream_quiet()
    with O.pic_probe([[k, 0.0]], steps) as pr:
        O.pic_run_multi(O.periodic_grid([TS_NX, TS_NY]), [s], 1.0, 1.0, TS_DT, steps, 3,
                        tol=1e-12, max_cycles=60)
    a = np.abs(pr.modes(steps)[:, 0])
    t = np.arange(steps) * TS_DT
    # linear phase: from 20x the seed (the growing root dominates) to 1e3x
    # (density perturbation ~0.1, well before trapping)
    sel = (a > 20 * a[0]) & (a < 1e3 * a[0])
    if sel.sum() < 20:
        return float("nan"), k
    gam = np.polyfit(t[sel], np.log(a[sel]), 1)[0]
    return gam, k


def test_two_stream_growth_rate():
    gam, k = _two_stream_growth()
    theory = two_stream_growth_theory(k)
    assert abs(gam / theory - 1) < 0.10, (gam, theory)


# --------------------------------------------------------------------------
# planted mistakes: each must fail at least one of the checks above


def _check_fails(name):
    """True when the named whole-step check fails on the oracle as it is now
    (with a planted mistake active)."""
    if name.startswith("one_step_"):
        kind, dim = name[len("one_step_"):].rsplit("_", 1)
        ep, ev = _run_case(kind, int(dim[0]))
        return not (ep < 1e-9 and ev < 1e-8)
    if name == "quiet_box":
        try:
            test_quiet_neutral_box_has_no_field()
            return False
        except AssertionError:
            return True
    if name in ("langmuir_frequency", "energy_drift"):
        h, mode = _c1_run()
        ferr, drift = _langmuir_checks(h, mode)
        if name == "langmuir_frequency":
            return not (ferr < LANGMUIR_FREQ_TOL)
        return not (drift < ENERGY_DRIFT_TOL) or not np.all(np.isfinite(h))
    if name == "two_stream_growth":
        gam, k = _two_stream_growth()
        return not (abs(gam / two_stream_growth_theory(k) - 1) < 0.10)
    raise ValueError(name)


# Which checks each planted mistake fails (measured: the lists are what the
# mistakes break, so each is caught by an independent check and the physics
# checks catch the mistakes that change the dynamics of a neutral plasma).
PLANTS = {
    1: ["one_step_periodic_2d", "one_step_walls_3d", "quiet_box", "langmuir_frequency",
        "two_stream_growth"],                          # deposit charge sign
    2: ["one_step_walls_2d", "one_step_walls_3d", "quiet_box"],   # rho_bg sign
    3: ["one_step_periodic_2d", "one_step_walls_2d", "langmuir_frequency",
        "energy_drift"],                               # push with last step's phi
    4: ["one_step_periodic_2d", "one_step_periodic_3d"],          # no neutralisation
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_step_mistakes_are_caught(plant):
    O.pic_plant(plant)
    try:
        missed = [c for c in PLANTS[plant] if not _check_fails(c)]
    finally:
        O.pic_plant(0)
    assert not missed, f"planted mistake {plant} passed {missed}"
