"""Monte Carlo collisions on the GPU (psum_hip_mcc through psum_amd.mcc)
against the oracle on the same particles, seeds and model.

Bar: per (particle, process) the collided channel equals the oracle's (the
cross sections and the collision test go through device log/exp, which may
round differently from the host's by an ulp; a decision could flip only
where a draw sits within ~1e-16 of its threshold, so none is allowed at
these sizes); every seed is bit-identical; velocities agree to
1e-12 relative to the particle's speed in FP64 and to one float ulp in
FP32; products are identical in count, order, positions and seeds, with
velocities to the same tolerance.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ME, MHE, QE = 9.10956e-31, 6.67e-27, 1.602e-19
X0 = 28.0          # ln(v2) of a ~2.3 eV electron


def _fit(a, b, c, x_on=None):
    """coefficients of exp(a + b (x-X0) + c (x-X0)^2); below x_on a segment
    with sigma = e^-200 (a threshold)"""
    row = [0.0, 0.0, c, b - 2 * c * X0, a - b * X0 + c * X0 * X0]
    if x_on is None:
        return dict(seg_hi=[], coef=[row])
    return dict(seg_hi=[x_on], coef=[[0, 0, 0, 0, -200.0], row])


def _model(dim_nodes, dens_field):
    """He-like electron model: against atoms (density field) elastic, two
    excitations and ionization; against ions (constant) elastic"""
    from psum_amd import mcc as M
    atom = dict(incident_mass=ME, target_mass=MHE, target_temperature=300.0, unit_charge=QE)
    chans = [("elastic", _fit(-44.0, -0.4, -0.02), 0.0),
             ("excitation", _fit(-46.0, 0.8, -0.15, x_on=X0 + 2.16), 19.82),
             ("excitation", _fit(-46.5, 0.7, -0.12, x_on=X0 + 2.2), 20.61),
             ("ionization", _fit(-45.0, 0.9, -0.1, x_on=X0 + 2.37), 24.587)]
    ion_ch = [("elastic", dict(sigma=3e-20), 0.0)]
    spec = [(chans, dens_field), (ion_ch, 2e20)]
    gpu, orc = [], []
    for chs, dens in spec:
        gch = [M.channel(k, **kw, threshold_ev=t) for k, kw, t in chs]
        och = [O.mcc_channel(k, **kw, threshold_ev=t) for k, kw, t in chs]
        gd = torch.tensor(dens, dtype=torch.float64, device="cuda") if isinstance(dens, np.ndarray) else dens
        gpu.append(M.Process(gch, gd, **atom))
        orc.append(O.mcc_process(och, density=0.0 if isinstance(dens, np.ndarray) else dens, **atom))
    fields = [dens_field, None]
    return gpu, orc, fields


def _particles(n, dim, L, seed, dtype):
    rng = np.random.default_rng(seed)
    pos = [rng.uniform(0, L, n).astype(dtype) for _ in range(dim)]
    pos[0][rng.random(n) < 0.01] = np.nan
    vel = [(rng.standard_normal(n) * 2.5e6).astype(dtype) for _ in range(3)]
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    return pos, vel, seeds


def _species(pos, vel, seeds, cap):
    from psum_amd import mcc as M
    n = seeds.size
    sp = M.Particles.empty(len(pos), cap, dtype=torch.float64 if pos[0].dtype == np.float64 else torch.float32)
    for d, a in enumerate(pos):
        sp.pos[d][:n] = torch.from_numpy(a)
    for c, a in enumerate(vel):
        sp.vel[c][:n] = torch.from_numpy(a)
    sp.seed[:n] = torch.from_numpy(seeds.view(np.int32))
    sp.n = n
    return sp


def _host(sp, lo, hi):
    return ([t[lo:hi].cpu().numpy() for t in sp.pos], [t[lo:hi].cpu().numpy() for t in sp.vel],
            sp.seed[lo:hi].cpu().numpy().view(np.uint32))


def _assert_vel_close(a, b, dtype):
    a = np.array(a, np.float64)
    b = np.array(b, np.float64)
    if dtype == np.float64:
        speed = np.sqrt((b ** 2).sum(axis=0)) + 1.0
        assert np.all(np.abs(a - b) <= 1e-12 * speed), np.max(np.abs(a - b) / speed)
    else:
        ulp = np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(a - b) <= ulp), np.max(np.abs(a - b) / ulp)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_mcc_matches_oracle(dtype, dim):
    from psum_amd import engine as E
    from psum_amd import mcc as M
    cells = {1: [64], 2: [24, 20], 3: [10, 9, 8]}[dim]
    L = [float(c) for c in cells]
    g = E.wall_grid(cells)
    go = O.wall_grid(cells)
    rng = np.random.default_rng(dim)
    field = rng.uniform(0.5e21, 1.5e21, O.num_nodes(go))
    gpu, orc, fields = _model(O.num_nodes(go), field)
    n = 60000
    pos, vel, seeds = _particles(n, dim, min(L), 100 + dim, dtype)
    dt = 1e-10
    print(f"mcc parity dim={dim} dtype={np.dtype(dtype).name} n={n}", flush=True)
    ref = O.mcc(go, orc, pos, vel, seeds, dt, density=fields)
    inc = _species(pos, vel, seeds, cap=n + n // 4)
    ions = M.Particles.empty(dim, 1000, dtype=inc.pos[0].dtype)
    made, out = M.collide(g, gpu, inc, dt, electrons=inc, ions=ions, outcome=True)
    torch.cuda.synchronize()
    oc = out.cpu().numpy().reshape(n, 2)
    flips = np.sum(np.any(oc != ref["outcome"], axis=1))
    assert flips == 0, f"{flips} outcome differences"
    assert np.sum(ref["outcome"] >= 0) > 3000 and np.sum(ref["outcome"][:, 0] == 3) > 1000
    p, v, s = _host(inc, 0, n)
    assert np.array_equal(s, ref["seed"])
    for d in range(dim):
        assert np.array_equal(p[d], ref["pos"][d], equal_nan=True)
    _assert_vel_close(v, ref["vel"], dtype)
    # products: the secondary electrons were appended to the incident arrays
    assert made == ref["made"] and inc.n == n + made and ions.n == made
    pe, ve, se = _host(inc, n, n + made)
    pi, vi, si = _host(ions, 0, made)
    assert np.array_equal(se, ref["e_seed"]) and np.array_equal(si, ref["i_seed"])
    for d in range(dim):
        assert np.array_equal(pe[d], ref["e_pos"][d]) and np.array_equal(pi[d], ref["i_pos"][d])
    _assert_vel_close(ve, ref["e_vel"], dtype)
    _assert_vel_close(vi, ref["i_vel"], dtype)


def test_mcc_capacity_overflow_leaves_state():
    from psum_amd import engine as E
    from psum_amd import mcc as M
    from psum_amd._lib import PsumError
    cells = [32]
    g = E.wall_grid(cells)
    go = O.wall_grid(cells)
    field = np.full(O.num_nodes(go), 2e21)
    gpu, orc, fields = _model(O.num_nodes(go), field)
    n = 20000
    pos, vel, seeds = _particles(n, 1, 32.0, 7, np.float64)
    inc = _species(pos, vel, seeds, cap=n + 2)
    ions = M.Particles.empty(1, 2, dtype=torch.float64)
    before = _host(inc, 0, n + 2)
    with pytest.raises(PsumError):
        M.collide(g, gpu, inc, 2e-11, electrons=inc, ions=ions, grow=False)
    after = _host(inc, 0, n + 2)
    for a, b in zip(before[0] + before[1] + [before[2]], after[0] + after[1] + [after[2]]):
        assert np.array_equal(a, b, equal_nan=True)
    assert inc.n == n and ions.n == 0
    # with growth the same call succeeds and matches the oracle's count
    made, _ = M.collide(g, gpu, inc, 2e-11, electrons=inc, ions=ions)
    ref = O.mcc(go, orc, pos, vel, seeds, 2e-11, density=fields)
    assert made == ref["made"] > 2 and ions.capacity >= made


def test_mcc_ion_processes_and_empty():
    from psum_amd import engine as E
    from psum_amd import mcc as M
    cells = [12, 12]
    g = E.wall_grid(cells)
    go = O.wall_grid(cells)
    T = 300.0
    gp = M.Process([M.channel("ion_isotropic", sigma=6e-19), M.channel("ion_backward", sigma=4e-19)], 1e21,
                   incident_mass=MHE, target_mass=MHE, target_temperature=T)
    op = O.mcc_process([O.mcc_channel("ion_isotropic", sigma=6e-19), O.mcc_channel("ion_backward", sigma=4e-19)],
                       density=1e21, incident_mass=MHE, target_mass=MHE, target_temperature=T)
    n = 30000
    rng = np.random.default_rng(9)
    pos = [rng.uniform(0, 12, n) for _ in range(2)]
    vel = [rng.standard_normal(n) * 2000.0 for _ in range(3)]
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    ref = O.mcc(go, [op], pos, vel, seeds, 1e-7)
    inc = _species(pos, vel, seeds, cap=n)
    made, out = M.collide(g, [gp], inc, 1e-7, outcome=True)
    assert made == 0
    oc = out.cpu().numpy().reshape(n, 1)
    assert np.array_equal(oc, ref["outcome"]) and np.sum(oc >= 0) > 1000
    p, v, s = _host(inc, 0, n)
    assert np.array_equal(s, ref["seed"])
    _assert_vel_close(v, ref["vel"], np.float64)
    empty = M.Particles.empty(2, 16, dtype=torch.float64)
    assert M.collide(g, [gp], empty, 1e-7)[0] == 0
