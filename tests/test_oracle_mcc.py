"""Pins of the oracle's Monte Carlo collisions (oracle/mcc.c).

Checked against what the mathematics and the published generator fix:
  * the LCG states from seed 0 are Numerical Recipes' ranqd1 sequence
    (1013904223, 1196435762, 3519870697, 2868466484, 1649599747,
    2670642822), R() = state / (2^32 - 1);
  * spherical vectors have unit length and isotropic first and second
    moments; Maxwell components are Gaussian with variance kT/m (second
    and fourth moments);
  * the fraction of particles that collide is sigma*n*|v|*dt and channel
    k takes the share sigma_k / sum; a zero density collides nothing;
  * elastic collisions keep |v|, excitation removes exactly 2*e*thr/m from
    |v|^2, ionization conserves energy (E = thr*e + E_incident + E_new
    electron), with the two electrons at equal energy, products at the
    parent's position and seeds below 2^31;
  * the seed after each process is the state after the process's own
    draws (two with a chosen channel, one when the cross sections sum to
    zero), whatever the action drew; processes chain;
  * a density field gives the same outcomes as the equal constant density;
  * planted mistakes (seed from the action's stream, unhalved ionization
    energy, Maxwell without its sqrt(-log) factor) fail these checks.
"""
import numpy as np
import pytest

import oracle_lib as O

ME, MHE, QE, KB = 9.10956e-31, 6.67e-27, 1.602e-19, 1.38065e-23


def lcg_state(seed, k):
    return O.lcg_draws(int(seed), k)[1]


def _particles(n, dim, speed, seed=0, dtype=np.float64, L=10.0):
    rng = np.random.default_rng(seed)
    pos = [rng.uniform(0.5, L - 0.5, n).astype(dtype) for _ in range(dim)]
    u = rng.standard_normal((3, n))
    u /= np.linalg.norm(u, axis=0)
    vel = [(speed * u[c]).astype(dtype) for c in range(3)]
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    return pos, vel, seeds


def _grid(dim, L=10.0):
    return O.wall_grid([int(L)] * dim)


def test_lcg_known_answers():
    expect = [1013904223, 1196435762, 3519870697, 2868466484, 1649599747, 2670642822]
    x = 0
    for e in expect:
        vals, x = O.lcg_draws(x, 1)
        assert x == e
        assert vals[0] == e / 4294967295.0
    vals, x = O.lcg_draws(0xFFFFFFFF, 1)
    assert x == (0xFFFFFFFF * 1664525 + 1013904223) % 2**32


def test_spherical_moments():
    rng = np.random.default_rng(1)
    N = 20000
    u = np.array([O.rand_spherical(int(s))[0] for s in rng.integers(0, 2**32, N, dtype=np.uint64)])
    assert np.max(np.abs(np.linalg.norm(u, axis=1) - 1)) < 4e-16
    assert np.all(np.abs(u.mean(axis=0)) < 5 * np.sqrt(1 / (3 * N)))
    m2 = (u ** 2).mean(axis=0)
    assert np.all(np.abs(m2 - 1 / 3) < 5 * np.sqrt(4 / 45 / N))      # var(u_i^2) = 1/5 - 1/9
    _, x = O.rand_spherical(12345)
    assert x == lcg_state(12345, 2)


@pytest.mark.parametrize("plant", [0, 3])
def test_maxwell_is_gaussian(plant):
    rng = np.random.default_rng(2)
    N = 20000
    T, m = 300.0, MHE
    O.mcc_plant(plant)
    try:
        v = np.array([O.rand_maxwell(int(s), T, m)[0] for s in rng.integers(0, 2**32, N, dtype=np.uint64)])
    finally:
        O.mcc_plant(0)
    s2 = KB * T / m
    m2 = (v ** 2).mean(axis=0)
    kurt = (v ** 4).mean(axis=0) / m2 ** 2
    assert np.all(np.abs(m2 / s2 - 1) < 5 * np.sqrt(2 / N))          # var(v^2) = 2 s2^2
    ok = bool(np.all(np.abs(kurt - 3) < 0.25))
    assert ok == (plant == 0)
    _, x = O.rand_maxwell(777, T, m)
    assert x == lcg_state(777, 6)


def _const_process(sig_el, sig_ex, dens, thr=19.82, **kw):
    return O.mcc_process([O.mcc_channel("elastic", sigma=sig_el),
                          O.mcc_channel("excitation", sigma=sig_ex, threshold_ev=thr)], density=dens, **kw)


def test_collision_fraction_and_channel_shares():
    N = 40000
    g = _grid(1)
    speed = 4e6                                  # ~45 eV electrons
    dt, dens = 1e-11, 1e21
    s_el, s_ex = 5e-20, 2.5e-20
    pos, vel, seeds = _particles(N, 1, speed, seed=3)
    r = O.mcc(g, [_const_process(s_el, s_ex, dens)], pos, vel, seeds, dt)
    p = (s_el + s_ex) * dens * speed * dt       # 0.3
    hit = r["outcome"][:, 0]
    frac = np.mean(hit >= 0)
    assert abs(frac - p) < 5 * np.sqrt(p * (1 - p) / N)
    share = np.mean(hit[hit >= 0] == 0)
    q = s_el / (s_el + s_ex)
    assert abs(share - q) < 5 * np.sqrt(q * (1 - q) / np.sum(hit >= 0))
    # every particle drew twice (a channel is always chosen when sum > 0)
    assert all(r["seed"][i] == lcg_state(seeds[i], 2) for i in range(0, N, 97))
    # zero density: nothing collides, seeds still advance by two
    r0 = O.mcc(g, [_const_process(s_el, s_ex, 0.0)], pos, vel, seeds, dt)
    assert np.all(r0["outcome"] == -1)
    assert np.array_equal(r0["seed"], r["seed"])
    # zero cross sections: one draw
    rz = O.mcc(g, [_const_process(0.0, 0.0, dens)], pos, vel, seeds, dt)
    assert all(rz["seed"][i] == lcg_state(seeds[i], 1) for i in range(0, N, 97))


def test_elastic_and_excitation_energy():
    N = 5000
    g = _grid(2)
    speed = 4e6
    pos, vel, seeds = _particles(N, 2, speed, seed=4)
    thr = 19.82
    r = O.mcc(g, [_const_process(1e-19, 1e-19, 1e21, thr=thr)], pos, vel, seeds, 1e-10)
    v2a = sum(v.astype(np.float64) ** 2 for v in vel)
    v2b = sum(v ** 2 for v in r["vel"])
    hit = r["outcome"][:, 0]
    el, ex = hit == 0, hit == 1
    assert el.sum() > 100 and ex.sum() > 100
    np.testing.assert_allclose(v2b[el], v2a[el], rtol=1e-14)
    loss = 2 * QE * thr / ME
    np.testing.assert_allclose(v2b[ex], v2a[ex] - loss, rtol=1e-12)
    assert np.array_equal(v2b[hit < 0], v2a[hit < 0])
    # the direction changed for collided particles
    cosang = sum(a * b for a, b in zip(vel, r["vel"]))[el] / v2a[el]
    assert np.mean(cosang) < 0.1


@pytest.mark.parametrize("plant", [0, 1, 2])
def test_ionization_conserves_energy(plant):
    N = 4000
    g = _grid(3)
    speed = 6e6                                  # ~100 eV
    thr = 24.587
    pos, vel, seeds = _particles(N, 3, speed, seed=5)
    P = O.mcc_process([O.mcc_channel("ionization", sigma=1e-19, threshold_ev=thr)], density=1e21)
    O.mcc_plant(plant)
    try:
        r = O.mcc(g, [P], pos, vel, seeds, 1e-10)
    finally:
        O.mcc_plant(0)
    hit = r["outcome"][:, 0] == 0
    made = r["made"]
    ok = made == hit.sum() and made > 100
    idx = np.nonzero(hit)[0]
    Ea = 0.5 * ME * sum(v[idx] ** 2 for v in vel)
    Eb = 0.5 * ME * sum(v[idx] ** 2 for v in r["vel"])
    Ee = 0.5 * ME * sum(v ** 2 for v in r["e_vel"])
    ok &= bool(np.allclose(Ea, thr * QE + Eb + Ee, rtol=1e-12, atol=0))
    ok &= bool(np.allclose(Eb, Ee, rtol=1e-12))
    for d in range(3):
        ok &= bool(np.array_equal(r["e_pos"][d], pos[d][idx]) and np.array_equal(r["i_pos"][d], pos[d][idx]))
    ok &= bool(np.all(r["e_seed"] < 2**31) and np.all(r["i_seed"] < 2**31))
    ok &= all(r["seed"][i] == lcg_state(seeds[i], 2) for i in idx[:50])
    vi2 = np.mean(sum(v ** 2 for v in r["i_vel"])) / 3
    ok &= abs(vi2 / (KB * 300 / MHE) - 1) < 5 * np.sqrt(2 / made)
    assert ok == (plant == 0)


def test_ion_actions():
    N = 6000
    g = _grid(1)
    T = 300.0
    rng = np.random.default_rng(6)
    s0 = np.sqrt(KB * T / MHE)
    pos = [rng.uniform(1, 9, N)]
    vel = [rng.standard_normal(N) * 5 * s0 + (2000.0 if c == 0 else 0.0) for c in range(3)]
    seeds = rng.integers(0, 2**32, N, dtype=np.uint64).astype(np.uint32)
    P = O.mcc_process([O.mcc_channel("ion_isotropic", sigma=1e-18), O.mcc_channel("ion_backward", sigma=1e-18)],
                      density=1e21, incident_mass=MHE, target_mass=MHE, target_temperature=T)
    r = O.mcc(g, [P], pos, vel, seeds, 1e-6)
    hit = r["outcome"][:, 0]
    bw, iso = hit == 1, hit == 0
    assert bw.sum() > 500 and iso.sum() > 500
    # charge exchange: the new velocity is the atom's Maxwellian
    vb = np.array([v[bw] for v in r["vel"]])
    assert abs(np.mean(vb ** 2) / s0 ** 2 - 1) < 5 * np.sqrt(2 / vb.size)
    # isotropic in the centre-of-mass frame with equal masses: each
    # collision keeps the pair's kinetic energy, so E_ion' = E_ion + E_atom
    # - E_atom' on average; the centre-of-mass velocity moves to half the
    # incident velocity (atoms at rest on average)
    vi = np.array([v[iso] for v in r["vel"]])
    v0 = np.array([v[iso] for v in vel])
    assert abs(np.mean(vi[0]) / np.mean(v0[0]) - 0.5) < 0.1


def test_processes_chain_and_field_density():
    N = 8000
    g = _grid(2)
    pos, vel, seeds = _particles(N, 2, 4e6, seed=7)
    A = _const_process(4e-20, 2e-20, 1e21)
    B = O.mcc_process([O.mcc_channel("elastic", sigma=3e-20)], density=5e20)
    r = O.mcc(g, [A, B], pos, vel, seeds, 1e-11)
    assert all(r["seed"][i] == lcg_state(seeds[i], 4) for i in range(0, N, 53))
    # the equal constant field: same outcomes (CIC weights of a constant
    # field sum to the constant up to rounding)
    nn = O.num_nodes(g)
    fA, fB = np.full(nn, 1e21), np.full(nn, 5e20)
    rf = O.mcc(g, [A, B], pos, vel, seeds, 1e-11, density=[fA, fB])
    assert np.mean(rf["outcome"] == r["outcome"]) > 0.9999
    assert np.array_equal(rf["seed"], r["seed"])
    # a field that vanishes on nodes x < 5: nothing in the cells below x = 4
    # collides (the cell [4, 5) interpolates towards node 5)
    xs = np.repeat(np.arange(11.0), 11)
    fhalf = np.where(xs >= 5, 1e21, 0.0)
    rh = O.mcc(g, [A], pos, vel, seeds, 1e-11, density=[fhalf])
    left = pos[0] < 4.0
    assert np.all(rh["outcome"][left, 0] == -1) and np.any(rh["outcome"][~left, 0] >= 0)


def test_invalid_particles_untouched():
    g = _grid(1)
    pos, vel, seeds = _particles(100, 1, 4e6, seed=8)
    pos[0][::3] = np.nan
    r = O.mcc(g, [_const_process(1e-18, 1e-18, 1e22)], pos, vel, seeds, 1e-9)
    bad = np.isnan(pos[0])
    assert np.all(r["outcome"][bad] == -2)
    assert np.array_equal(r["seed"][bad], seeds[bad])
    for c in range(3):
        assert np.array_equal(r["vel"][c][bad], vel[c][bad])


def test_piecewise_cross_section():
    coef = [[1e-4, -2e-3, 0.01, -0.3, -40.0], [0.0, 1e-3, -0.05, 0.2, -45.0], [0, 0, 0, -0.1, -41.0]]
    ch = O.mcc_channel("elastic", seg_hi=[28.0, 31.0], coef=coef, exp_form=True)
    for x in [25.0, 27.9, 28.0, 29.5, 31.0, 33.0]:
        v2 = np.exp(x)
        xx = np.log(v2)
        k = 0 if xx < 28.0 else (1 if xx < 31.0 else 2)
        expect = np.exp(np.polyval(coef[k], xx))
        assert abs(O.mcc_sigma(ch, v2) / expect - 1) < 1e-13
    half = O.mcc_channel("elastic", seg_hi=[28.0, 31.0], coef=coef, exp_form=True, arg_scale=0.5)
    assert O.mcc_sigma(half, 2 * np.exp(29.0)) == O.mcc_sigma(ch, np.exp(29.0))
    lin = O.mcc_channel("elastic", seg_hi=[], coef=[[0, 0, 0, 2.0, 1.0]], exp_form=False)
    assert O.mcc_sigma(lin, np.exp(3.0)) == pytest.approx(7.0, rel=1e-15)
    assert O.mcc_sigma(O.mcc_channel("elastic", sigma=2.5e-19), 1.0) == 2.5e-19
