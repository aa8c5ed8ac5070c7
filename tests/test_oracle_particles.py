"""Pins of the oracle's particle-side functions (CPU only).

Each check is chosen so a plausible mistake (wrong index order, dropped
wrap, swapped weight, sign error) fails: geometric definitions instead of
the formula, the reference's own known-answer properties
(test/field/test_simpleinterp.cpp, test_simpleinterpDiff.cpp,
test_field_interpDiff.cpp), closed forms, numpy brute force.
"""
import numpy as np
import pytest

import oracle_lib as O


def _rand_grid(rng, d, periodic=False):
    cells = list(rng.integers(3, 40, d))
    if periodic:
        n = [c + 1 for c in cells]
        dx = float(rng.uniform(0.3, 2.0))
        return O.periodic_grid(n, dx)
    lb = rng.uniform(-1, 1, d)
    ub = lb + np.exp(rng.uniform(-0.5, 1.0, d))
    return O.make_grid(list(lb), list(ub), cells)


def _node_pos(g, idx):
    """coordinates of linear node index (last dimension fastest)"""
    out = []
    for d in reversed(range(g.dim)):
        w = g.cells[d] + 1
        out.append(idx % w)
        idx //= w
    nd = list(reversed(out))
    return np.array([g.lb[d] + nd[d] * g.delta[d] for d in range(g.dim)]), nd


@pytest.mark.parametrize("d", [1, 2, 3])
def test_cell_key_geometric_definition(d):
    """The key's base node is the node with node <= x < node + delta in every
    dimension (cell containment), for points inside [lb, ub)."""
    rng = np.random.default_rng(d)
    for _ in range(5):
        g = _rand_grid(rng, d)
        n = 2000
        pos = [rng.uniform(g.lb[k], g.ub[k], n) for k in range(d)]
        keys = O.cell_keys(g, pos)
        for i in range(0, n, 37):
            p, nd = _node_pos(g, int(keys[i]))
            for k in range(d):
                assert p[k] - 1e-12 <= pos[k][i] < p[k] + g.delta[k] * (1 + 1e-12)
                assert 0 <= nd[k] < g.cells[k]


def test_cell_key_edges_periodic_band_and_invalid():
    g = O.periodic_grid([8, 8], 1.0)          # nodes at 0.5 + i
    x = np.array([0.5, 0.25, 7.75, 7.5, 0.0, np.nextafter(8.0, 0), np.nan, 1e300])
    y = np.full_like(x, 3.2)
    keys = O.cell_keys(g, [x, y])
    # row of node 0 for x on node 0; x in [0, 0.5) belongs to the wrap cell of node 7
    assert keys[0] == 0 * 8 + 2
    assert keys[1] == 7 * 8 + 2
    assert keys[2] == 7 * 8 + 2
    assert keys[3] == 7 * 8 + 2
    assert keys[4] == 7 * 8 + 2
    assert keys[5] == 7 * 8 + 2
    assert keys[6] == O.KEY_INVALID
    assert keys[7] == O.KEY_INVALID
    # FP32 inputs promote exactly
    k32 = O.cell_keys(g, [x.astype(np.float32), y.astype(np.float32)])
    assert np.array_equal(k32[:6], keys[:6])


def test_stable_sort_matches_numpy():
    rng = np.random.default_rng(0)
    for n, nk in ((0, 5), (1, 1), (1000, 7), (50_000, 4096)):
        keys = rng.integers(0, nk, n).astype(np.uint32)
        if n:
            keys[rng.integers(0, n, n // 7)] = O.KEY_INVALID
        perm = O.stable_sort(keys, nk)
        ref = np.argsort(np.where(keys == O.KEY_INVALID, nk, keys), kind="stable")
        assert np.array_equal(perm, ref)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_cic_weights_reference_kats(d):
    """test_simpleinterp.cpp:13-65 (unity, range, index range), :130-163
    (one-hot at nodes), :166-230 (linear fields reproduced inside a cell)."""
    rng = np.random.default_rng(10 + d)
    g = _rand_grid(rng, d)
    nn = O.num_nodes(g)
    a = rng.standard_normal(d)
    for _ in range(500):
        x = np.array([rng.uniform(g.lb[k], g.ub[k]) for k in range(d)])
        ok, idx, coef = O.cic_weights(g, x)
        assert ok
        assert np.all((idx >= 0) & (idx < nn))
        assert np.all((coef >= 0) & (coef <= 1))
        assert abs(coef.sum() - 1) < 1e-12
        # linear field f = a.x interpolates exactly
        f = sum(coef[i] * np.dot(a, _node_pos(g, int(idx[i]))[0]) for i in range(len(idx)))
        assert abs(f - np.dot(a, x)) < 1e-10
    # one-hot at nodes (interior nodes; x exactly on a node)
    for _ in range(50):
        nd = [int(rng.integers(0, g.cells[k])) for k in range(d)]
        x = np.array([g.lb[k] + nd[k] * g.delta[k] for k in range(d)])
        ok, idx, coef = O.cic_weights(g, x)
        lin = 0
        for k in range(d):
            lin = lin * (g.cells[k] + 1) + nd[k]
        big = coef > 1 - 1e-12
        assert big.sum() == 1 and idx[big][0] == lin
        assert np.all(coef[~big] < 1e-12)


def test_cic_support_order_closed_form():
    """2D point with s = (0.25, 0.75): support order (0,0),(0,1),(1,0),(1,1)
    with weights (1-s0)(1-s1), (1-s0)s1, s0(1-s1), s0 s1."""
    g = O.make_grid([0.0, 0.0], [4.0, 4.0], [4, 4])
    ok, idx, coef = O.cic_weights(g, np.array([1.25, 2.75]))
    assert ok
    assert list(idx) == [1 * 5 + 2, 1 * 5 + 3, 2 * 5 + 2, 2 * 5 + 3]
    assert np.allclose(coef, [0.75 * 0.25, 0.75 * 0.75, 0.25 * 0.25, 0.25 * 0.75], rtol=0, atol=0)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_gradient_weights_match_fd(d):
    """test_field_interpDiff.cpp:15-148: gradient of the interpolant vs central
    FD of interp with h = 1e-7, mean error < 1e-6."""
    rng = np.random.default_rng(20 + d)
    g = _rand_grid(rng, d)
    phi = rng.standard_normal(O.num_nodes(g))
    errs = []
    for _ in range(300):
        x = np.array([rng.uniform(g.lb[k] + 0.01, g.ub[k] - 0.01) for k in range(d)])
        ok, idx, gc = O.cic_grad_weights(g, x)
        grad = gc @ phi[idx]
        for k in range(d):
            h = 1e-7
            xp, xm = x.copy(), x.copy()
            xp[k] += h
            xm[k] -= h
            _, ip, cp = O.cic_weights(g, xp)
            _, im, cm = O.cic_weights(g, xm)
            fd = (cp @ phi[ip] - cm @ phi[im]) / (2 * h)
            errs.append(abs(fd - grad[k]))
    assert np.mean(errs) < 1e-6


def test_deposit_conservation_transpose_and_fixed_point():
    rng = np.random.default_rng(3)
    for g in (O.periodic_grid([16, 12]), O.wall_grid([7, 9, 5]), O.periodic_grid([6, 5, 4])):
        d = g.dim
        n = 5000
        if g.periodic[0]:
            pos = [rng.uniform(g.plo[k], g.plo[k] + g.plen[k], n) for k in range(d)]
        else:
            pos = [rng.uniform(g.lb[k], g.ub[k], n) for k in range(d)]
        w = -0.37
        acc, scale, skipped = O.deposit_fx(g, pos, w)
        assert skipped == 0
        rho_ref = O.deposit_ref(g, pos, w)
        # fixed point: each term rounded to 2^-S, at most 2^d terms per particle
        assert np.max(np.abs(acc / scale - rho_ref)) <= n * 2 ** d / scale
        # conservation
        assert abs(rho_ref.sum() - n * w) < 1e-10 * n
        # transpose identity <deposit(w), phi> = sum_p w interp(phi, x_p)
        phi = rng.standard_normal(O.num_nodes(g))
        lhs = np.dot(rho_ref, phi)
        rhs = 0.0
        for p in range(n):
            ok, idx, coef = O.cic_weights(g, np.array([pos[k][p] for k in range(d)]))
            rhs += w * np.dot(coef, phi[idx])
        assert abs(lhs - rhs) < 1e-9 * (abs(lhs) + 1)
        # order independence of the integer sum
        perm = rng.permutation(n)
        acc2, _, _ = O.deposit_fx(g, [p[perm] for p in pos], w)
        assert np.array_equal(acc, acc2)


def test_deposit_uniformity():
    """test_simpleinterp.cpp:68-127: many uniform points give a flat density."""
    rng = np.random.default_rng(4)
    g = O.periodic_grid([16, 16])
    n = 400_000
    pos = [rng.uniform(0, 16, n) for _ in range(2)]
    rho = O.deposit_ref(g, pos, 1.0)
    mean = n / 256
    assert np.all(np.abs(rho - mean) < 6 * np.sqrt(mean))


def test_push_uniform_field_closed_form():
    """phi = -E0 x on a wall grid: one step of kick-drift is closed form."""
    g = O.wall_grid([10, 10])
    shp = O.node_shape(g)
    E0 = 0.3
    X, _ = np.meshgrid(np.arange(shp[0], dtype=float), np.arange(shp[1], dtype=float),
                       indexing="ij")
    phi = (-E0 * X).ravel()
    x = [np.array([3.3, 5.01]), np.array([4.4, 2.5])]
    v = [np.array([0.1, -0.2]), np.array([0.0, 0.3])]
    qm, dt = -1.0, 0.1
    xr = [a.copy() for a in x]
    vr = [a.copy() for a in v]
    absorbed, diag = O.gather_push(g, xr, vr, phi, qm, dt, mass=2.0)
    assert absorbed == 0
    vx = v[0] + (E0 * qm) * dt
    assert np.allclose(vr[0], vx, atol=1e-14)
    assert np.allclose(vr[1], v[1], atol=1e-14)
    assert np.allclose(xr[0], x[0] + vx * dt, atol=1e-14)
    assert abs(diag[0] - sum(0.5 * 2.0 * (v[0][i] ** 2 + v[1][i] ** 2) for i in range(2))) < 1e-14
    assert abs(diag[1] - 2.0 * v[0].sum()) < 1e-14


def test_push_wrap_and_absorb():
    g = O.make_grid([0.5, 0.0], [7.5, 4.0], [7, 4], periodic=[1, 0], plo=[0.0, 0.0],
                    plen=[8.0, 0.0])
    phi = np.zeros(O.num_nodes(g))
    x = [np.array([7.95, 0.02, 3.0, 3.0]), np.array([1.0, 1.0, 3.99, 0.01])]
    v = [np.array([1.0, -1.0, 0.0, 0.0]), np.array([0.0, 0.0, 0.2, -0.2])]
    absorbed, _ = O.gather_push(g, x, v, phi, -1.0, 0.1)
    assert abs(x[0][0] - (7.95 + 0.1 - 8.0)) < 1e-14
    assert abs(x[0][1] - (0.02 - 0.1 + 8.0)) < 1e-14
    assert absorbed == 2 and np.isnan(x[0][2]) and np.isnan(x[0][3])


def test_compact_matches_mask():
    rng = np.random.default_rng(6)
    pos = [rng.standard_normal(1000) for _ in range(2)]
    vel = [rng.standard_normal(1000) for _ in range(2)]
    dead = rng.random(1000) < 0.4
    pos[0][dead] = np.nan
    op, ov = O.compact(pos, vel)
    for a, b in zip(op + ov, pos + vel):
        assert np.array_equal(a, b[~dead])


def test_fx_shift_bounds():
    for w in (1.0, 0.5, 0.75, 1 / 64, 3e-7, 12345.0):
        s = O.fx_shift(w)
        assert 2.0 ** 39 <= abs(w) * 2.0 ** s < 2.0 ** 40 or abs(w) * 2.0 ** s == 2.0 ** 39
