"""Pins of the oracle's higher-order (B-spline) coupling.

Checked against what the mathematics fixes, not against the formulas:
  * the centred B-spline values at integer and half-integer offsets;
  * partition of unity, reproduction of linear functions (zero first moment)
    and the second moment (order+1)/12 of a degree-`order` B-spline;
  * the gradient weights equal central differences of the interpolant, and
    sum to zero;
  * the interpolant is continuous where the base node switches;
  * the deposit conserves charge and is the transpose of the interpolation;
    the fixed-point deposit equals the FP64 one to 2^-S;
  * a linear potential gives the exact uniform field in gather + push;
  * planted mistakes (corner instead of nearest base node for odd orders,
    gradient without delta_r, a dropped quartic term) fail these checks.
"""
import numpy as np
import pytest

import oracle_lib as O

ORDERS = [2, 3, 4]


def shift(order):
    S = order + 1
    return -(S // 2) if S % 2 else -(S // 2 - 1)


def test_closed_form_values():
    c, _ = O.bspline_1d(2, 0.0)
    np.testing.assert_allclose(c, [1 / 8, 3 / 4, 1 / 8], rtol=0, atol=1e-16)
    c, _ = O.bspline_1d(2, -0.5)
    np.testing.assert_allclose(c, [0.5, 0.5, 0.0], rtol=0, atol=1e-16)
    c, _ = O.bspline_1d(3, 0.0)
    np.testing.assert_allclose(c, [1 / 6, 2 / 3, 1 / 6, 0.0], rtol=0, atol=1e-16)
    c, _ = O.bspline_1d(3, 0.5)
    np.testing.assert_allclose(c, [1 / 48, 23 / 48, 23 / 48, 1 / 48], rtol=0, atol=1e-16)
    c, _ = O.bspline_1d(4, 0.0)
    np.testing.assert_allclose(c, np.array([1, 76, 230, 76, 1]) / 384, rtol=0, atol=1e-16)
    c, _ = O.bspline_1d(4, -0.5)
    np.testing.assert_allclose(c, np.array([1, 11, 11, 1, 0]) / 24, rtol=0, atol=1e-16)


@pytest.mark.parametrize("order", ORDERS)
def test_moments_and_derivative_sums(order):
    rng = np.random.default_rng(order)
    lo, hi = (-0.5, 0.5) if (order + 1) % 2 else (0.0, 1.0)
    k = np.arange(order + 1) + shift(order)
    for s in rng.uniform(lo, hi, 200):
        c, dc = O.bspline_1d(order, s)
        assert abs(c.sum() - 1) < 1e-15 and c.min() >= 0
        assert abs(np.dot(c, k - s)) < 1e-14
        assert abs(np.dot(c, (k - s) ** 2) - (order + 1) / 12) < 1e-14
        assert abs(dc.sum()) < 1e-14
        assert abs(np.dot(dc, k) - 1) < 1e-14          # d/ds sum c_k k = d/ds s
        h = 1e-6
        cp, _ = O.bspline_1d(order, s + h)
        cm, _ = O.bspline_1d(order, s - h)
        if lo < s - h and s + h < hi:
            np.testing.assert_allclose((cp - cm) / (2 * h), dc, rtol=0, atol=1e-8)


def _field(g, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(O.num_nodes(g))


def _interp(g, order, phi, x):
    idx, coef, gc = O.bspline_weights(g, order, x)
    keep = np.abs(coef) > 1e-10
    val = np.sum(phi[idx[keep]] * coef[keep])
    grad = [np.sum(phi[idx] * np.where(np.abs(gc[d]) > 1e-10, gc[d], 0)) for d in range(g.dim)]
    return val, np.array(grad)


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("dim", [2, 3])
def test_gradient_is_derivative_of_interpolant(order, dim):
    g = O.periodic_grid([12] * dim) if dim == 2 else O.periodic_grid([8] * dim)
    phi = _field(g, order)
    rng = np.random.default_rng(10 + order)
    h = 1e-6
    for _ in range(20):
        x = rng.uniform(1.0, g.plen[0] - 2.0, dim)
        _, grad = _interp(g, order, phi, x)
        for d in range(dim):
            e = np.zeros(dim)
            e[d] = h
            vp, _ = _interp(g, order, phi, x + e)
            vm, _ = _interp(g, order, phi, x - e)
            assert abs((vp - vm) / (2 * h) - grad[d]) < 1e-6 * (1 + abs(grad[d]))


@pytest.mark.parametrize("order", ORDERS)
def test_continuous_across_base_switch(order):
    g = O.periodic_grid([16, 16])
    phi = _field(g, 3)
    # odd supports switch base node at half-integer offsets, even ones at nodes
    off = 0.5 if (order + 1) % 2 else 0.0
    x0 = g.lb[0] + 5 + off
    a, _ = _interp(g, order, phi, [np.nextafter(x0, -np.inf), 7.3])
    b, _ = _interp(g, order, phi, [x0, 7.3])
    assert abs(a - b) < 1e-12


@pytest.mark.parametrize("order", ORDERS)
def test_deposit_conserves_charge_and_is_gather_transpose(order):
    g = O.periodic_grid([16, 12])
    rng = np.random.default_rng(order)
    n = 500
    pos = [rng.uniform(0, g.plen[0], n), rng.uniform(0, g.plen[1], n)]
    w = -0.37
    rho = O.deposit_bs_ref(g, order, pos, w)
    # conserved up to the weights at or below the 1e-10 cut (field_interp.hpp:70)
    assert abs(rho.sum() - w * n) < (order + 1) ** 2 * 1e-10 * abs(w) * n
    phi = _field(g, 5)
    lhs = np.dot(rho, phi)
    rhs = sum(w * _interp(g, order, phi, [pos[0][p], pos[1][p]])[0] for p in range(n))
    assert abs(lhs - rhs) < 1e-11 * n
    scale = 2.0 ** 40
    acc, skipped = O.deposit_bs(g, order, pos, w, scale)
    assert skipped == 0
    assert np.max(np.abs(acc / scale - rho)) <= (order + 1) ** 2 * n * 2.0 ** -40
    acc32, _ = O.deposit_bs(g, order, [a.astype(np.float32) for a in pos], w, scale)
    acc64f, _ = O.deposit_bs(g, order, [a.astype(np.float32).astype(np.float64) for a in pos], w, scale)
    np.testing.assert_array_equal(acc32, acc64f)     # FP64 weights for both dtypes


@pytest.mark.parametrize("order", ORDERS)
def test_uniform_field_push(order):
    """phi = -E0 x on a wall grid: every B-spline reproduces linear functions,
    so the gathered field is E0 and one step is the closed-form kick/drift"""
    g = O.wall_grid([20, 10])
    X, Y = np.meshgrid(np.arange(21.0), np.arange(11.0), indexing="ij")
    E0 = 0.3
    phi = (-E0 * X).ravel()
    rng = np.random.default_rng(order)
    n = 50
    pos = [rng.uniform(5, 15, n), rng.uniform(4, 6, n)]
    vel = [rng.standard_normal(n), rng.standard_normal(n)]
    qm, dt = -1.0, 0.1
    p2, v2, ab, diag = O.gather_push_bs(g, order, pos, vel, phi, qm, dt)
    assert ab == 0
    # exact up to the gradient weights at or below the 1e-10 cut
    tol = (order + 1) ** 2 * 1e-10 * np.abs(phi).max() * abs(qm * dt) + 1e-13
    np.testing.assert_allclose(v2[0], vel[0] + E0 * qm * dt, rtol=0, atol=tol)
    np.testing.assert_allclose(v2[1], vel[1], rtol=0, atol=tol)
    np.testing.assert_allclose(p2[0], pos[0] + v2[0] * dt, rtol=0, atol=1e-12)
    assert abs(diag[0] - 0.5 * np.sum(vel[0] ** 2 + vel[1] ** 2)) < 1e-12 * n


@pytest.mark.parametrize("plant,order", [(1, 2), (1, 4), (2, 3), (3, 4)])
def test_planted_mistakes_fail_the_pins(plant, order):
    g = O.periodic_grid([12, 12], dx=0.5)
    phi = _field(g, 1)
    dx = g.delta[0]
    O.bspline_plant(plant)
    try:
        bad = False
        rng = np.random.default_rng(0)
        for _ in range(10):
            x = rng.uniform(1.0, 4.5, 2)
            idx, coef, gc = O.bspline_weights(g, order, x)
            # the interpolant of a linear function must be exact
            lin = (g.lb[0] + np.arange(12.0) * dx)[idx // 12]
            bad |= abs(np.dot(coef, lin) - x[0]) > 1e-9
            _, grad = _interp(g, order, phi, x)
            h = 1e-6
            vp, _ = _interp(g, order, phi, x + [h, 0])
            vm, _ = _interp(g, order, phi, x - [h, 0])
            bad |= abs((vp - vm) / (2 * h) - grad[0]) > 1e-5 * (1 + abs(grad[0]))
            bad |= abs(coef.sum() - 1) > 1e-12
            bad |= coef.min() < -1e-15                       # B-spline weights are >= 0
        for off in (0.0, 0.5):                              # continuity where any base may switch
            x0 = g.lb[0] + (5 + off) * dx
            a, _ = _interp(g, order, phi, [np.nextafter(x0, -np.inf), 3.3])
            b, _ = _interp(g, order, phi, [x0, 3.3])
            bad |= abs(a - b) > 1e-9
        assert bad
    finally:
        O.bspline_plant(0)
