"""Pins of the oracle's general Poisson assembly (orc_poisson_assemble_ex).

Each check is fixed by the mathematics or by the reference's own tests, not
by the assembly formula:
  * without mixed nodes it equals orc_poisson_coo (pinned elsewhere);
  * the second-order FD/FV scheme and the central-difference ghost equations
    are exact for quadratics, so A' phi = s + M g holds to rounding for a
    quadratic phi with its exact source s and exact boundary values g
    (Cartesian 2D/3D with Robin/Neumann sides and corners, a linear
    permittivity, and the cylindrical form with r^2 and z^2);
  * the reference's closed forms converge at second order:
    test/field_solver/test_psolver_2d.cpp:17-76 (cylindrical, Neumann sides,
    source c1 r^c2, axis potential c1 R^(c2+2) / (eps (c2+2)^2)),
    :147-176 (charged sphere inside a Dirichlet region, centre potential
    rho R^2 / (6 eps)), test_manufactured_solution_2d.cpp:23-93;
  * planted mistakes (ghost sign on the minus sides, unclamped axis radius,
    sign of M) fail these checks.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_lib as O

EPS0 = 8.854e-12


def node_coords(g):
    shp = O.node_shape(g)
    axes = [g.lb[d] + np.arange(shp[d]) * g.delta[d] for d in range(g.dim)]
    return np.meshgrid(*axes, indexing="ij")


def edge_nodes(g, axis, plus):
    shp = O.node_shape(g)
    idx = np.arange(O.num_nodes(g)).reshape(shp)
    sl = [slice(None)] * g.dim
    sl[axis] = shp[axis] - 1 if plus else 0
    return idx[tuple(sl)].ravel()


def dir_code(g, axis, plus):
    if g.dim == 2:
        return (0 if plus else 1) if axis == 1 else (2 if plus else 3)
    return 2 * axis + (0 if plus else 1)


def matrix(A, n):
    r, c, v = A
    return sp.csr_matrix((v, (r.astype(np.int64), c.astype(np.int64))), shape=(n, n))


def mixed_side(g, axis, plus, vcoef, gcoef, phi_f, grad_f):
    """mixed nodes on one side with exact values v*phi + g*dphi/dn"""
    nodes, vals = [], []
    X = [x.ravel() for x in node_coords(g)]
    for e in edge_nodes(g, axis, plus):
        p = [X[d][e] for d in range(g.dim)]
        dn = grad_f(p)[axis] * (1.0 if plus else -1.0)
        nodes.append((int(e), dir_code(g, axis, plus), vcoef, gcoef))
        vals.append(vcoef * phi_f(p) + gcoef * dn)
    return nodes, vals


def check_exact(g, phi_f, grad_f, src_f, sides, fixed_sides, eps=1.0, eps_fn=None,
                cylindrical=False, rtol=1e-9):
    n = O.num_nodes(g)
    X = [x.ravel() for x in node_coords(g)]
    pts = list(zip(*X))
    phi = np.array([phi_f(p) for p in pts])
    s = np.array([src_f(p) for p in pts])
    fixed = np.zeros(n, np.uint8)
    for axis, plus in fixed_sides:
        fixed[edge_nodes(g, axis, plus)] = 1
    nodes, vals = [], []
    for axis, plus, vc, gc in sides:
        nd, vl = mixed_side(g, axis, plus, vc, gc, phi_f, grad_f)
        nodes += nd
        vals += vl
    A, M = O.poisson_assemble_ex(g, eps, fixed, nodes, cylindrical, eps_fn)
    fi = np.nonzero(fixed)[0]
    b = O.poisson_rhs(s, M, vals, fi, phi[fi])
    Am = matrix(A, n)
    res = Am @ phi - b
    scale = abs(Am) @ abs(phi) + abs(b) + 1e-300
    return np.max(np.abs(res) / scale), (A, M, nodes, fixed)


def quad2():
    a, b, c, d, e, f = 0.7, -0.4, 0.3, 0.2, -0.5, 1.1
    phi = lambda p: a * p[0] ** 2 + b * p[1] ** 2 + c * p[0] * p[1] + d * p[0] + e * p[1] + f
    grad = lambda p: (2 * a * p[0] + c * p[1] + d, 2 * b * p[1] + c * p[0] + e)
    return phi, grad, (a, b)


def test_no_mixed_equals_coo():
    for g, fx in ((O.periodic_grid([12, 9]), "pin"), (O.wall_grid([10, 7]), "edges"),
                  (O.wall_grid([5, 4, 6]), "edges")):
        n = O.num_nodes(g)
        fixed = np.zeros(n, np.uint8)
        if fx == "pin":
            fixed[0] = 1
        else:
            for ax in range(g.dim):
                fixed[edge_nodes(g, ax, True)] = 1
                fixed[edge_nodes(g, ax, False)] = 1
        r, c, v = O.poisson_coo(g, 1.5, fixed)
        (r2, c2, v2), M = O.poisson_assemble_ex(g, 1.5, fixed)
        assert M[0].size == 0
        np.testing.assert_array_equal(r, r2)
        np.testing.assert_array_equal(c, c2)
        np.testing.assert_array_equal(v, v2)   # bit-exact


def test_constant_eps_function_equals_constant():
    g = O.make_grid([-1.0, 0.0], [1.0, 2.0], [8, 6])
    nodes = [(int(e), 0, 0.0, 1.0) for e in edge_nodes(g, 1, True)]
    A1, M1 = O.poisson_assemble_ex(g, 2.5, None, nodes)
    A2, M2 = O.poisson_assemble_ex(g, 0.0, None, nodes, eps_fn=lambda p: 2.5)
    for x, y in zip(A1 + M1, A2 + M2):
        np.testing.assert_array_equal(x, y)


def test_quadratic_exact_2d_robin_neumann_corners():
    g = O.make_grid([-1.0, -0.5], [1.0, 1.5], [8, 6])
    phi, grad, (a, b) = quad2()
    eps = 2.5
    err, (A, M, nodes, fixed) = check_exact(
        g, phi, grad, lambda p: -eps * (2 * a + 2 * b),
        sides=[(1, True, 0.0, 1.0),     # N Neumann
               (1, False, 2.0, 0.5),    # S Robin
               (0, True, 1.0, 3.0),     # E Robin
               (0, False, 1.0, 1.0)],   # W: covered by the fixed nodes
        fixed_sides=[(0, False)], eps=eps)
    assert err < 1e-12
    # ghost columns: every N/S/E node that is not fixed has one M entry,
    # the W nodes (fixed) none; corners (N,E) and (S,E) carry two
    mr = M[0]
    assert set(np.unique(mr).tolist()) == set(
        e for e, d, _, _ in nodes if d != 3 and not fixed[e])
    corner = int(edge_nodes(g, 0, True)[-1])
    assert np.count_nonzero(mr == corner) == 2


def test_quadratic_exact_3d():
    g = O.make_grid([0.0, -1.0, 0.5], [1.0, 0.5, 1.5], [5, 6, 4])
    a, b, c, d = 0.6, -0.3, 0.45, 0.2
    phi = lambda p: a * p[0] ** 2 + b * p[1] ** 2 + c * p[2] ** 2 + d * p[0] * p[2] + p[1] - 0.3
    grad = lambda p: (2 * a * p[0] + d * p[2], 2 * b * p[1] + 1.0, 2 * c * p[2] + d * p[0])
    err, _ = check_exact(
        g, phi, grad, lambda p: -1.5 * (2 * a + 2 * b + 2 * c),
        sides=[(0, True, 1.0, 2.0), (1, True, 0.0, 1.0), (1, False, 0.5, 1.0),
               (2, True, 0.0, 1.0), (2, False, 3.0, 0.25)],
        fixed_sides=[(0, False)], eps=1.5)
    assert err < 1e-12


def test_linear_eps_exact():
    """-div(eps grad phi) with eps linear and phi = a x^2 + b y^2 + d x + e y:
    the midpoint permittivity makes the flux differences exact"""
    g = O.make_grid([0.0, 0.0], [1.0, 1.0], [7, 9])
    al, be, ga = 2.0, 0.3, -0.2
    eps_f = lambda p: al + be * p[0] + ga * p[1]
    A2, B2, D2, E2 = 0.5, -0.8, 0.1, 0.4
    phi = lambda p: A2 * p[0] ** 2 + B2 * p[1] ** 2 + D2 * p[0] + E2 * p[1]
    grad = lambda p: (2 * A2 * p[0] + D2, 2 * B2 * p[1] + E2)
    src = lambda p: -(be * grad(p)[0] + eps_f(p) * 2 * A2 + ga * grad(p)[1] + eps_f(p) * 2 * B2)
    err, _ = check_exact(g, phi, grad, src, sides=[(1, True, 1.0, 2.0), (0, True, 0.0, 1.0)],
                         fixed_sides=[(0, False), (1, False)], eps_fn=eps_f)
    assert err < 1e-12


def cyl_case(neumann_top=False):
    g = O.make_grid([0.0, 0.0], [1.0, 1.2], [6, 8])
    a, b, c, d = 0.9, -0.4, 0.25, 0.7
    phi = lambda p: a * p[1] ** 2 + b * p[0] ** 2 + c * p[0] + d
    grad = lambda p: (2 * b * p[0] + c, 2 * a * p[1])
    sides = [(1, False, 0.0, 1.0),          # S: the axis, Neumann
             (0, True, 2.0, 0.5),           # E Robin
             (0, False, 0.0, 1.0)]          # W Neumann
    fixed_sides = [(1, True)]
    if neumann_top:
        sides.append((1, True, 0.0, 1.0))
        fixed_sides = [(0, False)]
    return check_exact(g, phi, grad, lambda p: -EPS0 * (4 * a + 2 * b), sides, fixed_sides,
                       eps=EPS0, cylindrical=True)


def test_cylindrical_quadratic_exact():
    err, (A, M, nodes, fixed) = cyl_case()
    assert err < 1e-12
    err2, _ = cyl_case(neumann_top=True)
    assert err2 < 1e-12


def test_cylindrical_axis_row_has_no_outer_neighbour():
    """the axis face has zero area: the S term is dropped, not wrapped"""
    g = O.make_grid([0.0, 0.0], [1.0, 1.2], [6, 8])
    (r, c, v), _ = O.poisson_assemble_ex(g, 1.0, None, (), True)
    shp = O.node_shape(g)
    row = 3 * shp[1] + 0            # (i=3, j=0) on the axis
    cols = set(c[r == row].tolist())
    assert cols == {row, row + 1, row + shp[1], row - shp[1]}


def solve(A, M, n, source, mixed_vals, fixed, fix_vals):
    fi = np.nonzero(fixed)[0]
    b = O.poisson_rhs(source, M, mixed_vals, fi, fix_vals)
    return spla.spsolve(matrix(A, n).tocsc(), b)


def c_rate_simple(k):
    """test_psolver_2d.cpp:17-76 (axis potential of a c1 r^c2 column)"""
    c1, c2, R = 0.01, 0.5, 1.67
    g = O.make_grid([0.0, 0.0], [1.0, R], [60 * k, 40 * k])
    n = O.num_nodes(g)
    nodes = []
    for axis, plus in ((0, True), (0, False), (1, False)):
        nodes += [(int(e), dir_code(g, axis, plus), 0.0, 1.0) for e in edge_nodes(g, axis, plus)]
    fixed = np.zeros(n, np.uint8)
    fixed[edge_nodes(g, 1, True)] = 1
    A, M = O.poisson_assemble_ex(g, EPS0, fixed, nodes, True)
    Z, Rr = node_coords(g)
    src = (c1 * Rr ** c2).ravel()
    phi = solve(A, M, n, src, np.zeros(len(nodes)), fixed, np.zeros(int(fixed.sum())))
    theory = R ** (c2 + 2) * c1 / (EPS0 * (c2 + 2) ** 2)
    at = int(edge_nodes(g, 0, True)[0])       # (z=1, r=0)
    return abs(phi[at] - theory) / theory


def test_reference_cylindrical_axis_closed_form_order2():
    errs = [c_rate_simple(k) for k in (1, 2, 4)]
    order = np.log(errs[0] / errs[-1]) / np.log(4)
    assert errs[0] > errs[1] > errs[2]
    assert 1.7 < order < 2.3, (errs, order)
    assert errs[-1] < 1e-3


def test_reference_charged_sphere():
    """test_psolver_2d.cpp:147-176: Dirichlet 0 outside the sphere
    (z-1)^2 + r^2 > 0.25, constant 0.01 inside; centre 0.25*0.01/(6 eps)"""
    g = O.make_grid([0.0, 0.0], [2.0, 2.0], [100, 200])
    n = O.num_nodes(g)
    Z, Rr = node_coords(g)
    fixed = (((Z - 1) ** 2 + Rr ** 2) > 0.25).ravel().astype(np.uint8)
    A, M = O.poisson_assemble_ex(g, EPS0, fixed, (), True)
    phi = solve(A, M, n, np.full(n, 0.01), np.zeros(0), fixed, np.zeros(int(fixed.sum())))
    theory = 0.25 * 0.01 / (6 * EPS0)
    at = 50 * 201 + 0
    assert abs(phi[at] - theory) / theory < 2e-2


def test_reference_manufactured_dirichlet_order2():
    """test_manufactured_solution_2d.cpp:23-93"""
    errs = []
    for k in (1, 2, 3):
        N = 40 * k
        g = O.make_grid([0.0, 0.0], [1.0, 1.0], [N, N])
        n = O.num_nodes(g)
        fixed = np.zeros(n, np.uint8)
        for ax in range(2):
            fixed[edge_nodes(g, ax, True)] = 1
            fixed[edge_nodes(g, ax, False)] = 1
        X, Y = node_coords(g)
        lap = -2 * np.pi ** 2 * np.sin(np.pi * X) * np.sin(np.pi * Y)
        A, M = O.poisson_assemble_ex(g, EPS0, fixed)
        phi = solve(A, M, n, (-EPS0 * lap).ravel(), np.zeros(0), fixed, np.zeros(int(fixed.sum())))
        ex = (np.sin(np.pi * X) * np.sin(np.pi * Y)).ravel()
        errs.append(np.sqrt(np.mean((phi - ex) ** 2)))
    order = np.log(errs[0] / errs[-1]) / np.log(3)
    assert 1.5 < order < 2.5 and errs[-1] < 1e-3


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_planted_mistakes_fail_the_pins(plant):
    O.poisson_plant(plant)
    try:
        if plant == 2:
            err, _ = cyl_case()
        else:
            g = O.make_grid([-1.0, -0.5], [1.0, 1.5], [8, 6])
            phi, grad, (a, b) = quad2()
            err, _ = check_exact(g, phi, grad, lambda p: -(2 * a + 2 * b),
                                 sides=[(1, True, 0.0, 1.0), (1, False, 2.0, 0.5),
                                        (0, True, 1.0, 3.0)], fixed_sides=[(0, False)])
        assert not (err < 1e-6)       # large or NaN
    finally:
        O.poisson_plant(0)


def test_bad_mixed_systems_are_rejected():
    g = O.wall_grid([4, 3])
    with pytest.raises(ValueError):     # grad_coef 0: no ghost term, D singular
        O.poisson_assemble_ex(g, 1.0, None, [(4, 0, 1.0, 0.0)])
    with pytest.raises(ValueError):     # direction out of range
        O.poisson_assemble_ex(g, 1.0, None, [(4, 7, 0.0, 1.0)])
