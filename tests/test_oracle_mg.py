"""Pins of the oracle's Poisson assembly and multigrid (CPU only).

  * assembly: closed-form 5/7-point coefficients, identity rows, column-major
    order (Poisson_solver_2d.hpp:221-349, 387-404)
  * transfers and Galerkin products: dense numpy restatement of
    multigrid_utilities.hpp:55-104 and R A^ P on a small grid
  * V-cycle: reference criterion (>= 1e5 reduction in 10 V(2,1) cycles,
    test_multigrid_core.cpp:142-209); exact discrete solution of a pinned
    periodic problem by FFT; manufactured Dirichlet problem vs a direct solve
  * mixed (32f) solve converges to the FP64 solution
  * field energy: summation by parts 1/2 sum phi (A phi) = field energy
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle_lib as O


def test_assembly_closed_form_2d_periodic():
    og = O.periodic_grid([6, 5], dx=0.5)
    nn = O.num_nodes(og)
    fixed = np.zeros(nn, np.uint8)
    fixed[0] = 1
    r, c, v = O.poisson_coo(og, 2.0, fixed)
    # column-major walk: columns ascending, rows ascending within a column
    key = c.astype(np.int64) * nn + r.astype(np.int64)
    assert np.all(np.diff(key) > 0)
    A = sp.coo_matrix((v, (r.astype(int), c.astype(int))), shape=(nn, nn)).toarray()
    assert A[0, 0] == 1.0 and np.count_nonzero(A[0]) == 1
    eps, dx = 2.0, 0.5
    for i in range(6):
        for j in range(5):
            row = i * 5 + j
            if row == 0:
                continue
            assert abs(A[row, row] - 4 * eps / dx ** 2) < 1e-12
            for (a, b) in (((i + 1) % 6, j), ((i - 1) % 6, j), (i, (j + 1) % 5), (i, (j - 1) % 5)):
                assert abs(A[row, a * 5 + b] + eps / dx ** 2) < 1e-12
            assert np.count_nonzero(A[row]) == 5


def test_assembly_3d_walls():
    og = O.wall_grid([4, 3, 5], dx=1.0)
    shp = O.node_shape(og)
    fixed = np.zeros(shp, np.uint8)
    fixed[0], fixed[-1] = 1, 1
    fixed[:, 0], fixed[:, -1] = 1, 1
    fixed[:, :, 0], fixed[:, :, -1] = 1, 1
    fixed = fixed.ravel()
    r, c, v = O.poisson_coo(og, 1.0, fixed)
    nn = O.num_nodes(og)
    A = sp.coo_matrix((v, (r.astype(int), c.astype(int))), shape=(nn, nn)).toarray()
    for i in np.nonzero(fixed)[0]:
        assert A[i, i] == 1 and np.count_nonzero(A[i]) == 1
    inner = np.nonzero(fixed == 0)[0]
    assert np.allclose(np.diag(A)[inner], 6.0)
    assert np.allclose(A[inner].sum(axis=1), 0.0)


def _transfers_numpy(n):
    """multigrid_utilities.hpp get_structured_transfers, one axis, dense."""
    xf = np.arange(n, dtype=float)
    xc = list(xf[::2])
    if xc[-1] != xf[-1]:
        xc.append(xf[-1])
    xc = np.array(xc)
    P = np.zeros((n, len(xc)))
    for i, x in enumerate(xf):
        hi = np.searchsorted(xc, x)
        if xc[hi] == x:
            P[i, hi] = 1
        else:
            t = (x - xc[hi - 1]) / (xc[hi] - xc[hi - 1])
            P[i, hi - 1], P[i, hi] = 1 - t, t
    vol = np.array([((xf[i] - xf[i - 1]) if i else 0) + ((xf[i + 1] - xf[i]) if i + 1 < n else 0)
                    for i in range(n)]) * 0.5
    return P, vol


def test_galerkin_hierarchy_matches_dense_restatement():
    """Coarse operators of the oracle (mode 0) reproduce a dense numpy
    R D^-1 A P on an 8 x 7 Dirichlet problem: solve one coarse problem
    through the V-cycle's coarsest dense solve and compare."""
    I, J = 8, 7
    og = O.wall_grid([I - 1, J - 1])
    fixed = np.zeros((I, J), np.uint8)
    fixed[0], fixed[-1], fixed[:, 0], fixed[:, -1] = 1, 1, 1, 1
    r, c, v = O.poisson_coo(og, 1.0, fixed.ravel())
    n = I * J
    A = sp.coo_matrix((v, (r.astype(int), c.astype(int))), shape=(n, n)).toarray()
    Pi, vi = _transfers_numpy(I)
    Pj, vj = _transfers_numpy(J)
    P = np.kron(Pi, Pj)
    fv = np.kron(vi, vj)
    cv = P.T @ fv
    R = (P.T * fv[None, :]) / cv[:, None]
    Ah = A / np.diag(A)[:, None]
    Ac = R @ Ah @ P
    mg = O.MG(n, r, c, v, I, J, 1, depth=1, pre=0, post=0, mode=0)
    assert mg.levels() == [n, Ac.shape[0]]
    # with no smoothing one V-cycle from x = 0 is x = P Ac^-1 D_c^-1 R D^-1 b
    rng = np.random.default_rng(0)
    b = rng.standard_normal(n)
    x = np.zeros(n)
    mg.vcycles(b, x, 1)
    bc = R @ (b / np.diag(A))
    xc = np.linalg.solve(Ac / np.diag(Ac)[:, None], bc / np.diag(Ac))
    assert np.max(np.abs(x - P @ xc)) < 1e-12 * np.max(np.abs(x))


def test_reference_residual_reduction_2d():
    I = J = 256
    rows, cols, vals = [], [], []
    for j in range(J):
        for i in range(I):
            idx = i * J + j
            if i in (0, I - 1) or j in (0, J - 1):
                rows.append(idx); cols.append(idx); vals.append(1.0)
            else:
                rows.append(idx); cols.append(idx); vals.append(4.0)
                for a, bb in ((i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)):
                    rows.append(idx); cols.append(a * J + bb); vals.append(-1.0)
    for mode in (0, 1):
        mg = O.MG(I * J, rows, cols, vals, I, J, 1, depth=4, pre=2, post=1, mode=mode)
        b = np.ones(I * J)
        x = np.zeros(I * J)
        r0 = mg.residual_norm(b, x)
        mg.vcycles(b, x, 10)
        assert mg.residual_norm(b, x) / r0 < 1e-5


def test_reference_residual_reduction_3d():
    n = 33
    og = O.wall_grid([n - 1] * 3)
    fixed = np.zeros((n,) * 3, np.uint8)
    fixed[0], fixed[-1], fixed[:, 0], fixed[:, -1], fixed[:, :, 0], fixed[:, :, -1] = 1, 1, 1, 1, 1, 1
    r, c, v = O.poisson_coo(og, 1.0, fixed.ravel())
    mg = O.MG(n ** 3, r, c, v, n, n, n, depth=3, pre=2, post=1)
    b = np.ones(n ** 3)
    x = np.zeros(n ** 3)
    r0 = mg.residual_norm(b, x)
    mg.vcycles(b, x, 25)
    assert mg.residual_norm(b, x) / r0 < 1e-5


def _periodic(n, seed):
    og = O.periodic_grid([n, n])
    nn = O.num_nodes(og)
    fixed = np.zeros(nn, np.uint8)
    fixed[0] = 1
    r, c, v = O.poisson_coo(og, 1.0, fixed)
    rng = np.random.default_rng(seed)
    rho = rng.standard_normal(nn)
    rho -= rho.mean()
    b = rho.copy()
    b[0] = 0.0
    return og, nn, r, c, v, rho, b


def _fft_solution(rho, n):
    """-Lap_h phi = rho (5-point, periodic), gauge phi[0] = 0."""
    k = 2 * np.pi * np.fft.fftfreq(n)
    lam = (2 - 2 * np.cos(k))[:, None] + (2 - 2 * np.cos(k))[None, :]
    lam[0, 0] = 1.0
    ph = np.fft.fft2(rho.reshape(n, n)) / lam
    ph[0, 0] = 0.0
    phi = np.real(np.fft.ifft2(ph)).ravel()
    return phi - phi[0]


@pytest.mark.parametrize("n,depth", [(64, 4), (128, 6)])
def test_periodic_mode_converges_to_fft_solution(n, depth):
    og, nn, r, c, v, rho, b = _periodic(n, n)
    mg = O.MG(nn, r, c, v, n, n, 1, depth=depth)
    assert mg.singular() == (True, 0)
    x = np.zeros(nn)
    r0 = np.linalg.norm(b)
    for k in range(40):
        mg.vcycles(b, x, 1)
        if mg.residual_norm(b, x) <= 1e-12 * r0:
            break
    assert k < 30
    assert x[0] == 0.0
    phi = _fft_solution(rho, n)
    assert np.max(np.abs(x - phi)) < 1e-10 * np.max(np.abs(phi))


def test_reference_mode_fails_on_pinned_periodic():
    """The reference algorithm (mode 0) does not converge on this matrix,
    which is why the build's backends use mode 1."""
    og, nn, r, c, v, rho, b = _periodic(128, 1)
    mg = O.MG(nn, r, c, v, 128, 128, 1, depth=4, mode=0)
    x = np.zeros(nn)
    r0 = np.linalg.norm(b)
    mg.vcycles(b, x, 12)
    assert mg.residual_norm(b, x) > 1e-2 * r0


def test_mixed_converges_to_fp64():
    og, nn, r, c, v, rho, b = _periodic(64, 3)
    mg = O.MG(nn, r, c, v, 64, 64, 1, depth=4)
    x64 = np.zeros(nn)
    mg.vcycles(b, x64, 30)
    x32 = np.zeros(nn)
    mg.solve_mixed(b, x32, 40, 4)
    assert np.max(np.abs(x32 - x64)) < 1e-9 * np.max(np.abs(x64))


def test_manufactured_dirichlet_vs_direct():
    """test_native_vs_mg_2d_manufactured.cpp:21-129: sin(pi x) sin(pi y) on
    64^2 with phi = 0 walls; MG agrees with a direct sparse solve."""
    n = 65
    og = O.make_grid([0.0, 0.0], [1.0, 1.0], [n - 1, n - 1])
    fixed = np.zeros((n, n), np.uint8)
    fixed[0], fixed[-1], fixed[:, 0], fixed[:, -1] = 1, 1, 1, 1
    fixed = fixed.ravel()
    r, c, v = O.poisson_coo(og, 1.0, fixed)
    xs = np.linspace(0, 1, n)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    b = (2 * np.pi ** 2 * np.sin(np.pi * X) * np.sin(np.pi * Y)).ravel()
    b[fixed == 1] = 0.0
    A = sp.csr_matrix((v, (r.astype(int), c.astype(int))), shape=(n * n, n * n))
    direct = sla.spsolve(A.tocsc(), b)
    mg = O.MG(n * n, r, c, v, n, n, 1, depth=4)
    x = np.zeros(n * n)
    mg.vcycles(b, x, 25)
    assert np.max(np.abs(x - direct)) < 1e-9
    exact = (np.sin(np.pi * X) * np.sin(np.pi * Y)).ravel()
    assert np.max(np.abs(direct - exact)) < 1e-3       # O(h^2) discretisation error


def test_field_energy_summation_by_parts():
    n = 32
    og, nn, r, c, v, rho, b = _periodic(n, 5)
    phi = _fft_solution(rho, n)
    # 1/2 sum phi * (A phi) with the periodic operator = 1/2 sum |grad phi|^2
    k = phi.reshape(n, n)
    Aphi = (4 * k - np.roll(k, 1, 0) - np.roll(k, -1, 0) - np.roll(k, 1, 1) - np.roll(k, -1, 1))
    lhs = 0.5 * np.sum(k * Aphi)
    assert abs(O.field_energy(og, phi, 1.0) - lhs) < 1e-10 * abs(lhs)
