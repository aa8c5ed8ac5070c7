"""GPU parity of the HIP multigrid backends behind PSuM's func_table ABI.

The COO matrix comes from the oracle's restatement of the reference Poisson
frontend (Poisson_solver_{2,3}d::_init_impl).  The HIP backend and the
oracle's CSR restatement of multigrid_core must agree per V-cycle to
rounding, and both must converge to the exact discrete solution.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _E():
    from psum_amd import engine
    return engine


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _periodic_problem(n, seed=0):
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
    return og, nn, r, c, v, b


def _wall_problem_3d(n, seed=0):
    og = O.wall_grid([n - 1] * 3)
    shp = O.node_shape(og)
    nn = O.num_nodes(og)
    fixed = np.zeros(shp, np.uint8)
    fixed[0, :, :] = fixed[-1, :, :] = fixed[:, 0, :] = fixed[:, -1, :] = 1
    fixed[:, :, 0] = fixed[:, :, -1] = 1
    fixed = fixed.ravel()
    r, c, v = O.poisson_coo(og, 1.0, fixed)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(nn)
    b[fixed == 1] = 0.0
    return og, nn, r, c, v, b, fixed


def test_registry_names():
    E = _E()
    for name in ("hip_multigrid_gpu", "hip_multigrid32f_gpu"):
        t = E.get_solver_funcs(name)
        assert t.init and t.set_matrix and t.solve and t.set_options and t.free
    assert not E.get_solver_funcs("no_such_backend").init


@pytest.mark.parametrize("n,depth,mode", [(64, 3, 1), (128, 5, 1), (64, 2, 0), (66, 3, 1)])
def test_vcycle_matches_oracle_periodic(n, depth, mode):
    """Per-cycle agreement with the oracle: mode 1 (periodic-aware, the
    default) and mode 0 (the reference algorithm exactly; 64^2 at depth 2 is
    a case where it stays bounded)."""
    E = _E()
    og, nn, r, c, v, b = _periodic_problem(n, seed=n)
    mg = O.MG(nn, r, c, v, n, n, 1, depth=depth, pre=2, post=1, mode=mode)
    assert mg.singular()[0] == (mode == 1)
    s = E.SolverBackend("hip_multigrid_gpu")
    s.set_options(f"I {n} J {n} depth {depth} pre_relax 2 post_relax 1 max_cycles 1 "
                  f"periodic_mode {mode}")
    s.set_matrix(nn, nn, r, c, v)
    x_ref = np.zeros(nn)
    x = torch.zeros(nn, dtype=torch.float64, device="cuda")
    for cyc in range(6):
        mg.vcycles(b, x_ref, 1)
        bd = _dev(b)
        s.solve(bd, x)
        got = x.cpu().numpy()
        err = np.max(np.abs(got - x_ref)) / np.max(np.abs(x_ref))
        assert err < 1e-10, (cyc, err)


def test_converges_to_exact_periodic_solution():
    """Exact discrete solution of the pinned periodic problem via FFT."""
    E = _E()
    n = 128
    og, nn, r, c, v, b = _periodic_problem(n, seed=7)
    s = E.SolverBackend("hip_multigrid_gpu")
    s.set_options(f"I {n} J {n} depth 5 pre_relax 2 post_relax 1 max_cycles 60 tol 1e-12")
    s.set_matrix(nn, nn, r, c, v)
    s.set_source_replace([0], [0.0])
    rho = b.copy()
    x = torch.zeros(nn, dtype=torch.float64, device="cuda")
    s.solve(_dev(b), x)
    cyc, rel = s.stats()
    assert rel <= 1e-12 and cyc < 60
    # FFT: -Lap phi = rho with the 5-point symbol 4 - 2cos - 2cos, gauge phi[0] = 0
    rho2 = rho.reshape(n, n).copy()
    rho2[0, 0] = 0.0
    rho2[0, 0] = -rho2.sum()        # row 0 is implied by the others: restore neutrality
    k = 2 * np.pi * np.fft.fftfreq(n)
    lam = (2 - 2 * np.cos(k))[:, None] + (2 - 2 * np.cos(k))[None, :]
    lam[0, 0] = 1.0
    ph = np.fft.fft2(rho2) / lam
    ph[0, 0] = 0.0
    phi = np.real(np.fft.ifft2(ph))
    phi -= phi[0, 0]
    got = x.cpu().numpy().reshape(n, n)
    assert np.max(np.abs(got - phi)) <= 1e-9 * np.max(np.abs(phi))


def test_mixed_backend_converges_and_matches_fp64():
    E = _E()
    n = 128
    og, nn, r, c, v, b = _periodic_problem(n, seed=9)
    s64 = E.SolverBackend("hip_multigrid_gpu")
    s64.set_options(f"I {n} J {n} depth 5 max_cycles 60 tol 1e-11")
    s64.set_matrix(nn, nn, r, c, v)
    s32 = E.SolverBackend("hip_multigrid32f_gpu")
    s32.set_options(f"I {n} J {n} depth 5 max_cycles 60 tol 1e-10")
    s32.set_matrix(nn, nn, r, c, v)
    x64 = torch.zeros(nn, dtype=torch.float64, device="cuda")
    x32 = torch.zeros(nn, dtype=torch.float64, device="cuda")
    s64.solve(_dev(b), x64)
    s32.solve(_dev(b), x32)
    c32, rel32 = s32.stats()
    assert rel32 <= 1e-10, rel32
    a, m = x64.cpu().numpy(), x32.cpu().numpy()
    assert np.max(np.abs(a - m)) <= 1e-8 * np.max(np.abs(a))
    # reference 32f semantics (fixed count, correction_interval 4) vs the oracle
    mg = O.MG(nn, r, c, v, n, n, 1, depth=5, pre=2, post=1)
    x_ref = np.zeros(nn)
    mg.solve_mixed(b, x_ref, 8, 4)
    s32b = E.SolverBackend("hip_multigrid32f_gpu")
    s32b.set_options(f"I {n} J {n} depth 5 max_cycles 8 correction_interval 4")
    s32b.set_matrix(nn, nn, r, c, v)
    xb = torch.zeros(nn, dtype=torch.float64, device="cuda")
    s32b.solve(_dev(b), xb)
    assert np.max(np.abs(xb.cpu().numpy() - x_ref)) <= 1e-5 * np.max(np.abs(x_ref))


@pytest.mark.parametrize("n", [33, 49])
def test_vcycle_matches_oracle_3d_walls(n):
    """n = 49 (117649 nodes) runs the fine level's sweeps and residuals in
    the chunk kernels."""
    E = _E()
    og, nn, r, c, v, b, fixed = _wall_problem_3d(n, seed=3)
    mg = O.MG(nn, r, c, v, n, n, n, depth=3, pre=2, post=1)
    s = E.SolverBackend("hip_multigrid_gpu")
    s.set_options(f"I {n} J {n} K {n} depth 3 max_cycles 4")
    s.set_matrix(nn, nn, r, c, v)
    x_ref = np.zeros(nn)
    mg.vcycles(b, x_ref, 4)
    x = torch.zeros(nn, dtype=torch.float64, device="cuda")
    s.solve(_dev(b), x)
    got = x.cpu().numpy()
    assert np.max(np.abs(got - x_ref)) <= 1e-12 * np.max(np.abs(x_ref))


def test_reference_residual_reduction_2d():
    """multigrid_core reduces the residual >= 1e5 in 10 V(2,1) cycles on the
    reference's 256^2 boundary-identity problem (test_multigrid_core.cpp:142-209)."""
    E = _E()
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
    s = E.SolverBackend("hip_multigrid_gpu")
    s.set_options(f"I {I} J {J} depth 4 pre_relax 2 post_relax 1 max_cycles 10")
    s.set_matrix(I * J, I * J, rows, cols, vals)
    b = np.ones(I * J)
    x = torch.zeros(I * J, dtype=torch.float64, device="cuda")
    s.solve(_dev(b), x)
    mg = O.MG(I * J, rows, cols, vals, I, J, 1, depth=4)
    r0 = mg.residual_norm(b, np.zeros(I * J))
    r = mg.residual_norm(b, x.cpu().numpy())
    assert r / r0 < 1e-5


def test_source_addback_and_replace():
    E = _E()
    n = 64
    og, nn, r, c, v, b = _periodic_problem(n, seed=1)
    s = E.SolverBackend("hip_multigrid_gpu")
    s.set_options(f"I {n} J {n} depth 3 max_cycles 3")
    s.set_matrix(nn, nn, r, c, v)
    s.set_source_addback([5, 9, 5], [1.0, 2.0, 0.5])
    s.set_source_replace([0, 17], [0.0, 3.0])
    bd = _dev(b)
    x = torch.zeros(nn, dtype=torch.float64, device="cuda")
    s.solve(bd, x)
    bb = b.copy()
    bb[5] += 1.5
    bb[9] += 2.0
    bb[0] = 0.0
    bb[17] = 3.0
    assert np.array_equal(bd.cpu().numpy(), bb)


def test_solver_errors():
    E = _E()
    s = E.SolverBackend("hip_multigrid_gpu")
    with pytest.raises(E.PsumError):
        s.solve(torch.zeros(4, dtype=torch.float64, device="cuda"),
                torch.zeros(4, dtype=torch.float64, device="cuda"))
    s.set_options("I 4 J 4 depth 2")
    with pytest.raises(E.PsumError):
        # not a stencil on the 4x4 grid
        s.set_matrix(16, 16, [0], [10], [1.0])
    with pytest.raises(E.PsumError):
        E.SolverBackend("no_such_backend")


def _wall_problem_2d(n0, n1, seed=0):
    og = O.wall_grid([n0 - 1, n1 - 1])
    shp = O.node_shape(og)
    nn = O.num_nodes(og)
    fixed = np.zeros(shp, np.uint8)
    fixed[0, :] = fixed[-1, :] = fixed[:, 0] = fixed[:, -1] = 1
    fixed = fixed.ravel()
    r, c, v = O.poisson_coo(og, 1.0, fixed)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(nn)
    b[fixed == 1] = 0.0
    return og, nn, r, c, v, b


@pytest.mark.parametrize("backend", ["hip_multigrid_gpu", "hip_multigrid32f_gpu"])
@pytest.mark.parametrize("case", ["periodic128", "periodic96", "walls_70x45"])
@pytest.mark.parametrize("prepost", [(3, 3), (2, 1), (5, 0), (1, 4)])
@pytest.mark.parametrize("stop", ["max_cycles 3", "max_cycles 30 tol 1e-9"])
def test_fused_2d_kernels_are_bit_identical(backend, case, prepost, stop):
    """The fused 2D tile smoother (several sweeps per launch, coarse-grid
    correction folded in) and the fused residual + restriction perform the
    unfused kernels' operations in the same order: equal bits."""
    E = _E()
    if case.startswith("periodic"):
        n = int(case[8:])
        og, nn, r, c, v, b = _periodic_problem(n, seed=3)
        I, J = n, n
    else:
        I, J = 70, 45
        og, nn, r, c, v, b = _wall_problem_2d(I, J, seed=4)
    pre, post = prepost
    out = []
    for fuse in (0, 1):
        s = E.SolverBackend(backend)
        s.set_options(f"I {I} J {J} depth 4 pre_relax {pre} post_relax {post} {stop} "
                      f"fuse {fuse}")
        s.set_matrix(nn, nn, r, c, v)
        x = torch.zeros(nn, dtype=torch.float64, device="cuda")
        s.solve(_dev(b), x)
        s.solve(_dev(b * 1.01), x)
        out.append((x.cpu().numpy(), s.stats()))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1][0] == out[1][1][0]
    assert np.array_equal(np.float64(out[0][1][1]), np.float64(out[1][1][1]), equal_nan=True)


@pytest.mark.parametrize("backend", ["hip_multigrid_gpu", "hip_multigrid32f_gpu"])
@pytest.mark.parametrize("case", ["periodic256_d6", "periodic128_d4", "walls_129x97_d5"])
@pytest.mark.parametrize("prepost", [(3, 3), (2, 1), (0, 2)])
def test_vcycle_tail_is_bit_identical(backend, case, prepost):
    """The one-workgroup tail (levels from the first that fits in LDS down to
    the coarsest, in one launch) performs the per-level kernels' operations
    in the same order: equal bits with the tail switched off."""
    E = _E()
    if case.startswith("periodic"):
        n = int(case[8:case.index("_")])
        og, nn, r, c, v, b = _periodic_problem(n, seed=5)
        I, J = n, n
    else:
        I, J = 129, 97
        og, nn, r, c, v, b = _wall_problem_2d(I, J, seed=6)
    depth = int(case[case.rindex("d") + 1:])
    pre, post = prepost
    out = []
    for tail in (0, 1):
        s = E.SolverBackend(backend)
        s.set_options(f"I {I} J {J} depth {depth} pre_relax {pre} post_relax {post} "
                      f"max_cycles 30 tol 1e-9 tail {tail}")
        s.set_matrix(nn, nn, r, c, v)
        x = torch.zeros(nn, dtype=torch.float64, device="cuda")
        s.solve(_dev(b), x)
        s.solve(_dev(b * 1.01), x)
        out.append((x.cpu().numpy(), s.stats()))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1][0] == out[1][1][0]


@pytest.mark.parametrize("backend", ["hip_multigrid_gpu", "hip_multigrid32f_gpu"])
@pytest.mark.parametrize("case", [(33, 4), (37, 4), (65, 5), (17, 3)])
@pytest.mark.parametrize("prepost", [(3, 3), (2, 1), (0, 2)])
def test_vcycle_tail_3d_is_bit_identical(backend, case, prepost):
    """3D walls: the one-workgroup tail ("tail 1") runs levels of at most 9^3 nodes down
    to a coarsest level of at most 5^3 (33^3 depth 4 and 65^3 depth 5: 9, 5;
    37^3 depth 4: 10^3 is too large, so no tail) with the 27-point rows, 3D
    restriction, dense solve and prolongation of the per-level kernels:
    equal bits and cycle counts with the tail switched off."""
    E = _E()
    n, depth = case
    og, nn, r, c, v, b, fixed = _wall_problem_3d(n, seed=11)
    pre, post = prepost
    out = []
    for tail in (0, 1):
        s = E.SolverBackend(backend)
        s.set_options(f"I {n} J {n} K {n} depth {depth} pre_relax {pre} post_relax {post} "
                      f"max_cycles 30 tol 1e-9 tail {tail}")
        s.set_matrix(nn, nn, r, c, v)
        x = torch.zeros(nn, dtype=torch.float64, device="cuda")
        s.solve(_dev(b), x)
        s.solve(_dev(b * 1.01), x)
        out.append((x.cpu().numpy(), s.stats()))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1][0] == out[1][1][0]
    assert np.array_equal(np.float64(out[0][1][1]), np.float64(out[1][1][1]), equal_nan=True)


@pytest.mark.parametrize("backend", ["hip_multigrid_gpu", "hip_multigrid32f_gpu"])
@pytest.mark.parametrize("n", [49, 68, 97])
@pytest.mark.parametrize("prepost", [(3, 3), (2, 1)])
def test_chunked_3d_kernels_are_bit_identical(backend, n, prepost):
    """On 7-point 3D levels of 2^16 nodes or more the sweeps, residuals and
    the mixed solver's outer residual run four nodes per thread (interior
    nodes of the dominant class) plus a per-node pass over the listed rest,
    and prolongation to 3D levels of 2^15 nodes or more runs over planes; the
    mixed solver takes its first level-0 sweep from the outer residual.  The
    terms and their order are the separate kernels', so the bits equal those
    with "line 0".
    n = 49 gives lines whose length is not a multiple of four, 68 one that is
    (chunks then never straddle a line); at n = 97 the second level (49^3,
    27-point Galerkin rows) is chunked too."""
    E = _E()
    og, nn, r, c, v, b, fixed = _wall_problem_3d(n, seed=7)
    pre, post = prepost
    out = []
    for line in (0, 1):
        s = E.SolverBackend(backend)
        s.set_options(f"I {n} J {n} K {n} depth 4 pre_relax {pre} post_relax {post} "
                      f"max_cycles 30 tol 1e-9 line {line}")
        s.set_matrix(nn, nn, r, c, v)
        x = torch.zeros(nn, dtype=torch.float64, device="cuda")
        s.solve(_dev(b), x)
        s.solve(_dev(b * 1.01), x)
        out.append((x.cpu().numpy(), s.stats()))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1][0] == out[1][1][0]
    assert np.array_equal(np.float64(out[0][1][1]), np.float64(out[1][1][1]), equal_nan=True)
