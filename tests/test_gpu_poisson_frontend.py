"""GPU solves of the general Poisson frontend systems (mixed Robin/Neumann
sides through ghost elimination, the cylindrical form, a position-dependent
permittivity) on the hip_multigrid_gpu backend.

The product path is psum_poisson_assemble_ex + poisson_addback +
set_source_addback/replace + the multigrid solve; the checker is a direct
sparse solve of the oracle's matrix and right-hand side (the two matrices are
bit-identical, tests/test_abi.py), plus the reference's closed form for the
cylindrical case (test/field_solver/test_psolver_2d.cpp:17-76).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_lib as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS0 = 8.854e-12


def _E():
    from psum_amd import engine
    return engine


def edge(g, axis, plus):
    shp = O.node_shape(g)
    idx = np.arange(O.num_nodes(g)).reshape(shp)
    sl = [slice(None)] * g.dim
    sl[axis] = shp[axis] - 1 if plus else 0
    return idx[tuple(sl)].ravel()


def code(g, axis, plus):
    if g.dim == 2:
        return (0 if plus else 1) if axis == 1 else (2 if plus else 3)
    return 2 * axis + (0 if plus else 1)


def coords(g):
    shp = O.node_shape(g)
    return np.meshgrid(*[g.lb[d] + np.arange(shp[d]) * g.delta[d] for d in range(g.dim)], indexing="ij")


def solve_both(g, eps, fixed, fix_val, sides, source, cylindrical=False, eps_fn=None, depth=4):
    """sides: (axis, plus, value_coef, grad_coef, value); returns (gpu, ref)"""
    E = _E()
    n = O.num_nodes(g)
    nodes, vals = [], []
    for axis, plus, vc, gc, val in sides:
        for e in edge(g, axis, plus):
            nodes.append((int(e), code(g, axis, plus), vc, gc))
            vals.append(val)
    fi = np.nonzero(fixed)[0]
    # checker: oracle matrix and right-hand side, direct solve
    A, M = O.poisson_assemble_ex(g, eps, fixed, nodes, cylindrical, eps_fn)
    b = O.poisson_rhs(source, M, vals, fi, fix_val)
    ref = spla.spsolve(sp.csc_matrix((A[2], (A[0].astype(np.int64), A[1].astype(np.int64))),
                                     shape=(n, n)), b)
    # product: C-ABI assembly, addback, multigrid
    pg = E.make_grid([g.lb[d] for d in range(g.dim)], [g.ub[d] for d in range(g.dim)],
                     [g.cells[d] for d in range(g.dim)])
    (r, c, v), Mp = E.poisson_assemble(pg, eps, fixed, nodes, cylindrical, eps_fn)
    rows, sums = E.poisson_addback(Mp, vals)
    s = E.SolverBackend("hip_multigrid_gpu")
    shp = O.node_shape(g)
    dims = f"I {shp[0]} J {shp[1]}" + (f" K {shp[2]}" if g.dim == 3 else "")
    s.set_options(f"{dims} depth {depth} pre_relax 2 post_relax 2 max_cycles 300 tol 1e-13")
    s.set_matrix(n, n, r, c, v)
    s.set_source_addback(rows, sums)
    s.set_source_replace(fi.astype(np.uint64), fix_val)
    bd = torch.from_numpy(np.ascontiguousarray(source, np.float64)).cuda()
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    s.solve(bd, x)
    torch.cuda.synchronize()
    cyc, rel = s.stats()
    s.close()
    return x.cpu().numpy(), ref, cyc, rel


def test_cylindrical_neumann_axis_column():
    """test_psolver_2d.cpp:17-76 at k=2: cylindrical (z, r), Neumann at
    z=0, z=1 and the axis, phi=0 at r=R, source c1 r^c2"""
    c1, c2, R = 0.01, 0.5, 1.67
    g = O.make_grid([0.0, 0.0], [1.0, R], [120, 80])
    n = O.num_nodes(g)
    fixed = np.zeros(n, np.uint8)
    fixed[edge(g, 1, True)] = 1
    Z, Rr = coords(g)
    src = (c1 * Rr ** c2).ravel()
    x, ref, cyc, rel = solve_both(g, EPS0, fixed, np.zeros(int(fixed.sum())),
                                  [(0, True, 0.0, 1.0, 0.0), (0, False, 0.0, 1.0, 0.0),
                                   (1, False, 0.0, 1.0, 0.0)], src, cylindrical=True)
    print(f"cylindrical: {cyc} cycles, rel residual {rel:.2e}", flush=True)
    assert np.max(np.abs(x - ref)) <= 1e-9 * np.max(np.abs(ref))
    theory = R ** (c2 + 2) * c1 / (EPS0 * (c2 + 2) ** 2)
    at = int(edge(g, 0, True)[0])
    assert abs(x[at] - theory) / theory < 5e-4     # the discretisation error at this size is 1.6e-4


def test_cartesian_robin_neumann_dirichlet_boxes():
    """test_psolver_2d.cpp:78-131 at k=2: Neumann N/S/E, Robin (1, 1) = 1e9 on
    W, two Dirichlet boxes, source cos(10x) cos(12y)"""
    g = O.make_grid([-1.0, -1.0], [1.0, 1.0], [80, 64])
    n = O.num_nodes(g)
    X, Y = coords(g)
    tol = 1e-10
    box1 = (X >= 0.5 - tol) & (Y >= 0.5 - tol)
    box2 = (X <= -0.5 + tol) & (Y <= -0.5 + tol)
    fixed = (box1 | box2).ravel().astype(np.uint8)
    fv_full = np.where(box2, 1e9, 0.0).ravel()
    fi = np.nonzero(fixed)[0]
    src = (np.cos(10 * X) * np.cos(12 * Y)).ravel()
    x, ref, cyc, rel = solve_both(g, EPS0, fixed, fv_full[fi],
                                  [(1, True, 0.0, 1.0, 0.0), (1, False, 0.0, 1.0, 0.0),
                                   (0, True, 0.0, 1.0, 0.0), (0, False, 1.0, 1.0, 1e9)], src)
    print(f"robin/neumann: {cyc} cycles, rel residual {rel:.2e}", flush=True)
    assert np.max(np.abs(x - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_3d_mixed_faces():
    g = O.make_grid([0.0, 0.0, 0.0], [2.0, 1.0, 1.0], [32, 16, 16])
    n = O.num_nodes(g)
    fixed = np.zeros(n, np.uint8)
    fixed[edge(g, 0, False)] = 1
    X, Y, Zc = coords(g)
    src = (np.sin(3 * X) * np.cos(2 * Y) + Zc).ravel()
    x, ref, cyc, rel = solve_both(g, 1.0, fixed, np.full(int(fixed.sum()), 0.5),
                                  [(0, True, 0.0, 1.0, 0.2), (1, True, 2.0, 1.0, 1.0),
                                   (1, False, 2.0, 1.0, -1.0), (2, True, 0.0, 1.0, 0.0),
                                   (2, False, 0.0, 1.0, 0.0)], src, depth=3)
    print(f"3D mixed: {cyc} cycles, rel residual {rel:.2e}", flush=True)
    assert np.max(np.abs(x - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_variable_permittivity():
    g = O.make_grid([0.0, 0.0], [1.0, 1.0], [64, 64])
    n = O.num_nodes(g)
    fixed = np.zeros(n, np.uint8)
    for ax in (0,):
        fixed[edge(g, ax, True)] = 1
        fixed[edge(g, ax, False)] = 1
    X, Y = coords(g)
    src = np.exp(-20 * ((X - 0.4) ** 2 + (Y - 0.6) ** 2)).ravel()
    eps_fn = lambda p: 1.0 + 0.8 * np.tanh(8 * (p[0] - 0.5))
    x, ref, cyc, rel = solve_both(g, 0.0, fixed, np.zeros(int(fixed.sum())),
                                  [(1, True, 0.0, 1.0, 0.0), (1, False, 0.0, 1.0, 0.0)], src,
                                  eps_fn=eps_fn)
    print(f"variable eps: {cyc} cycles, rel residual {rel:.2e}", flush=True)
    assert np.max(np.abs(x - ref)) <= 1e-9 * np.max(np.abs(ref))
