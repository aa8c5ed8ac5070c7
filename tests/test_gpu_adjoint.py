"""Reverse mode on the HIP engine (adjoint push / gather / deposit kernels,
self-adjoint MG solve, checkpointed recompute) against the oracle's FP64
adjoint and tangent-linear runs, and against finite differences of the HIP
forward loss in crossing-free directions."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _oracle_grid(eg):
    g = O.Grid()
    for f, _ in O.Grid._fields_:
        v = getattr(eg, f)
        if hasattr(v, "__len__"):
            getattr(g, f)[:] = list(v)
        else:
            setattr(g, f, v)
    return g


def two_stream(n=32, ppc=8, seed=1, v0=0.5, vth=0.05):
    from psum_amd import engine as E
    rng = np.random.default_rng(seed)
    L = float(n)
    N = n * n * ppc
    x = rng.uniform(0, L, N)
    y = rng.uniform(0, L, N)
    vx = np.where(rng.random(N) < 0.5, v0, -v0) + vth * rng.standard_normal(N)
    vy = vth * rng.standard_normal(N)
    g = E.periodic_grid([n, n])
    return g, [dict(q=-1.0, m=1.0, w=L * L / N, pos=[x, y], vel=[vx, vy])]


def _stepper(g, sp, dtype="f64", **kw):
    from psum_amd import engine as E
    np_dt = np.float64 if dtype == "f64" else np.float32
    spc = [dict(s, pos=[a.astype(np_dt) for a in s["pos"]], vel=[a.astype(np_dt) for a in s["vel"]])
           for s in sp]
    args = dict(dtype=dtype, depth=3, tol=1e-12, max_cycles=80, dt=0.1, rho_bg=1.0)
    args.update(kw)
    return E.Stepper(g, spc, **args)


def _rel(a, b):
    a = np.concatenate([np.ravel(t) for t in a])
    b = np.concatenate([np.ravel(t) for t in b])
    return np.abs(a - b).max() / np.abs(b).max()


def test_adjoint_fp64_matches_oracle_periodic():
    g, sp = two_stream()
    steps = 12
    pic = O.Pic(_oracle_grid(g), sp, rho_bg=1.0, dt=0.1, steps=steps, depth=3, tol=1e-13)
    L_ref, gx_ref, gv_ref = pic.adjoint()
    st = _stepper(g, sp)
    L, gx, gv, tf, tb = st.adjoint(steps, ckpt=5)
    assert abs(L - L_ref) <= 1e-9 * abs(L_ref)
    assert _rel(gx[0], gx_ref[0]) < 1e-7
    assert _rel(gv[0], gv_ref[0]) < 1e-7
    # the state is restored: a forward loss from it reproduces L
    assert abs(st.forward_loss(steps) - L) <= 1e-9 * abs(L)


def test_checkpoint_interval_leaves_gradients_bit_identical():
    """Recompute from checkpoints replays the forward bit for bit (integer
    deposit, deterministic solves), so the checkpoint interval cannot change
    the gradient."""
    g, sp = two_stream(n=16, ppc=8, seed=2)
    outs = []
    for ck in (1, 4, 20):
        st = _stepper(g, sp, tol=1e-10)
        outs.append(st.adjoint(10, ckpt=ck))
        st.close()
    for L, gx, gv, _, _ in outs[1:]:
        assert L == outs[0][0]
        for a, b in zip(gx[0] + gv[0], outs[0][1][0] + outs[0][2][0]):
            assert np.array_equal(a, b)


def test_adjoint_walls_3d_two_species_matches_oracle():
    from psum_amd import engine as E
    rng = np.random.default_rng(3)
    cells, ppc = 8, 3
    L = float(cells)
    N = cells ** 3 * ppc
    sp = []
    for q, m, vth in ((-1.0, 1.0, 1.0), (1.0, 100.0, 0.05)):
        pos = [rng.uniform(1.0, L - 1.0, N) for _ in range(3)]
        vel = [rng.standard_normal(N) * vth for _ in range(3)]
        sp.append(dict(q=q, m=m, w=1.0 / ppc, pos=pos, vel=vel))
    g = E.wall_grid([cells] * 3)
    fixed = np.zeros((cells + 1,) * 3, np.uint8)
    fixed[0], fixed[-1] = 1, 1
    fixed[:, 0], fixed[:, -1] = 1, 1
    fixed[:, :, 0], fixed[:, :, -1] = 1, 1
    fixed = fixed.ravel()
    steps = 6
    pic = O.Pic(_oracle_grid(g), sp, rho_bg=0.0, dt=0.1, steps=steps, depth=2, tol=1e-13,
                fixed=fixed)
    L_ref, gx_ref, gv_ref = pic.adjoint()
    st = _stepper(g, sp, depth=2, rho_bg=0.0, fixed=fixed)
    L, gx, gv, _, _ = st.adjoint(steps, ckpt=4)
    assert abs(L - L_ref) <= 1e-9 * abs(L_ref)
    for s in range(2):
        assert _rel(gx[s], gx_ref[s]) < 1e-7
        assert _rel(gv[s], gv_ref[s]) < 1e-7


def test_fp32_adjoint_close_to_fp64_oracle():
    """FP32 particles with mixed-precision V-cycles: the gradient of the
    FP32 forward, compared with the FP64 oracle gradient."""
    g, sp = two_stream(n=32, ppc=8, seed=4)
    steps = 10
    pic = O.Pic(_oracle_grid(g), sp, rho_bg=1.0, dt=0.1, steps=steps, depth=3, tol=1e-13)
    L_ref, gx_ref, gv_ref = pic.adjoint()
    st = _stepper(g, sp, dtype="f32", mixed_mg=True, tol=1e-8)
    L, gx, gv, _, _ = st.adjoint(steps, ckpt=5)
    assert abs(L - L_ref) <= 1e-5 * abs(L_ref)
    assert _rel(gx[0], gx_ref[0]) < 1e-3
    assert _rel(gv[0], gv_ref[0]) < 1e-3


def test_gpu_finite_differences_in_crossing_free_directions():
    """Central FD (h = 1e-4) of the HIP FP64 forward loss; a coordinate is
    used when the oracle's +-h runs keep every particle in its cells."""
    g, sp = two_stream(n=16, ppc=4, seed=5)
    steps = 8
    pic = O.Pic(_oracle_grid(g), sp, rho_bg=1.0, dt=0.1, steps=steps, depth=3, tol=1e-13)
    st = _stepper(g, sp)
    L, gx, gv, _, _ = st.adjoint(steps, ckpt=3)
    gmax = max(np.abs(a).max() for a in gx[0] + gv[0])
    h, checked, worst = 1e-4, 0, 0.0
    for p in (5, 300, 777, 1001):
        for kind in ("x", "v"):
            for d in range(2):
                vals = []
                hashes = []
                for sgn in (1, -1):
                    pos = [a.copy() for a in sp[0]["pos"]]
                    vel = [a.copy() for a in sp[0]["vel"]]
                    (pos if kind == "x" else vel)[d][p] += sgn * h
                    hashes.append(pic.loss([pos], [vel])[1])
                    s2 = _stepper(g, [dict(sp[0], pos=pos, vel=vel)])
                    vals.append(s2.forward_loss(steps))
                    s2.close()
                if hashes[0] != hashes[1]:
                    continue
                fd = (vals[0] - vals[1]) / (2 * h)
                ad = (gx if kind == "x" else gv)[0][d][p]
                worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-3 * gmax))
                checked += 1
    assert checked >= 12
    assert worst < 1e-4, worst


def test_pypsum_autograd_mode_amplitudes_and_velocities():
    """L(a, v0) through ModeDisplace and PicFieldEnergy; torch.autograd gives
    dL/da_k and dL/dv0, compared with the oracle adjoint projected on the
    modes and with the oracle's tangent-linear run in the mode directions."""
    from psum_amd import engine as E
    from psum_amd.pypsum import ModeDisplace, PicFieldEnergy
    n, ppc, steps = 32, 8, 10
    g, sp = two_stream(n=n, ppc=ppc, seed=6)
    Lb = float(n)
    xi = sp[0]["pos"][0]
    nm = 4
    a = torch.tensor([0.01, -0.02, 0.005, 0.01], dtype=torch.float64, device="cuda",
                     requires_grad=True)
    xi_t = torch.from_numpy(xi).cuda()
    y_t = torch.from_numpy(sp[0]["pos"][1]).cuda()
    vx_t = torch.from_numpy(sp[0]["vel"][0]).cuda().requires_grad_(True)
    vy_t = torch.from_numpy(sp[0]["vel"][1]).cuda()
    st = _stepper(g, sp)
    x_t = ModeDisplace.apply(xi_t, a, 0.0, Lb)
    loss = PicFieldEnergy.apply(st, steps, 4, x_t, y_t, vx_t, vy_t)
    ga, gvx = torch.autograd.grad(loss, [a, vx_t])
    # oracle at the displaced positions
    ks = np.arange(1, nm + 1)
    an = a.detach().cpu().numpy()
    x0 = np.mod(xi + sum(an[i] * Lb / (2 * np.pi * k) * np.sin(2 * np.pi * k * xi / Lb)
                         for i, k in enumerate(ks)), Lb)
    assert np.max(np.abs(x_t.detach().cpu().numpy() - x0)) < 1e-12
    spo = [dict(sp[0], pos=[x0, sp[0]["pos"][1]])]
    pic = O.Pic(_oracle_grid(g), spo, rho_bg=1.0, dt=0.1, steps=steps, depth=3, tol=1e-13)
    L_ref, gx_ref, gv_ref = pic.adjoint()
    assert abs(loss.item() - L_ref) <= 1e-9 * abs(L_ref)
    basis = [Lb / (2 * np.pi * k) * np.sin(2 * np.pi * k * xi / Lb) for k in ks]
    ga_ref = np.array([float((gx_ref[0][0] * b).sum()) for b in basis])
    assert np.max(np.abs(ga.cpu().numpy() - ga_ref)) < 1e-7 * np.abs(ga_ref).max()
    for i, b in enumerate(basis):
        dL, _ = pic.tangent([[b, np.zeros_like(b)]], [[np.zeros_like(b)] * 2])
        assert abs(ga[i].item() - dL) < 1e-7 * np.abs(ga_ref).max()
    assert _rel([gvx.cpu().numpy()], [gv_ref[0][0]]) < 1e-7


def test_mode_project_matches_closed_form():
    from psum_amd.pypsum import ModeDisplace
    rng = np.random.default_rng(7)
    N, Lb, nm = 100_003, 256.0, 32
    xi = torch.from_numpy(rng.uniform(0, Lb, N)).cuda()
    a = torch.from_numpy(rng.uniform(-0.05, 0.05, nm)).cuda().requires_grad_(True)
    x = ModeDisplace.apply(xi, a, 0.0, Lb)
    w = torch.from_numpy(rng.standard_normal(N)).cuda()
    (ga,) = torch.autograd.grad((x * w).sum(), [a])
    xin = xi.cpu().numpy()
    ks = np.arange(1, nm + 1)
    B = Lb / (2 * np.pi * ks[:, None]) * np.sin(2 * np.pi * ks[:, None] * xin[None, :] / Lb)
    ref = B @ w.cpu().numpy()
    assert np.max(np.abs(ga.cpu().numpy() - ref)) < 1e-9 * np.abs(ref).max()
    xr = np.mod(xin + a.detach().cpu().numpy() @ B, Lb)
    assert np.max(np.abs(x.detach().cpu().numpy() - xr)) < 1e-10


def test_field_energy_gradient_matches_quadratic_form():
    """W is quadratic in phi: the directional central difference is exact
    up to rounding."""
    from psum_amd import engine as E
    from psum_amd._lib import check, lib
    import ctypes as C
    rng = np.random.default_rng(8)
    for g in (E.periodic_grid([24, 20]), E.wall_grid([9, 7, 5])):
        nn = E.num_nodes(g)
        phi = rng.standard_normal(nn)
        d = rng.standard_normal(nn)
        out = torch.empty(nn, dtype=torch.float64, device="cuda")
        ph = torch.from_numpy(phi).cuda()
        check(lib().psum_hip_field_energy_grad(C.byref(g), C.c_void_p(ph.data_ptr()), 1.5,
                                               C.c_void_p(out.data_ptr()), None), "grad")
        torch.cuda.synchronize()
        og = _oracle_grid(g)
        h = 1e-3
        fd = (O.field_energy(og, phi + h * d, 1.5) - O.field_energy(og, phi - h * d, 1.5)) / (2 * h)
        ad = float(out.cpu().numpy() @ d)
        assert abs(fd - ad) <= 1e-10 * abs(ad)
