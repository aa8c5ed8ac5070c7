"""GPU parity of the particle kernels against the CPU oracle.

Bit-exact: cell keys, sort permutation, permutation, compaction, the
fixed-point deposit, gather + push (positions and velocities) in FP64 and
FP32.  Diagnostics sums (different reduction trees) within 1e-12 relative.
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


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _grids():
    """(name, engine grid, oracle grid) pairs: 2D periodic, 2D walls, 3D walls,
    3D periodic, an odd non-unit grid."""
    E = _E()
    out = []
    specs = [
        ("p2", lambda m: m.periodic_grid([32, 24])),
        ("w2", lambda m: m.wall_grid([20, 17])),
        ("w3", lambda m: m.wall_grid([9, 10, 11])),
        ("p3", lambda m: m.periodic_grid([8, 12, 10])),
        ("odd2", lambda m: m.make_grid([-0.3, 0.1], [1.7, 2.6], [13, 7], periodic=[0, 1],
                                       plo=[0.0, 0.1 - 0.5 * 2.5 / 7], plen=[0.0, 2.5 + 2.5 / 7])),
    ]
    for name, f in specs:
        out.append((name, f(E), f(O)))
    return out


def _positions(og, n, rng, spread=0.1):
    """Random positions covering the particle domain plus edge cases."""
    d = og.dim
    pos = []
    for k in range(d):
        if og.periodic[k]:
            lo, hi = og.plo[k], og.plo[k] + og.plen[k]
        else:
            lo, hi = og.lb[k], og.ub[k]
        x = rng.uniform(lo, hi, n)
        # exact node positions, the lower bound, just below the upper bound
        m = max(1, n // 20)
        nodes = og.lb[k] + rng.integers(0, og.cells[k] + 1, m) * og.delta[k]
        x[:m] = np.clip(nodes, lo, np.nextafter(hi, lo))
        x[m] = lo
        x[m + 1] = np.nextafter(hi, lo)
        pos.append(x)
    return pos


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cell_keys_bit_exact(dtype):
    E = _E()
    rng = np.random.default_rng(11)
    for name, eg, og in _grids():
        pos = [p.astype(dtype) for p in _positions(og, 5000, rng)]
        pos[0][7] = np.nan
        pos[0][8] = np.inf
        ref = O.cell_keys(og, pos)
        got = _host(E.cell_keys(eg, [_dev(p) for p in pos])).view(np.uint32)
        assert np.array_equal(ref, got), name
        assert ref[7] == O.KEY_INVALID and ref[8] == O.KEY_INVALID


@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 4097, 100_000])
def test_stable_sort_bit_exact(n):
    E = _E()
    rng = np.random.default_rng(n + 3)
    nkeys = 70_000
    keys = rng.integers(0, nkeys, n).astype(np.uint32)
    if n > 10:
        keys[rng.integers(0, n, n // 10)] = O.KEY_INVALID
        keys[: n // 5] = keys[0]          # long runs of equal keys
    ref = O.stable_sort(keys, nkeys)
    want = np.argsort(np.where(keys == O.KEY_INVALID, nkeys, keys), kind="stable").astype(np.uint32)
    assert np.array_equal(ref, want)
    if n == 0:
        return
    sk, perm = E.stable_sort(_dev(keys.view(np.int32)), nkeys)
    assert np.array_equal(_host(perm).view(np.uint32), ref)
    assert np.array_equal(_host(sk).view(np.uint32), keys[ref])


def test_sort_small_key_space_and_permute():
    E = _E()
    rng = np.random.default_rng(5)
    for nkeys in (1, 2, 255, 256, 257, 65536, 1 << 24):
        keys = rng.integers(0, nkeys, 9000).astype(np.uint32)
        sk, perm = E.stable_sort(_dev(keys.view(np.int32)), nkeys)
        ref = O.stable_sort(keys, nkeys)
        assert np.array_equal(_host(perm).view(np.uint32), ref), nkeys
    a = rng.standard_normal(9000)
    b = rng.standard_normal(9000).astype(np.float32)
    pa = E.permute([_dev(a)], perm)[0]
    pb = E.permute([_dev(b)], perm)[0]
    assert np.array_equal(_host(pa), a[ref])
    assert np.array_equal(_host(pb), b[ref])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 3, 4, 5, 4097, 9001])
def test_permute_vector_and_scalar_paths(dtype, n):
    """psum_hip_permute takes four outputs per thread when perm and every
    output are 16-byte aligned (tail of n % 4 one by one) and one per thread
    otherwise (an output offset by one element): both equal a[perm]."""
    import ctypes as C
    E = _E()
    from psum_amd import _lib
    rng = np.random.default_rng(n + 11)
    perm = rng.permutation(n).astype(np.uint32)
    src = [rng.standard_normal(n).astype(dtype) for _ in range(3)]
    want = [a[perm] for a in src]
    got = E.permute([_dev(a) for a in src], _dev(perm.view(np.int32)))
    for w, g in zip(want, got):
        assert np.array_equal(w, _host(g))
    # unaligned outputs: views one element into larger buffers
    ds = [_dev(a) for a in src]
    big = [torch.empty(n + 1, dtype=d.dtype, device="cuda") for d in ds]
    dst = [b[1:] for b in big]
    dp = _dev(perm.view(np.int32))
    code = _lib.PSUM_F64 if dtype == np.float64 else _lib.PSUM_F32
    arr = C.c_void_p * 3
    rc = _lib.lib().psum_hip_permute(code, 3, arr(*[d.data_ptr() for d in ds]), arr(*[d.data_ptr() for d in dst]),
                                     C.c_void_p(dp.data_ptr()), n, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    for w, d in zip(want, dst):
        assert np.array_equal(w, _host(d))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 1023, 1024, 1025, 50_000])
def test_compact_bit_exact(dtype, n):
    E = _E()
    rng = np.random.default_rng(n)
    pos = [rng.standard_normal(n).astype(dtype) for _ in range(3)]
    vel = [rng.standard_normal(n).astype(dtype) for _ in range(3)]
    dead = rng.random(n) < 0.3
    pos[0][dead] = np.nan
    rp, rv = O.compact(pos, vel)
    gp, gv = E.compact([_dev(p) for p in pos], [_dev(v) for v in vel])
    for a, b in zip(rp + rv, gp + gv):
        assert np.array_equal(a, _host(b), equal_nan=True)
    assert len(rp[0]) == int((~dead).sum())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_deposit_bit_exact_and_order_independent(dtype):
    E = _E()
    rng = np.random.default_rng(21)
    for name, eg, og in _grids():
        pos = [p.astype(dtype) for p in _positions(og, 20_000, rng)]
        w = -1.0 / 64
        ref, scale, _ = O.deposit_fx(og, pos, w)
        acc, s2 = E.deposit(eg, [_dev(p) for p in pos], w)
        assert s2 == scale
        got = _host(acc)
        assert np.array_equal(ref, got), name
        # sorted order: the LDS tile path; same integers
        keys = O.cell_keys(og, pos)
        order = O.stable_sort(keys, O.num_nodes(og))
        acc2, _ = E.deposit(eg, [_dev(p[order]) for p in pos], w)
        assert np.array_equal(ref, _host(acc2)), name
        # FP64 value agrees with the reference arithmetic to the fixed-point step
        if dtype == np.float64:
            rho_ref = O.deposit_ref(og, pos, w)
            rho = ref.astype(np.float64) / scale
            cnt = np.bincount(np.zeros(1, int), minlength=1)
            assert np.max(np.abs(rho - rho_ref)) <= 4 * 1e3 / scale + 1e-14


def test_deposit_lds_fallback_counter():
    """Randomly ordered particles over a large grid exceed the LDS tile; the
    global-atomic path is taken and the result is still exact."""
    E = _E()
    eg, og = E.periodic_grid([512, 512]), O.periodic_grid([512, 512])
    rng = np.random.default_rng(2)
    pos = [rng.uniform(0, 512, 200_000) for _ in range(2)]
    counters = torch.zeros(8, dtype=torch.int64, device="cuda")
    acc, _ = E.deposit(eg, [_dev(p) for p in pos], 0.5, counters=counters)
    ref, _, _ = O.deposit_fx(og, pos, 0.5)
    assert np.array_equal(ref, _host(acc))
    assert _host(counters)[1] > 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_deposit_overflow_flag_counts_int64_wraps(dtype):
    """counters[3] is the int64 accumulator's overflow flag.  Particles stacked
    on one node each add a term of 2^50 (w = 2^10, scale = 2^40): 2^12 of them
    sum to 2^62 (exact, flag clear), 2^14 pass 2^63 (flagged).  The CIC
    deposit, the fused push + deposit and the quadratic B-spline deposit
    (centre weight 9/16: 2^14 terms reach 1.125 * 2^63) all count it."""
    E = _E()
    eg = E.periodic_grid([16, 16])
    x0 = [5.5, 7.5]                               # node (5, 7): coefficient 1
    for n, wraps in ((1 << 12, False), (1 << 14, True)):
        print(f"[overflow] dtype={np.dtype(dtype).name} n={n}", flush=True)
        pos = [_dev(np.full(n, x, dtype)) for x in x0]
        cnt = torch.zeros(4, dtype=torch.int64, device="cuda")
        acc, _ = E.deposit(eg, pos, 1024.0, scale=2.0**40, counters=cnt)
        c = _host(cnt)
        assert (c[3] > 0) == wraps, c
        if not wraps:
            a = _host(acc)
            assert a[5 * 16 + 7] == n * 2**50 and np.count_nonzero(a) == 1
        cnt.zero_()
        vel = [_dev(np.zeros(n, dtype)) for _ in range(2)]
        phi = _dev(np.zeros(256, dtype))
        E.gather_push(eg, pos, vel, phi, -1.0, 0.1, 1.0, w_dep=1024.0, scale=2.0**40, counters=cnt)
        assert (_host(cnt)[3] > 0) == wraps
        cnt.zero_()
        E.deposit_bspline(eg, pos, 1024.0, 2, scale=2.0**40, counters=cnt)
        assert (_host(cnt)[3] > 0) == wraps
    # a term above 2^50 could wrap an LDS tile entry: refused
    with pytest.raises(Exception, match="2\\^50"):
        E.deposit(eg, pos, 2048.0, scale=2.0**40)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gather_push_bit_exact(dtype):
    E = _E()
    rng = np.random.default_rng(33)
    for name, eg, og in _grids():
        n = 30_000
        pos = [p.astype(dtype) for p in _positions(og, n, rng)]
        vel = [(rng.standard_normal(n) * 0.7).astype(dtype) for _ in range(og.dim)]
        pos[0][5] = np.nan
        phi = rng.standard_normal(O.num_nodes(og)).astype(dtype)
        qm, dt, mass = -1.0, 0.1, 0.25
        rp = [p.copy() for p in pos]
        rv = [v.copy() for v in vel]
        tq = np.float32 if dtype == np.float32 else float
        absorbed, rdiag = O.gather_push(og, rp, rv, phi, tq(qm), tq(dt), mass)
        gp = [_dev(p) for p in pos]
        gv = [_dev(v) for v in vel]
        wdep = 0.75
        counters = torch.zeros(8, dtype=torch.int64, device="cuda")
        acc, gdiag = E.gather_push(eg, gp, gv, _dev(phi), qm, dt, mass, w_dep=wdep,
                                   counters=counters)
        for a, b in zip(rp + rv, gp + gv):
            assert np.array_equal(a, _host(b), equal_nan=True), name
        assert int(_host(counters)[2]) == absorbed, name
        gd = _host(gdiag)
        for k in range(og.dim + 1):
            assert abs(gd[k] - rdiag[k]) <= 1e-12 * (abs(rdiag[k]) + 1e-300) + 1e-12, (name, k)
        # fused deposit of the new positions == oracle deposit of the oracle's new positions
        ref_acc, _, _ = O.deposit_fx(og, rp, wdep)
        assert np.array_equal(ref_acc, _host(acc)), name


def test_field_energy():
    E = _E()
    rng = np.random.default_rng(3)
    for name, eg, og in _grids():
        phi = rng.standard_normal(O.num_nodes(og))
        ref = O.field_energy(og, phi, 0.7)
        got = float(_host(E.field_energy(eg, _dev(phi), 0.7)))
        assert abs(got - ref) <= 1e-12 * abs(ref), name
