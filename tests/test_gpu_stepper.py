"""Whole-step parity: the native HIP step engine against the oracle's
restatement of the reference user loop (deposit -> MG solve -> push)."""
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


def test_langmuir_history_fp64_matches_oracle():
    from psum_amd import engine as E
    from psum_amd.workloads import langmuir_2d
    wl = langmuir_2d(n=64, ppc=16, seed=4)
    sp = wl["species"][0]
    steps = 25
    og = _oracle_grid(wl["grid"])
    rp = [a.copy() for a in sp["pos"]]
    rv = [a.copy() for a in sp["vel"]]
    hist_ref, phi_ref, cyc_ref = O.pic_run(og, rp, rv, sp["q"], sp["m"], sp["w"], wl["rho_bg"],
                                           1.0, wl["dt"], steps, depth=3, tol=1e-10,
                                           max_cycles=40)
    st = E.Stepper(wl["grid"], wl["species"], dtype="f64", depth=3, tol=1e-10, max_cycles=40,
                   dt=wl["dt"], rho_bg=wl["rho_bg"])
    hist = st.run(steps)
    phi = st.phi()
    # field energy, kinetic energy: relative 1e-9; momentum: absolute vs total |p| scale
    scale_p = np.abs(hist_ref[:, 1]).max() ** 0.5 * 10
    assert np.max(np.abs(hist[:, 0] - hist_ref[:, 0]) / np.abs(hist_ref[:, 0])) < 1e-8
    assert np.max(np.abs(hist[:, 1] - hist_ref[:, 1]) / np.abs(hist_ref[:, 1])) < 1e-10
    assert np.max(np.abs(hist[:, 2:4] - hist_ref[:, 2:4])) < 1e-9 * scale_p
    # final particle state
    gp, gv = st.state(0)
    for a, b in zip(rp + rv, gp + gv):
        assert np.max(np.abs(a - b)) < 1e-9
    assert np.max(np.abs(phi - phi_ref)) < 1e-9 * np.abs(phi_ref).max()


def test_mixed_vs_fp64_field_energy():
    from psum_amd import engine as E
    from psum_amd.workloads import langmuir_2d
    wl = langmuir_2d(n=64, ppc=16, seed=5)
    steps = 30
    st64 = E.Stepper(wl["grid"], wl["species"], dtype="f64", depth=3, tol=1e-10, max_cycles=40,
                     dt=wl["dt"], rho_bg=wl["rho_bg"])
    h64 = st64.run(steps)
    st32 = E.Stepper(wl["grid"], wl["species"], dtype="f32", mixed_mg=True, depth=3, tol=1e-8,
                     max_cycles=40, dt=wl["dt"], rho_bg=wl["rho_bg"])
    h32 = st32.run(steps)
    rel = np.abs(h32[:, 0] - h64[:, 0]) / np.abs(h64[:, 0])
    assert rel.max() < 1e-5, rel.max()
    relk = np.abs(h32[:, 1] - h64[:, 1]) / np.abs(h64[:, 1])
    assert relk.max() < 1e-5


def test_sorting_does_not_change_history():
    """rho is an exact integer sum, so re-sorting every step leaves the FP64
    history bit-identical."""
    from psum_amd import engine as E
    from psum_amd.workloads import langmuir_2d
    wl = langmuir_2d(n=32, ppc=16, seed=6)
    a = E.Stepper(wl["grid"], wl["species"], dtype="f64", depth=2, tol=1e-10, max_cycles=40,
                  dt=wl["dt"], rho_bg=wl["rho_bg"], sort_every=0).run(10)
    b = E.Stepper(wl["grid"], wl["species"], dtype="f64", depth=2, tol=1e-10, max_cycles=40,
                  dt=wl["dt"], rho_bg=wl["rho_bg"], sort_every=1).run(10)
    # particle order changes only the diagnostic sums' rounding
    assert np.array_equal(a[:, 0], b[:, 0])
    assert np.max(np.abs(a[:, 1] - b[:, 1]) / np.abs(a[:, 1])) < 1e-12


def test_bounded_3d_two_species_matches_oracle():
    """3D box, phi = 0 on all faces, absorbing walls, electrons + heavy ions."""
    from psum_amd import engine as E
    from psum_amd.workloads import bounded_3d
    wl = bounded_3d(cells=16, ppc=4, seed=8, dtype=np.float64)
    steps = 8
    og = _oracle_grid(wl["grid"])
    ref_sp = [dict(s, pos=[a.copy() for a in s["pos"]], vel=[a.copy() for a in s["vel"]])
              for s in wl["species"]]
    h_ref, phi_ref, _ = O.pic_run_multi(og, ref_sp, 0.0, 1.0, wl["dt"], steps, depth=3, tol=1e-10,
                                        max_cycles=40, fixed=wl["fixed"])
    st = E.Stepper(wl["grid"], wl["species"], dtype="f64", depth=3, tol=1e-10, max_cycles=40,
                   dt=wl["dt"], rho_bg=0.0, fixed=wl["fixed"])
    h = st.run(steps)
    assert np.max(np.abs(h[:, 0] - h_ref[:, 0]) / np.abs(h_ref[:, 0])) < 1e-8
    assert np.max(np.abs(h[:, 1] - h_ref[:, 1]) / np.abs(h_ref[:, 1])) < 1e-10
    for s in range(2):
        gp, gv = st.state(s)
        for a, b in zip(ref_sp[s]["pos"] + ref_sp[s]["vel"], gp + gv):
            assert np.array_equal(np.isnan(a), np.isnan(b))
            ok = ~np.isnan(a)
            assert np.max(np.abs(a[ok] - b[ok])) < 1e-9
    assert st.stats()["absorbed"] > 0


def test_bounded_3d_sort_compacts_absorbed():
    from psum_amd import engine as E
    from psum_amd.workloads import bounded_3d
    wl = bounded_3d(cells=12, ppc=4, seed=9, dtype=np.float32)
    st = E.Stepper(wl["grid"], wl["species"], dtype="f32", mixed_mg=True, depth=3, tol=1e-8,
                   max_cycles=40, dt=wl["dt"], rho_bg=0.0, fixed=wl["fixed"], sort_every=3)
    st.run(10)
    absorbed = st.stats()["absorbed"]
    gp, _ = st.state(0)
    gi, _ = st.state(1)
    n0 = sum(len(s["pos"][0]) for s in wl["species"])
    alive = len(gp[0]) + len(gi[0])
    # particles absorbed in the last (unsorted) steps are still in the arrays as NaN
    nan_left = int(np.isnan(gp[0]).sum() + np.isnan(gi[0]).sum())
    assert alive - nan_left == n0 - absorbed
    assert absorbed > 0 and alive < n0


def test_per_species_sort_intervals():
    """Ions sorted every 6 steps, electrons every 2: the deposit is an exact
    integer sum, so phi and the field energy are bit-identical to sorting
    both species every 2 steps; each species holds the same particles (as a
    set) with the same absorbed count.  set_schedule keeps a species' own
    interval; the tuner gives every species one of its candidates."""
    from psum_amd import engine as E
    from psum_amd.workloads import bounded_3d
    wl = bounded_3d(cells=12, ppc=4, seed=9, dtype=np.float32)
    kw = dict(dtype="f32", mixed_mg=True, depth=3, tol=1e-8, max_cycles=40, dt=wl["dt"], rho_bg=0.0,
              fixed=wl["fixed"], sort_every=2)
    a = E.Stepper(wl["grid"], wl["species"], **kw)
    b = E.Stepper(wl["grid"], wl["species"], **kw)
    b.set_species_sort(1, 6)
    assert (b.species_sort(0), b.species_sort(1)) == (2, 6)
    b.set_schedule(0, 3, 0)
    assert (b.species_sort(0), b.species_sort(1)) == (3, 6)
    b.set_schedule(0, 2, 0)
    ha, hb = a.run(12), b.run(12)
    assert np.array_equal(ha[:, 0], hb[:, 0])
    assert np.max(np.abs(ha[:, 1] - hb[:, 1]) / np.abs(ha[:, 1])) < 1e-12
    assert np.array_equal(a.phi(), b.phi())
    assert a.stats()["absorbed"] == b.stats()["absorbed"]

    def alive_sorted(st, s):
        p, v = st.state(s)
        m = ~np.isnan(p[0])
        rows = np.stack([x[m] for x in p + v], axis=1)
        return rows[np.lexsort(rows.T[::-1])]

    for s in range(2):
        assert np.array_equal(alive_sorted(a, s), alive_sorted(b, s))
    with pytest.raises(E.PsumError):
        b.set_species_sort(2, 4)
    tun = a.tune(warmup=3, probe=4)
    every = tun["species_sort_every"]
    assert len(every) == 2 and all(k in (5, 10, 20, 50, 100, 200) for k in every)
    assert tun["sort_every"] == min(every)
    assert [a.species_sort(q) for q in range(2)] == every
    a.close()
    b.close()


def test_sort_order_is_cell_then_velocity_signs():
    """psum_stepper_sort: stable order by cell key * 2^D + velocity-sign bits
    (dimension 0 the most significant), restated with the oracle's keys and
    stable sort."""
    from psum_amd import engine as E
    from psum_amd.workloads import langmuir_2d
    wl = langmuir_2d(n=32, ppc=16, seed=8)
    st = E.Stepper(wl["grid"], wl["species"], dtype="f64", depth=2, dt=wl["dt"], rho_bg=wl["rho_bg"])
    st.run(3)                                  # unsorted, moved state
    p0, v0 = st.state(0)
    og = _oracle_grid(wl["grid"])
    ck = O.cell_keys(og, p0)
    keys = (ck.astype(np.uint64) << 2) | ((v0[0] >= 0).astype(np.uint64) << 1) | (v0[1] >= 0).astype(np.uint64)
    keys = np.where(ck == O.KEY_INVALID, np.uint64(O.KEY_INVALID), keys)
    order = O.stable_sort(keys.astype(np.uint32), O.num_nodes(og) * 4)
    st.sort()
    p1, v1 = st.state(0)
    for a, b in zip(p0 + v0, p1 + v1):
        np.testing.assert_array_equal(a[order], b)


def test_charge_accumulator_overflow_fails_the_run():
    """2^24 ions stacked on one node add 2^24 terms of 2^39 (the stepper's
    scale keeps one term in [2^39, 2^40)), 2^63 in all, one past the int64
    range (a negative charge would end exactly on -2^63, still in range):
    the run reports the accumulator's overflow instead of stepping on a
    wrapped charge density, and stats() shows the wrapped adds."""
    from psum_amd import engine as E
    n = 1 << 24
    g = E.periodic_grid([8, 8])
    sp = [dict(q=1.0, m=1.0, w=1.0 / 16, pos=[np.full(n, 2.5), np.full(n, 3.5)],
               vel=[np.zeros(n), np.zeros(n)])]
    st = E.Stepper(g, sp, dtype="f64", depth=2, max_cycles=4, tol=1e-6, dt=0.1, rho_bg=-n / 16 / 64)
    with pytest.raises(Exception, match="overflow"):
        st.run(1)
    assert st.stats()["acc_wraps"] > 0
    # a fresh load clears the flag; a harmless state runs
    small = 1 << 12
    st.load(0, [np.full(small, 2.5), np.full(small, 3.5)], [np.zeros(small), np.zeros(small)])
    st.run(1)
    assert st.stats()["acc_wraps"] == 0


def test_set_smoothing_equals_creation_options_and_tuner_keeps_a_candidate():
    """set_smoothing(pre, post) after creation steps exactly as a stepper
    created with those sweeps (captured cycles are dropped); the smoothing
    autotuner returns one of its candidates with a time for each, and the
    stepper keeps stepping with it; bad counts are refused."""
    from psum_amd import engine as E
    from psum_amd.workloads import bounded_3d
    wl = bounded_3d(cells=16, ppc=4, seed=8, dtype=np.float32)
    kw = dict(dtype="f32", mixed_mg=True, depth=3, tol=1e-8, max_cycles=40, dt=wl["dt"], rho_bg=0.0,
              fixed=wl["fixed"])
    a = E.Stepper(wl["grid"], wl["species"], pre=5, post=4, **kw)
    ha = a.run(6)
    b = E.Stepper(wl["grid"], wl["species"], pre=2, post=1, **kw)
    b.set_smoothing(5, 4)
    hb = b.run(6)
    assert np.array_equal(ha, hb)
    # mid-run: the solves replayed from captured cycles take the new sweeps
    c = E.Stepper(wl["grid"], wl["species"], pre=1, post=1, **kw)
    c.run(3)
    c0 = c.stats()["cycles"]
    c.run(3)
    per_11 = c.stats()["cycles"] - c0
    c.set_smoothing(6, 6)
    c1 = c.stats()["cycles"]
    c.run(3)
    per_66 = c.stats()["cycles"] - c1
    assert per_66 < per_11, (per_11, per_66)
    best, ms = b.tune_smoothing(((2, 2), (3, 3), (4, 4)), steps=2)
    assert best in ((2, 2), (3, 3), (4, 4)) and len(ms) == 3 and all(v > 0 for v in ms.values())
    assert np.all(np.isfinite(b.run(3)))
    with pytest.raises(Exception, match="smoothing"):
        b.set_smoothing(0, 0)
    with pytest.raises(Exception, match="smoothing"):
        b.set_smoothing(-1, 2)
