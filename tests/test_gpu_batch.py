"""Batched ensemble: independent instances stepped by one replayed hipGraph
must reproduce, bit for bit, each instance stepped alone with the same
solver and sort schedule; the autotuner's choices must keep the solve at
its tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _instance(n, ppc, k, amp, seed, dtype="f32"):
    from psum_amd import engine as E
    rng = np.random.default_rng(seed)
    L = float(n)
    N = n * n * ppc
    xi = rng.uniform(0, L, N)
    x = np.mod(xi + amp * L / (2 * np.pi * k) * np.sin(2 * np.pi * k * xi / L), L)
    y = rng.uniform(0, L, N)
    vx = rng.standard_normal(N)
    vy = rng.standard_normal(N)
    npd = np.float32 if dtype == "f32" else np.float64
    sp = [dict(q=-1.0, m=1.0, w=1.0 / ppc, pos=[x.astype(npd), y.astype(npd)],
               vel=[vx.astype(npd), vy.astype(npd)])]
    depth = max(2, int(np.log2(n)) - 3)
    return dict(g=E.periodic_grid([n, n]), species=sp, dtype=dtype, mixed_mg=(dtype == "f32"),
                depth=depth, pre=2, post=2, max_cycles=40, tol=1e-8, dt=0.1, rho_bg=1.0)


@pytest.mark.parametrize("dtype,lanes", [("f32", (1, 1)), ("f64", (1, 1)), ("f32", (3, 3)),
                                         ("f64", (2, 2)), ("f32", (1, 3))])
def test_batch_matches_instances_stepped_alone(dtype, lanes):
    """lanes: the lane count for the first and second half of the steps (a
    change between runs captures the lane graphs again)"""
    from psum_amd import engine as E
    specs = [_instance(32, 16, 1, 0.05, 1, dtype), _instance(64, 4, 3, 0.02, 2, dtype),
             _instance(48, 9, 2, 0.01, 3, dtype)]
    sched = [(6, 5, 1), (7, 0, 0), (5, 3, 2)]
    steps = 12
    b = E.Batch(specs)
    for i, (c, k, sh) in enumerate(sched):
        b.configure(i, c, k, sh)
    b.reset_history(steps)
    b.set_lanes(lanes[0])
    b.run(steps // 2)
    b.set_lanes(lanes[1])
    b.run(steps - steps // 2)
    for i, spec in enumerate(specs):
        a = dict(spec)
        g, species = a.pop("g"), a.pop("species")
        st = E.Stepper(g, species, **a)
        st.set_schedule(*sched[i])
        h_ref = st.run(steps)
        h = b.history(i)
        assert h.shape == (steps, 5)
        assert np.array_equal(h, h_ref), (i, np.abs(h - h_ref).max())
        p_ref, v_ref = st.state(0)
        p, v = b.instances[i].state(0)
        for x, y in zip(p + v, p_ref + v_ref):
            assert np.array_equal(x, y)
        st.close()
    b.close()


def test_autotuned_batch_holds_the_tolerance():
    """After psum_batch_tune, fixed cycle counts must still meet the
    tolerance the warm-up calibrated them on: the field energy of the
    batched run matches a tolerance-controlled run of each instance."""
    from psum_amd import engine as E
    specs = [_instance(32, 16, 1, 0.05, 4), _instance(64, 16, 2, 0.03, 5),
             _instance(128, 4, 5, 0.02, 6)]
    b = E.Batch(specs)
    tun = b.tune(warmup=10, probe=6)
    for t in tun:
        assert 2 <= t["fixed_cycles"] <= 41
        assert t["sort_every"] in (5, 10, 20, 50, 100, 200)
        assert t["sort_shift"] in (0, 1, 2)
        assert t["sort_ms"] > 0 and t["push_ms"] > 0
    # tuning advanced every instance; restart all from the initial state
    for i, spec in enumerate(specs):
        b.instances[i].load(0, spec["species"][0]["pos"], spec["species"][0]["vel"])
    steps = 20
    b.reset_history(steps)
    b.run(steps)
    for i, spec in enumerate(specs):
        a = dict(spec)
        g, species = a.pop("g"), a.pop("species")
        a["tol"] = 1e-10
        ref = E.Stepper(g, species, **a).run(steps)
        h = b.history(i)
        rel = np.abs(h[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])
        assert rel.max() < 1e-4, (i, rel.max())
    b.close()


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_batch_with_vcycle_tail_matches_instances_stepped_alone(dtype):
    """The one-workgroup V-cycle tail switched on in every instance (its
    LDS size and kernel attribute are fixed when the hierarchy is built,
    outside the batch's stream capture): the batch still reproduces each
    instance stepped alone, and both reproduce the instances without the
    tail bit for bit."""
    from psum_amd import engine as E
    # 32^2 at depth 2 ends at 8^2 and 64^2 at depth 3 ends at 8^2, so both
    # run their coarse levels in the tail; 48^2 ends at 12^2 (no tail)
    specs = [_instance(32, 16, 1, 0.05, 11, dtype), _instance(64, 4, 3, 0.02, 12, dtype),
             _instance(48, 9, 2, 0.01, 13, dtype)]
    sched = [(6, 5, 1), (7, 0, 0), (5, 3, 2)]
    steps = 8
    b = E.Batch(specs)
    for i, (c, k, sh) in enumerate(sched):
        b.configure(i, c, k, sh)
        b.instances[i].mg_tuning("tail 1")
    b.reset_history(steps)
    b.run(steps)
    for i, spec in enumerate(specs):
        a = dict(spec)
        g, species = a.pop("g"), a.pop("species")
        hs = []
        for tail in (1, 0):
            st = E.Stepper(g, species, **a)
            st.mg_tuning(f"tail {tail}")
            st.set_schedule(*sched[i])
            hs.append(st.run(steps))
            st.close()
        h = b.history(i)
        assert np.array_equal(h, hs[0]), (i, np.abs(h - hs[0]).max())
        assert np.array_equal(hs[0], hs[1]), (i, np.abs(hs[0] - hs[1]).max())
    b.close()


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_batched_solver_launches_match_per_instance_launches(dtype):
    """With solver batching (the default with one lane), every phase of the
    instances' V-cycles runs as one launch over all instances; the results
    equal, bit for bit, the step with each solver launching its own kernels
    and each instance stepped alone."""
    from psum_amd import engine as E
    specs = [_instance(32, 16, 1, 0.05, 21, dtype), _instance(64, 4, 3, 0.02, 22, dtype),
             _instance(48, 9, 2, 0.01, 23, dtype), _instance(128, 2, 4, 0.03, 24, dtype)]
    sched = [(6, 5, 1), (7, 0, 0), (5, 3, 2), (4, 2, 1)]
    steps = 9
    hists = []
    for batching in (True, False):
        b = E.Batch(specs)
        for i, (c, k, sh) in enumerate(sched):
            b.configure(i, c, k, sh)
        b.set_lanes(1)
        b.set_solver_batching(batching)
        b.reset_history(steps)
        b.run(steps)
        launches, ops = b.solver_launches()
        if batching:
            # one launch per phase instead of one per instance and phase
            assert 0 < launches < ops
        else:
            assert launches == 0
        hists.append([b.history(i) for i in range(len(specs))])
        if batching:
            states = [b.instances[i].state(0) for i in range(len(specs))]
        b.close()
    for i, spec in enumerate(specs):
        assert np.array_equal(hists[0][i], hists[1][i]), i
        a = dict(spec)
        g, species = a.pop("g"), a.pop("species")
        st = E.Stepper(g, species, **a)
        st.set_schedule(*sched[i])
        h_ref = st.run(steps)
        assert np.array_equal(hists[0][i], h_ref), i
        p_ref, v_ref = st.state(0)
        for x, y in zip(states[i][0] + states[i][1], p_ref + v_ref):
            assert np.array_equal(x, y), i
        st.close()
