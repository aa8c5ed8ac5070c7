"""Benchmark: particle-steps/s of the electrostatic PIC step, MG Poisson solve
included, mixed precision vs FP64, on one MI355X.

Headline workload (`value`): the 3D bounded box, the largest configuration
one GPU holds: 256^3 cells with grounded (phi = 0) walls, 32 electrons + 32
ions per cell (1,073,741,824 particles, mass ratio 1836, Te/Ti = 10),
absorbing particle walls with compaction, 7-point multigrid, FP32 particle
state with FP32 V-cycles inside FP64 residual refinement to
||r||/||b|| <= 1e-8, dt 0.1.  The autotuner picks the sort bin and a sort
interval per species (ions keep their order far longer than electrons).
Every timed window starts on a step where every species sorts; the default
window (200 steps) is a whole number of every species' sort periods, so each
sort is counted at its share (`sorts_in_window`, per species).  A shorter
--steps still times exactly that many steps and counts the sorts at its
start whole, which can only overstate the cost per step.  The same inputs
(cast to FP64) run through the plain FP64 path for the mixed/FP64 ratio.

Beside it: the 2D two-stream case (1024^2 periodic nodes, 100 e-/cell,
mixed and FP64), the plain FP64 2D Langmuir case (128^2, 64 e-/cell), the
differentiable two-stream case (adjoint / forward), the batched ensemble,
Monte Carlo collisions and the triangle-mesh boundary.

One step = deposit (fused into the previous push) + MG solve + field energy
+ gather/push/boundaries + diagnostics, all on the device;
`value` = N_p * K / (wall time of K steps) for the mixed path.

Prints ONE JSON line.  `python bench.py --gpus N` with N > 1 exits: this
engine runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s incl. MG Poisson, mixed vs FP64 on 1 GPU; adjoint/fwd time"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _sort_period(intervals):
    """Least common multiple of the positive sort intervals (0: none)."""
    period = 0
    for k in intervals:
        if k > 0:
            period = k if period == 0 else period * k // math.gcd(period, k)
    return period


def _align_to_sort(st, period):
    """Untimed steps until the step index is a multiple of `period`, the step
    where every species sorts, so that a timed window starts there; returns
    how many were run."""
    if period <= 0:
        return 0
    extra = (-st.stats()["steps"]) % period
    if extra:
        st.run(extra, history=False)
    return extra


def timed_run(st, n, steps, sort_every, absorbed_at_load=0):
    """`steps` timed steps of stepper st (n particles when loaded, the
    absorbed counter then at absorbed_at_load) from a step where every
    species sorts; sort_every: one interval or one per species.  Wall clock
    around the loop, HIP events around every push launch.  Particle-steps
    count the particles alive: K * (alive at the window's start + at its
    end) / 2 (absorption is steady over a window)."""
    import torch
    every = list(sort_every) if isinstance(sort_every, (list, tuple)) else [sort_every]
    extra = _align_to_sort(st, _sort_period(every))
    torch.cuda.synchronize()
    s0 = st.stats()
    st.set_timing(True)
    t0 = time.perf_counter()
    hist = st.run(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    push_ms, launches, pushed = st.push_timing()
    s1 = st.stats()
    sorts = [0 if k <= 0 else (steps + k - 1) // k for k in every]
    sorts = sorts[0] if len(sorts) == 1 else sorts
    alive = n - 0.5 * ((s0["absorbed"] - absorbed_at_load) + (s1["absorbed"] - absorbed_at_load))
    return dict(wall_s=wall, ms_per_step=1e3 * wall / steps, particle_steps_per_s=alive * steps / wall,
                particles_alive_avg=alive,
                push_ms_avg=push_ms / max(1, launches), push_launches=launches,
                particles_per_launch=pushed / max(1, launches),
                cycles_per_step=(s1["cycles"] - s0["cycles"]) / steps,
                last_rel_res=s1["last_rel_res"],
                tile_fallback_blocks_in_window=s1["tile_fallback_blocks"] - s0["tile_fallback_blocks"],
                absorbed_in_window=s1["absorbed"] - s0["absorbed"], sort_every=sort_every, sorts_in_window=sorts,
                align_steps=extra, hist=hist)


def run_engine(E, wl, dtype, mixed, steps, warmup, tol, depth, sort_every, pre=2, post=1,
               graph=True, extrapolate=False, label=""):
    sp = wl["species"]
    t0 = time.time()
    st = E.Stepper(wl["grid"], sp, dtype=dtype, mixed_mg=mixed, depth=depth, pre=pre, post=post,
                   max_cycles=40, tol=tol, dt=wl["dt"], rho_bg=wl["rho_bg"],
                   sort_every=sort_every, mg_graph=graph, extrapolate=extrapolate,
                   fixed=wl.get("fixed"))
    log(f"[bench] {label} {dtype}{'+mixed MG' if mixed else ''}: setup {time.time() - t0:.1f}s")
    if warmup:
        st.run(warmup, history=False)
    n = sum(len(s["pos"][0]) for s in sp)
    res = timed_run(st, n, steps, sort_every)
    st.close()
    log(f"[bench] {label} {dtype}: {res['ms_per_step']:.3f} ms/step, "
        f"{res['particle_steps_per_s']:.3e} particle-steps/s, push {res['push_ms_avg']:.3f} ms, "
        f"{res['cycles_per_step']:.2f} V-cycles/step, {res['sorts_in_window']} sorts in window, "
        f"tile fallbacks {res['tile_fallback_blocks_in_window']}")
    return res


def fe_rel_err(h_mixed, h_fp64):
    """max_k |W_k(mixed) - W_k(FP64)| / |W_k(FP64)| over the steps both
    histories hold (W = field energy), with the step count."""
    k = min(len(h_mixed), len(h_fp64))
    err = np.abs(h_mixed[:k, 0] - h_fp64[:k, 0]) / np.abs(h_fp64[:k, 0])
    return {"max": float(err.max()), "steps": int(k)}


def profiled_traffic(kernel, grid, ppc):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summaries (copy-calibrated) under profiles/, for
    the workload they were measured on (a line "traffic_bytes_per_launch
    <kernel> <grid> <ppc> <bytes>"); None when no summary covers it."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*hbm_traffic*.md"))):
        try:
            for line in open(path):
                f = line.split()
                if len(f) == 5 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel \
                        and int(f[2]) == grid and int(f[3]) == ppc:
                    return float(f[4])
        except OSError:
            pass
    return None


def run_adjoint(E, grid, ppc, steps, ckpt, depth, pre, post):
    """Differentiable two-stream: L = field energy after `steps` steps of the
    FP32 mixed-precision step, differentiated with respect to 32 initial
    Fourier amplitudes and every initial velocity (ModeDisplace +
    PicFieldEnergy through torch.autograd), checkpoint every `ckpt` steps.
    The FP64 path gives the reference gradient of the same set-up."""
    import torch
    from psum_amd.pypsum import ModeDisplace, PicFieldEnergy
    from psum_amd.workloads import two_stream_modes
    wl = two_stream_modes(n=grid, ppc=ppc, nm=32, amp=1e-3, dtype=np.float32)
    sp = wl["species"][0]
    Lb = float(grid)
    n = len(sp["pos"][0])
    out = {"workload": f"2D two-stream {grid}x{grid} periodic nodes, {ppc} e-/cell ({n} particles), "
                       f"{steps} steps, dt 0.05, loss = final field energy, d/d(32 Fourier "
                       f"amplitudes, a_k = 1e-3) and d/d(initial velocities), checkpoint every "
                       f"{ckpt} steps", "particles": n, "steps": steps, "ckpt": ckpt}
    grads = {}
    for dtype in ("f32", "f64"):
        np_dt = np.float32 if dtype == "f32" else np.float64
        tdt = torch.float32 if dtype == "f32" else torch.float64
        spc = [dict(sp, pos=[a.astype(np_dt) for a in sp["pos"]],
                    vel=[a.astype(np_dt) for a in sp["vel"]])]
        st = E.Stepper(wl["grid"], spc, dtype=dtype, mixed_mg=(dtype == "f32"), depth=depth,
                       pre=pre, post=post, max_cycles=40, tol=1e-8, dt=wl["dt"],
                       rho_bg=wl["rho_bg"], sort_every=0)
        xi = torch.from_numpy(spc[0]["pos"][0]).cuda()
        y = torch.from_numpy(spc[0]["pos"][1]).cuda()
        vx = torch.from_numpy(spc[0]["vel"][0]).cuda().requires_grad_(True)
        vy = torch.from_numpy(spc[0]["vel"][1]).cuda().requires_grad_(True)
        amps = torch.from_numpy(wl["amps"]).cuda().requires_grad_(True)
        x = ModeDisplace.apply(xi, amps, 0.0, Lb)
        # plain forward of the same run (diagnostics on), timed by wall clock
        st.load(0, [x.detach(), y], [vx.detach(), vy.detach()])
        st.run(2, history=False)
        st.load(0, [x.detach(), y], [vx.detach(), vy.detach()])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st.run(steps, history=False)
        torch.cuda.synchronize()
        fwd_ms = 1e3 * (time.perf_counter() - t0)
        t0 = time.perf_counter()
        loss = PicFieldEnergy.apply(st, steps, ckpt, x, y, vx, vy)
        torch.cuda.synchronize()
        rec_ms = 1e3 * (time.perf_counter() - t0)
        t0 = time.perf_counter()
        ga, gvx, gvy = torch.autograd.grad(loss, [amps, vx, vy])
        torch.cuda.synchronize()
        rev_ms = 1e3 * (time.perf_counter() - t0)
        grads[dtype] = (ga.cpu().numpy(), gvx.double().cpu().numpy(), gvy.double().cpu().numpy())
        out[dtype] = {"loss": loss.item(), "forward_ms_per_step": fwd_ms / steps,
                      "record_ms_per_step": rec_ms / steps, "reverse_ms_per_step": rev_ms / steps,
                      "reverse_over_forward": rev_ms / fwd_ms,
                      "record_plus_reverse_over_forward": (rec_ms + rev_ms) / fwd_ms}
        log(f"[bench] adjoint {dtype}: forward {fwd_ms / steps:.3f} ms/step, record "
            f"{rec_ms / steps:.3f}, reverse {rev_ms / steps:.3f} ms/step "
            f"(reverse/forward {rev_ms / fwd_ms:.2f}), L = {loss.item():.6e}")
        st.close()
        del x, xi, y, vx, vy
    g32, g64 = grads["f32"], grads["f64"]
    dv = np.concatenate([a - b for a, b in zip(g32[1:], g64[1:])])
    v64 = np.concatenate(g64[1:])
    out["fp32_vs_fp64"] = {
        "grad_amps_rel_err": float(np.abs(g32[0] - g64[0]).max() / np.abs(g64[0]).max()),
        "grad_v_rel_l2_err": float(np.linalg.norm(dv) / np.linalg.norm(v64)),
        "grad_v_p99_abs_err_over_max": float(np.quantile(np.abs(dv), 0.99) / np.abs(v64).max()),
        "grad_v_max_abs_err_over_max": float(np.abs(dv).max() / np.abs(v64).max()),
        "loss_rel_err": abs(out["f32"]["loss"] - out["f64"]["loss"]) / abs(out["f64"]["loss"])}
    out["adjoint_over_forward"] = out["f32"]["reverse_over_forward"]
    return out


def run_batch(E, n_inst, steps, lanes=0, seed=11):
    """Batched ensemble: n_inst independent 2D periodic instances (grids
    from {64^2, 128^2, 256^2, 512^2}, 16-128 e-/cell, perturbation mode
    1..8, amplitude 1e-3..5e-2), FP32 with FP64 MG refinement, autotuned
    (cycle count, sort bin, sort interval), one step of all instances
    replayed as per-lane hipGraphs on concurrent streams (lanes=0: the
    lane count in {1, 2, 4, 8} with the shortest 20-step replay).  With one
    lane the solver cycles of all instances run as batched launches (one per
    phase and kernel variant).  Also times the same instances stepped one
    after another without the graph."""
    from psum_amd.workloads import ensemble
    t0 = time.time()
    specs = ensemble(n_inst, seed, device="cuda")
    meta = [s.pop("meta") for s in specs]
    total = sum(m["particles"] for m in meta)
    b = E.Batch(specs)
    for s in specs:
        s["species"][0]["pos"] = s["species"][0]["vel"] = None
    del specs
    import torch
    torch.cuda.empty_cache()
    log(f"[bench] batch: {n_inst} instances, {total} particles, set up in {time.time() - t0:.1f}s")
    t0 = time.time()
    tun = b.tune(warmup=20, probe=10)
    log(f"[bench] batch: tuned in {time.time() - t0:.1f}s")
    sweep = {}
    if lanes == 0:
        for L in (1, 2, 4, 8):
            b.set_lanes(L)
            b.run(2)              # capture + first replay
            sweep[L] = b.run(20) / 20
        lanes = min(sweep, key=sweep.get)
        log("[bench] batch: lanes " + ", ".join(f"{L}: {t:.2f} ms/step" for L, t in sweep.items()))
    b.set_lanes(lanes)
    b.reset_history(steps + 2)
    b.run(2)                      # capture + first replay
    solver_launches, solver_ops = b.solver_launches()
    ms = b.run(steps)
    # the same instances, same schedules, stepped one after another without the graph
    seq_steps = max(1, min(steps, 20))
    t0 = time.perf_counter()
    for st in b.instances:
        st.run(seq_steps, history=False)
    seq_ms = 1e3 * (time.perf_counter() - t0) / seq_steps
    hist_ok = all(np.all(np.isfinite(b.history(i)[:, 0])) for i in range(len(b)))
    out = {
        "workload": f"{n_inst} independent 2D periodic instances, grids from 64^2..512^2 nodes, "
                    f"16-128 e-/cell, mode 1..8, amplitude 1e-3..5e-2, FP32 + FP64 MG refinement "
                    f"(1e-8 calibration), {steps} steps, {lanes} hipGraph lanes per step",
        "instances": n_inst, "particles": total, "steps": steps, "lanes": lanes,
        "lane_sweep_ms_per_step": {str(k): v for k, v in sweep.items()},
        "ms_per_step": ms / steps, "particle_steps_per_s": total * steps / (ms * 1e-3),
        "solver_launches_per_step": solver_launches, "solver_ops_batched": solver_ops,
        "sequential_ms_per_step": seq_ms, "graph_speedup": seq_ms / (ms / steps),
        "histories_finite": bool(hist_ok),
        "tuning": {
            "fixed_cycles": [t["fixed_cycles"] for t in tun],
            "sort_every": [t["sort_every"] for t in tun],
            "sort_shift": [t["sort_shift"] for t in tun]},
    }
    log(f"[bench] batch: {ms / steps:.3f} ms/step ({out['particle_steps_per_s']:.3e} particle-steps/s), "
        f"sequential {seq_ms:.3f} ms/step, graph speedup {out['graph_speedup']:.2f}")
    b.close()
    return out


def run_mcc(E, n, calls):
    """Monte Carlo collisions of n FP64 electrons (1D3V, 2 eV-ish core with
    a 50 eV tail) against He-like atoms at a constant density on a 1000-cell
    wall grid: elastic + two excitations (one pass), then the same plus
    ionization (counting pass, scan, host check, apply pass; the product
    counts are reset before each call so every call appends into the same
    slots).  Cross sections are synthetic exp-quadratic fits in ln(v2) of
    the reference's HeCCS form.  Times `calls` calls of each with HIP events
    on the launching stream."""
    import torch
    from psum_amd import mcc as M
    me, mhe, x0 = 9.10956e-31, 6.67e-27, 28.0

    def fit(a, b, c, x_on=None):
        row = [0.0, 0.0, c, b - 2 * c * x0, a - b * x0 + c * x0 * x0]
        if x_on is None:
            return dict(seg_hi=[], coef=[row])
        return dict(seg_hi=[x_on], coef=[[0, 0, 0, 0, -200.0], row])
    ch = [M.channel("elastic", **fit(-44.0, -0.4, -0.02)),
          M.channel("excitation", **fit(-46.0, 0.8, -0.15, x0 + 2.16), threshold_ev=19.82),
          M.channel("excitation", **fit(-46.5, 0.7, -0.12, x0 + 2.2), threshold_ev=20.61)]
    ion = M.channel("ionization", **fit(-45.0, 0.9, -0.1, x0 + 2.37), threshold_ev=24.587)
    kw = dict(incident_mass=me, target_mass=mhe, target_temperature=300.0)
    no_ion = [M.Process(ch, 3e21, **kw)]
    with_ion = [M.Process(ch + [ion], 3e21, **kw)]
    g = E.wall_grid([1000], dx=0.067 / 1000)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    cap = n + n // 20
    sp = M.Particles.empty(1, cap)
    sp.pos[0][:n] = torch.rand(n, generator=gen, device="cuda", dtype=torch.float64) * 0.067
    tail = (torch.rand(n, generator=gen, device="cuda") < 0.1).double()
    for c in range(3):
        sp.vel[c][:n] = torch.randn(n, generator=gen, device="cuda", dtype=torch.float64) * (6e5 + 2.4e6 * tail)
    sp.seed[:n] = torch.randint(-2**31, 2**31 - 1, (n,), generator=gen, device="cuda", dtype=torch.int32)
    sp.n = n
    ions = M.Particles.empty(1, n // 20)
    del tail
    dt = 1e-11
    out = {"workload": f"{n} FP64 electrons (1D3V) against He-like atoms at 3e21 m^-3, 1000-cell "
                       f"wall grid, dt 1e-11, synthetic HeCCS-form cross sections", "particles": n}
    for name, procs in (("no_ionization", no_ion), ("with_ionization", with_ion)):
        made = 0
        for _ in range(2):
            sp.n, ions.n = n, 0
            made, _ = M.collide(g, procs, sp, dt, electrons=sp, ions=ions, grow=False)
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(calls):
            sp.n, ions.n = n, 0
            made, _ = M.collide(g, procs, sp, dt, electrons=sp, ions=ions, grow=False)
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / calls
        # algorithmic bytes: read x, v (3), seed; write seed (velocities of
        # the few collided particles and the products are not counted)
        algo = n * (8 + 24 + 4 + 4)
        out[name] = {"ms_per_call": ms, "particles_per_s": n / (ms * 1e-3),
                     "algorithmic_GBps": algo / (ms * 1e-3) / 1e9,
                     "hbm_frac": algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "products_per_call": made}
        log(f"[bench] mcc {name}: {ms:.3f} ms/call for {n} electrons ({n / (ms * 1e-3):.3e}/s, "
            f"{out[name]['algorithmic_GBps']:.0f} GB/s algorithmic), {made} product pairs")
    del sp, ions
    torch.cuda.empty_cache()
    return out


def run_mesh(E, n, steps):
    """Triangle-mesh boundary: n FP64 particles (3D) in free flight through
    a scene of the reference's router demonstration (an absorbing box, a
    specular 64-segment nozzle, a diffuse plane; 100^3 partition).  Each step moves x += v dt with torch and calls the router;
    only the router calls are timed (HIP events on the launching stream)."""
    import torch
    from psum_amd import mesh as Mh
    noz = Mh.revolve_zr_curve([(-4, 1.0), (-3.5, 1.0), (-3, 0.9), (-2.5, 0.8), (-2, 0.6), (-1.5, 0.5), (-1, 0.6),
                               (-0.5, 0.8), (0, 1.0), (0.5, 1.2), (1, 1.4), (1.5, 1.6), (2, 1.8), (2.5, 1.9),
                               (3, 2.0), (3.5, 2.0), (4, 2.0)], 64)
    scene = [(Mh.axis_aligned_cube(-5, -5, -5, 5, 5, 10), "absorb"), (noz, "specular"),
             (Mh.axis_aligned_plane(-5, -5, -4.0, 5, 5, -4.0), "diffuse")]
    router = Mh.BoundaryRouter().set(scene, 100, 100, 100)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(9)
    r = torch.sqrt(torch.rand(n, generator=gen, device="cuda", dtype=torch.float64)) * 0.7
    th = 2 * np.pi * torch.rand(n, generator=gen, device="cuda", dtype=torch.float64)
    pos = [r * torch.cos(th), r * torch.sin(th),
           -4 + 0.1 * torch.rand(n, generator=gen, device="cuda", dtype=torch.float64)]
    vel = [torch.randn(n, generator=gen, device="cuda", dtype=torch.float64) for _ in range(3)]
    seed = torch.randint(-2**31, 2**31 - 1, (n,), generator=gen, device="cuda", dtype=torch.int32)
    del r, th
    dt = 0.005
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    old = [torch.empty_like(p) for p in pos]
    hits = 0
    for k in range(steps):
        for c in range(3):
            old[c].copy_(pos[c])
            pos[c].add_(vel[c] * dt)
        ev[k][0].record()
        h, _ = router.deal(old, pos, vel, seed)
        ev[k][1].record()
        hits += h
    torch.cuda.synchronize()
    ms = sum(a.elapsed_time(b) for a, b in ev[1:]) / (steps - 1)
    alive = int(torch.isfinite(pos[0]).sum().item())
    out = {"workload": f"{n} FP64 particles, router scene (box/nozzle/plane, {len(router.trigger.tri)} triangles, "
                       f"100^3 cells), dt 0.005, {steps} steps", "particles": n,
           "ms_per_call": ms, "segments_per_s": n / (ms * 1e-3), "hits": hits, "alive_at_end": alive,
           "algorithmic_GBps": n * 48 / (ms * 1e-3) / 1e9}
    log(f"[bench] mesh: {ms:.3f} ms per router call for {n} particles ({n / (ms * 1e-3):.3e} segments/s), "
        f"{hits} hits over {steps} steps, {alive} alive")
    router.trigger.close()
    del pos, vel, old, seed
    torch.cuda.empty_cache()
    return out


def c3_workload(cells, ppc):
    from psum_amd.workloads import bounded_3d
    return bounded_3d(cells=cells, ppc=ppc, device="cuda")


# (pre, post) smoothing sweeps the 3D box's autotuner chooses from
SMOOTHING_CANDIDATES = ((3, 3), (4, 4), (5, 5), (6, 6))


def run_c3(E, wl, dtype, steps, warmup, depth, pre, post, schedule=None, tune_probe=20,
           extrapolate=False, tune_smoothing=True):
    """3D bounded box (phi = 0 on the faces, absorbing walls, compaction by
    the sort), electrons + ions (mass ratio 1836, Te/Ti = 10), 7-point MG.
    dtype "f32": FP32 particles with FP32 V-cycles in FP64 refinement to
    1e-8; "f64": the plain FP64 path on the same inputs.  schedule = (sort
    every, sort shift) or None to run the autotuner first; tune_smoothing:
    the smoothing autotuner picks (pre, post) from SMOOTHING_CANDIDATES,
    else (pre, post) as given."""
    import torch
    t0 = time.time()
    tdt = torch.float32 if dtype == "f32" else torch.float64
    species = [dict(s, pos=[a.to(tdt) for a in s["pos"]], vel=[a.to(tdt) for a in s["vel"]])
               for s in wl["species"]]
    n = sum(len(s["pos"][0]) for s in species)
    st = E.Stepper(wl["grid"], species, dtype=dtype, mixed_mg=(dtype == "f32"), depth=depth,
                   pre=pre, post=post, max_cycles=40, tol=1e-8, dt=wl["dt"], rho_bg=0.0,
                   sort_every=20, fixed=wl["fixed"], extrapolate=extrapolate)
    del species
    torch.cuda.empty_cache()
    log(f"[bench] c3 {dtype}: {n} particles, set up in {time.time() - t0:.1f}s")
    tun = None
    if schedule is None:
        t0 = time.time()
        tun = st.tune(warmup=3, probe=tune_probe)
        schedule = (tun["sort_every"], tun["sort_shift"], tun["species_sort_every"])
        log(f"[bench] c3: tuned in {time.time() - t0:.1f}s: sort every {schedule[2]} steps (per species) "
            f"on {2 ** schedule[1]}^3-cell bins")
    smoothing = (pre, post)
    smoothing_ms = None
    if tune_smoothing:
        t0 = time.time()
        smoothing, smoothing_ms = st.tune_smoothing(SMOOTHING_CANDIDATES, steps=8)
        log(f"[bench] c3 {dtype}: smoothing tuned in {time.time() - t0:.1f}s: pre/post {smoothing} ("
            + ", ".join(f"{c}: {v:.2f} ms" for c, v in smoothing_ms.items()) + ")")
    absorbed0 = 0
    if tun is not None or tune_smoothing:
        # tuning advanced the state: start again from the initial particles,
        # so that the mixed and FP64 runs time the same physical steps
        for i, s in enumerate(wl["species"]):
            st.load(i, [a.to(tdt) for a in s["pos"]], [a.to(tdt) for a in s["vel"]])
        absorbed0 = st.stats()["absorbed"]
    st.set_schedule(0, schedule[0], schedule[1])
    every = [schedule[0]] * len(wl["species"])
    if len(schedule) > 2:
        every = list(schedule[2])
        for q, k in enumerate(every):
            st.set_species_sort(q, k)
    if warmup:
        st.run(warmup, history=False)
    res = timed_run(st, n, steps, every, absorbed0)
    res["particles"] = n
    res["schedule"] = {"sort_every": every, "sort_shift": schedule[1]}
    res["smoothing"] = list(smoothing)
    if smoothing_ms is not None:
        res["smoothing_probe_ms_per_step"] = {f"{c[0]},{c[1]}": v for c, v in smoothing_ms.items()}
    res["field_energy_last"] = float(res["hist"][-1, 0])
    st.close()
    torch.cuda.empty_cache()
    log(f"[bench] c3 {dtype}: {res['ms_per_step']:.2f} ms/step, {res['particle_steps_per_s']:.3e} "
        f"particle-steps/s, push {res['push_ms_avg']:.3f} ms, {res['cycles_per_step']:.2f} "
        f"V-cycles/step, {res['sorts_in_window']} sorts in window, absorbed {res['absorbed_in_window']}, "
        f"tile fallbacks {res['tile_fallback_blocks_in_window']} (in the window)")
    return res, schedule


def cpu_baseline_c3(cells, ppc, steps):
    """The CPU oracle (FP64 restatement of the reference step, OpenMP over the
    host cores) on a bounded sample of the 3D bounded-box workload: the same
    particles per cell, species, temperatures, walls and solver on a smaller
    box, `steps` steps, reported per particle-step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from psum_amd.workloads import bounded_3d
    wl = bounded_3d(cells=cells, ppc=ppc, dtype=np.float64)
    og = O.wall_grid([cells] * 3)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    species = [dict(q=s["q"], m=s["m"], w=s["w"], pos=list(s["pos"]), vel=list(s["vel"]))
               for s in wl["species"]]
    n = sum(len(s["pos"][0]) for s in species)
    t0 = time.perf_counter()
    O.pic_run_multi(og, species, 0.0, 1.0, wl["dt"], steps, 5, pre=3, post=3, tol=1e-8,
                    max_cycles=40, fixed=wl["fixed"])
    t = max(1e-9, time.perf_counter() - t0)
    log(f"[bench] cpu oracle: {cells}^3 box, {n} particles, {steps} steps in {t:.1f}s "
        f"(MG set-up inside), {cores} threads")
    return {"value": n * steps / t, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} steps of the 3D bounded-box workload on a {cells}^3-cell box "
                      f"({ppc} e- + {ppc} ions per cell, {n} particles; the headline box is 256^3), "
                      f"FP64 oracle restatement of the reference step, OpenMP, MG set-up included"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200,
                    help="timed steps of the headline (3D box) run (200: whole sort periods of every species)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--pre", type=int, default=3)
    ap.add_argument("--post", type=int, default=3)
    ap.add_argument("--c3-cells", type=int, default=256)
    ap.add_argument("--c3-ppc", type=int, default=32)
    ap.add_argument("--c3-depth", type=int, default=6)
    ap.add_argument("--c3-no-fp64", action="store_true")
    ap.add_argument("--c3-fixed-smoothing", action="store_true",
                    help="3D box: --pre/--post sweeps instead of the smoothing autotuner")
    ap.add_argument("--c3-extrapolate", action="store_true",
                    help="warm-start each 3D solve from 2 phi_n - phi_(n-1)")
    ap.add_argument("--c3-schedule", default=None,
                    help="'every,shift': sort schedule of the 3D box instead of the autotuner's")
    ap.add_argument("--grid", type=int, default=1024, help="2D two-stream grid")
    ap.add_argument("--ppc", type=int, default=100)
    ap.add_argument("--depth", type=int, default=5)
    ap.add_argument("--c2-steps", type=int, default=100)
    ap.add_argument("--sort-every", type=int, default=25, help="2D two-stream sort interval")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--c1-steps", type=int, default=100)
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--extrapolate", action="store_true")
    ap.add_argument("--cpu-cells", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-grid", type=int, default=256)
    ap.add_argument("--adjoint-ppc", type=int, default=64)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--batch-instances", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--batch-lanes", type=int, default=1,
                    help="concurrent graph lanes for the batch (0: pick from 1, 2, 4, 8)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--mcc-particles", type=int, default=100_000_000)
    ap.add_argument("--mcc-calls", type=int, default=20)
    ap.add_argument("--no-mcc", action="store_true")
    ap.add_argument("--mesh-particles", type=int, default=20_000_000)
    ap.add_argument("--mesh-steps", type=int, default=20)
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the headline run, copy one buffer of the 3D push kernel's "
                         "algorithmic size so that HBM counters can be scaled by a plain copy "
                         "of known size")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from psum_amd import engine as E
    from psum_amd.workloads import langmuir_2d, two_stream_2d
    assert torch.cuda.is_available(), "bench.py needs a GPU"

    # ---- headline: 3D bounded box, mixed precision, and the FP64 path on the same inputs
    t0 = time.time()
    wl3 = c3_workload(args.c3_cells, args.c3_ppc)
    log(f"[bench] c3 workload {wl3['name']}: generated in {time.time() - t0:.1f}s")
    fixed_sched = None if args.c3_schedule is None else tuple(int(v) for v in args.c3_schedule.split(","))
    c3, sched = run_c3(E, wl3, "f32", args.steps, args.warmup, args.c3_depth, args.pre, args.post,
                       schedule=fixed_sched, extrapolate=args.c3_extrapolate,
                       tune_smoothing=not args.c3_fixed_smoothing)
    if args.calib_copy:
        nb = int(c3["particles_per_launch"]) * 2 * 3
        src = torch.ones(nb, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {src.numel() * 4} bytes read, {dst.numel() * 4} written")
        del src, dst
    c3_64 = None
    if not args.c3_no_fp64:
        c3_64, _ = run_c3(E, wl3, "f64", args.steps, args.warmup, args.c3_depth, args.pre,
                          args.post, schedule=sched, extrapolate=args.c3_extrapolate,
                          tune_smoothing=not args.c3_fixed_smoothing)
    del wl3
    torch.cuda.empty_cache()

    # dominant kernel: the fused 3D gather + push + deposit; algorithmic bytes
    # per launch = particles per launch * 2 * (2D) * 4 B (read + write FP32
    # position and velocity once)
    kern = "k_push_deposit<float,3,true,true,2>"
    bytes_per_launch = c3["particles_per_launch"] * 2 * (2 * 3) * 4
    achieved = bytes_per_launch / (c3["push_ms_avg"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": profiled_traffic(kern, args.c3_cells, args.c3_ppc),
                "kernel": kern, "bytes_per_launch": bytes_per_launch,
                "avg_launch_ms": c3["push_ms_avg"]}

    # ---- 2D two-stream (mixed and FP64) and the plain FP64 Langmuir case
    c2 = None
    if not args.no_c2:
        t0 = time.time()
        wl = two_stream_2d(n=args.grid, ppc=args.ppc, dtype=np.float32)
        log(f"[bench] c2 workload {wl['name']}: generated in {time.time() - t0:.1f}s")
        m2 = run_engine(E, wl, "f32", True, args.c2_steps, args.warmup, 1e-8, args.depth,
                        args.sort_every, args.pre, args.post, not args.no_graph, args.extrapolate,
                        label="c2")
        wl64 = dict(wl)
        wl64["species"] = [dict(s, pos=[a.astype(np.float64) for a in s["pos"]],
                                vel=[a.astype(np.float64) for a in s["vel"]]) for s in wl["species"]]
        f2 = run_engine(E, wl64, "f64", False, args.c2_steps, args.warmup, 1e-8, args.depth,
                        args.sort_every, args.pre, args.post, not args.no_graph, args.extrapolate,
                        label="c2")
        del wl, wl64
        c2 = {"workload": f"2D two-stream, {args.grid}x{args.grid} periodic nodes, {args.ppc} e-/cell, "
                          f"dt 0.05; mixed: FP32 state, FP32 V-cycles in FP64 refinement to 1e-8",
              "particles": args.grid * args.grid * args.ppc, "steps": args.c2_steps,
              "mixed": _leg(m2), "fp64": _leg(f2),
              "mixed_over_fp64": m2["particle_steps_per_s"] / f2["particle_steps_per_s"],
              "field_energy_rel_err_mixed_vs_fp64": fe_rel_err(m2["hist"], f2["hist"]),
              "push_GBps": m2["particles_per_launch"] * 32 / (m2["push_ms_avg"] * 1e-3) / 1e9}
    c1 = None
    if not args.no_c1:
        wl = langmuir_2d(n=128, ppc=64, seed=1)
        f1 = run_engine(E, wl, "f64", False, args.c1_steps, args.warmup, 1e-10, 4, 0, 2, 1,
                        not args.no_graph, label="c1")
        wl32 = dict(wl)
        wl32["species"] = [dict(s, pos=[a.astype(np.float32) for a in s["pos"]],
                                vel=[a.astype(np.float32) for a in s["vel"]]) for s in wl["species"]]
        m1 = run_engine(E, wl32, "f32", True, args.c1_steps, args.warmup, 1e-10, 4, 0, 2, 1,
                        not args.no_graph, label="c1")
        c1 = {"workload": "2D Langmuir, 128x128 periodic nodes, 64 e-/cell (1048576 particles), "
                          "density 1 + 0.01 cos(2 pi x/L), dt 0.1, MG to 1e-10",
              "steps": args.c1_steps, "fp64": _leg(f1), "mixed": _leg(m1),
              "mixed_over_fp64": m1["particle_steps_per_s"] / f1["particle_steps_per_s"],
              "field_energy_rel_err_mixed_vs_fp64": fe_rel_err(m1["hist"], f1["hist"])}
    adjoint = None
    if not args.no_adjoint:
        adjoint = run_adjoint(E, args.adjoint_grid, args.adjoint_ppc, args.adjoint_steps, 20,
                              args.depth, args.pre, args.post)
    batch = None
    if not args.no_batch:
        batch = run_batch(E, args.batch_instances, args.batch_steps, args.batch_lanes)
    mcc = None
    if not args.no_mcc:
        mcc = run_mcc(E, args.mcc_particles, args.mcc_calls)
    mesh = None
    if not args.no_mesh:
        mesh = run_mesh(E, args.mesh_particles, args.mesh_steps)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_c3(args.cpu_cells, args.c3_ppc, args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": c3["particle_steps_per_s"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": c3["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"3D bounded box, {args.c3_cells}^3 cells, {args.c3_ppc} e- + "
                               f"{args.c3_ppc} ions per cell ({c3['particles']} particles, mass "
                               f"ratio 1836, Te/Ti 10), grounded walls, absorbing boundary + "
                               f"compaction, 7-point MG, FP32 state, FP32 V-cycles in FP64 "
                               f"refinement to 1e-8, dt 0.1",
                   "cells": [args.c3_cells] * 3, "particles": c3["particles"],
                   "mg_depth": args.c3_depth, "mg_pre_post": c3["smoothing"],
                   "sort_every": c3["schedule"]["sort_every"], "sort_bin_cells": 2 ** sched[1],
                   "sorts_in_window": c3["sorts_in_window"], "parallelism": "single GPU"},
        "fp64": None if c3_64 is None else _leg(c3_64),
        "mixed_over_fp64": None if c3_64 is None else
            c3["particle_steps_per_s"] / c3_64["particle_steps_per_s"],
        "field_energy_rel_err_mixed_vs_fp64": None if c3_64 is None else
            fe_rel_err(c3["hist"], c3_64["hist"]),
        "headline": _leg(c3),
        "c2": c2,
        "c1": c1,
        "adjoint": adjoint,
        "batch": batch,
        "mcc": mcc,
        "mesh": mesh,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def _leg(r):
    """The JSON fields of one timed run."""
    keys = ("ms_per_step", "particle_steps_per_s", "push_ms_avg", "push_launches",
            "particles_per_launch", "particles_alive_avg", "cycles_per_step", "last_rel_res",
            "tile_fallback_blocks_in_window",
            "absorbed_in_window", "sort_every", "sorts_in_window", "align_steps")
    return {k: r[k] for k in keys if k in r}


if __name__ == "__main__":
    main()
