"""Seeded synthetic initial conditions of the benchmark workloads.

Normalised units: omega_p = 1, lambda_D = dx = 1, v_th = 1, electron
q/m = -1, eps = 1.  Each generator returns a dict with the grid, the species
list (q, m, w, pos, vel as numpy arrays) and the step parameters, so the
same arrays feed the HIP engine and the CPU oracle.
"""
import numpy as np

from .engine import periodic_grid, wall_grid


def _inverse_cdf_cos(u, L, amp, k=1):
    """x in [0, L) with density (1 + amp*cos(2 pi k x / L)) / L, by Newton on
    F(x) = x + amp L/(2 pi k) sin(2 pi k x / L) = u L."""
    x = u * L
    c = amp * L / (2 * np.pi * k)
    for _ in range(30):
        f = x + c * np.sin(2 * np.pi * k * x / L) - u * L
        fp = 1 + amp * np.cos(2 * np.pi * k * x / L)
        x = x - f / fp
    return np.mod(x, L)


def langmuir_2d(n=128, ppc=64, amp=0.01, seed=1, dtype=np.float64):
    """Plain 2D electrostatic PIC: n x n periodic nodes, ppc electrons per
    cell over an immobile uniform ion background, Maxwellian v (v_th = 1),
    density 1 + amp cos(2 pi x / L), CIC, dt = 0.1."""
    rng = np.random.default_rng(seed)
    N = n * n * ppc
    L = float(n)
    x = _inverse_cdf_cos(rng.random(N), L, amp)
    y = rng.random(N) * L
    vx = rng.standard_normal(N)
    vy = rng.standard_normal(N)
    sp = dict(q=-1.0, m=1.0, w=1.0 / ppc, pos=[x.astype(dtype), y.astype(dtype)],
              vel=[vx.astype(dtype), vy.astype(dtype)])
    return dict(grid=periodic_grid([n, n]), species=[sp], dt=0.1, rho_bg=1.0, eps=1.0,
                name=f"langmuir2d_{n}x{n}_ppc{ppc}")


def two_stream_2d(n=1024, ppc=100, drift=5.0, noise=1e-3, seed=2, dtype=np.float32):
    """2D two-stream instability: n x n periodic nodes, ppc electrons per cell
    on a sqrt(ppc) x sqrt(ppc) lattice in every cell plus `noise` dx seeded
    jitter, counter-drifting beams v_x = +-drift + N(0,1) (alternating),
    v_y = N(0,1), dt = 0.05."""
    rng = np.random.default_rng(seed)
    m = int(round(np.sqrt(ppc)))
    assert m * m == ppc
    N = n * n * ppc
    L = float(n)
    # lattice: cell (i, j), sub-position (a, b)
    sub = (np.arange(m, dtype=np.float64) + 0.5) / m
    ci = np.repeat(np.arange(n, dtype=np.float64), n * ppc)
    cj = np.tile(np.repeat(np.arange(n, dtype=np.float64), ppc), n)
    sa = np.tile(np.repeat(sub, m), n * n)
    sb = np.tile(sub, n * n * m)
    x = ci + sa + noise * rng.standard_normal(N)
    y = cj + sb + noise * rng.standard_normal(N)
    x = np.mod(x, L)
    y = np.mod(y, L)
    sign = np.where(np.arange(N) % 2 == 0, 1.0, -1.0)
    vx = sign * drift + rng.standard_normal(N)
    vy = rng.standard_normal(N)
    sp = dict(q=-1.0, m=1.0, w=1.0 / ppc, pos=[x.astype(dtype), y.astype(dtype)],
              vel=[vx.astype(dtype), vy.astype(dtype)])
    return dict(grid=periodic_grid([n, n]), species=[sp], dt=0.05, rho_bg=1.0, eps=1.0,
                name=f"two_stream2d_{n}x{n}_ppc{ppc}")


def bounded_3d(cells=256, ppc=32, mass_ratio=1836.0, te_ti=10.0, seed=3, dtype=np.float32,
               device=None):
    """3D bounded box: cells^3 cells, phi = 0 on all faces, absorbing walls,
    ppc electrons + ppc ions per cell, uniform, Maxwellian (Te/Ti = te_ti),
    ion mass mass_ratio, dt = 0.1.  device="cuda" draws the particles on
    the GPU (torch generator) and returns CUDA tensors."""
    rng = np.random.default_rng(seed)
    N = cells ** 3 * ppc
    L = float(cells)
    species = []
    for si, (q, m, vth) in enumerate(((-1.0, 1.0, 1.0),
                                      (1.0, mass_ratio, np.sqrt(1.0 / te_ti / mass_ratio)))):
        if device is None:
            pos = [(rng.random(N) * L).astype(dtype) for _ in range(3)]
            vel = [(rng.standard_normal(N) * vth).astype(dtype) for _ in range(3)]
        else:
            import torch
            tdt = torch.float32 if dtype == np.float32 else torch.float64
            gen = torch.Generator(device=device).manual_seed(seed * 10 + si)
            pos = [torch.rand(N, generator=gen, device=device, dtype=tdt) * L for _ in range(3)]
            vel = [torch.randn(N, generator=gen, device=device, dtype=tdt) * vth for _ in range(3)]
        species.append(dict(q=q, m=m, w=1.0 / ppc, pos=pos, vel=vel))
    g = wall_grid([cells] * 3)
    shp = (cells + 1,) * 3
    fixed = np.zeros(shp, np.uint8)
    fixed[0], fixed[-1] = 1, 1
    fixed[:, 0], fixed[:, -1] = 1, 1
    fixed[:, :, 0], fixed[:, :, -1] = 1, 1
    return dict(grid=g, species=species, dt=0.1, rho_bg=0.0, eps=1.0, fixed=fixed.ravel(),
                name=f"bounded3d_{cells}^3_ppc{ppc}x2")


def mode_basis(xi, nm, L):
    """B[k-1, p] = L/(2 pi k) sin(2 pi k xi_p / L): the displacement field of
    Fourier perturbation mode k (x = xi + sum_k a_k B[k-1])."""
    k = np.arange(1, nm + 1, dtype=np.float64)[:, None]
    return L / (2 * np.pi * k) * np.sin(2 * np.pi * k * np.asarray(xi, np.float64)[None, :] / L)


def two_stream_modes(n=256, ppc=64, nm=32, amp=1e-3, seed=5, dtype=np.float32):
    """Differentiable two-stream set-up: the two_stream_2d beams on n x n
    periodic nodes, with x = xi + sum_{k<=nm} a_k L/(2 pi k) sin(2 pi k xi/L)
    (a_k = amp for every k) applied on the device, so that the loss can be
    differentiated with respect to the nm amplitudes and the initial
    velocities.  Returns the two_stream_2d dict plus xi, amps, nm."""
    wl = two_stream_2d(n=n, ppc=ppc, seed=seed, dtype=dtype)
    wl["xi"] = wl["species"][0]["pos"][0]
    wl["amps"] = np.full(nm, amp, np.float64)
    wl["nm"] = nm
    wl["name"] = f"two_stream2d_modes_{n}x{n}_ppc{ppc}_nm{nm}"
    return wl


def ensemble(n_inst=64, seed=11, dtype=np.float32, sizes=(64, 128, 256, 512), ppc_range=(16, 128),
             k_range=(1, 8), amp_range=(1e-3, 5e-2), device=None):
    """Batched-ensemble instances: n_inst independent 2D periodic problems,
    grid edge drawn uniformly from `sizes` (nodes), electrons per cell from
    ppc_range, a perturbation of mode k and amplitude amp (k, amp drawn from
    k_range / amp_range) imposed by the displacement
    x = xi + amp L/(2 pi k) sin(2 pi k xi / L) of uniform xi (density
    1 - amp cos(2 pi k x / L) to first order), Maxwellian v (v_th = 1) over a
    uniform ion background, dt = 0.1, FP32 particles with FP32 V-cycles in
    FP64 refinement to 1e-8.  device="cuda" draws the particles on the GPU
    (torch generator seeded per instance) and returns CUDA tensors.
    Returns a list of Stepper keyword dicts (g, species, ...) with a 'meta'
    dict each."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inst):
        n = int(rng.choice(sizes))
        ppc = int(rng.integers(ppc_range[0], ppc_range[1] + 1))
        k = int(rng.integers(k_range[0], k_range[1] + 1))
        amp = float(rng.uniform(*amp_range))
        N = n * n * ppc
        L = float(n)
        if device is None:
            r = np.random.default_rng(seed * 1000 + i)
            xi = r.random(N) * L
            x = np.mod(xi + amp * L / (2 * np.pi * k) * np.sin(2 * np.pi * k * xi / L), L).astype(dtype)
            y = (r.random(N) * L).astype(dtype)
            vx = r.standard_normal(N).astype(dtype)
            vy = r.standard_normal(N).astype(dtype)
        else:
            import torch
            tdt = torch.float32 if dtype == np.float32 else torch.float64
            gen = torch.Generator(device=device).manual_seed(seed * 1000 + i)
            xi = torch.rand(N, generator=gen, device=device, dtype=torch.float64) * L
            x = torch.remainder(xi + amp * L / (2 * np.pi * k) * torch.sin(2 * np.pi * k * xi / L), L)
            x = x.to(tdt)
            del xi
            y = (torch.rand(N, generator=gen, device=device, dtype=torch.float64) * L).to(tdt)
            vx = torch.randn(N, generator=gen, device=device, dtype=tdt)
            vy = torch.randn(N, generator=gen, device=device, dtype=tdt)
        sp = dict(q=-1.0, m=1.0, w=1.0 / ppc, pos=[x, y], vel=[vx, vy])
        depth = max(2, int(np.log2(n)) - 4)
        out.append(dict(g=periodic_grid([n, n]), species=[sp],
                        dtype="f32" if dtype == np.float32 else "f64",
                        mixed_mg=dtype == np.float32, depth=depth, pre=2, post=2, max_cycles=40,
                        tol=1e-8, dt=0.1, rho_bg=1.0, sort_every=50,
                        meta=dict(n=n, ppc=ppc, k=k, amp=amp, particles=N)))
    return out
