"""Stepper checkpoints in the MAS format: save, restore into a fresh stepper,
and the continued run is bit-identical to an uninterrupted one (particles,
phi and the diagnostic history).  The file is read back with psum_amd.mas and
the particle blocks equal the state at the time of the save."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("dtype,extrap,sort_every", [("f64", False, 4), ("f32", True, 3)])
def test_restore_continues_bit_identically(tmp_path, dtype, extrap, sort_every):
    from psum_amd import engine as E
    from psum_amd.mas import MasFile, load_particles
    from psum_amd.workloads import langmuir_2d
    wl = langmuir_2d(n=32, ppc=16, seed=11)
    kw = dict(dtype=dtype, mixed_mg=dtype == "f32", depth=2, tol=1e-9, max_cycles=40, dt=wl["dt"],
              rho_bg=wl["rho_bg"], sort_every=sort_every, extrapolate=extrap)
    print(f"checkpoint case {dtype} extrapolate={extrap} sort_every={sort_every}", flush=True)
    a = E.Stepper(wl["grid"], wl["species"], **kw)
    a.run(7)
    path = tmp_path / "ck.mas"
    a.save(path)
    p0, v0 = a.state(0)
    ha = a.run(6)
    pa, va = a.state(0)
    phia = a.phi()
    b = E.Stepper(wl["grid"], wl["species"], **kw)
    b.restore(path)
    hb = b.run(6)
    pb, vb = b.state(0)
    for x, y in zip(pa + va, pb + vb):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(phia, b.phi())
    np.testing.assert_array_equal(ha, hb)
    m = MasFile(path)
    pos, vel = load_particles(m, "species0")
    np.testing.assert_array_equal(pos, np.stack(p0, axis=1))
    np.testing.assert_array_equal(vel, np.stack(v0, axis=1))
    assert int(m.get_data("psum.step").ravel()[0]) == 7


def test_restore_rejects_a_mismatched_stepper(tmp_path):
    from psum_amd import engine as E
    from psum_amd._lib import PsumError
    from psum_amd.workloads import langmuir_2d
    wl = langmuir_2d(n=32, ppc=8, seed=12)
    a = E.Stepper(wl["grid"], wl["species"], dtype="f64", depth=2, dt=wl["dt"], rho_bg=wl["rho_bg"])
    a.run(2)
    a.save(tmp_path / "a.mas")
    b = E.Stepper(wl["grid"], wl["species"], dtype="f32", depth=2, dt=wl["dt"], rho_bg=wl["rho_bg"])
    with pytest.raises(PsumError):
        b.restore(tmp_path / "a.mas")
