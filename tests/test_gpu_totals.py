"""FP64 conservation totals on the device (mass, momentum, kinetic energy,
mass moment) against FP64 sums of the downloaded state, single scene and
per scene of an ensemble; the per-step ring of a tracked run."""
import numpy as np
import pytest

import sphx
from helpers import options

pytestmark = pytest.mark.gpu

RHO = [1000.0, 1.2]
NU = [1e-6, 1.5e-5]


def host_totals(d, c, g, scenes=1):
    """FP64 sums of the downloaded state, and the same sums of absolute
    values (the scale of each total's rounding error)."""
    fl = (d["type"] & 0xFF == 0) & (g["mgcell"] != 0xFFFFFFFF)
    sc = (d["type"] >> 8) & 0xFFFF
    out = np.zeros((scenes, 8))
    mag = np.zeros((scenes, 8))
    for s in range(scenes):
        m = fl & (sc == s)
        mass = c["mass"][m].astype(np.float64)
        v = c["vel"][m].astype(np.float64)
        x = d["pos"][m].astype(np.float64)
        out[s, 0] = mass.sum()
        out[s, 1:4] = (mass[:, None] * v).sum(0)
        out[s, 4] = 0.5 * (mass * (v * v).sum(1)).sum()
        out[s, 5:8] = (mass[:, None] * x).sum(0)
        mag[s, 0], mag[s, 4] = out[s, 0], out[s, 4]
        mag[s, 1:4] = (mass[:, None] * np.abs(v)).sum(0)
        mag[s, 5:8] = (mass[:, None] * np.abs(x)).sum(0)
    return out, mag


def close(a, want):
    tot, mag = want
    return np.all(np.abs(a - tot) <= 1e-12 * mag + 1e-300)


def bubbles(n=8000, L=0.2, seed=3):
    s = (0.8 * L ** 3 / n) ** (1.0 / 3.0)
    p, disp, calc, alpha, info = sphx.scene_bubbles(n_fluid=n, L=(L, L, L), fill=0.8,
                                                    h=1.0944 * s, n_bubbles=12, rmin=0.03,
                                                    rmax=0.05, layers=2, seed=seed)
    p.dt = 1e-3
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1, omega=0.5,
                clamp_pressure=1, tol=1e-3, max_iters=200, drift_relax=2e-11,
                drift_diffusion=1e-4, boussinesq=1, clamp_to_grid=1)
    return p, disp, calc, alpha, o


def test_totals_match_host_sums_and_track_steps(gpu):
    p, disp, calc, alpha, o = bubbles()
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.mf_setup(RHO, NU)
        ctx.mf_upload(alpha)
        ctx.call("grid_build")
        t0 = ctx.conservation_totals()
        d, c = ctx.download()
        assert close(t0, host_totals(d, c, ctx.download_grid()))
        assert t0[0, 0] > 0 and np.abs(t0[0, 5:8]).max() > 0
        ctx.conservation_track(True)
        for _ in range(6):
            ctx.mf_step()
        h = ctx.conservation_history()
        t1 = ctx.conservation_totals()
        d, c = ctx.download()
        want = host_totals(d, c, ctx.download_grid())
        t2 = ctx.conservation_totals()
    assert h.shape == (6, 1, 8)
    np.testing.assert_array_equal(h[-1], t1)        # the last step's ring entry
    np.testing.assert_array_equal(t1, t2)           # reproducible bit for bit
    assert close(t1, want)
    # fixed particle masses: the total mass does not change over the steps
    np.testing.assert_allclose(h[:, 0, 0], t0[0, 0], rtol=1e-13)
    # gravity pulls the momentum down while the fluid settles
    assert np.all(np.isfinite(h))


def test_per_scene_totals_of_an_ensemble(gpu):
    sizes = [3000, 5000, 4000]
    p, disp, calc, alpha, first, info = sphx.scene_ensemble(sizes, seed=11)
    p.dt = 1e-3
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1, omega=0.5,
                clamp_pressure=1, tol=1e-3, max_iters=200, drift_relax=2e-11,
                drift_diffusion=1e-4, boussinesq=1, clamp_to_grid=1)
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.mf_setup(RHO, NU)
        ctx.mf_upload(alpha)
        ctx.ensemble_setup(len(sizes))
        ctx.call("grid_build")
        ctx.conservation_track(True)
        for _ in range(3):
            ctx.mf_step()
        t = ctx.conservation_totals()
        d, c = ctx.download()
        want = host_totals(d, c, ctx.download_grid(), len(sizes))
        h = ctx.conservation_history()
    assert t.shape == (3, 8) and h.shape == (3, 3, 8)
    assert close(t, want)
    np.testing.assert_allclose(h[-1], t, rtol=1e-12, atol=1e-300)
    for s, n in enumerate(sizes):
        assert t[s, 0] > 0
