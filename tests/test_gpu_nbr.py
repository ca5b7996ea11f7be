"""The neighbour list's fallbacks on the GPU: particles whose list would
overflow (more neighbours than its capacity, or an entry offset beyond its
14 bits) are served by the cell scan, which visits a superset of the list
in the same order.  The IISPH passes on such scenes match the oracle (which
always scans cells) at the usual tolerances."""
import numpy as np
import pytest

import sphx
from helpers import assert_close, options, state_from_gpu

pytestmark = pytest.mark.gpu

DENS_TOL = 1e-5
FORCE_TOL = 1e-4
UNDEF = 0xFFFFFFFF


def run_iisph_passes(p, disp, calc, o):
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.call("grid_build")
        ctx.call("boundary_volume")
        s32 = state_from_gpu(ctx, p, np.float32, o)
        s64 = state_from_gpu(ctx, p, np.float64, o)
        for s in (s32, s64):
            s.ii_density()
        ctx.call("iisph_density")
        d, c = ctx.download()
        m = (s64["type"] == 0) & (s64["mgcell"] != UNDEF)
        assert_close("rho", c["dens"], s32["dens"], s64["dens"], DENS_TOL, m)
        s32 = state_from_gpu(ctx, p, np.float32, o)
        s64 = state_from_gpu(ctx, p, np.float64, o)
        for s in (s32, s64):
            s.ii_dii()
            s.ii_aii()
        ctx.L.sphx_iisph_predict_advection(ctx.h)
        ctx.sync()
        b = ctx.download_iisph()
        assert_close("dii", b["dii"], s32["dii"], s64["dii"], FORCE_TOL, m[:, None] & np.ones((1, 3), bool))
        assert_close("aii", b["aii"], s32["aii"], s64["aii"], FORCE_TOL, m)
        s32 = state_from_gpu(ctx, p, np.float32, o)
        s64 = state_from_gpu(ctx, p, np.float64, o)
        s32.ii_iteration()
        s64.ii_iteration()
        ctx.pressure_iteration()
        b = ctx.download_iisph()
        assert_close("p", b["press_l"], s32["press_l"], s64["press_l"], FORCE_TOL, m)


def test_crowded_particles_fall_back_to_cell_scan(gpu):
    """A clump packed at half the lattice spacing has ~8x the neighbours the
    list capacity (from the spacing) allows."""
    s = h = 0.01
    p, disp, calc, info = sphx.scene_dambreak(10, 10, 10, s, (0.16, 0.16, 0.16), h, 2, 1e-4, 3, 1)
    p.dt = 1e-3
    fl = np.where(disp["type"] == 0)[0]
    pos = disp["pos"][fl]
    c = pos.mean(axis=0)
    near = fl[np.max(np.abs(pos - c), axis=1) < 0.021]
    disp["pos"][near] = c + (disp["pos"][near] - c) * 0.5
    rng = np.random.default_rng(2)
    calc["vel"][fl] = rng.normal(0, 0.05, (len(fl), 3))
    calc["pressure"][fl] = rng.uniform(0, 200.0, len(fl))
    print(f"crowded clump: {len(near)} of {len(fl)} fluid particles at half spacing", flush=True)
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1, omega=0.5, clamp_pressure=1)
    run_iisph_passes(p, disp, calc, o)


def test_long_rows_fall_back_to_cell_scan(gpu):
    """A 42 m channel of 4 x 4 particles in cross-section: one z-row of cells
    holds about 17k particles, so a neighbour's offset from the start of its
    layer's window exceeds 14 bits for particles near the channel's start."""
    s = h = 0.01
    p, disp, calc, info = sphx.scene_dambreak(4200, 4, 4, s, (42.04, 0.06, 0.06), h, 1, 1e-4, 4, 1)
    p.dt = 1e-3
    fl = np.where(disp["type"] == 0)[0]
    rng = np.random.default_rng(3)
    calc["vel"][fl] = rng.normal(0, 0.05, (len(fl), 3))
    calc["pressure"][fl] = rng.uniform(0, 200.0, len(fl))
    print(f"long channel: {len(disp)} particles, grid {p.gridres.x}x{p.gridres.y}x{p.gridres.z}",
          flush=True)
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1, omega=0.5, clamp_pressure=1)
    run_iisph_passes(p, disp, calc, o)
