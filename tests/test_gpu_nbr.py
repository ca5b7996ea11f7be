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


def brute_counts(pos, typ, valid, ss, h):
    """Neighbours j != i with |x_ij ss|^2 < 4 h^2 (FP32, as the list build
    tests), bracketed by the counts at 4 h^2 (1 -+ 1e-5) for pairs whose
    FMA-contracted distance may round either way."""
    pos = pos.astype(np.float32)
    rc2 = np.float32(4.0) * np.float32(h) * np.float32(h)
    lo = np.zeros(len(pos), np.int64)
    hi = np.zeros(len(pos), np.int64)
    for i in np.where(valid)[0]:
        d = (pos[i] - pos) * np.float32(ss)
        dsq = np.sum(d * d, axis=1, dtype=np.float32)
        dsq[i] = np.inf
        lo[i] = np.count_nonzero(dsq < rc2 * np.float32(1 - 1e-5))
        hi[i] = np.count_nonzero(dsq < rc2 * np.float32(1 + 1e-5))
    return lo, hi


def test_list_counts_match_brute_force(gpu):
    """Per-particle list counts (boundary neighbours included) against an
    all-pairs count on a jittered block; boundary and out-of-grid particles
    report 0."""
    s = h = 0.01
    p, disp, calc, info = sphx.scene_dambreak(12, 10, 9, s, (0.18, 0.16, 0.15), h, 2, 1e-4, 5, 1)
    fl = np.where(disp["type"] == 0)[0]
    rng = np.random.default_rng(6)
    disp["pos"][fl] += rng.uniform(-0.3 * s, 0.3 * s, (len(fl), 3)).astype(np.float32)
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1)
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.call("grid_build")
        cnt, cap = ctx.nbr_counts()
        d, _ = ctx.download()
        g = ctx.download_grid()
    valid = (d["type"] == 0) & (g["mgcell"] != UNDEF)
    lo, hi = brute_counts(d["pos"], d["type"], valid, p.simscale, p.smoothradius)
    print(f"{valid.sum()} fluid particles, cap {cap}, counts {cnt[valid].min()}..{cnt[valid].max()} "
          f"mean {cnt[valid].mean():.2f}; {np.sum(hi > lo)} with a pair at the radius", flush=True)
    assert np.all(cnt[~valid] == 0)
    assert np.all((cnt[valid] >= lo[valid]) & (cnt[valid] <= hi[valid]))
    assert cnt[valid].max() > 20


def test_crowded_counts_report_overflow(gpu):
    """Particles of the half-spacing clump have more neighbours than the
    capacity and report -1 (cell scan); the others report their count."""
    s = h = 0.01
    p, disp, calc, info = sphx.scene_dambreak(10, 10, 10, s, (0.16, 0.16, 0.16), h, 2, 1e-4, 3, 1)
    fl = np.where(disp["type"] == 0)[0]
    pos = disp["pos"][fl]
    c = pos.mean(axis=0)
    near = fl[np.max(np.abs(pos - c), axis=1) < 0.021]
    disp["pos"][near] = c + (disp["pos"][near] - c) * 0.5
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1)
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.call("grid_build")
        cnt, cap = ctx.nbr_counts()
        d, _ = ctx.download()
        g = ctx.download_grid()
    valid = (d["type"] == 0) & (g["mgcell"] != UNDEF)
    lo, hi = brute_counts(d["pos"], d["type"], valid, p.simscale, p.smoothradius)
    over = valid & (lo > cap)
    fits = valid & (hi <= cap)
    print(f"cap {cap}: {over.sum()} over, {np.sum(cnt == -1)} report -1", flush=True)
    assert over.sum() > 0
    assert np.all(cnt[over] == -1)
    assert np.all((cnt[fits] >= lo[fits]) & (cnt[fits] <= hi[fits]))
