"""Drift-flux mixture stages on the GPU vs the oracle (FP32 restatement and
FP64 twin), each stage from identical inputs, after three two-phase steps
so pressures, velocities and fractions are non-trivial.

Tolerances (normwise, infinity norm, vs FP64): drift velocity 1e-4,
fractions 1e-5, mixture acceleration 1e-4 (or 4x the CPU-FP32 error)."""
import numpy as np
import pytest

import orc
import sphx
from helpers import assert_close, options, state_from_gpu

pytestmark = pytest.mark.gpu

RHO = [1000.0, 1.2]
NU = [1e-6, 1.5e-5]
UNDEF = 0xFFFFFFFF


def setup(n=6000, L=0.2):
    # h / spacing as in the 1M tank (0.01 / 0.0091376)
    s = (0.8 * L * L * L / n) ** (1.0 / 3.0)
    h = 1.0944 * s
    p, disp, calc, alpha, info = sphx.scene_bubbles(n_fluid=n, L=(L, L, L), fill=0.8, h=h,
                                                    n_bubbles=12, rmin=0.03, rmax=0.05,
                                                    layers=2, seed=3)
    p.dt = 1e-3
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1, omega=0.5,
                clamp_pressure=1, tol=1e-3, max_iters=200, drift_relax=2e-11,
                drift_diffusion=1e-4, boussinesq=1, clamp_to_grid=1)
    return p, disp, calc, alpha, o


def oracle_mf(ctx, p, o, dtype, alpha_pre, mforce, drift=None):
    st = state_from_gpu(ctx, p, dtype, o)
    mf = orc.MF(st, RHO, NU, o)
    rho = np.array([np.float32(r) for r in RHO], np.float64)
    mf.set_alpha(alpha_pre, alpha_pre @ rho)
    mf.a["mforce"][:] = mforce
    if drift is not None:
        mf.a["drift"][:, :2, :] = drift
    return st, mf


def test_mf_stages_match_oracle(gpu):
    p, disp, calc, alpha, o = setup()
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.mf_setup(RHO, NU)
        ctx.mf_upload(alpha)
        ctx.call("grid_build")
        for _ in range(3):
            ctx.mf_step()
        ctx.call("grid_build")
        ctx.call("iisph_density")
        pre = ctx.mf_download()
        s32, m32 = oracle_mf(ctx, p, o, np.float32, pre["alpha"], pre["mforce"])
        s64, m64 = oracle_mf(ctx, p, o, np.float64, pre["alpha"], pre["mforce"])
        fl = (s64["type"] == 0) & (s64["mgcell"] != UNDEF)
        mixed = fl & (pre["alpha"][:, 1] > 1e-6) & (pre["alpha"][:, 1] < 1 - 1e-6)
        print(f"fluid {fl.sum()} mixed {mixed.sum()}", flush=True)
        assert mixed.sum() > 10
        # drift velocity
        m32.drift()
        m64.drift()
        ctx.call("mf_drift_velocity")
        g = ctx.mf_download()
        assert_close("drift", g["drift"], m32["drift"][:, :2], m64["drift"][:, :2], 1e-4,
                     fl[:, None, None] & np.ones((1, 2, 3), bool))
        # fraction advection + correction, from the GPU's drift
        s32, m32 = oracle_mf(ctx, p, o, np.float32, pre["alpha"], pre["mforce"], g["drift"])
        s64, m64 = oracle_mf(ctx, p, o, np.float64, pre["alpha"], pre["mforce"], g["drift"])
        for m in (m32, m64):
            m.alpha_advance()
            m.correction()
        ctx.call("mf_alpha_advance")
        ctx.call("mf_correction")
        g2 = ctx.mf_download()
        d, c = ctx.download()
        assert_close("alpha", g2["alpha"], m32["alpha_new"][:, :2], m64["alpha_new"][:, :2], 1e-5,
                     fl[:, None] & np.ones((1, 2), bool))
        assert_close("rho_m", g2["rho_m"], m32["alpha_new"][:, 3], m64["alpha_new"][:, 3], 1e-5, fl)
        assert_close("nu_m", c["visc"], s32["visc"], s64["visc"], 1e-5, fl)
        assert np.allclose(g2["alpha"][fl].sum(1), 1.0, atol=1e-6)
        # mixture acceleration (drift stress + buoyancy) on the advected velocity
        ctx.call("iisph_predict_dii")
        d0, c0 = ctx.download()
        s32, m32 = oracle_mf(ctx, p, o, np.float32, pre["alpha"], pre["mforce"], g["drift"])
        s64, m64 = oracle_mf(ctx, p, o, np.float64, pre["alpha"], pre["mforce"], g["drift"])
        rho = np.array([np.float32(r) for r in RHO], np.float64)
        for m in (m32, m64):
            m.a["alpha_new"][:, :2] = g2["alpha"]
            m.a["alpha_new"][:, 3] = g2["rho_m"]
            m.force()
        ctx.call("mf_force")
        d1, c1 = ctx.download()
        a_gpu = (c1["veleval"].astype(np.float64) - c0["veleval"]) / p.dt
        assert_close("a_mf", a_gpu, m32["amf"], m64["amf"], 1e-4, fl[:, None] & np.ones((1, 3), bool))


def test_mf_steps_keep_fractions_and_converge(gpu):
    """Ten two-phase steps: fractions stay in [0,1] and sum to 1, the solve
    converges every step, and the air rises (buoyancy)."""
    p, disp, calc, alpha, o = setup()
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.mf_setup(RHO, NU)
        ctx.mf_upload(alpha)
        ctx.call("grid_build")
        d0, c0 = ctx.download()
        a0 = ctx.mf_download()["alpha"]
        fl0 = d0["type"] == 0
        y_air0 = np.average(d0["pos"][fl0, 1], weights=a0[fl0, 1])
        its = []
        for s in range(10):
            it, err = ctx.mf_step()
            assert np.isfinite(err)
            if s >= 3:   # after the start-up transient of the lattice
                assert err <= o.tol, (s, it, err)
            its.append(it)
        d, c = ctx.download()
        a = ctx.mf_download()["alpha"]
    fl = d["type"] == 0
    assert np.all(a[fl] >= 0) and np.all(a[fl] <= 1 + 1e-6)
    assert np.allclose(a[fl].sum(1), 1.0, atol=1e-5)
    y_air = np.average(d["pos"][fl, 1], weights=np.maximum(a[fl, 1], 0) + 1e-30)
    print("iterations", its, "air centroid", y_air0, "->", y_air)
    assert y_air > y_air0


def test_force_from_fused_alpha_pass_equals_own_pass(gpu):
    """The force stage adds the drift stress the fraction-advection pass
    summed; when anything it depends on changed since (here: densities
    recomputed, to the same values), it runs its own neighbour pass.  Both
    give the same mixture acceleration (1e-6 normwise: the compiler may
    contract the same sums into FMAs differently in the two kernels)."""
    p, disp, calc, alpha, o = setup()
    out = []
    for redo_density in (False, True):
        with sphx.Context(p, o) as ctx:
            ctx.upload(disp, calc)
            ctx.mf_setup(RHO, NU)
            ctx.mf_upload(alpha)
            ctx.call("grid_build")
            for _ in range(3):
                ctx.mf_step()
            ctx.call("grid_build")
            ctx.call("iisph_density")
            for s in ["mf_drift_velocity", "mf_alpha_advance", "mf_correction"]:
                ctx.call(s)
            if redo_density:
                ctx.call("iisph_density")
            ctx.call("iisph_predict_dii")
            d0, c0 = ctx.download()
            ctx.call("mf_force")
            d1, c1 = ctx.download()
            out.append((c1["veleval"].astype(np.float64) - c0["veleval"]) / p.dt)
            fl = (d1["type"] == 0)
    a, b = out
    err = np.max(np.abs(a[fl] - b[fl])) / np.max(np.abs(b[fl]))
    print(f"fused vs own pass: {err:.3e} (max |a| {np.max(np.abs(b[fl])):.3e})", flush=True)
    assert np.max(np.abs(b[fl])) > 0
    assert err <= 1e-6


def test_density_dii_pass_equals_two_passes(gpu):
    """The composite steps compute density, advected velocity and d_ii in one
    neighbour pass at the density's place; the fraction passes in between
    touch none of their inputs.  The one pass gives the values of
    iisph_density ... iisph_predict_dii (1e-6 relative: the compiler may
    contract the same sums into FMAs differently in the two kernels)."""
    p, disp, calc, alpha, o = setup()
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.mf_setup(RHO, NU)
        ctx.mf_upload(alpha)
        ctx.call("grid_build")
        for _ in range(3):
            ctx.mf_step()
        ctx.call("grid_build")
        ctx.call("iisph_density")
        for s in ["mf_drift_velocity", "mf_alpha_advance", "mf_correction"]:
            ctx.call(s)
        ctx.call("iisph_predict_dii")
        d0, c0 = ctx.download()
        b0 = ctx.download_iisph()
        ctx.call("iisph_density_dii")
        d1, c1 = ctx.download()
        b1 = ctx.download_iisph()
    fl = d0["type"] == 0
    for name, a, b in [("density", c1["dens"], c0["dens"]),
                       ("veleval", c1["veleval"], c0["veleval"]),
                       ("dii", b1["dii"], b0["dii"])]:
        a = np.asarray(a, np.float64)[fl]
        b = np.asarray(b, np.float64)[fl]
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(f"{name}: one pass vs two {err:.3e} (exact {np.array_equal(a, b)})", flush=True)
        assert np.max(np.abs(b)) > 0
        assert err <= 1e-6, name


def test_mf_step_equals_reference_call_order(gpu):
    """sphx_mf_step fuses density with d_ii and the correction with the
    force add; two steps of it give bitwise the state of two steps of the
    reference's call order (grid, density, drift, alpha, correction,
    predict_dii, force, predict_aii, solve, integrate)."""
    p, disp, calc, alpha, o = setup()
    out = []
    for fused in (True, False):
        with sphx.Context(p, o) as ctx:
            ctx.upload(disp, calc)
            ctx.mf_setup(RHO, NU)
            ctx.mf_upload(alpha)
            ctx.call("grid_build")
            its = []
            for _ in range(2):
                if fused:
                    its.append(ctx.mf_step()[0])
                else:
                    for s in ["grid_build", "iisph_density", "mf_drift_velocity",
                              "mf_alpha_advance", "mf_correction", "iisph_predict_dii",
                              "mf_force", "iisph_predict_aii"]:
                        ctx.call(s)
                    its.append(ctx.pressure_solve()[0])
                    ctx.call("iisph_integrate")
            d, c = ctx.download()
            out.append((its, d, c, ctx.mf_download()))
    (ia, da, ca, ma), (ib, db, cb, mb) = out
    print(f"iterations fused {ia} calls {ib}; max |dx| {np.max(np.abs(da['pos'] - db['pos'])):.3e}",
          flush=True)
    assert ia == ib
    np.testing.assert_array_equal(da["pos"], db["pos"])
    np.testing.assert_array_equal(ca["vel"], cb["vel"])
    np.testing.assert_array_equal(ca["pressure"], cb["pressure"])
    np.testing.assert_array_equal(ma["alpha"], mb["alpha"])
