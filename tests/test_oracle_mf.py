"""Pins on the oracle's drift-flux mixture stages (no GPU): identities of
the drift-flux model that a dropped term or a wrong coefficient breaks."""
import numpy as np
import pytest

import orc
import sphx
from helpers import options

UNDEF = 0xFFFFFFFF
# the phase constants cross the C ABI as float: use their float32 values
RHO = [float(np.float32(1000.0)), float(np.float32(1.2))]
NU = [float(np.float32(1e-6)), float(np.float32(1.5e-5))]


def mf_state(mixed=True, seed=3, still=False):
    p, disp, calc, alpha, info = sphx.scene_bubbles(n_fluid=6000, L=(0.2, 0.2, 0.2), fill=0.8,
                                                    h=0.012, n_bubbles=12, rmin=0.02, rmax=0.04,
                                                    layers=2, seed=seed)
    p.dt = 1e-3
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1, drift_relax=2e-11,
                drift_diffusion=1e-4, boussinesq=1)
    rng = np.random.default_rng(seed)
    fl = disp["type"] == 0
    if mixed:
        a1 = np.clip(alpha[:, 1] + rng.uniform(0, 0.3, len(alpha)), 0, 1) * fl
        alpha = np.stack([(1 - a1) * fl, a1], 1).astype(np.float32)
    if not still:
        calc["vel"][fl] = rng.normal(0, 0.05, (fl.sum(), 3))
    calc["pressure"][fl] = rng.uniform(0, 2000.0, fl.sum())
    calc["visc"][fl] = alpha[fl] @ np.array(NU, np.float32)
    st = orc.State(disp, calc, p, np.float64)
    st.opts = o
    st.grid_build()
    # alpha travels with the sort
    src = np.where(st["mgcell"] != UNDEF, st["midsort"], np.arange(len(disp)))
    alpha = alpha[src]
    st.ii_density()
    mf = orc.MF(st, RHO, NU, o)
    mf.set_alpha(alpha, alpha @ np.array(RHO))
    mf.a["mforce"][:] = [0.0, -9.81, 0.0]
    return st, mf, alpha


def fluid(st):
    return (st["type"] == 0) & (st["mgcell"] != UNDEF)


def test_mass_weighted_drift_sums_to_zero():
    """sum_k c_k u_mk = 0 with c_k = alpha_k rho_k / rho_m (drift-flux
    identity: every term of multiphase.cu:131-163 cancels in the sum)."""
    st, mf, alpha = mf_state()
    mf.drift()
    fl = fluid(st)
    a = alpha[fl].astype(np.float64)
    c = a * np.array(RHO) / (a @ np.array(RHO))[:, None]
    terms = c[:, :, None] * mf["drift"][fl][:, :2, :]
    s = terms.sum(1)
    scale = np.abs(terms).max(1)
    assert scale.max() > 0
    # cancellation to FP64 rounding of the individual terms
    floor = 1e-14 * np.abs(mf["drift"][fl]).max()
    assert np.all(np.abs(s) <= 1e-10 * scale + floor)


def test_drift_flux_conserves_phase_volume_at_rest():
    """With v = 0 the fraction change is the drift flux alone, which is
    antisymmetric per pair: sum_i V_i alpha_k,i is conserved."""
    st, mf, alpha = mf_state(still=True)
    mf.drift()
    mf.alpha_advance()
    fl = fluid(st)
    V = st["mass"] / st["dens"]
    for k in range(2):
        before = np.sum(V[fl] * mf["alpha"][fl, k])
        after = np.sum(V[fl] * mf["alpha_new"][fl, k])
        change = np.abs(mf["alpha_new"][fl, k] - mf["alpha"][fl, k]).sum()
        assert change > 0
        assert abs(after - before) <= 1e-9 * max(abs(before), 1e-30) + 1e-12 * change


def test_single_phase_limit_keeps_fractions_one_hot():
    st, mf, alpha = mf_state(mixed=False)
    pure = np.zeros_like(alpha)
    pure[:, 0] = (st["type"] == 0)
    mf.set_alpha(pure, pure @ np.array(RHO))
    mf.drift()
    mf.alpha_advance()
    mf.correction()
    fl = fluid(st)
    # x * (1/x) may round to 1 - ulp
    assert np.allclose(mf["alpha_new"][fl, 0], 1.0, rtol=0, atol=4e-16)
    assert np.all(mf["alpha_new"][fl, 1] == 0.0)
    assert np.allclose(mf["alpha_new"][fl, 3], 1000.0)


def test_correction_normalises_and_mixes():
    st, mf, alpha = mf_state()
    mf.drift()
    mf.alpha_advance()
    mf.a["alpha_new"][:5, 0] = -0.2          # negative fractions are clamped
    mf.correction()
    fl = fluid(st)
    a = mf["alpha_new"][fl]
    assert np.all(a[:, :2] >= 0)
    assert np.allclose(a[:, :2].sum(1), 1.0, atol=1e-12)
    assert np.allclose(a[:, 3], a[:, :2] @ np.array(RHO))
    assert np.allclose(st["visc"][fl], a[:, :2] @ np.array(NU))


def test_buoyancy_lifts_air():
    """Boussinesq term: a = (rho_m - rho_ref)/rho_ref g points up for air."""
    st, mf, alpha = mf_state(mixed=False)
    mf.a["drift"][:] = 0.0
    mf.a["alpha_new"][:] = mf.a["alpha"]
    ve0 = st["veleval"].copy()
    mf.force()
    fl = fluid(st)
    air = fl & (mf["alpha"][:, 1] > 0.5)
    water = fl & (mf["alpha"][:, 0] > 0.5)
    assert air.any() and water.any()
    dv = (st["veleval"] - ve0)[:, 1] / st.params.dt
    assert np.allclose(dv[air], -9.81 * (1.2 - 1000.0) / 1000.0, rtol=1e-6)
    assert np.allclose(dv[water], 0.0, atol=1e-9)
