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


# -- closed forms of the drift velocity on a regular lattice -----------------
# The drift-flux model of multiphase.cu:19-166 (Manninen et al. 1996, in the
# SPH form of Ren et al. 2014) gives, per phase k,
#   u_mk = tau (rho_k - sum_l c_l rho_l)(g - a)
#          - tau cpress (grad p_k - sum_l c_l grad p_l)
#          - cdiff sigma (grad alpha_k / alpha_k - sum_l c_l grad alpha_l / alpha_l)
# with c_l = alpha_l rho_l / rho_m, p_k = alpha_k p and tau = relax / nu_i.
# The SPH gradient of a linear field f is exactly M grad f with
# M_i = sum_j V_j gradW_ij (x) (x_j - x_i); M is summed here by brute force
# from the cubic spline's closed-form derivative, independently of the
# oracle's cell traversal.

def f32(x):
    """The drift constants cross the C ABI as float."""
    return float(np.float32(x))


def cubic_dwdr(r, h, s):
    """dW/dr of the cubic spline of normalisation s = 1 / (pi h^3) (pinned
    by the lattice normalisation test), support 2h."""
    q = r / h
    return np.where(q <= 1.0, s / h * (-3.0 * q + 2.25 * q * q),
                    np.where(q <= 2.0, -0.75 * s / h * (2.0 - q) ** 2, 0.0))


def lattice_mf(n=12, s=0.01, h=0.012, **opt):
    p, disp, calc, info = sphx.scene_dambreak(n, n, n, s, (n * s,) * 3, h, 0, 0.0, 1, 1)
    o = options(wcsph_kernel=1, full_support=1, **opt)
    st = orc.State(disp, calc, p, np.float64)
    st.opts = o
    st.grid_build()
    assert np.all(st["type"] == 0) and np.all(st["mgcell"] != UNDEF)
    x = st["pos"]
    lo, hi = x.min(0), x.max(0)
    inner = np.all((x > lo + 2 * h + 1e-9) & (x < hi - 2 * h - 1e-9), axis=1)
    assert inner.sum() >= 100
    st.a["mass"][:] = 1e-3
    st.a["dens"][:] = 1003.0
    V = st["mass"] / st["dens"]
    h, ks = p.smoothradius, p.kspline      # the float32 values the oracle uses
    assert ks == pytest.approx(1.0 / (np.pi * h ** 3), rel=1e-6)
    M = np.zeros((st.n, 3, 3))
    for i in np.where(inner)[0]:
        d = x[i] - x                         # x_i - x_j
        r = np.sqrt((d * d).sum(1))
        nb = (r > 0) & (r < 2 * h)
        gw = (cubic_dwdr(r[nb], h, ks) / r[nb])[:, None] * d[nb]        # grad_i W_ij
        M[i] = (V[nb, None, None] * gw[:, :, None] * (-d[nb])[:, None, :]).sum(0)
    return p, st, o, inner, M


def test_drift_pressure_and_buoyancy_terms_closed_form():
    """Uniform fractions and a linear pressure: the diffusion term vanishes
    and u_mk = tau (rho_k - sum c rho)(g - a) - tau cpress (alpha_k - sum c
    alpha) M grad p."""
    relax, nu, cpress = f32(3e-11), 2e-6, f32(0.8)
    p, st, o, inner, M = lattice_mf(drift_relax=relax, drift_pressure_scale=cpress,
                                    drift_diffusion_scale=1.0, drift_diffusion=1e-4)
    gp = np.array([120.0, -9000.0, 300.0])
    st.a["pressure"][:] = 5000.0 + st["pos"] @ gp
    st.a["visc"][:] = nu
    alpha = np.tile([0.7, 0.3], (st.n, 1))
    rho = np.array(RHO)
    mf = orc.MF(st, RHO, NU, o)
    mf.set_alpha(alpha, alpha @ rho)
    acc = np.array([0.4, -9.81, -0.2])
    mf.a["mforce"][:] = acc
    mf.drift()
    tau = relax / nu
    rho_m = alpha[0] @ rho
    c = alpha[0] * rho / rho_m
    got = mf["drift"][inner][:, :2, :]
    Mg = M[inner] @ gp                       # per particle (uniform on a lattice)
    for k in range(2):
        wk = (tau * (rho[k] - c @ rho) * acc[None, :]
              - tau * cpress * (alpha[0, k] - c @ alpha[0]) * Mg)
        scale = np.abs(wk).max()
        assert scale > 0
        assert np.abs(got[:, k, :] - wk).max() <= 1e-9 * scale, k
    # both terms matter for the air phase: neither alone gives its drift
    air_g = tau * (rho[1] - c @ rho) * acc
    assert np.abs(got[:, 1, :] - air_g).max() > 0.1 * np.abs(air_g).max()
    # the interior M is close to the identity (SPH gradient of a linear field)
    assert np.allclose(M[inner], np.eye(3), atol=0.05)


def test_drift_diffusion_term_closed_form():
    """Linear fractions, no pressure and no acceleration: only the
    diffusion term remains, u_mk = -cdiff sigma (M grad alpha_k / alpha_k -
    sum_l c_l M grad alpha_l / alpha_l)."""
    cdiff, sigma, nu = f32(1.3), f32(2e-4), 1e-6
    p, st, o, inner, M = lattice_mf(drift_relax=2e-11, drift_pressure_scale=1.0,
                                    drift_diffusion_scale=cdiff, drift_diffusion=sigma)
    st.a["pressure"][:] = 0.0
    st.a["visc"][:] = nu
    ga = np.array([2.0, 1.5, -1.0])
    a1 = 0.3 + (st["pos"] - 0.06) @ ga
    assert a1.min() > 1e-3 and a1.max() < 0.999
    alpha = np.stack([1.0 - a1, a1], 1)
    rho = np.array(RHO)
    mf = orc.MF(st, RHO, NU, o)
    mf.set_alpha(alpha, alpha @ rho)
    mf.a["mforce"][:] = 0.0
    mf.drift()
    got = mf["drift"][inner][:, :2, :]
    a = alpha[inner]
    c = a * rho / (a @ rho)[:, None]
    grads = np.stack([M[inner] @ (-ga), M[inner] @ ga], 1)        # M grad alpha_k
    st_k = grads / a[:, :, None]
    want = -cdiff * sigma * (st_k - (c[:, :, None] * st_k).sum(1, keepdims=True))
    scale = np.abs(want).max()
    assert scale > 0
    assert np.abs(got - want).max() <= 1e-9 * scale


# -- fraction advection: exact moment identity for linear fields -------------
# mfComputeAlphaAdvance (multiphase.cu:169-272) changes alpha_k by
#   dt * -sum_j V_j [ (alpha_k,j + alpha_k,i) / 2 (v_j - v_i)
#                     + (alpha_k,j u_k,j + alpha_k,i u_k,i) ] . gradW_ij.
# For fields linear in x (alpha_k = a + g.x, v = A x + v0, u_k = B x + u0)
# and D = x_j - x_i this is, for ANY particle arrangement, a combination of
# the geometric moments G = sum V gradW, M_ab = sum V gradW_a D_b and
# T_abc = sum V gradW_a D_b D_c (all summed by brute force here):
#   term1 = alpha_i A:M + 1/2 A_ab g_c T_abc
#   term2 = 2 alpha_i u_i.G + alpha_i B:M + u_i,a g_c M_ac + B_ab g_c T_abc
# (G != 0 on a jittered lattice, so the alpha_i u_i part is pinned too;
# physically -(term1 + term2) -> -(alpha div v + div(alpha u)) as M -> I).

def test_alpha_advection_moment_identity():
    n, s, h = 9, 0.01, 0.012
    p, disp, calc, info = sphx.scene_dambreak(n, n, n, s, (n * s,) * 3, h, 0, 0.0, 1, 1)
    rng = np.random.default_rng(4)
    disp["pos"] += rng.uniform(-0.2, 0.2, disp["pos"].shape) * s
    p.dt = 1e-3
    o = options(wcsph_kernel=1, full_support=1)
    st = orc.State(disp, calc, p, np.float64)
    st.opts = o
    st.grid_build()
    assert np.all(st["type"] == 0) and np.all(st["mgcell"] != UNDEF)
    st.a["mass"][:] = 1e-3 * rng.uniform(0.9, 1.1, st.n)
    st.a["dens"][:] = 1003.0 * rng.uniform(0.98, 1.02, st.n)
    x = st["pos"].astype(np.float64)
    V = st["mass"] / st["dens"]
    hh, ks = float(p.smoothradius), float(p.kspline)
    G = np.zeros((st.n, 3)); M = np.zeros((st.n, 3, 3)); T = np.zeros((st.n, 3, 3, 3))
    for i in range(st.n):
        e = x[i] - x                          # x_i - x_j
        r = np.sqrt((e * e).sum(1))
        nb = (r > 0) & (r < 2 * hh)
        gw = (cubic_dwdr(r[nb], hh, ks) / r[nb])[:, None] * e[nb] * V[nb, None]   # V_j gradW_ij
        D = -e[nb]
        G[i] = gw.sum(0)
        M[i] = np.einsum("ja,jb->ab", gw, D)
        T[i] = np.einsum("ja,jb,jc->abc", gw, D, D)
    assert np.abs(G).max() > 1e-3 * np.abs(M).max()     # the jitter breaks the symmetry
    x0 = x.mean(0)
    g1 = np.array([1.5, -2.0, 0.8])
    a1 = 0.35 + (x - x0) @ g1
    alpha = np.stack([1.0 - a1, a1], 1)
    A = np.array([[0.3, 0.1, -0.2], [0.05, -0.4, 0.1], [0.2, 0.0, 0.25]])
    st.a["vel"][:] = (x - x0) @ A.T + np.array([0.01, -0.02, 0.03])
    B = [np.array([[0.02, 0.0, 0.01], [0.0, -0.03, 0.0], [0.01, 0.0, 0.02]]),
         np.array([[0.1, -0.05, 0.0], [0.02, 0.3, 0.04], [0.0, 0.06, -0.15]])]
    u0 = [np.array([0.001, -0.002, 0.0]), np.array([-0.01, 0.05, 0.02])]
    rho = np.array(RHO)
    mf = orc.MF(st, RHO, NU, o)
    mf.set_alpha(alpha, alpha @ rho)
    for k in range(2):
        mf.a["drift"][:, k, :] = u0[k] + (x - x0) @ B[k].T
    mf.alpha_advance()
    gks = [-g1, g1]
    dt = float(p.dt)
    for k in range(2):
        ak = alpha[:, k]
        u = u0[k] + (x - x0) @ B[k].T
        t1 = ak * np.einsum("ab,iab->i", A, M) + 0.5 * np.einsum("ab,c,iabc->i", A, gks[k], T)
        t2 = (2 * ak * (u * G).sum(1) + ak * np.einsum("ab,iab->i", B[k], M)
              + np.einsum("ia,c,iac->i", u, gks[k], M) + np.einsum("ab,c,iabc->i", B[k], gks[k], T))
        want = ak - dt * (t1 + t2)
        got = mf["alpha_new"][:, k]
        d = np.abs(got - want).max() / np.abs(t1 + t2).max() / dt
        print(f"phase {k}: relative error of the change {d:.2e}")
        assert d < 1e-9, k


# -- drift stress: exact moment identity for linear drifts --------------------
# contributeForce_projectu's drift-stress term (multiphase.cu:555-562) adds
#   sum_j sum_k rho_k / rho_i V_j [alpha_k,j (u_k,j . gradW_ij) u_k,j
#                                  + alpha_k,i (u_k,i . gradW_ij) u_k,i].
# For uniform fractions and linear drifts u_k = u_k,i + B_k D (D = x_j - x_i)
# this is, for any arrangement, with the moments of the advection test:
#   sum_k rho_k / rho_i alpha_k [2 (u.G) u + (B:M) u + B M^T u + Q],
#   Q_a = B_bc B_ad T_bcd.
def test_drift_stress_moment_identity():
    n, s, h = 9, 0.01, 0.012
    p, disp, calc, info = sphx.scene_dambreak(n, n, n, s, (n * s,) * 3, h, 0, 0.0, 1, 1)
    rng = np.random.default_rng(8)
    disp["pos"] += rng.uniform(-0.2, 0.2, disp["pos"].shape) * s
    p.dt = 1e-3
    o = options(wcsph_kernel=1, full_support=1, boussinesq=0)
    st = orc.State(disp, calc, p, np.float64)
    st.opts = o
    st.grid_build()
    st.a["mass"][:] = 1e-3 * rng.uniform(0.9, 1.1, st.n)
    st.a["dens"][:] = 1003.0 * rng.uniform(0.98, 1.02, st.n)
    st.a["veleval"][:] = 0.0
    x = st["pos"].astype(np.float64)
    V = st["mass"] / st["dens"]
    hh, ks = float(p.smoothradius), float(p.kspline)
    G = np.zeros((st.n, 3)); M = np.zeros((st.n, 3, 3)); T = np.zeros((st.n, 3, 3, 3))
    for i in range(st.n):
        e = x[i] - x
        r = np.sqrt((e * e).sum(1))
        nb = (r > 0) & (r < 2 * hh)
        gw = (cubic_dwdr(r[nb], hh, ks) / r[nb])[:, None] * e[nb] * V[nb, None]
        D = -e[nb]
        G[i] = gw.sum(0)
        M[i] = np.einsum("ja,jb->ab", gw, D)
        T[i] = np.einsum("ja,jb,jc->abc", gw, D, D)
    x0 = x.mean(0)
    alpha = np.tile([0.7, 0.3], (st.n, 1))
    rho = np.array(RHO)
    B = [np.array([[0.02, 0.0, 0.01], [0.0, -0.03, 0.0], [0.01, 0.0, 0.02]]),
         np.array([[0.1, -0.05, 0.0], [0.02, 0.3, 0.04], [0.0, 0.06, -0.15]])]
    u0 = [np.array([0.001, -0.002, 0.0]), np.array([-0.01, 0.05, 0.02])]
    mf = orc.MF(st, RHO, NU, o)
    mf.set_alpha(alpha, alpha @ rho)
    mf.a["alpha_new"][:, :2] = alpha
    mf.a["alpha_new"][:, 3] = alpha @ rho
    for k in range(2):
        mf.a["drift"][:, k, :] = u0[k] + (x - x0) @ B[k].T
    mf.force()
    want = np.zeros((st.n, 3))
    for k in range(2):
        u = u0[k] + (x - x0) @ B[k].T
        Bk = B[k]
        term = (2 * (u * G).sum(1)[:, None] * u + np.einsum("ab,iab->i", Bk, M)[:, None] * u
                + np.einsum("ad,ibd,ib->ia", Bk, M, u) + np.einsum("bc,ad,ibcd->ia", Bk, Bk, T))
        want += (rho[k] / st["dens"])[:, None] * alpha[:, k:k + 1] * term
    got = mf["amf"]
    d = np.abs(got - want).max() / np.abs(want).max()
    print(f"drift stress: relative error {d:.2e}")
    assert d < 1e-9
