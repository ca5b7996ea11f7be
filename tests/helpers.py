"""Shared helpers for the parity tests: seeded scenes, GPU -> oracle state
transfer, and error reporting against the FP64 oracle."""
from __future__ import annotations

import numpy as np

import orc
import sphx


def shuffled_scene(nx=12, ny=10, nz=8, s=0.01, L=(0.24, 0.2, 0.16), h=0.02, seed=7,
                   kernel_family=1, n_out=0, layers=2):
    """Dam-break scene with the particle order shuffled (so the sort has work)
    and `n_out` particles moved outside the valid cells."""
    p, disp, calc, info = sphx.scene_dambreak(nx, ny, nz, s, L, h, layers, 1e-4, seed,
                                              kernel_family)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(disp))
    disp, calc = disp[perm].copy(), calc[perm].copy()
    calc["bornid"] = np.arange(len(calc))
    if n_out:
        idx = rng.choice(np.where(disp["type"] == 0)[0], n_out, replace=False)
        far = np.array([[-5.0, 0.1, 0.1], [0.1, 50.0, 0.1], [0.1, 0.1, -0.5]], np.float32)
        disp["pos"][idx] = far[np.arange(n_out) % 3]
    return p, disp, calc, info


def options(**kw):
    o = sphx.default_options()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def state_from_gpu(ctx: sphx.Context, params, dtype=np.float32, opts=None, iisph=True):
    """Oracle state holding exactly the GPU's current particle state, grid and
    IISPH buffers (so one kernel can be compared on identical inputs)."""
    disp, calc = ctx.download()
    st = orc.State(disp, calc, params, dtype)
    g = ctx.download_grid()
    for k in ["mgcell", "midsort", "mgndx", "gridoff", "gridcnt", "MFidTable"]:
        st.a[k][:] = g[k]
    if iisph:
        b = ctx.download_iisph()
        st.a["dii"][:] = b["dii"]
        st.a["dijpj"][:] = b["dijpj"]
        st.a["aii"][:] = b["aii"]
        st.a["rho_adv"][:] = b["rho_adv"]
        st.a["press_l"][:] = b["press_l"]
        st.a["press_l1"][:] = b["press_l1"]
        st.a["residue"][:] = b["residue"]
    if opts is not None:
        st.opts = opts
    return st


def err_report(name, gpu_val, f32_val, f64_val, mask=None):
    g = np.asarray(gpu_val, np.float64)
    a = np.asarray(f32_val, np.float64)
    d = np.asarray(f64_val, np.float64)
    if mask is not None:
        g, a, d = g[mask], a[mask], d[mask]
    e_gpu = orc.relerr(g, d)
    e_f32 = orc.relerr(a, d)
    print(f"{name}: gpu-vs-fp64 {e_gpu:.3e}  cpu32-vs-fp64 {e_f32:.3e}")
    return e_gpu, e_f32


def assert_close(name, gpu_val, f32_val, f64_val, tol, mask=None, factor=4.0):
    """GPU within max(tol, factor * cpu-fp32 error) of the FP64 oracle."""
    e_gpu, e_f32 = err_report(name, gpu_val, f32_val, f64_val, mask)
    bound = max(tol, factor * e_f32)
    assert e_gpu <= bound, f"{name}: gpu error {e_gpu:.3e} > bound {bound:.3e}"
    return e_gpu


# -- the two-phase step on the oracle (sphx_mf_step's stage order) ----------
UNDEF = 0xFFFFFFFF


def oracle_mf_init(disp, calc, alpha, params, opts, rho, nu, dtype=np.float32):
    """Oracle state and fractions as sphx_mf_upload leaves them: fractions
    with rho_m = sum alpha_k rho_k and nu = sum alpha_k nu_k in FP32, and
    g - a = g for the first drift pass."""
    st = orc.State(disp, calc, params, dtype)
    st.opts = opts
    mf = orc.MF(st, rho, nu, opts)
    a = np.asarray(alpha, np.float32)
    r32 = np.asarray(rho, np.float32)
    n32 = np.asarray(nu, np.float32)
    rm = np.zeros(len(a), np.float32)
    nm = np.zeros(len(a), np.float32)
    for k in range(a.shape[1]):
        rm = rm + a[:, k] * r32[k]
        nm = nm + a[:, k] * n32[k]
    mf.set_alpha(a, rm)
    mf.a["mforce"][:] = params.gravity.tolist()
    fl = (st["type"] & 0xFF) == 0
    st.a["visc"][fl] = nm[fl]
    return st, mf


def oracle_mf_step(st, mf, params, opts):
    """One two-phase step: grid build (fractions and g - a carried with the
    permutation), density, drift, fraction advection, correction, d_ii,
    mixture force, a_ii, pressure solve, integration; then the clamp into
    the valid cells (SolverOptions.clamp_to_grid) and g - a of the step, as
    the library's integration does.  Returns (iterations, error)."""
    st.grid_build()
    src = np.where(st["mgcell"] != UNDEF, st["midsort"], np.arange(st.n)).astype(np.int64)
    for k in ("alpha", "mforce"):
        mf.a[k][:] = mf.a[k][src]
    st.ii_density()
    mf.drift()
    mf.alpha_advance()
    mf.correction()
    st.ii_dii()
    mf.force()
    st.ii_aii()
    it, err = st.ii_solve()
    v_old = st["vel"].copy()
    st.ii_integrate()
    live = ((st["type"] & 0xFF) != 1) & (st["mgcell"] != UNDEF)
    f = np.float32 if st.dtype == np.float32 else np.float64
    if opts.clamp_to_grid:
        gm = np.array(params.gridmin.tolist(), f)
        gf = np.array(params.gridIdfac.tolist(), f)
        gr = np.array([params.gridres.x, params.gridres.y, params.gridres.z], f)
        lo = gm + f(1.0001) / gf
        hi = gm + (gr - f(1.0001)) / gf
        x, v = st.a["pos"], st.a["vel"]
        for a in range(3):
            low = live & (x[:, a] < lo[a])
            x[low, a] = lo[a]
            v[low, a] = np.maximum(v[low, a], 0)
            high = live & (x[:, a] > hi[a])
            x[high, a] = hi[a]
            v[high, a] = np.minimum(v[high, a], 0)
    g = np.array(params.gravity.tolist(), f)
    inv = f(1) / f(params.dt)
    mf.a["mforce"][live] = g - (st["vel"][live] - v_old[live]) * inv
    mf.a["alpha"][:] = mf.a["alpha_new"]
    return it, err
