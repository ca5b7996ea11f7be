"""The batched ensemble at its benchmark shape: 64 two-phase tanks with
fluid counts log-uniform in [20k, 400k] (seeded as the benchmark draws
them), all scenes in one grid and one launch per stage.

* Every scene's box is separated from every other by more than the 2h
  support, so no particle has a neighbour in another scene.
* Eager and captured steps: every scene's solve reaches the tolerance, the
  fractions stay a partition of unity, and every scene's FP64 total mass
  (sum m over its fluid particles) is unchanged step to step.
* The three smallest scenes, cut out of the ensemble after two steps: the
  density, drift velocity and fraction advection of the ensemble launches
  against the FP64 oracle run on each scene alone (FP32 oracle for the
  bounds), as the single-scene parity tests do.
"""
import numpy as np
import pytest

import orc
import sphx
from helpers import assert_close, options

pytestmark = pytest.mark.gpu

RHO = [1000.0, 1.2]
NU = [1e-6, 1.5e-5]
UNDEF = 0xFFFFFFFF


def c5_sizes(S=64):
    rng = np.random.default_rng(20261207)
    return np.exp(rng.uniform(np.log(2e4), np.log(4e5), S)).astype(np.int64)


def c5_options():
    return options(wcsph_kernel=1, full_support=1, boundary_pressure=1, omega=0.5,
                   clamp_pressure=1, tol=1e-3, min_iters=2, max_iters=300, clamp_to_grid=1,
                   warm_start=0.5, drift_relax=2e-11, drift_diffusion=1e-4, boussinesq=1,
                   rho_ref=1000.0)


@pytest.fixture(scope="module")
def c5_scene():
    sizes = c5_sizes()
    p, disp, calc, alpha, first, info = sphx.scene_ensemble(sizes.tolist(), seed=20261208)
    p.dt = 1e-3
    print(f"C5: {len(sizes)} scenes, {int(info.n_fluid)} fluid + {int(info.n_boundary)} boundary "
          f"particles, sizes {sizes.min()}..{sizes.max()}", flush=True)
    return sizes, p, disp, calc, alpha, first, info


def make(p, disp, calc, alpha, S, o):
    c = sphx.Context(p, o)
    c.upload(disp, calc)
    c.mf_setup(RHO, NU)
    c.mf_upload(alpha)
    c.ensemble_setup(S)
    c.call("grid_build")
    return c


def test_c5_scenes_are_isolated_and_steps_keep_invariants(gpu, c5_scene):
    sizes, p, disp, calc, alpha, first, info = c5_scene
    S = len(sizes)
    sc = (disp["type"] >> 8) & 0xFFFF
    lo = np.array([disp["pos"][sc == s].min(0) for s in range(S)])
    hi = np.array([disp["pos"][sc == s].max(0) for s in range(S)])
    support = 2.0 * p.smoothradius
    for s in range(S):
        gap = np.maximum(lo - hi[s], lo[s] - hi).max(axis=1)   # > 0: separated on some axis
        gap[s] = np.inf
        assert gap.min() > support, (s, gap.min())
    o = c5_options()
    with make(p, disp, calc, alpha, S, o) as ctx:
        m0 = ctx.conservation_totals()[:, 0]
        assert m0.shape == (S,) and np.all(m0 > 0)
        for step in range(2):
            it, err = ctx.mf_step()
            its, errs = ctx.scene_stats()
            print(f"eager step {step}: most iterations {it}, worst scene error {err:.2e}",
                  flush=True)
            assert np.all(errs <= o.tol), errs
        ctx.graph_build(4)
        for step in range(4):
            it, err = ctx.graph_step()
            its, errs = ctx.scene_stats()
            print(f"graph step {step}: most iterations {it}, per-scene iterations "
                  f"{its.min()}..{its.max()}, worst error {errs.max():.2e}", flush=True)
            assert np.all(errs <= o.tol), errs
            m = ctx.conservation_totals()[:, 0]
            np.testing.assert_allclose(m, m0, rtol=1e-12)
        d, c = ctx.download()
        a = ctx.mf_download()["alpha"]
    fl = (d["type"] & 0xFF) == 0
    assert np.all(a[fl] >= 0) and np.all(a[fl] <= 1 + 1e-6)
    np.testing.assert_allclose(a[fl].sum(1), 1.0, atol=1e-5)
    assert np.all(np.isfinite(d["pos"])) and np.all(np.isfinite(c["vel"]))


def test_c5_smallest_scenes_match_oracle(gpu, c5_scene):
    sizes, p, disp, calc, alpha, first, info = c5_scene
    S = len(sizes)
    o = c5_options()
    small = np.argsort(sizes)[:3]
    with make(p, disp, calc, alpha, S, o) as ctx:
        for _ in range(2):
            ctx.mf_step()
        ctx.call("grid_build")
        ctx.call("iisph_density")
        d0, c0 = ctx.download()
        pre = ctx.mf_download()
        ctx.call("mf_drift_velocity")
        g = ctx.mf_download()
        ctx.call("mf_alpha_advance")
        ctx.call("mf_correction")
        g2 = ctx.mf_download()
    rho = np.array([np.float32(r) for r in RHO], np.float64)
    for s in small:
        sel = ((d0["type"] >> 8) & 0xFFFF) == s
        ds, cs = d0[sel].copy(), c0[sel].copy()
        ds["type"] &= 0xFF
        q = type(p).from_buffer_copy(p)
        q.num = q.maxNum = int(sel.sum())
        print(f"scene {s}: {int((ds['type'] == 0).sum())} fluid particles", flush=True)

        def oracle(dtype, drift=None):
            st = orc.State(ds, cs, q, dtype)
            st.opts = o
            st.grid_build()
            src = np.where(st["mgcell"] != UNDEF, st["midsort"], np.arange(st.n)).astype(np.int64)
            mf = orc.MF(st, RHO, NU, o)
            a = pre["alpha"][sel][src]
            mf.set_alpha(a, a @ rho)
            mf.a["mforce"][:] = pre["mforce"][sel][src]
            if drift is not None:
                mf.a["drift"][:, :2, :] = drift[sel][src]
            return st, mf, src

        s32, m32, src = oracle(np.float32)
        s64, m64, _ = oracle(np.float64)
        fl = (s64["type"] == 0) & (s64["mgcell"] != UNDEF)
        for st in (s32, s64):
            st.ii_density()
        assert_close(f"scene {s} rho", cs["dens"][src], s32["dens"], s64["dens"], 1e-5, fl)
        # the GPU's drift and fractions of the scene's particles, in the
        # oracle's (re-sorted) order
        for st in (s32, s64):
            st.a["dens"][:] = cs["dens"][src]
        m32.drift()
        m64.drift()
        assert_close(f"scene {s} drift", g["drift"][sel][src], m32["drift"][:, :2],
                     m64["drift"][:, :2], 1e-4, fl[:, None, None] & np.ones((1, 2, 3), bool))
        s32, m32, _ = oracle(np.float32, g["drift"])
        s64, m64, _ = oracle(np.float64, g["drift"])
        for st in (s32, s64):
            st.a["dens"][:] = cs["dens"][src]
        for m in (m32, m64):
            m.alpha_advance()
            m.correction()
        assert_close(f"scene {s} alpha", g2["alpha"][sel][src], m32["alpha_new"][:, :2],
                     m64["alpha_new"][:, :2], 1e-5, fl[:, None] & np.ones((1, 2), bool))
