"""The legacy drop-in entry points (fluid_system_host.cuh:44-107) driven in
the reference's call orders through the C ABI, against the same stages
called on a context: results must be bitwise equal.

* IISPH: RunIISPH (fluid_system.cpp:566-588) plus the Integration() the
  reference leaves commented out there.
* Multiphase: the declared-only Mf* entry points around the IISPH solve.
* MPM: MpmAllocateBufferCUDA, IndexMPMSortCUDA, then RunMPM
  (fluid_system.cpp:541-564) twice, then MpmGetMomentumCUDA,
  MpmColorTestCUDA and ComputeMpmForce (to 1e-5 normwise: the scatter's
  sums are order-nondeterministic in their last bits).
* InitializeSolid_CUDA (host.cu:492-501) = the three grid calls.
"""
import ctypes as C

import numpy as np
import pytest

import sphx
from helpers import options

pytestmark = pytest.mark.gpu

RHO = [1000.0, 1.2]
NU = [1e-6, 1.5e-5]


class Legacy:
    """The process-wide legacy context for the duration of a test."""

    def __init__(self, p, o=None):
        self.L = sphx.lib()
        self.L.sphx_legacy_context.restype = C.c_void_p
        self.L.FluidSetupCUDA(C.byref(p))
        self.h = C.c_void_p(self.L.sphx_legacy_context())
        assert self.h.value, sphx.last_error()
        if o is not None:
            sphx.check(self.L.sphx_set_options(self.h, C.byref(o)), "set_options")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.L.FluidClearCUDA()
        assert not self.L.sphx_legacy_context()

    def __call__(self, *names):
        for f in names:
            print("legacy", f, flush=True)
            getattr(self.L, f)()

    def upload(self, disp, calc):
        self.L.sphx_legacy_upload(sphx.ptr(disp), sphx.ptr(calc), 0, len(disp))

    def download(self, n):
        d = np.zeros(n, sphx.display_dtype)
        c = np.zeros(n, sphx.calc_dtype)
        self.L.sphx_legacy_download(sphx.ptr(d), sphx.ptr(c), n)
        return d, c


def assert_same(a, b, fields):
    for k in fields:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def assert_close_all(a, b, fields, tol=1e-5):
    for k in fields:
        x = np.asarray(a[k], np.float64)
        y = np.asarray(b[k], np.float64)
        np.testing.assert_allclose(x, y, rtol=0, atol=tol * max(np.abs(y).max(), 1e-30),
                                   err_msg=k)


GRID = ("GetParticleIndexCUDA", "GetGridListCUDA", "RearrageDataCUDA")
CALC = ("vel", "veleval", "pressure", "dens", "accel")


def iisph_scene():
    p, disp, calc, info = sphx.scene_dambreak(10, 10, 10, 0.01, (0.2, 0.15, 0.15), 0.01, 2, 1e-4,
                                              20261209, 1)
    p.dt = 1e-3
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1, omega=0.5,
                clamp_pressure=1, tol=1e-3, max_iters=200)
    return p, disp, calc, o


def test_legacy_iisph_chain_matches_context(gpu):
    p, disp, calc, o = iisph_scene()
    n = len(disp)
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.call("grid_build")
        ctx.call("boundary_volume")
        for _ in range(2):
            ctx.call("grid_build")
            ctx.call("iisph_density")
            ctx.call("iisph_predict_advection")
            ctx.pressure_solve()
            ctx.call("iisph_integrate")
        wd, wc = ctx.download()
    with Legacy(p, o) as leg:
        leg.upload(disp, calc)
        leg(*GRID, "ComputeBoundaryDensity")
        for _ in range(2):
            leg(*GRID, "ComputeDensityIISPH_CUDA", "PredictAdvection", "PressureSolve",
                "Integration")
        gd, gc = leg.download(n)
    assert_same(gd, wd, ("pos", "type"))
    assert_same(gc, wc, CALC)
    assert np.abs(wc["pressure"]).max() > 0


def test_legacy_multiphase_chain_matches_context(gpu):
    s = (0.8 * 0.2 ** 3 / 5000) ** (1.0 / 3.0)
    p, disp, calc, alpha, info = sphx.scene_bubbles(n_fluid=5000, L=(0.2, 0.2, 0.2), fill=0.8,
                                                    h=1.0944 * s, n_bubbles=10, rmin=0.03,
                                                    rmax=0.05, layers=2, seed=5)
    p.dt = 1e-3
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1, omega=0.5,
                clamp_pressure=1, tol=1e-3, max_iters=200, drift_relax=2e-11,
                drift_diffusion=1e-4, boussinesq=1, clamp_to_grid=1)
    n = len(disp)
    steps = ("iisph_density", "mf_drift_velocity", "mf_alpha_advance", "mf_correction",
             "iisph_predict_advection", "mf_force")
    legacy = ("ComputeDensityIISPH_CUDA", "MfComputeDriftVelCUDA", "MfComputeAlphaAdvanceCUDA",
              "MfComputeCorrectionCUDA", "PredictAdvection", "ComputeForceCUDA_ProjectU")
    rho = np.ascontiguousarray(RHO, np.float32)
    nu = np.ascontiguousarray(NU, np.float32)
    a32 = np.ascontiguousarray(alpha, np.float32)
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.mf_setup(RHO, NU)
        ctx.mf_upload(alpha)
        for _ in range(2):
            ctx.call("grid_build")
            for s_ in steps:
                ctx.call(s_)
            ctx.pressure_solve()
            ctx.call("iisph_integrate")
        wd, wc = ctx.download()
        wm = ctx.mf_download()
    with Legacy(p, o) as leg:
        leg.upload(disp, calc)
        L = leg.L
        sphx.check(L.sphx_mf_setup(leg.h, 2, sphx.ptr(rho), sphx.ptr(nu)), "mf_setup")
        sphx.check(L.sphx_mf_upload(leg.h, sphx.ptr(a32), 0, len(a32)), "mf_upload")
        for _ in range(2):
            leg(*GRID, *legacy, "PressureSolve", "Integration")
        gd, gc = leg.download(n)
        ga = np.zeros((n, 2), np.float32)
        sphx.check(L.sphx_mf_download(leg.h, sphx.ptr(ga), None, None, None), "mf_download")
    assert_same(gd, wd, ("pos",))
    assert_same(gc, wc, CALC)
    np.testing.assert_array_equal(ga, wm["alpha"])
    assert np.abs(wm["drift"]).max() > 0


def test_legacy_mpm_chain_matches_context(gpu):
    p, disp, calc, info = sphx.scene_sand(n=4096, lo=0.3, hi=0.6, seed=14)
    rng = np.random.default_rng(14)
    calc["vel"] = rng.normal(0, 0.2, (len(disp), 3))
    p.dt = 1e-4
    n = len(disp)
    with sphx.Context(p) as ctx:
        ctx.upload(disp, calc)
        ctx.mpm_setup(0, 0)
        for _ in range(2):
            ctx.call("grid_build")
            ctx.mpm_p2g(1)
            ctx.call("mpm_stress")
            ctx.call("mpm_node_update")
            ctx.mpm_g2p(1)
        ctx.call("mpm_get_momentum")
        ctx.call("mpm_color_test")
        wd, wc = ctx.download()
        wv, wmass = ctx.mpm_nodes()
    with Legacy(p) as leg:
        leg.upload(disp, calc)
        leg.L.MpmAllocateBufferCUDA(C.byref(p))
        leg("IndexMPMSortCUDA")
        for _ in range(2):
            leg(*GRID, "MpmParticleToGrid_APIC_CUDA", "MpmParticleStress_CUDA",
                "MpmNodeUpdate_CUDA", "MpmParticleUpdate_APIC_CUDA")
        leg("MpmGetMomentumCUDA", "MpmColorTestCUDA", "ComputeMpmForce")
        gd, gc = leg.download(n)
        N = p.mpmRes.x * p.mpmRes.y * p.mpmRes.z
        gv = np.zeros((N, 3), np.float32)
        gm = np.zeros(N, np.float32)
        sphx.check(leg.L.sphx_mpm_download_nodes(leg.h, sphx.ptr(gv), sphx.ptr(gm)), "nodes")
    # the particle-to-grid scatter ranks a cell's particles with LDS atomics,
    # so its FP32 sums (and everything after them) agree between two runs
    # to rounding, not bit for bit
    assert_close_all(gd, wd, ("pos",))
    assert_same(gd, wd, ("color",))
    assert_close_all(gc, wc, ("vel", "stress", "B", "dens"))
    np.testing.assert_allclose(gv, wv, rtol=0, atol=1e-5 * np.abs(wv).max())
    np.testing.assert_allclose(gm, wmass, rtol=0, atol=1e-5 * np.abs(wmass).max())
    assert np.abs(wv).max() > 0


def test_legacy_null_params_are_refused(gpu):
    """MpmAllocateBufferCUDA(NULL) and FluidSetupCUDA(NULL) return without
    touching the context (the reference takes a reference; a C caller can
    pass NULL)."""
    p, disp, calc, o = iisph_scene()
    with Legacy(p, o) as leg:
        leg.upload(disp, calc)
        leg.L.MpmAllocateBufferCUDA(None)
        leg.L.FluidSetupCUDA(None)
        assert leg.L.sphx_legacy_context() == leg.h.value
        leg(*GRID)


def test_initialize_solid_is_the_grid_build(gpu):
    p, disp, calc, o = iisph_scene()
    n = len(disp)
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.call("grid_build")
        wd, wc = ctx.download()
        wg = ctx.download_grid()
    with Legacy(p, o) as leg:
        leg.upload(disp, calc)
        leg("InitializeSolid_CUDA")
        gd, gc = leg.download(n)
        gg = {k: np.zeros(n, np.uint32) for k in ("mgcell", "midsort", "mgndx")}
        gg.update({k: np.zeros(p.gridtotal, np.int32) for k in ("gridoff", "gridcnt")})
        mf = np.zeros(n, np.int32)
        sphx.check(leg.L.sphx_download_grid(leg.h, *[sphx.ptr(gg[k]) for k in (
            "mgcell", "midsort", "mgndx", "gridoff", "gridcnt")], sphx.ptr(mf)), "grid")
    assert_same(gd, wd, ("pos", "type"))
    assert_same(gc, wc, ("bornid",))
    for k in ("mgcell", "midsort", "gridoff", "gridcnt"):
        np.testing.assert_array_equal(gg[k], wg[k], err_msg=k)
