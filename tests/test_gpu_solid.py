"""SPH elastic solid (ComputeSolidTensor, ComputeSolidForce) and surface
tension on the GPU vs the oracle (FP32 restatement and FP64 twin) on
identical inputs: a shuffled dam-break scene (poly6/spiky kernels, support
h) whose lower half is elastic, with random velocities and deformation
gradients near the identity.

Tolerances (normwise, infinity norm, vs FP64; or 4x the CPU-FP32
restatement's own error when larger): deformation gradient 1e-6; stress
1e-4 (it is solidK (F F^T - I), a difference of O(1) terms of about 1e-2);
accelerations 1e-4; surface-tension accelerations 1e-5."""
import numpy as np
import pytest

import sphx
from helpers import assert_close, options, shuffled_scene, state_from_gpu

pytestmark = pytest.mark.gpu

ELASTIC = 2
UNDEF = 0xFFFFFFFF


def solid_scene(seed=11):
    p, disp, calc, info = shuffled_scene(seed=seed, kernel_family=0, h=0.021)
    fl = disp["type"] == 0
    ylim = np.median(disp["pos"][fl, 1])
    el = fl & (disp["pos"][:, 1] < ylim)
    disp["type"][el] = ELASTIC
    rng = np.random.default_rng(seed)
    calc["vel"][fl] = rng.normal(0, 0.05, (np.count_nonzero(fl), 3))
    F = np.tile(np.eye(3, dtype=np.float32).reshape(9), (len(disp), 1))
    F[el] += rng.normal(0, 0.01, (np.count_nonzero(el), 9)).astype(np.float32)
    calc["deformGrad"] = F
    p.dt = 2e-3
    p.solidK = 2e4
    p.surfaceTensionK = 2.0
    p.surfaceTensionFluidC = 0.3
    print(f"solid scene n={len(disp)} elastic={el.sum()}", flush=True)
    return p, disp, calc


def test_solid_tensor_force_and_surface_tension_match_oracle(gpu):
    p, disp, calc = solid_scene()
    o = options(wcsph_kernel=0)
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.call("grid_build")
        ctx.call("boundary_volume")
        ctx.call("wcsph_density_pressure")     # dens = 1/rho, as the reference's RunSolid
        # strain and stress
        d0, c0 = ctx.download()
        s32 = state_from_gpu(ctx, p, np.float32, o, iisph=False)
        s64 = state_from_gpu(ctx, p, np.float64, o, iisph=False)
        F32 = np.ascontiguousarray(c0["deformGrad"], np.float32)
        F64 = np.ascontiguousarray(c0["deformGrad"], np.float64)
        S32 = np.zeros_like(F32)
        S64 = np.zeros_like(F64)
        s32.solid_tensor(F32, S32)
        s64.solid_tensor(F64, S64)
        ctx.call("solid_tensor")
        d1, c1 = ctx.download()
        el = (s64["type"] == ELASTIC) & (s64["mgcell"] != UNDEF)
        assert el.sum() > 100
        m9 = el[:, None] & np.ones((1, 9), bool)
        assert np.max(np.abs(F64[el] - c0["deformGrad"][el])) > 0   # F moved
        assert_close("deformGrad", c1["deformGrad"], F32, F64, 1e-6, m9)
        assert_close("stress", c1["stress"], S32, S64, 1e-4, m9)
        # non-elastic particles keep their matrices
        np.testing.assert_array_equal(c1["deformGrad"][~el], c0["deformGrad"][~el])
        # force, from the GPU's stresses
        s32 = state_from_gpu(ctx, p, np.float32, o, iisph=False)
        s64 = state_from_gpu(ctx, p, np.float64, o, iisph=False)
        S = c1["stress"]
        s32.solid_force(np.ascontiguousarray(S, np.float32))
        s64.solid_force(np.ascontiguousarray(S, np.float64))
        ctx.call("solid_force")
        d2, c2 = ctx.download()
        m3 = el[:, None] & np.ones((1, 3), bool)
        assert np.max(np.abs(s64["accel"][el])) > 0
        assert_close("solid accel", c2["accel"], s32["accel"], s64["accel"], 1e-4, m3)
        # surface tension adds to accel of every particle in the grid
        s32 = state_from_gpu(ctx, p, np.float32, o, iisph=False)
        s64 = state_from_gpu(ctx, p, np.float64, o, iisph=False)
        for s in (s32, s64):
            s.a["accel"][:] = 0.0
            s.surface_tension()
        ctx.call("surface_tension")
        d3, c3 = ctx.download()
        ing = s64["mgcell"] != UNDEF
        inc = c3["accel"].astype(np.float64) - c2["accel"]
        assert_close("surface tension", inc, s32["accel"], s64["accel"], 1e-5,
                     ing[:, None] & np.ones((1, 3), bool))


def test_legacy_solid_entry_points(gpu):
    """ComputeSolidTensorCUDA / ComputeSolidForceCUDA run on the legacy
    context like the context calls; the reference's empty _X kernels are
    no-ops."""
    p, disp, calc = solid_scene(seed=12)
    o = options(wcsph_kernel=0)
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        for s in ("grid_build", "boundary_volume", "wcsph_density_pressure", "solid_tensor",
                  "solid_force"):
            ctx.call(s)
        want_d, want_c = ctx.download()
    L = sphx.lib()
    L.sphx_legacy_context.restype = sphx.C.c_void_p
    L.FluidSetupCUDA(sphx.C.byref(p))
    try:
        n = len(disp)
        L.sphx_legacy_upload(sphx.ptr(disp), sphx.ptr(calc), 0, n)
        L.sphx_set_options(L.sphx_legacy_context(), sphx.C.byref(o))
        for f in ("GetParticleIndexCUDA", "GetGridListCUDA", "RearrageDataCUDA",
                  "ComputeBoundaryDensity", "MfComputePressureCUDA", "ComputeSolidTensorCUDA",
                  "ComputeSolidTensorX_CUDA", "ComputeSolidForceCUDA", "ComputeSolidForceX_CUDA"):
            getattr(L, f)()
        got_d = np.zeros(n, sphx.display_dtype)
        got_c = np.zeros(n, sphx.calc_dtype)
        L.sphx_legacy_download(sphx.ptr(got_d), sphx.ptr(got_c), n)
    finally:
        L.FluidClearCUDA()
    np.testing.assert_array_equal(got_c["deformGrad"], want_c["deformGrad"])
    np.testing.assert_array_equal(got_c["stress"], want_c["stress"])
    np.testing.assert_array_equal(got_c["accel"], want_c["accel"])
