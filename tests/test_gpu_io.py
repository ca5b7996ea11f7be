"""Checkpoint / resume on the GPU: a state saved after some IISPH steps
and restored into a fresh context continues bit-identically to the
uninterrupted run (the AoS state plus the parameter block is the whole
state of a single-phase step: the warm start reads the previous pressure
from the calculationPack)."""
import numpy as np
import pytest

import sphx
from helpers import options

pytestmark = pytest.mark.gpu


def test_checkpoint_resume_is_bit_identical(gpu, tmp_path):
    s = h = 0.01
    p, disp, calc, info = sphx.scene_dambreak(8, 12, 8, s, (0.1, 0.2, 0.1), h, 2, 1e-4, 5, 1)
    p.dt = 1e-3
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1, omega=0.5,
                clamp_pressure=1, tol=1e-3, max_iters=200)
    ckp = tmp_path / "step3.ckp"
    print(f"resume test n={len(disp)}", flush=True)
    with sphx.Context(p, o) as A:
        A.upload(disp, calc)
        it_a = [A.iisph_step() for _ in range(6)]
        da, ca = A.download()
    with sphx.Context(p, o) as B:
        B.upload(disp, calc)
        it_b = [B.iisph_step() for _ in range(3)]
        B.checkpoint_save(ckp)
    q, dq, cq = sphx.checkpoint_read(ckp)
    assert q.num == len(disp) and len(dq) == len(disp)
    with sphx.Context(p, o) as C:
        C.checkpoint_restore(ckp)
        it_c = [C.iisph_step() for _ in range(3)]
        dc, cc = C.download()
    print("iterations", [i for i, _ in it_a], [i for i, _ in it_b + it_c], flush=True)
    assert [i for i, _ in it_a] == [i for i, _ in it_b + it_c]
    assert da.tobytes() == dc.tobytes()
    for k in ("vel", "veleval", "pressure", "dens", "accel", "bornid"):
        np.testing.assert_array_equal(ca[k], cc[k], err_msg=k)


def test_learned_two_phase_resume_is_bit_identical(gpu, tmp_path):
    """A two-phase run with the learned start and the drift closure on:
    the checkpoint carries the fractions, g - a of the last step, the
    network (weights, Adam moments, solve counter) and each particle's
    history, so a restored context continues bit-identically."""
    RHO, NU = [1000.0, 1.2], [1e-6, 1.5e-5]
    n = 6000
    s = (0.8 * 0.2 ** 3 / n) ** (1.0 / 3.0)
    p, disp, calc, alpha, info = sphx.scene_bubbles(n_fluid=n, L=(0.2, 0.2, 0.2), fill=0.8,
                                                    h=1.0944 * s, n_bubbles=12, rmin=0.03,
                                                    rmax=0.05, layers=2, seed=8)
    p.dt = 1e-3
    o = options(wcsph_kernel=1, full_support=1, boundary_pressure=1, omega=0.5,
                clamp_pressure=1, tol=1e-3, max_iters=200, drift_relax=2e-11,
                drift_diffusion=1e-4, boussinesq=1, clamp_to_grid=1)
    rng = np.random.default_rng(8)
    ckp = tmp_path / "mf3.ckp"

    def start():
        c = sphx.Context(p, o)
        c.upload(disp, calc)
        c.mf_setup(RHO, NU)
        c.mf_upload(alpha)
        c.call("grid_build")
        c.mlp_init(3, 1000.0 * 9.81 * 0.16, 0.05)
        w = c.mlp_weights()
        w += rng_w
        c.mlp_set_weights(w)
        c.mlp_mode(warm=True, drift=True)
        return c

    rng_w = rng.normal(0, 0.02, sphx.lib().sphx_mlp_nparam()).astype(np.float32)
    A = start()
    it_a = [A.mf_step()[0] for _ in range(6)]
    da, ca = A.download()
    ma = A.mf_download()
    A.close()
    B = start()
    it_b = [B.mf_step()[0] for _ in range(3)]
    B.checkpoint_save(ckp)
    B.close()
    with sphx.Context(p, o) as C:
        C.checkpoint_restore(ckp)
        C.nphase = 2
        it_c = [C.mf_step()[0] for _ in range(3)]
        dc, cc = C.download()
        mc = C.mf_download()
    print("iterations", it_a, it_b + it_c, flush=True)
    assert it_a == it_b + it_c
    assert da.tobytes() == dc.tobytes()
    for k in ("vel", "veleval", "pressure", "dens", "accel", "visc"):
        np.testing.assert_array_equal(ca[k], cc[k], err_msg=k)
    np.testing.assert_array_equal(ma["alpha"], mc["alpha"])
    np.testing.assert_array_equal(ma["mforce"], mc["mforce"])
