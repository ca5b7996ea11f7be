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
