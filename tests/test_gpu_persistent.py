"""The persistent solve (the whole pressure solve as one cooperative launch,
grid barriers between each iteration's passes) against the per-iteration
launches: the same arithmetic per particle, so the same pressures and
positions bit for bit and the same iteration counts; the FP64 error sums
are grouped differently (per co-resident workgroup), so errors agree to
1e-12.  Single scene, batched ensemble through the step graphs with no host
read-back, and the learned start's fallback done on the device."""
import numpy as np
import pytest

import sphx
from test_gpu_ensemble import by_born, make as make_ensemble, options as ens_options
from test_gpu_mf import NU, RHO, setup

pytestmark = pytest.mark.gpu


def single(p, disp, calc, alpha, o):
    c = sphx.Context(p, o)
    c.upload(disp, calc)
    c.mf_setup(RHO, NU)
    c.mf_upload(alpha)
    c.call("grid_build")
    return c


def same_state(A, B):
    dA, cA = by_born(A)
    dB, cB = by_born(B)
    np.testing.assert_array_equal(dA["pos"], dB["pos"])
    np.testing.assert_array_equal(cA["vel"], cB["vel"])
    np.testing.assert_array_equal(cA["pressure"], cB["pressure"])


def test_persistent_solve_single_scene(gpu):
    p, disp, calc, alpha, o = setup()
    with single(p, disp, calc, alpha, o) as A, single(p, disp, calc, alpha, o) as B:
        B.persistent_solve(True)
        for step in range(5):
            ia, ea = A.mf_step()
            ib, eb = B.mf_step()
            print(f"step {step}: launches {ia} it err {ea:.3e}; persistent {ib} it err {eb:.3e}",
                  flush=True)
            assert ia == ib
            assert eb == pytest.approx(ea, rel=1e-12, abs=1e-300)
        same_state(A, B)


@pytest.mark.parametrize("learned", ["plain", "fallback"])
def test_persistent_ensemble_graph_steps(gpu, learned):
    """Eager steps (host checks, host fallback) against captured steps with
    the persistent solve and no read-back between steps; a poor learned
    start that misses the tolerance in every scene exercises the device
    fallback."""
    sizes = [2500, 5000, 3500]
    p, disp, calc, alpha, first, info = sphx.scene_ensemble(sizes, seed=6)
    with make_ensemble(p, disp, calc, alpha, 3) as A, \
            make_ensemble(p, disp, calc, alpha, 3) as B:
        if learned == "fallback":
            rng = np.random.default_rng(3)
            w = rng.normal(0, 0.05, 6020).astype(np.float32)
            w[6016] = 30.0       # a start far above the solution
            o = ens_options()
            o.max_iters = 6
            for c in (A, B):
                c.set_options(o)
                c.mlp_init(1, 7848.0, 1e-3)
                c.mlp_set_weights(w)
                c.mlp_mode(warm=True)
        A.mf_step()
        B.mf_step()
        B.persistent_solve(True)
        B.graph_build(2)
        its = []
        for step in range(4):
            ia, ea = A.mf_step()
            B.graph_step_async()
            its.append(ia)
        A.sync()
        B.sync()
        print("eager iterations", its, flush=True)
        same_state(A, B)
        ia, ea = A.mf_step()
        ib, eb = B.graph_step()
        print(f"last step: eager {ia} it err {ea:.3e}; persistent graph {ib} it err {eb:.3e}",
              flush=True)
        assert ia == ib
        assert eb == pytest.approx(ea, rel=1e-12, abs=1e-300)
        ea_s, eb_s = A.scene_stats(), B.scene_stats()
        np.testing.assert_array_equal(ea_s[0], eb_s[0])
        same_state(A, B)
        if learned == "fallback":
            assert B.mlp_stats()["fallbacks"] >= 1


def test_persistent_solve_and_online_refresh_exclude_each_other(gpu):
    """The persistent solve is refused while the learned start's online
    refresh is on, and the online refresh while the persistent solve is on
    (in either order the context stays as it was)."""
    p, disp, calc, alpha, o = setup()
    with single(p, disp, calc, alpha, o) as c:
        c.mlp_init(1, 7848.0, 1e-3)
        c.mlp_reserve(2 * c.num)
        c.mlp_online(5, 2, 1024, 1e-3, (1.0, 0.0, 0.0, 0.0))
        with pytest.raises(sphx.SphxError, match="online refresh"):
            c.persistent_solve(True)
        c.mlp_online(0, 0, 1024, 1e-3, (1.0, 0.0, 0.0, 0.0))
        c.persistent_solve(True)
        with pytest.raises(sphx.SphxError, match="persistent solve"):
            c.mlp_online(5, 2, 1024, 1e-3, (1.0, 0.0, 0.0, 0.0))
        c.persistent_solve(False)
        c.mlp_online(5, 2, 1024, 1e-3, (1.0, 0.0, 0.0, 0.0))
