"""Batched ensemble and step graphs on the GPU.

1. A scene advanced inside the ensemble equals the same scene advanced alone
   on the same grid, bit for bit (every launch is shared, nothing leaks
   between scenes, each scene's solve converges on its own).
2. Steps through the captured graphs equal eager steps bit for bit, replay
   after replay, with eager work and synchronisations between replays; with
   a small iteration budget (the host finishes the solve) and with a learned
   start whose poor predictions exercise the per-scene fallback."""
import numpy as np
import pytest

import sphx
from test_gpu_mf import NU, RHO, setup

pytestmark = pytest.mark.gpu


def options():
    return setup(n=2000, L=0.15)[4]


def make(p, disp, calc, alpha, nscenes):
    q = type(p).from_buffer_copy(p)
    q.num = q.maxNum = len(disp)
    ctx = sphx.Context(q, options())
    ctx.upload(disp, calc)
    ctx.mf_setup(RHO, NU)
    ctx.mf_upload(alpha)
    ctx.ensemble_setup(nscenes)
    ctx.call("grid_build")
    return ctx


def by_born(ctx):
    d, c = ctx.download()
    o = np.argsort(c["bornid"])
    return d[o], c[o]


def scene_state(ctx, lo, hi, base):
    """The scene's particles (bornid in [lo, hi)) in bornid order, with their
    fractions; bornids renumbered from `base`."""
    d, c = ctx.download()
    mf = ctx.mf_download()
    sel = (c["bornid"] >= lo) & (c["bornid"] < hi)
    o = np.argsort(c["bornid"][sel])
    return d[sel][o], c[sel][o], mf["alpha"][sel][o]


def test_scene_inside_ensemble_equals_scene_alone(gpu):
    sizes = [3000, 6000, 4500]
    p, disp, calc, alpha, first, info = sphx.scene_ensemble(sizes, seed=11)
    s = 1
    a, b = first[s], first[s + 1]
    calc_s = calc[a:b].copy()
    calc_s["bornid"] -= a           # the scene alone numbers its particles from 0
    with make(p, disp, calc, alpha, 3) as E, make(p, disp[a:b], calc_s, alpha[a:b], 3) as S:
        for step in range(4):
            E.mf_step()
            S.mf_step()
            ie, _ = E.scene_stats()
            is_, _ = S.scene_stats()
            dE, cE, aE = scene_state(E, a, b, a)
            dS, cS, aS = scene_state(S, 0, b - a, 0)
            diffs = {name: int(np.sum(np.any((x != y).reshape(len(x), -1), axis=1)))
                     for name, x, y in [("pos", dE["pos"], dS["pos"]), ("vel", cE["vel"], cS["vel"]),
                                        ("pressure", cE["pressure"], cS["pressure"]),
                                        ("dens", cE["dens"], cS["dens"]), ("alpha", aE, aS)]}
            print(f"step {step}: iterations ensemble {ie.tolist()} alone {is_[s]}; "
                  f"particles differing {diffs}", flush=True)
            assert ie[s] == is_[s]
            assert all(v == 0 for v in diffs.values()), diffs
        types = E.download()[0]["type"]
        assert set(np.unique(types >> 8)) == {0, 1, 2}


@pytest.mark.parametrize("learned", ["plain", "learned", "fallback"])
def test_graph_steps_equal_eager_steps(gpu, learned):
    sizes = [2500, 5000]
    p, disp, calc, alpha, first, info = sphx.scene_ensemble(sizes, seed=5)
    with make(p, disp, calc, alpha, 2) as A, make(p, disp, calc, alpha, 2) as B:
        if learned != "plain":
            rng = np.random.default_rng(2)
            w = rng.normal(0, 0.05, 6020).astype(np.float32)
            if learned == "fallback":
                # a start far above the solution that the solver cannot
                # repair within max_iters: every step falls back to the
                # plain start
                w[6016] = 30.0
            for c in (A, B):
                if learned == "fallback":
                    o = options()
                    o.max_iters = 6
                    c.set_options(o)
                c.mlp_init(1, 7848.0, 1e-3)
                c.mlp_set_weights(w)
                c.mlp_mode(warm=True)
        A.mf_step()
        B.mf_step()
        B.graph_build(1)            # budget below min_iters: the host finishes every solve
        for step in range(6):
            ia, ea = A.mf_step()
            ib, eb = B.graph_step()
            A.sync()
            dA, cA = by_born(A)     # eager work and synchronisation between replays
            dB, cB = by_born(B)
            print(f"step {step}: eager {ia} it err {ea:.2e}, graph {ib} it err {eb:.2e}", flush=True)
            assert ia == ib
            np.testing.assert_array_equal(dA["pos"], dB["pos"])
            np.testing.assert_array_equal(cA["vel"], cB["vel"])
            np.testing.assert_array_equal(cA["pressure"], cB["pressure"])
        if learned != "plain":
            sa, sb = A.mlp_stats(), B.mlp_stats()
            print("fallbacks eager", sa, "graph", sb)
            assert sa["fallbacks"] == sb["fallbacks"]
            if learned == "fallback":
                assert sb["fallbacks"] >= 6


def test_block_sizes_leave_results_unchanged_and_autotune(gpu):
    sizes = [3000, 4000]
    p, disp, calc, alpha, first, info = sphx.scene_ensemble(sizes, seed=9)
    with make(p, disp, calc, alpha, 2) as A, make(p, disp, calc, alpha, 2) as B:
        for k in B.TUNE_KERNELS:
            B.tune_set(k, 64)
        for _ in range(3):
            ia, _ = A.mf_step()
            ib, _ = B.mf_step()
            assert ia == ib
        dA, cA = by_born(A)
        dB, cB = by_born(B)
        np.testing.assert_array_equal(dA["pos"], dB["pos"])
        np.testing.assert_array_equal(cA["pressure"], cB["pressure"])
        B.graph_build(2)
        t = B.autotune(1)
        print("tuned", t)
        assert set(t.values()) <= {64, 128, 256}
        B.graph_step()
        B.sync()


def test_tune_cache_round_trip(gpu, tmp_path):
    """The autotuner's JSON cache, keyed by (architecture, size bucket =
    floor(log2 particles), kernel): a context of the same bucket reads back
    the saved sizes, one of another bucket reads none, and saving from it
    keeps the first bucket's entries."""
    import json
    path = tmp_path / "tune.json"
    p, disp, calc, alpha, first, info = sphx.scene_ensemble([3000, 4000], seed=9)
    n = len(disp)
    picks = {k: (64, 128, 256)[t % 3] for t, k in enumerate(sphx.Context.TUNE_KERNELS)}
    with make(p, disp, calc, alpha, 2) as A:
        assert A.tune_cache_load(path) == 0          # no file yet
        for k, b in picks.items():
            A.tune_set(k, b)
        A.tune_cache_save(path)
    with make(p, disp, calc, alpha, 2) as B:
        assert B.tune_cache_load(path) == len(picks)
        assert B.tuning() == picks
    entries = json.loads(path.read_text())
    bucket = int(np.floor(np.log2(n)))
    assert {e["bucket"] for e in entries} == {bucket}
    assert all(e["arch"].startswith("gfx") for e in entries)
    # another size bucket: nothing applies; its save adds its own entries
    p2, d2, c2, a2, _, _ = sphx.scene_ensemble([9000, 9000], seed=9)
    assert int(np.floor(np.log2(len(d2)))) != bucket
    with make(p2, d2, c2, a2, 2) as C:
        before = C.tuning()
        assert C.tune_cache_load(path) == 0
        assert C.tuning() == before
        C.tune_cache_save(path)
    entries = json.loads(path.read_text())
    assert len(entries) == 2 * len(picks)
    assert {e["bucket"] for e in entries} == {bucket, int(np.floor(np.log2(len(d2))))}
