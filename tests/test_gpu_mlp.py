"""Learned warm start on the GPU: features, FP16 inference and training
gradients against the FP64 oracle, then what the trained network gains
over the plain warm start on a two-phase tank.

Tolerances (normwise, infinity norm): features 1e-4 per column vs the
oracle on the same FP32 state; FP16-weight inference 1e-5 vs the FP64
forward pass with the same FP16-rounded weights; gradient 1e-5 and loss
1e-6 vs the FP64 oracle."""
import numpy as np
import pytest

import orc
import sphx
from helpers import state_from_gpu
from test_gpu_mf import NU, RHO, setup

pytestmark = pytest.mark.gpu

UNDEF = 0xFFFFFFFF


def scales(p, fill_height):
    g = float(np.linalg.norm(p.gravity.tolist()))
    return 1000.0 * g * fill_height, 1e-3


def staged_to_aii(ctx):
    """The stages of one two-phase step up to the a_ii pass; returns the
    fractions the step uses."""
    ctx.call("grid_build")
    ctx.call("iisph_density")
    pre = ctx.mf_download()
    for s in ["mf_drift_velocity", "mf_alpha_advance", "mf_correction", "iisph_predict_dii",
              "mf_force", "iisph_predict_aii"]:
        ctx.call(s)
    return pre


def test_features_match_oracle(gpu):
    p, disp, calc, alpha, o = setup()
    P, U = scales(p, 0.16)
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.mf_setup(RHO, NU)
        ctx.mf_upload(alpha)
        ctx.call("grid_build")
        for _ in range(3):
            ctx.mf_step()
        pre = staged_to_aii(ctx)
        ctx.mlp_init(1, P, U)
        Xg, pp, dpl = ctx.mlp_features()
        st = state_from_gpu(ctx, p, np.float32, o)
        a4 = np.zeros((st.n, 4), np.float32)
        a4[:, :2] = pre["alpha"]
        a4[:, 3] = pre["rho_m"]
        Xo = orc.mlp_features(st, a4, P, dpl)
        fl = (st["type"] == 0) & (st["mgcell"] != UNDEF)
        print(f"fluid {fl.sum()}", flush=True)
        np.testing.assert_array_equal(pp, st["pressure"])
        worst = 0.0
        for q in range(orc.MLP_NF):
            e = orc.relerr(Xg[fl, q], Xo[fl, q])
            worst = max(worst, e)
            assert e <= 1e-4, (q, e)
        assert np.all(Xg[~fl] == 0)
        print(f"features: worst column error {worst:.2e}")


def test_fp16_inference_matches_oracle(gpu):
    p, disp, calc, alpha, o = setup(n=2000, L=0.15)
    rng = np.random.default_rng(11)
    w = (rng.normal(0, 0.25, orc.MLP_NP)).astype(np.float32)
    X = rng.normal(0, 1, (5000, orc.MLP_NF)).astype(np.float32)
    with sphx.Context(p, o) as ctx:
        ctx.mlp_init(1, 1.0, 1.0)
        ctx.mlp_set_weights(w)
        Yg = ctx.mlp_eval(X)
    w16 = w.astype(np.float16).astype(np.float64)
    Y16, _, _ = orc.mlp_forward(w16, X.astype(np.float64))
    Y32, _, _ = orc.mlp_forward(w.astype(np.float64), X.astype(np.float64))
    e16 = orc.relerr(Yg, Y16)
    e32 = orc.relerr(Yg, Y32)
    print(f"inference: vs fp16-weight fp64 {e16:.2e}, vs fp32-weight fp64 {e32:.2e}")
    assert e16 <= 1e-5
    assert e32 <= 1e-2


def test_training_gradient_matches_oracle(gpu):
    p, disp, calc, alpha, o = setup(n=2000, L=0.15)
    rng = np.random.default_rng(12)
    w = (rng.normal(0, 0.2, orc.MLP_NP)).astype(np.float32)
    n = 3000
    X = rng.normal(0, 1, (n, orc.MLP_NF)).astype(np.float32)
    T = rng.normal(0, 1, (n, orc.MLP_NO)).astype(np.float32)
    M = (rng.uniform(size=n) > 0.2).astype(np.float32)
    lw = np.array([1.0, 0.25, 0.5, 0.125], np.float32)
    with sphx.Context(p, o) as ctx:
        ctx.mlp_init(1, 1.0, 1.0)
        ctx.mlp_set_weights(w)
        loss_g, grad_g = ctx.mlp_grad(X, T, lw, M)
        # all-valid call as well (mask NULL)
        loss_a, grad_a = ctx.mlp_grad(X[:257], T[:257], lw)
    v = M > 0
    loss_o, grad_o = orc.mlp_loss_grad(w.astype(np.float64), X[v], T[v], lw)
    e = orc.relerr(grad_g, grad_o)
    print(f"gradient: {e:.2e}, loss {loss_g:.6e} vs {loss_o:.6e}")
    assert e <= 1e-5
    assert loss_g == pytest.approx(loss_o, rel=1e-6)
    loss_o2, grad_o2 = orc.mlp_loss_grad(w.astype(np.float64), X[:257], T[:257], lw)
    assert orc.relerr(grad_a, grad_o2) <= 1e-5
    assert loss_a == pytest.approx(loss_o2, rel=1e-6)


def test_training_epoch_visits_every_sample_once(gpu):
    """An epoch visits the recorded samples in a keyed order computed on the
    device.  With one batch per epoch, the epoch's loss (summed before the
    Adam update) is the loss of the starting weights over the samples the
    order visited; it equals the loss over all valid samples (the gradient
    call, which takes them in stored order) only if each is visited exactly
    once.  The sample count is far from a power of 4, so the order walks
    past the end of its domain; two seeds."""
    p, disp, calc, alpha, o = column(n=3000, seed=5)
    P, _ = scales(p, 0.8)
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.mf_setup(RHO, NU)
        ctx.mf_upload(alpha)
        ctx.call("grid_build")
        ctx.mf_step()
        ctx.mlp_init(7, P, drift_scale(ctx))
        ctx.mlp_reserve(4 * ctx.num)
        ctx.mlp_mode(record=True)
        run_steps(ctx, 3)
        total, valid = ctx.mlp_samples()
        X, T, M = ctx.mlp_download_samples(0, total)
        lw = np.array([1.0, 0.5, 0.25, 0.125], np.float32)
        w0 = ctx.mlp_weights()
        loss_ref, _ = ctx.mlp_grad(X, T, lw, M)
        n4 = 4 ** int(np.ceil(np.log(valid) / np.log(4)))
        print(f"{valid} valid of {total} samples, permutation domain {n4}", flush=True)
        assert valid < 0.75 * n4
        for seed in (1, 2):
            ctx.mlp_set_weights(w0)
            loss = ctx.mlp_train(epochs=1, batch=total + 1, lr=1e-3, lw=lw, seed=seed)
            assert loss[0] == pytest.approx(loss_ref, rel=1e-9), seed


def drift_scale(ctx):
    """RMS of the phase-1 drift over the fluid: the drift normalisation."""
    d = ctx.mf_download()["drift"][:, 1, :]
    disp, _ = ctx.download()
    fl = disp["type"] == 0
    return max(float(np.sqrt(np.mean(np.sum(d[fl].astype(np.float64) ** 2, axis=1)))), 1e-9)


def loss_weights(T, M):
    """Each output's squared error over its target variance, the pressure
    increment weighted 4x."""
    var = np.var(T[M > 0].astype(np.float64), axis=0) + 1e-12
    return np.array([4.0 / var[0], 1.0 / var[1], 1.0 / var[2], 1.0 / var[3]], np.float32)


def run_steps(ctx, n):
    its, errs = [], []
    for _ in range(n):
        it, err = ctx.mf_step()
        its.append(it)
        errs.append(err)
    return np.array(its), np.array(errs)


def column(n=12000, seed=5):
    """A 0.1 x 1 x 0.1 m column filled to 0.8 m (the 1M tank's depth, spacing
    ratio h / s and bubble density) so the solver needs as many iterations
    as in the large tank."""
    Lx, Ly, Lz = 0.1, 1.0, 0.1
    s = (0.8 * Lx * Ly * Lz / n) ** (1.0 / 3.0)
    h = 1.0944 * s
    p, disp, calc, alpha, info = sphx.scene_bubbles(n_fluid=n, L=(Lx, Ly, Lz), fill=0.8, h=h,
                                                    n_bubbles=6, rmin=0.01, rmax=0.025,
                                                    layers=2, seed=seed)
    p.dt = 1e-3
    from test_gpu_mf import setup as _setup   # same solver options
    o = _setup(n=2000, L=0.15)[4]
    return p, disp, calc, alpha, o


def r2(Y, T, M):
    v = M > 0
    res = ((Y[v] - T[v]) ** 2).sum(axis=0)
    tot = ((T[v] - T[v].mean(axis=0)) ** 2).sum(axis=0) + 1e-30
    return 1.0 - res / tot


def test_learned_warm_start_cuts_iterations(gpu):
    """The benchmark's protocol on a deep column (the 1M tank's depth): the
    network is trained on the GPU on a separate column (other bubble layout)
    advanced with p0 = p(t-1); then, from the same state at t = 0 of the
    test column, the plain start (the reference's p0 = 0.5 p(t-1)) and the
    learned start (guard factor 10) run the same steps.  Every solve meets
    the tolerance, and the learned start needs at least 1.5x fewer iterations
    over the run and 2x fewer over its last 20 steps (the first steps of the
    resting column sit at the solver's minimum of 2 iterations)."""
    rec, steps = 20, 40
    p_t, disp_t, calc_t, alpha_t, o_t = column(seed=6)
    o_t.warm_start = 1.0
    P, _ = scales(p_t, 0.8)
    with sphx.Context(p_t, o_t) as T:
        T.upload(disp_t, calc_t)
        T.mf_setup(RHO, NU)
        T.mf_upload(alpha_t)
        T.call("grid_build")
        T.mf_step()
        U = drift_scale(T)
        T.mlp_init(7, P, U)
        T.mlp_reserve(rec * T.num)
        T.mlp_mode(record=True)
        it_t, _ = run_steps(T, rec)
        total, valid = T.mlp_samples()
        X, Tg, M = T.mlp_download_samples(0, total)
        assert valid >= 0.99 * rec * int((disp_t["type"] == 0).sum())
        assert np.all(np.isfinite(X)) and np.all(np.isfinite(Tg))
        lw = loss_weights(Tg, M)
        v = M > 0
        loss0 = float(np.mean(np.sum(lw * Tg[v].astype(np.float64) ** 2, axis=1)))
        loss = T.mlp_train(epochs=20, batch=2048, lr=2e-3, lw=lw, seed=3)
        print(f"training column iterations {it_t.tolist()}; train loss (zero predictor "
              f"{loss0:.3e}): " + " ".join(f"{x:.3e}" for x in loss), flush=True)
        print("in-sample R2 per output", r2(T.mlp_eval(X[v]), Tg[v], M[v]).round(3), flush=True)
        assert loss[-1] < 0.5 * loss0
        w = T.mlp_weights()
    p, disp, calc, alpha, o = column(seed=5)
    tol = o.tol
    with sphx.Context(p, o) as A, sphx.Context(p, o) as B:
        for c in (A, B):
            c.upload(disp, calc)
            c.mf_setup(RHO, NU)
            c.mf_upload(alpha)
            c.call("grid_build")
        A.mlp_init(7, P, U)
        A.mlp_set_weights(w)
        A.mlp_mode(warm=True)
        A.mlp_guard(10.0)
        ia, ea = run_steps(A, steps)
        ib, eb = run_steps(B, steps)
        st = A.mlp_stats()
        print(f"iterations learned {ia.mean():.2f} {ia.tolist()} fallbacks {st['fallbacks']} "
              f"guard {A.mlp_guard_stats()}")
        print(f"iterations plain   {ib.mean():.2f} {ib.tolist()}")
        assert np.all(ea <= tol) and np.all(eb <= tol)
        assert st["learned"] == steps
        # the first ~20 steps of the resting column need only the minimum
        # of 2 iterations either way; past them the learned start needs at
        # least 2x fewer
        late = slice(steps - 20, steps)
        print(f"last 20 steps: learned {ia[late].mean():.2f}, plain {ib[late].mean():.2f}")
        assert ib.mean() >= 1.5 * ia.mean()
        assert ib[late].mean() >= 2.0 * ia[late].mean()


def test_online_refresh_async_matches_sync(gpu):
    """The asynchronous online refresh (Adam steps on a second stream) trains
    the same weights as the in-stream refresh from the same samples; both
    solves keep the tolerance while the refreshed weights are adopted.
    Weights agree to 1e-5 (the weight gradients are summed with atomics, so
    their order is not fixed)."""
    p, disp, calc, alpha, o = setup()
    P, U = scales(p, 0.16)
    lw = np.array([1.0, 0.0, 0.0, 0.0], np.float32)
    with sphx.Context(p, o) as A, sphx.Context(p, o) as B:
        for c in (A, B):
            c.upload(disp, calc)
            c.mf_setup(RHO, NU)
            c.mf_upload(alpha)
            c.call("grid_build")
            c.mlp_init(11, P, U)
            c.mlp_reserve(4 * c.num)
            c.mlp_mode(warm=True)
            c.mlp_online(2, 3, 512, 1e-3, lw)
        B.mlp_online_async(True)
        w0 = A.mlp_weights()
        ea = [A.mf_step()[1] for _ in range(2)]
        eb = [B.mf_step()[1] for _ in range(2)]
        wa, wb = A.mlp_weights(), B.mlp_weights()
        d = np.max(np.abs(wa - wb)) / np.max(np.abs(wa))
        print(f"refreshed weights: change {np.max(np.abs(wa - w0)):.3e}, async vs sync {d:.3e}",
              flush=True)
        assert np.max(np.abs(wa - w0)) > 0
        assert d <= 1e-5
        ea += [A.mf_step()[1] for _ in range(4)]
        eb += [B.mf_step()[1] for _ in range(4)]
        B.sync()
        print("errors sync", np.round(ea, 6), "async", np.round(eb, 6), flush=True)
        assert max(ea) <= o.tol and max(eb) <= o.tol


def test_strided_online_recording(gpu):
    """Online recording with stride 4 keeps slots phase + 4 t of each solve,
    the phase rotating with the solve count, bit-identical to the same slots
    of a full recording of the same solves."""
    p, disp, calc, alpha, o = setup()
    P, U = scales(p, 0.16)
    lw = np.array([1.0, 0.0, 0.0, 0.0], np.float32)
    R = 4
    with sphx.Context(p, o) as A, sphx.Context(p, o) as B:
        for c in (A, B):
            c.upload(disp, calc)
            c.mf_setup(RHO, NU)
            c.mf_upload(alpha)
            c.call("grid_build")
            c.mlp_init(5, P, U)
            c.mlp_reserve(3 * c.num)
            c.mlp_mode(warm=True)
            c.mlp_online(1000, 1, 64, 1e-3, lw)   # ring recording, no refresh yet
        B.mlp_record_stride(R)
        n = A.num
        for _ in range(2):
            A.mf_step()
            B.mf_step()
        ta, _ = A.mlp_samples()
        tb, _ = B.mlp_samples()
        XA, TA, MA = A.mlp_download_samples(0, ta)
        XB, TB, MB = B.mlp_download_samples(0, tb)
        n0 = (n + R - 1) // R
        n1 = (n - 1 + R - 1) // R
        print(f"n {n}: full {ta} samples, strided {tb} ({n0} + {n1})", flush=True)
        assert ta == 2 * n and tb == n0 + n1
        sel = np.concatenate([np.arange(0, n, R), n + np.arange(1, n, R)])
        np.testing.assert_array_equal(XB, XA[sel])
        np.testing.assert_array_equal(TB, TA[sel])
        np.testing.assert_array_equal(MB, MA[sel])


def test_features_from_aii_pass_equal_feature_pass(gpu):
    """With the network in use, the a_ii pass computes the features in its
    own neighbour loop; the stand-alone feature pass recomputes them from the
    same state.  Same values (1e-6 normwise per column: the compiler may
    contract the same sums differently in the two kernels)."""
    p, disp, calc, alpha, o = setup()
    P, U = scales(p, 0.16)
    with sphx.Context(p, o) as ctx:
        ctx.upload(disp, calc)
        ctx.mf_setup(RHO, NU)
        ctx.mf_upload(alpha)
        ctx.call("grid_build")
        for _ in range(3):
            ctx.mf_step()
        ctx.mlp_init(1, P, U)
        ctx.mlp_mode(warm=True)
        staged_to_aii(ctx)
        Xf, pf = ctx.mlp_download_features()
        Xs, ps, _ = ctx.mlp_features()
    np.testing.assert_array_equal(pf, ps)
    worst = 0.0
    for q in range(Xs.shape[1]):
        scale = max(np.max(np.abs(Xs[:, q])), 1e-30)
        worst = max(worst, np.max(np.abs(Xf[:, q] - Xs[:, q])) / scale)
    print(f"a_ii-pass vs feature-pass features: worst column {worst:.3e}", flush=True)
    assert worst <= 1e-6
