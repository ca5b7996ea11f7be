#!/usr/bin/env python3
"""Headline benchmark: particle-steps/s of the two-phase bubble tank.

Workload: 1,048,576 fluid particles (water 1000 / air 1.2, 512 seeded
bubbles, radii U[0.01, 0.04] m) filling a 1 m tank to 0.8 m, h = 0.01 m,
IISPH pressure solve to a mean compression error of 0.1 % (FP64 reduction),
dt = 1e-3 s.  A step is the whole hot path: grid build, neighbour list,
density, drift velocity, fraction advection and correction, advected
velocity with mixture forces, a_ii, the pressure iterations and the
integration.  Inputs are resident in HBM before the timed region.

Prints one JSON line (rank 0, N = 1).  `value` counts fluid particles only.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "sph-multiphase_amd"))

RHO = [1000.0, 1.2]
NU = [1e-6, 1.5e-5]
HBM_PEAK_GBS = 8000.0

# HBM bytes per working launch measured with rocprofv3: FETCH_SIZE x 2.00 +
# WRITE_SIZE x 1.00, the factors a 256 MiB device copy reads on the same
# counters (profiles/iisph_mf_step_1m_hbm_traffic.md; TCC_EA0_RDREQ_128B x
# 128 B gives the same read bytes)
TRAFFIC_BYTES = {
    "iisph_iter": 212.3e6,
    "iisph_dp": 177.6e6,
    "mf_alpha": 398.5e6,
    "mf_drift": 279.1e6,
    "iisph_aii": 381.7e6,    # with the learned start
    "iisph_density_dii": 218.7e6,
    "nbr_build": 165.6e6,
    "mf_correction_force": 158.4e6,
}
# MPM sand (2M particles), same method (profiles/mpm_sand_2m_kernel_stats.md)
MPM_TRAFFIC_BYTES = {
    "mpm_p2g": 337.4e6,
    "mpm_force": 301.1e6,
    "mpm_stress": 536.9e6,
}

# Algorithmic bytes per fluid particle per launch (fields read or written
# once, natural sizes: float3 = 12 B); neighbour records re-read from L2 and
# the neighbour list a loop reads are not counted (the list build counts
# the list it writes).
ALGO_BYTES = {
    "iisph_dp": 68,        # x 12, p 4, m 4, d_ii 12 -> dijpj 12, m q 12, a_p 12
    "iisph_iter": 76,      # x, m, rho, rho0, a_ii, dijpj, G, H, rho_adv, p -> p, res, s
    "iisph_density": 36,   # x 12, m 4, rho0 4, type 4 (+ boundary V 4) -> rho 4
    "iisph_dii": 60,       # x, v, rho, m, rho0 -> v_adv 12, d_ii 12
    # density, v_adv and d_ii in one pass (the composite steps)
    "iisph_density_dii": 64,   # x, v, m, rho0, type -> rho 4, v_adv 12, d_ii 12
    # a_ii pass with the learned start: x, v_adv, d_ii, m, rho, p, rho0,
    # alpha, last increment (6 x 16 B) -> rho_adv, p0, s, G+H 16, a_ii
    # record 16, p(t-1), and the recorder's feature rows (every 8th
    # particle's 96 B)
    "iisph_aii": 156,
    "mf_drift": 64,        # x, p, m, rho, alpha(2), nu, g-a -> u_m (2x12)
    "mf_alpha": 84,        # x, v, m, rho, alpha, u_m -> alpha_new, drift stress
    "mf_force": 52,        # drift stress, rho_m, v_adv -> v_adv, a_mf
    # correction and force add in one pass (the composite steps): alpha_new,
    # alpha_pre, type, key, stress, v_adv -> alpha_new, nu, v_adv, a_mf
    "mf_correction_force": 124,
    "radix_sort": 32,      # 2 passes x (key+val read, key+val write)
    # per slot: key, index and 5 x 16 B + 2 x 4 B gathered (96 B), the same
    # written plus the born-id map entry (92 B)
    "reorder": 188,
    # x and key read 20, count 4, ps and the Jacobi record seeded 40, and the list
    # itself: 16 B per 8 entries (about 43 neighbours within 2h: 6 groups)
    "nbr_build": 160,
    "mlp_features": 184,   # x, phys, d_ii, v_adv, alpha, rho_adv -> 24 features, p(t-1)
    "mlp_infer": 140,      # 24 features, x, phys, key -> p0, s
}


def options(sphx):
    o = sphx.default_options()
    o.wcsph_kernel = 1          # cubic spline family (boundary volumes)
    o.full_support = 1          # pressure loops over the whole 2h support
    o.boundary_pressure = 1     # boundary terms of kern.cu:763,833,1000
    o.omega = 0.5
    o.clamp_pressure = 1
    o.tol = 1e-3
    o.min_iters = 2
    o.max_iters = 300
    o.clamp_to_grid = 1
    o.warm_start = 0.5          # reference warm start p0 = 0.5 p(t-1)
    o.drift_relax = 2e-11       # tau = relax / nu: 2e-5 s in water
    o.drift_diffusion = 1e-4
    o.boussinesq = 1
    o.rho_ref = 1000.0
    return o


def build_scene(sphx, n_fluid, L=(1.0, 1.0, 1.0), seed=20261203):
    p, disp, calc, alpha, info = sphx.scene_bubbles(n_fluid=n_fluid, L=L, fill=0.8, h=0.01,
                                                    n_bubbles=max(1, int(512 * n_fluid / (1 << 20))),
                                                    rmin=0.01, rmax=0.04, layers=2, seed=seed)
    p.dt = 1e-3
    return p, disp, calc, alpha, info


TRAIN_SEED = 20261213   # bubble layout of the training tank (the bench tank: 20261203)


def make_ctx(sphx, p, disp, calc, alpha, o):
    c = sphx.Context(p, o)
    c.upload(disp, calc)
    c.mf_setup(RHO, NU)
    c.mf_upload(alpha)
    c.call("grid_build")
    return c


def train_network(sphx, args):
    """The learned start's training run: a separate tank of the same
    configuration with another bubble layout (seed TRAIN_SEED), advanced
    with the reference solver started from p(t-1) (warm_start = 1) so the
    recorded converged pressures keep the hydrostatic load; every particle
    of `record_steps` steps is recorded and the network trained on the GPU."""
    p, disp, calc, alpha, info = build_scene(sphx, args.n, seed=TRAIN_SEED)
    o = options(sphx)
    o.warm_start = 1.0
    t = make_ctx(sphx, p, disp, calc, alpha, o)
    del disp, calc
    t.mf_step()
    g = float(np.linalg.norm(p.gravity.tolist()))
    P = 1000.0 * g * 0.8
    d = t.mf_download()["drift"][:, 1, :]
    dl, _ = t.download()
    fl = dl["type"] == 0
    U = max(float(np.sqrt(np.mean(np.sum(d[fl].astype(np.float64) ** 2, axis=1)))), 1e-9)
    del d, dl, fl
    t.mlp_init(20261202, P, U)
    t.mlp_reserve(args.record_steps * t.num)
    t.mlp_mode(record=True)
    its = []
    for _ in range(args.record_steps):
        its.append(t.mf_step()[0])
    t.sync()
    total, valid = t.mlp_samples()
    _, T, M = t.mlp_download_samples(0, min(total, 1 << 20))
    var = np.var(T[M > 0], axis=0) + 1e-12
    lw = (4.0 / var[0], 1.0 / var[1], 1.0 / var[2], 1.0 / var[3])
    t0 = time.perf_counter()
    loss = t.mlp_train(epochs=args.epochs, batch=args.batch, lr=args.lr, lw=lw, seed=args.train_seed)
    t.sync()
    train_s = time.perf_counter() - t0
    w = t.mlp_weights()
    # held out: the next step's samples, which the fit has not seen
    t.mlp_clear_samples()
    t.mf_step()
    total, _ = t.mlp_samples()
    X, T, M = t.mlp_download_samples(0, total)
    v = M > 0
    Y = t.mlp_eval(X[v])
    res = ((Y - T[v]) ** 2).sum(axis=0)
    tot = ((T[v] - T[v].mean(axis=0)) ** 2).sum(axis=0) + 1e-30
    r2 = (1.0 - res / tot).tolist()
    t.close()
    print(f"# training tank (seed {TRAIN_SEED}, p0 = p(t-1)): {args.record_steps} steps, "
          f"iterations {its}; {valid} samples; {args.epochs} epochs in {train_s:.2f} s, loss "
          + " ".join(f"{x:.3e}" for x in loss) + "; held-out R2 (pressure increment, drift "
          "correction x y z) " + " ".join(f"{x:.3f}" for x in r2), file=sys.stderr, flush=True)
    return dict(w=w, P=P, U=U, lw=lw, samples=int(valid), train_s=train_s, epochs=args.epochs,
                r2=r2)


def learned_ctx(sphx, args, p, disp, calc, alpha, net):
    o = options(sphx)   # warm_start 0.5: the plain start the guard and fallback use
    c = make_ctx(sphx, p, disp, calc, alpha, o)
    c.mlp_init(20261202, net["P"], net["U"])
    c.mlp_set_weights(net["w"])
    c.mlp_reserve(args.ring_steps * c.num)
    c.mlp_mode(warm=True, drift=args.drift, record=False)
    c.mlp_guard(args.guard)
    if args.online_every > 0:
        # online refresh trains the pressure output only
        c.mlp_online(args.online_every, args.online_iters, args.online_batch, args.online_lr,
                     (net["lw"][0], 0.0, 0.0, 0.0))
        c.mlp_online_async(bool(args.online_async))
        c.mlp_record_stride(args.record_stride)
    return c


NAMES = ["hash", "radix_sort", "cell_start", "reorder", "nbr_build", "iisph_density_dii",
         "mf_drift", "mf_alpha", "mf_correction_force", "iisph_aii", "mlp_features", "mlp_infer",
         "iisph_dp", "iisph_iter", "iisph_publish", "iisph_integrate", "mlp_record", "mlp_train", "totals"]


def run_method(sphx, args, make, label, only=None, track=False):
    """A fresh context from t = 0: `warmup` untimed steps, then `steps`
    timed steps (HIP events around `only`'s launches, or none)."""
    c = make()
    for _ in range(args.warmup):
        c.mf_step()
    if track:
        c.conservation_track(True)
    el, its, errs, kt = timed_segment(c, args.steps, NAMES, only=only)
    out = dict(elapsed=el, iters=its, errs=errs, kt=kt, ctx=c)
    print(f"# {label}: {1e3 * el / args.steps:.3f} ms/step, pressure iterations {its} "
          f"(mean {np.mean(its):.2f}), max error {max(errs):.2e}", file=sys.stderr, flush=True)
    return out


def bench_bubbles(args, sphx):
    """The headline: the 1M two-phase tank, three warm starts of the same
    solver, each a fresh context from the same seeded state at t = 0 over
    the same steps: the reference p0 = 0.5 p(t-1) (plain), p0 = p(t-1)
    (persistence, reported), and the learned start (trained beforehand on a
    separate tank).  `value` is the learned run's throughput."""
    p, disp, calc, alpha, info = build_scene(sphx, args.n)
    print(f"# scene: {info.n_fluid} fluid + {info.n_boundary} boundary particles, "
          f"grid {p.gridres.x}x{p.gridres.y}x{p.gridres.z}", file=sys.stderr, flush=True)
    net = train_network(sphx, args)
    plain = run_method(sphx, args, lambda: make_ctx(sphx, p, disp, calc, alpha, options(sphx)),
                       "plain (p0 = 0.5 p(t-1))")
    plain["ctx"].close()

    def persist_ctx():
        o = options(sphx)
        o.warm_start = 1.0
        return make_ctx(sphx, p, disp, calc, alpha, o)
    persist = run_method(sphx, args, persist_ctx, "persistence (p0 = p(t-1))")
    persist["ctx"].close()
    # the learned run, profiled (every kernel timed) to find the dominant
    # kernel, then again from t = 0 with events around that kernel only
    prof = run_method(sphx, args, lambda: learned_ctx(sphx, args, p, disp, calc, alpha, net),
                      "learned, profiled")
    dom = kernel_table(prof["kt"], NAMES, f"learned run, every kernel timed, {args.steps} steps "
                                          f"after {args.warmup} warm-up steps")
    prof["ctx"].close()
    run = run_method(sphx, args, lambda: learned_ctx(sphx, args, p, disp, calc, alpha, net),
                     "learned", only=dom, track=True)
    c = run["ctx"]
    stats = c.mlp_stats()
    guard = c.mlp_guard_stats()
    hist = c.conservation_history()[:, 0, :]
    c.close()
    nf = info.n_fluid
    elapsed, iters, errs, kt = run["elapsed"], run["iters"], run["errs"], run["kt"]
    t_dom, n_dom = kt[dom]
    # Launches of the pressure passes queued past convergence exit on the
    # device flag at once and move no bytes: the algorithmic bytes of the
    # iterations the solves ran are spread over every launch, so the
    # per-launch average and the launch time are both averages over all
    # launches (as a kernel trace reports them).
    n_iter = int(np.sum(iters))
    if dom in ("iisph_iter", "iisph_dp"):
        work_bytes = ALGO_BYTES[dom] * nf * n_iter   # guard probes count as iterations
    else:
        work_bytes = ALGO_BYTES.get(dom, 0) * nf * n_dom
    avg_ms = t_dom / max(n_dom, 1)
    algo = work_bytes / max(n_dom, 1)
    achieved = algo / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    print(f"# timed steps: {dom} {avg_ms * 1e3:.1f} us/launch x {n_dom} launches, "
          f"{algo / 1e6:.1f} algorithmic MB per launch", file=sys.stderr)
    m0 = hist[0, 0]
    cons = {"mass_rel_change": float(np.abs(hist[:, 0] - m0).max() / m0),
            "momentum_y_first_last": [float(hist[0, 2]), float(hist[-1, 2])],
            "kinetic_energy_first_last": [float(hist[0, 4]), float(hist[-1, 4])],
            "com_y_first_last": [float(hist[0, 6] / m0), float(hist[-1, 6] / hist[-1, 0])]}
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args.cpu_sample, float(np.mean(plain["iters"])))
        except Exception as e:   # the baseline is a report, not the measurement
            print(f"# cpu baseline failed: {e}", file=sys.stderr)
    ip, ipers, il = (float(np.mean(x["iters"])) for x in (plain, persist, run))
    line = {
        "metric": "particle-steps/s at 1M two-phase particles; pressure iters learned vs plain @0.1% err",
        "value": nf * args.steps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "two-phase water/air bubble tank, drift-flux mixture + IISPH to 0.1% error",
            "n_fluid": nf,
            "n_boundary": info.n_boundary,
            "bubbles": 512,
            "h": 0.01,
            "dt": 1e-3,
            "steps_timed": f"steps {args.warmup}..{args.warmup + args.steps - 1} from t = 0, "
                           "each method on a fresh context from the same seeded state",
            "pressure_iters_plain_mean": ip,
            "pressure_iters_persistence_mean": ipers,
            "pressure_iters_learned_mean": il,
            "pressure_iters_ratio": ip / max(il, 1e-9),
            "pressure_iters_ratio_vs_persistence": ipers / max(il, 1e-9),
            "max_density_error_plain": float(max(plain["errs"])),
            "max_density_error_persistence": float(max(persist["errs"])),
            "max_density_error_learned": float(max(errs)),
            "plain_ms_per_step": 1e3 * plain["elapsed"] / args.steps,
            "persistence_ms_per_step": 1e3 * persist["elapsed"] / args.steps,
            "learned_fallbacks": int(stats["fallbacks"]),
            "guard_probes": guard["probes"],
            "guard_kept_plain": guard["plain_picks"],
            "warm_start": "learned MLP 24-64-64-4 (FP16 weights) with a device-side guard against "
                          "0.5 p(t-1); fallback 0.5 p(t-1)",
            "training": f"separate tank (bubble seed {TRAIN_SEED}), {args.record_steps} steps "
                        f"with p0 = p(t-1), {net['samples']} samples, {net['epochs']} epochs "
                        f"on the GPU in {net['train_s']:.2f} s (untimed)",
            "drift_closure": bool(args.drift),
            "heldout_r2_pressure_increment": net["r2"][0],
            "heldout_r2_drift_correction": net["r2"][1:],
            "online_refresh": (f"every {args.online_every} steps: {args.online_iters} Adam steps "
                               f"of {args.online_batch} samples (inside the timed steps"
                               + (", on a second stream, adopted two steps later)"
                                  if args.online_async else ")")
                               + f"; recording every {args.record_stride}th particle of a step"
                               if args.online_every > 0 else None),
            "conservation_fp64": cons,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": TRAFFIC_BYTES.get(dom),
            "traffic_source": "rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 per working "
                              "launch, copy-calibrated (profiles/iisph_mf_step_1m_hbm_traffic.md)",
            "algorithmic_bytes_per_launch": algo,
            "avg_launch_ms": avg_ms,
            "launches": n_dom,
            "solve_iterations": n_iter,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(sample_fluid=131072, mean_iters=10.0, timed_iters=10, reps=5):
    """The oracle (FP32 restatement of the reference step, OpenMP) on a
    bounded sample of the workload: a tank of the same spacing, h and bubble
    density holding `sample_fluid` particles.  Each of `reps` steps times
    its fixed stages and `timed_iters` pressure iterations apart; a step's
    time is taken at the GPU run's mean iteration count, t_fixed + mean_iters
    * t_iteration, and the median over the steps is reported."""
    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "tests"))
    import sphx
    import orc
    side = (sample_fluid / (1 << 20)) ** (1.0 / 3.0)
    p, disp, calc, alpha, info = build_scene(sphx, sample_fluid, L=(side, side, side))
    o = options(sphx)
    st = orc.State(disp, calc, p, np.float32)
    st.opts = o
    mf = orc.MF(st, RHO, NU, o)
    rho = np.array([np.float32(r) for r in RHO], np.float64)
    mf.set_alpha(alpha, alpha @ rho)
    mf.a["mforce"][:] = [0.0, -9.81, 0.0]
    fl = st["type"] == 0
    st.a["visc"][fl] = (alpha[fl] @ np.array(NU, np.float32))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    steps = []
    for _ in range(reps):
        t0 = time.perf_counter()
        st.grid_build()
        src = np.where(st["mgcell"] != 0xFFFFFFFF, st["midsort"], np.arange(st.n))
        for k in ("alpha", "mforce"):
            mf.a[k][:] = mf.a[k][src]
        st.ii_density()
        mf.drift()
        mf.alpha_advance()
        mf.correction()
        st.ii_dii()
        mf.force()
        st.ii_aii()
        t1 = time.perf_counter()
        for _ in range(timed_iters):
            st.ii_iteration()
        t2 = time.perf_counter()
        st.ii_integrate()
        mf.a["alpha"][:] = mf.a["alpha_new"]
        t3 = time.perf_counter()
        t_fixed = (t1 - t0) + (t3 - t2)
        t_iter = (t2 - t1) / timed_iters
        steps.append((t_fixed + mean_iters * t_iter, t_fixed, t_iter))
    steps.sort()
    step, t_fixed, t_iter = steps[len(steps) // 2]
    return {"value": info.n_fluid / step, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"median of {reps} two-phase steps of a {info.n_fluid}-particle tank (same "
                      f"spacing, h and bubble density; {info.n_boundary} boundary particles), "
                      f"FP32 oracle on {threads} threads: fixed stages {t_fixed:.2f} s + "
                      f"{timed_iters} pressure iterations at {t_iter:.3f} s each, taken at "
                      f"{mean_iters:.1f} iterations/step (the plain GPU run's mean)"}


DAMBREAK_ALGO_BYTES = {
    # per particle: pos 16, vel 16, phys 16 -> rho, p (8)
    "wcsph_density": 56,
    # pos, vel, phys -> accel 16
    "wcsph_force": 64,
    # pos, vel, accel -> pos, vel
    "wcsph_advance": 80,
}


def dambreak_scene(sphx):
    """32^3 fluid lattice, spacing 0.01 m, jitter +-1e-4 m, in the corner
    of a 0.64 x 0.32 x 0.32 m box of boundary particles; cubic spline
    h = 0.02 m, Tait EOS (gamma 7, c0 20 m/s), artificial viscosity 0.01,
    dt = 1e-4 s."""
    p, disp, calc, info = sphx.scene_dambreak(32, 32, 32, 0.01, (0.64, 0.32, 0.32), 0.02, 2, 1e-4,
                                              20261202, 1)
    p.dt = 1e-4
    o = sphx.default_options()
    o.wcsph_kernel = 1
    o.eos_gamma = 7.0
    o.eos_B = 1000.0 * 20.0 ** 2 / 7.0
    o.art_visc_alpha = 0.01
    o.sound_speed = 20.0
    return p, disp, calc, info, o


def aggregates(pos, vel, fluid):
    x = pos[fluid].astype(np.float64)
    v = vel[fluid].astype(np.float64)
    return x.mean(0), 0.5 * (v * v).sum(1).mean()


def bench_dambreak(args, sphx):
    """The reference's CPU-runnable configuration: the single-phase WCSPH
    dam break, 100 steps from t = 0, on the GPU (sphx_wcsph_step) and on the
    reference step restated on the CPU (the oracle, FP32, OpenMP), each
    timed over the whole 100 steps; the two trajectories' centre of mass
    and kinetic energy are compared at the end."""
    steps = args.steps if args.steps != 20 else 100
    p, disp, calc, info, o = dambreak_scene(sphx)
    nf = info.n_fluid
    print(f"# dam break: {nf} fluid + {info.n_boundary} boundary particles, {steps} steps",
          file=sys.stderr, flush=True)

    def fresh():
        c = sphx.Context(p, o)
        c.upload(disp, calc)
        c.call("grid_build")
        c.call("boundary_volume")
        return c
    names = ["hash", "radix_sort", "cell_start", "reorder", "wcsph_density",
             "wcsph_force", "wcsph_advance"]
    step = None
    # profile pass (every kernel timed), then the timed pass from t = 0 with
    # events around the dominant kernel only
    c = fresh()
    _, _, _, kt_prof = timed_segment(c, steps, names, step=lambda: c.call("wcsph_step"))
    c.close()
    dom = kernel_table(kt_prof, names, f"dam break, every kernel timed, {steps} steps")
    c = fresh()
    elapsed, _, _, kt = timed_segment(c, steps, names, only=dom,
                                      step=lambda: c.call("wcsph_step"))
    d, cc = c.download()
    c.close()
    t_dom, n_dom = kt[dom]
    avg_ms = t_dom / max(n_dom, 1)
    algo = DAMBREAK_ALGO_BYTES.get(dom, 0) * nf
    achieved = algo / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    fl = d["type"] == 0
    com_g, ke_g = aggregates(d["pos"], cc["vel"], fl)
    cpu, cmp = None, None
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "oracle"))
        import orc
        st = orc.State(disp, calc, p, np.float32)
        st.opts = o
        st.grid_build()
        st.boundary_volume()
        per = []
        for _ in range(steps):
            t0 = time.perf_counter()
            st.grid_build()
            st.density_pressure()
            st.force()
            st.advance()
            per.append(time.perf_counter() - t0)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        com_c, ke_c = aggregates(st["pos"], st["vel"], st["type"] == 0)
        cmp = {"com_gpu": com_g.tolist(), "com_cpu": com_c.tolist(),
               "com_max_diff_m": float(np.abs(com_g - com_c).max()),
               "ke_gpu": float(ke_g), "ke_cpu": float(ke_c),
               "ke_rel_diff": float(abs(ke_g - ke_c) / max(ke_c, 1e-30))}
        print(f"# CPU reference step: {sum(per):.2f} s for {steps} steps (median step "
              f"{np.median(per) * 1e3:.1f} ms); GPU vs CPU after {steps} steps: {cmp}",
              file=sys.stderr, flush=True)
        cpu = {"value": nf * steps / sum(per), "unit": "particle-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"the whole workload: {steps} steps of the {nf}-particle dam break, "
                         f"FP32 oracle (the reference step restated) on {threads} threads, "
                         f"{sum(per):.2f} s (median step {np.median(per) * 1e3:.1f} ms)"}
    line = {
        "metric": "particle-steps/s, single-phase WCSPH dam break (32^3 particles, 100 steps)",
        "value": nf * steps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": 1e3 * elapsed / steps,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "WCSPH dam break, cubic spline h = 0.02, Tait gamma 7 c0 20, "
                               "artificial viscosity 0.01, dt 1e-4, steps from t = 0",
                   "n_fluid": nf, "n_boundary": info.n_boundary, "gpu_vs_cpu": cmp},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": algo, "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


MPM_ALGO_BYTES = {
    "mpm_p2g": 88,       # per particle: x, v, m (phys), B, key
    "mpm_gather": 16,    # per active node: {v, m} w
    "mpm_stress": 120,   # x, phys, sigma r/w, G/K, perm (node gathers served by L2)
    "mpm_force": 72,     # per particle: x, phys, sigma, key
    "mpm_g2p": 100,      # x r/w, v r/w, B w
    "mpm_node": 32,      # per active node: {v, m} r/w
    "mpm_clear": 4,      # per active block: flag w
    "reorder": 188,      # per slot, as ALGO_BYTES
}


def bench_sand(args, sphx):
    """MPM sand: 2^21 particles uniform in a 0.5 m cube inside [0,1]^3 over
    a 256^3 node grid (dx = 1/256, one particle per cell), quadratic
    B-spline, APIC, rate-form stress with Drucker-Prager (E = 3.5e5,
    nu = 0.3, phi = 30 deg), dt = 5e-5; a step is one substep of
    RunMPM (sort, P2G, stress, node update, G2P)."""
    n = args.n if args.n != (1 << 20) else (1 << 21)
    p, disp, calc, info = sphx.scene_sand(n=n, lo=0.25, hi=0.75, dx=1.0 / 256, res=256)
    p.dt = 5e-5
    print(f"# sand: {n} particles, grid {p.mpmRes.x}^3, sort grid {p.gridres.x}^3",
          file=sys.stderr, flush=True)
    ctx = sphx.Context(p)
    ctx.upload(disp, calc)
    ctx.mpm_setup(1, 1)
    ctx.call("grid_build")
    for _ in range(args.warmup):
        ctx.call("mpm_substep")
    ctx.sync()
    names = ["hash", "radix_sort", "cell_start", "reorder", "mpm_clear", "mpm_p2g",
             "mpm_gather", "mpm_stress", "mpm_force", "mpm_node", "mpm_g2p"]
    substep = lambda: ctx.call("mpm_substep")   # noqa: E731
    # profile segment (every kernel timed), then the timed steps with events
    # around the dominant kernel only
    _, _, _, kt_prof = timed_segment(ctx, args.profile_steps, names, step=substep)
    dom = kernel_table(kt_prof, names, f"profile segment, {args.profile_steps} substeps")
    elapsed, _, _, kt = timed_segment(ctx, args.steps, names, only=dom, step=substep)
    active = ctx.mpm_active_blocks()
    d, c = ctx.download()
    ok = bool(np.all(np.isfinite(d["pos"])) and np.all(np.isfinite(c["stress"])))
    ctx.close()
    nodes_active = active * 64
    t_dom, n_dom = kt[dom]
    avg_ms = t_dom / max(n_dom, 1)
    units = (nodes_active if dom in ("mpm_node", "mpm_gather") else
             active if dom == "mpm_clear" else n)
    algo = MPM_ALGO_BYTES.get(dom, 0) * units
    achieved = algo / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline_sand(sphx)
        except Exception as e:
            print(f"# cpu baseline failed: {e}", file=sys.stderr)
    line = {
        "metric": "particle-substeps/s, MPM sand (2M particles, 256^3 grid)",
        "value": n * args.steps / elapsed,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "MPM sand block: quadratic B-spline APIC, Drucker-Prager, "
                               "LDS-tile scatter", "n": n, "grid": 256, "dt": 5e-5,
                   "active_node_blocks": active, "finite": ok},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MPM_TRAFFIC_BYTES.get(dom),
                     "traffic_source": "rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, "
                                       "copy-calibrated (profiles/mpm_sand_2m_kernel_stats.md)",
                     "algorithmic_bytes_per_launch": algo, "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline_sand(sphx, sample=131072, reps=5):
    """The oracle's FP32 MPM stages (OpenMP over particles, FP64 atomic node
    sums) on a 131072-particle cube of the same spacing and material over a
    128^3 grid: the median of `reps` substeps."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import orc
    side = 0.5 * (sample / (1 << 21)) ** (1.0 / 3.0)
    p, disp, calc, info = sphx.scene_sand(n=sample, lo=0.25 - side / 2, hi=0.25 + side / 2,
                                          dx=1.0 / 256, res=128)
    p.dt = 5e-5
    st = orc.State(disp, calc, p, np.float32)
    mp = orc.MPM(st, p, 1, 1, calc)
    st.grid_build()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        st.grid_build()
        ok = st["mgcell"] != 0xFFFFFFFF
        src = np.where(ok, st["midsort"], np.arange(st.n))
        for k in ("B", "stress", "G", "K"):
            mp.a[k][:] = mp.a[k][src]
        mp.p2g(1)
        mp.stress()
        mp.node_update()
        mp.g2p(1)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": sample / t, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"median of {reps} substeps of a {sample}-particle cube (same dx, material, "
                      f"dt) over a 128^3 grid, FP32 oracle MPM stages on {threads} threads "
                      f"({t:.2f} s per substep)"}


def rpy_spheres(n, side, seed):
    """Centres uniform in a cube, rejection-sampled to a gap >= 0.1 max(a_i,
    a_j); radii uniform in [0.01, 0.03] m."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 3))
    rad = np.zeros(n)
    k = 0
    while k < n:
        x = rng.uniform(0, side, 3)
        a = rng.uniform(0.01, 0.03)
        if k:
            r = np.linalg.norm(pos[:k] - x, axis=1)
            if np.any(r - rad[:k] - a < 0.1 * np.maximum(rad[:k], a)):
                continue
        pos[k], rad[k] = x, a
        k += 1
    return pos.astype(np.float32), rad.astype(np.float32)


def rpy_dense_fp64(pos, rad, mu, F, block=256):
    """U = M F in FP64 with numpy, M built in row blocks of `block` spheres."""
    n = len(pos)
    p = pos.astype(np.float64)
    a = rad.astype(np.float64)
    F64 = F.astype(np.float64)
    U = np.zeros_like(F64)
    for i0 in range(0, n, block):
        i1 = min(n, i0 + block)
        d = p[None, :, :] - p[i0:i1, None, :]
        r2 = (d ** 2).sum(-1)
        idx = np.arange(i0, i1)
        r2[idx - i0, idx] = 1.0
        r = np.sqrt(r2)
        s2 = a[i0:i1, None] ** 2 + a[None, :] ** 2
        c0 = 1.0 / (8 * np.pi * mu * r)
        ci = c0 * (1 + s2 / (3 * r2))
        cr = c0 * (1 - s2 / r2) / r2
        M = ci[:, :, None, None] * np.eye(3) + cr[:, :, None, None] * d[:, :, :, None] * d[:, :, None, :]
        M[idx - i0, idx] = np.eye(3)[None] / (6 * np.pi * mu * a[i0:i1, None, None])
        M = M.transpose(0, 2, 1, 3).reshape(3 * (i1 - i0), 3 * n)
        U[3 * i0:3 * i1] = M @ F64
    return U


def blas_threads():
    """Threads numpy's BLAS runs on (the runtime's thread-count variables)."""
    for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        if os.environ.get(v, "").isdigit():
            return int(os.environ[v])
    return os.cpu_count()


def bench_mobility(args, sphx):
    """Stokesian mobility: N = 4096 polydisperse spheres in a 2 m cube, U =
    M F for 64 right-hand sides F ~ N(0, 1), bf16x3 + second-order split
    terms on MFMA; a step is one apply (19.33 GFLOP of the dense product)."""
    n, K, mu = 4096, 64, 1.0
    pos, rad = rpy_spheres(n, 2.0, 20261204)
    rng = np.random.default_rng(64)
    F = rng.standard_normal((3 * n, K)).astype(np.float32)
    flop = 2.0 * (3 * n) ** 2 * K
    import ctypes as C
    L = sphx.lib()
    dF = sphx.DeviceArray(F)
    dU = sphx.DeviceArray(np.zeros_like(F))
    res = {}
    with sphx.Mobility(n, K) as mob:
        mob.set_spheres(pos, rad, mu)
        if args.rpy_split:
            mob.set_split(args.rpy_split)
        for mode in (6, 3, 1):
            for _ in range(args.warmup):
                mob.apply_device(dF.ptr.value, dU.ptr.value, mode)
            mob.sync()
            mob.timing(True)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                mob.apply_device(dF.ptr.value, dU.ptr.value, mode)
            mob.sync()
            el = (time.perf_counter() - t0) / args.steps
            tile_ms, tile_n = mob.timing_get()
            mob.timing(False)
            res[mode] = (el, dU.get(), tile_ms / max(tile_n, 1) * 1e-3)
    t1 = time.perf_counter()
    U64 = rpy_dense_fp64(pos, rad, mu, F)
    t_cpu = time.perf_counter() - t1
    errs = {m: float(np.abs(u - U64).max() / np.abs(U64).max()) for m, (t, u, k) in res.items()}
    for m, (t, u, k) in res.items():
        print(f"# mode {m}: {t * 1e6:.1f} us/apply (k_rpy_tile {k * 1e6:.1f} us), "
              f"{flop / t / 1e12:.1f} TFLOP/s per apply, {flop / k / 1e12:.1f} TFLOP/s in the "
              f"tile kernel, error vs FP64 {errs[m]:.2e}", file=sys.stderr)
    t6, k6 = res[6][0], res[6][2]
    achieved = flop / k6 / 1e12
    line = {
        "metric": "RPY mobility applies/s (N=4096 spheres, 64 RHS, bf16x3 MFMA + refinement)",
        "value": 1.0 / t6,
        "unit": "applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t6 * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "bf16 (x3 split, FP32 accumulate)",
        "data": "synthetic",
        "config": {"workload": "polydisperse RPY U = M F, matrix-free on MFMA", "n": n, "nrhs": K,
                   "error_vs_fp64": errs,
                   "us_per_apply": {m: res[m][0] * 1e6 for m in res},
                   "us_tile_kernel": {m: res[m][2] * 1e6 for m in res}},
        "roofline": {"bound": "mfma", "kernel": "k_rpy_tile", "achieved": achieved,
                     "peak": 2500.0, "unit": "TFLOP/s", "frac": achieved / 2500.0,
                     "traffic": None, "algorithmic_flop_per_launch": flop,
                     "avg_launch_ms": k6 * 1e3,
                     "note": "achieved counts the 19.33 GFLOP of the dense product once, over "
                             "the tile kernel's HIP-event time; mode 6 issues 6 MFMA products "
                             "per tile"},
        "cpu_baseline": {"value": 1.0 / t_cpu, "unit": "applies/s", "cores": blas_threads(),
                         "kind": "port",
                         "sample": f"one FP64 apply with numpy (dense M in 256-sphere row blocks, "
                                   f"BLAS threads): {t_cpu:.2f} s"},
    }
    print(json.dumps(line), flush=True)


def ensemble_scene(sphx, S, seed):
    rng = np.random.default_rng(20261207)
    sizes = np.exp(rng.uniform(np.log(2e4), np.log(4e5), S)).astype(np.int64)
    p, disp, calc, alpha, first, info = sphx.scene_ensemble(sizes.tolist(), seed=seed)
    p.dt = 1e-3
    return sizes, p, disp, calc, alpha, info


def ensemble_ctx(sphx, S, p, disp, calc, alpha, warm_start=0.5):
    o = options(sphx)
    o.warm_start = warm_start
    c = sphx.Context(p, o)
    c.upload(disp, calc)
    c.mf_setup(RHO, NU)
    c.mf_upload(alpha)
    c.ensemble_setup(S)
    c.call("grid_build")
    return c


def bench_ensemble(args, sphx):
    """Batched ensemble: 64 two-phase tanks with fluid counts log-uniform in
    [20k, 400k] (seeded), bubble layouts seeded per scene, all in one grid so
    every stage is one launch.  A step is graph A (sort .. the learned start
    and a budget of launched pressure iterations), the persistent solve (one
    cooperative launch that continues any scene still open and returns at
    once when none is; with the learned start, the device-side fallback and
    a second launch) and graph B (integration), with nothing read back
    between steps.  Three warm starts from the same
    state: plain 0.5 p(t-1), persistence p(t-1) (reported), and the learned
    start, trained on a separate ensemble (other bubble seeds) advanced with
    p0 = p(t-1)."""
    S = 64
    # training ensemble: same sizes, other bubble layouts
    sizes, p, disp, calc, alpha, info = ensemble_scene(sphx, S, 20261218)
    t = ensemble_ctx(sphx, S, p, disp, calc, alpha, warm_start=1.0)
    del disp, calc, alpha
    t.mf_step()
    g = float(np.linalg.norm(p.gravity.tolist()))
    t.mlp_init(20261202, 1000.0 * g * 0.8, 0.05)
    # as many recorded steps as the single tank's training run, every 8th
    # particle of each (about as many samples)
    rec, stride = args.record_steps, 8
    t.mlp_record_stride(stride)
    t.mlp_reserve(rec * ((t.num + stride - 1) // stride))
    t.mlp_mode(record=True)
    for _ in range(rec):
        t.mf_step()
    total, valid = t.mlp_samples()
    _, T, M = t.mlp_download_samples(0, min(total, 1 << 21))
    var = np.var(T[M > 0], axis=0) + 1e-12
    lw = (4.0 / var[0], 1.0 / var[1], 1.0 / var[2], 1.0 / var[3])
    t0 = time.perf_counter()
    loss = t.mlp_train(epochs=args.epochs, batch=args.batch, lr=args.lr, lw=lw, seed=args.train_seed)
    t.sync()
    train_s = time.perf_counter() - t0
    w = t.mlp_weights()
    t.close()
    print(f"# training ensemble: {valid} samples in {rec} steps, {args.epochs} epochs in "
          f"{train_s:.1f} s, loss {loss[-1]:.3e}", file=sys.stderr, flush=True)
    # the benchmark ensemble
    sizes, p, disp, calc, alpha, info = ensemble_scene(sphx, S, 20261208)
    nf = int(info.n_fluid)
    print(f"# ensemble: {S} scenes, {nf} fluid + {info.n_boundary} boundary particles, grid "
          f"{p.gridres.x}x{p.gridres.y}x{p.gridres.z}", file=sys.stderr, flush=True)

    def run(label, ws, learned):
        c = ensemble_ctx(sphx, S, p, disp, calc, alpha, warm_start=ws)
        if learned:
            c.mlp_init(20261202, 1000.0 * g * 0.8, 0.05)
            c.mlp_set_weights(w)
            c.mlp_mode(warm=True)
        for _ in range(args.warmup):
            c.mf_step()
        c.persistent_solve(True)     # restarts the solve log
        c.graph_build(args.graph_budget)
        c.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            c.graph_step_async()
        c.sync()
        el = time.perf_counter() - t0
        its, errs = c.solve_history(args.steps)
        stats = c.mlp_stats() if learned else None
        kt = None
        if learned:
            # the Jacobi pass timed with HIP events on eager steps of the same
            # run (launches inside a graph carry no events)
            c.persistent_solve(False)
            _, eit, _, kt = timed_segment(c, 5, ["iisph_iter"], only="iisph_iter")
            kt = dict(t=kt["iisph_iter"][0], n=kt["iisph_iter"][1], iters=int(np.sum(eit)))
        c.close()
        print(f"# {label}: {1e3 * el / args.steps:.2f} ms/step, most iterations of any scene "
              f"{its.tolist()} (mean {its.mean():.2f}), worst error {errs.max():.2e}",
              file=sys.stderr, flush=True)
        return dict(el=el, its=its, errs=errs, kt=kt, stats=stats)

    plain = run("plain (p0 = 0.5 p(t-1))", 0.5, False)
    pers = run("persistence (p0 = p(t-1))", 1.0, False)
    learn = run("learned", 0.5, True)
    kt = learn["kt"]
    # the Jacobi pass moves 76 B per fluid particle per iteration; launches
    # queued past convergence exit at once and move none, so bytes and time
    # are both averaged over all launches
    avg_ms = kt["t"] / max(kt["n"], 1)
    algo_per_launch = ALGO_BYTES["iisph_iter"] * nf * kt["iters"] / max(kt["n"], 1)
    achieved = algo_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    print(f"# eager steps of the learned run: iisph_iter {avg_ms * 1e3:.1f} us/launch x {kt['n']}",
          file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(65536, float(np.mean(plain["its"])))
        except Exception as e:
            print(f"# cpu baseline failed: {e}", file=sys.stderr)
    ip, ipers, il = (float(np.mean(x["its"])) for x in (plain, pers, learn))
    line = {
        "metric": "aggregate particle-steps/s, 64-scene two-phase ensemble (log-uniform 20k-400k)",
        "value": nf * args.steps / learn["el"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * learn["el"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "64 bubble tanks in one grid, one launch per stage; step = graph A "
                               f"(with {args.graph_budget} launched pressure iterations), the "
                               "persistent cooperative solve continuing any scene still open "
                               "(+ device fallback), graph B; no host read-back between steps",
                   "scenes": S, "n_fluid": nf, "n_boundary": int(info.n_boundary),
                   "pressure_iters_plain_mean": ip,
                   "pressure_iters_persistence_mean": ipers,
                   "pressure_iters_learned_mean": il,
                   "pressure_iters_ratio": ip / max(il, 1e-9),
                   "iteration_count": "most iterations of any scene, per step",
                   "plain_ms_per_step": 1e3 * plain["el"] / args.steps,
                   "persistence_ms_per_step": 1e3 * pers["el"] / args.steps,
                   "max_density_error_learned": float(learn["errs"].max()),
                   "max_density_error_plain": float(plain["errs"].max()),
                   "train_samples": int(valid), "train_seconds": train_s,
                   "online_refresh": None},
        "roofline": {"bound": "hbm", "kernel": "iisph_iter",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": algo_per_launch, "avg_launch_ms": avg_ms,
                     "launches": kt["n"],
                     "note": "HIP events around the Jacobi launches of 5 eager steps after the "
                             "timed graph steps of the learned run; 76 B x fluid particles x "
                             "iterations, averaged over launches"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def timed_segment(ctx, steps, names, only=None, step=None):
    """`steps` timed steps between two device synchronisations; returns the
    wall time, per-step iterations and errors, and per-kernel event times.
    With `only`, HIP events bracket only that kernel's launches (the timed
    steps then carry two events per launch of it, not of every kernel)."""
    step = step or ctx.mf_step
    ctx.sync()
    ctx.timing_only(only)
    ctx.timing(True)
    ctx.timing_reset()
    iters, errs = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        r = step()
        if r is not None:
            iters.append(r[0])
            errs.append(r[1])
    ctx.sync()
    elapsed = time.perf_counter() - t0
    kt = {k: ctx.timing_get(k) for k in names}
    ctx.timing(False)
    ctx.timing_only(None)
    return elapsed, iters, errs, kt


def kernel_table(kt, names, label):
    """Print the per-kernel event table of a profile segment; returns the
    kernel with the largest total time."""
    total_k = sum(t for t, n in kt.values()) or 1.0
    print(f"# per-kernel HIP-event times ({label})", file=sys.stderr)
    for k in names:
        t, n = kt[k]
        if n:
            print(f"# {k:16s} {t / n * 1e3:9.1f} us/launch x {n:5d}  ({100 * t / total_k:5.1f}% of kernel time)",
                  file=sys.stderr)
    return max(kt, key=lambda k: kt[k][0])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--profile-steps", type=int, default=40,
                    help="steps with every kernel timed, before the timed steps")
    ap.add_argument("--n", type=int, default=1 << 20, help="fluid particles")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--record-steps", type=int, default=20,
                    help="steps of the training tank recorded for the offline fit")
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--train-seed", type=int, default=1, help="seed of the training sample order")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--drift", type=int, default=0, help="learned drift closure on")
    ap.add_argument("--guard", type=float, default=10.0,
                    help="guard factor of the learned start (probe 0.5 p(t-1) when the first "
                         "iteration's error exceeds factor * tol; 0: off)")
    ap.add_argument("--ring-steps", type=int, default=12, help="steps of samples kept")
    ap.add_argument("--online-every", type=int, default=5)
    ap.add_argument("--online-iters", type=int, default=8)
    ap.add_argument("--online-batch", type=int, default=8192)
    ap.add_argument("--online-lr", type=float, default=1e-3)
    ap.add_argument("--record-stride", type=int, default=8,
                    help="online recording keeps every n-th particle of a step")
    ap.add_argument("--online-async", type=int, default=1,
                    help="refresh on a second stream, adopted two solves later")
    ap.add_argument("--workload", default="bubbles",
                    choices=["bubbles", "sand", "mobility", "ensemble", "dambreak"],
                    help="bubbles: two-phase tank (headline); sand: MPM 2M particles; "
                         "mobility: RPY 4096 spheres x 64 RHS")
    ap.add_argument("--rpy-split", type=int, default=0)
    ap.add_argument("--graph-budget", type=int, default=12,
                    help="ensemble: pressure iterations launched inside the step graph before "
                         "the persistent solve takes over")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): this build measures "
              "one GPU's share at its real shape; the multi-GPU decomposition is not built",
              file=sys.stderr)
        sys.exit(2)
    import sphx
    if sphx.device_count() < 1:
        print("no HIP device visible: bench.py needs an MI355X", file=sys.stderr)
        sys.exit(1)

    if args.workload == "ensemble":
        bench_ensemble(args, sphx)
        return
    if args.workload == "mobility":
        bench_mobility(args, sphx)
        return
    if args.workload == "sand":
        bench_sand(args, sphx)
        return
    if args.workload == "dambreak":
        bench_dambreak(args, sphx)
        return
    bench_bubbles(args, sphx)


if __name__ == "__main__":
    main()
