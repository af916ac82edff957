#!/usr/bin/env python3
"""Benchmark of the MI355X CDD stepper: 3D 256^3 fp64 glide + climb
(Lambda = 0.9, i.e. climb mobility 0.1) under uniaxial strain-rate loading,
TVD-RK3, CFL 0.25, mu = 1, nu = 0.3 -- the workload the headline metric
(cell-updates/s + HBM roofline %) is quoted on.

A step is one full TVD-RK3 step (three stress solves + three fused
flux/stage kernels + the device-side CFL reduction) over the whole grid;
inputs are resident in HBM before the timed region starts.  After the
headline measurement the same line carries the benchmark's two other
metrics, measured compactly: "adjoint_vs_forward" (the 2D 512^2 x 400-step
energy gradient and the 3D 256^3 100-step stress window) and
"batched_vs_serial" (48 uneven 2D realisations); --no-sub skips them.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "cuda-plasticity_amd")]

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E datasheet peak
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix) datasheet peak


def spectral_bytes(n, d=3):
    """Algorithmic HBM bytes per cell of each kernel kind (fp64, complex = 16 B,
    half spectrum h = (n/2+1)/n; 6 symmetric stress components)."""
    h = (n // 2 + 1) / n
    return {
        # 9 real beta planes in, 6 half-spectrum planes out; the stage-2 launch
        # also reads L0 and writes F0 = u + dt L0 (one launch in three)
        "r2c": 72 + 96 * h + (72 + 72) / 3.0,
        "cols_fwd": 192 * h,          # 6 complex planes in + out
        "contract": 192 * h,
        "cols_inv": 192 * h,
        "c2r": 96 * h + 48,           # 6 half-spectrum planes in, 6 real G planes out
        "flux_stage1": 72 + 48 + 72,  # beta, G in; L0 out
        "flux_stage2": 72 + 48 + 72 + 72,  # F0, G, u in; F1 out
        "flux_stage3": 72 + 48 + 72 + 72,  # F1, G, u in; u' out
    }


def canonical_bytes_per_cell_update(n, d=3, order=3):
    """B/cell-update of the canonical fused pipeline: per stage
    B_fft = 96 + 384 h (d - 1); flux + update 408 B (stage 1) and 264 B
    (later stages)."""
    h = (n // 2 + 1) / n
    bfft = 96 + 384 * h * (d - 1)
    return order * bfft + 408 + 264 * (order - 1)


def cpu_baseline_numpy():
    """The NumPy oracle (the reference Python path restated) on a bounded
    sample of the same dynamics: 3D 64^3, 2 RK3 steps, on one core (every
    thread pool limited to one thread)."""
    from threadpoolctl import threadpool_limits
    from oracle import cdd_oracle as O
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    n, steps = 64, 2
    b = SpectralGaussianField((n,) * 3, 6.0, 0.01, seed=0)
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.9, 1.0, 0.3,
                     O.Load("uniaxial", init_strain=float(b[8].mean()), uniaxial_rate=0.01))
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        O.run(dyn, b, 0.0, steps, 0.25)
        dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "NumPy oracle, 3D 64^3 glide+climb Lambda=0.9 uniaxial load, 2 TVD-RK3 steps (%.1f s), "
                      "one thread" % dt}


def cpu_baseline(n=256):
    """The C/OpenMP restatement of the oracle on every host thread it is given
    (OMP_NUM_THREADS, else os.cpu_count()): one RK3 step of the benchmark's
    own workload, 3D n^3 glide + climb under the uniaxial control."""
    from oracle import cdd_oracle as O
    from oracle import cdd_oracle_c as OC
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    if not os.path.exists(OC.LIB):
        return None
    steps = 1
    b = SpectralGaussianField((n,) * 3, 6.0, 0.01, seed=0)
    dyn = O.Dynamics(O.GLIDE_CLIMB, 0.9, 1.0, 0.3,
                     O.Load("uniaxial", init_strain=float(b[8].mean()), uniaxial_rate=0.01))
    OC.flux(dyn, 0.0, b)  # warm-up: the oracle's kept workspaces and twiddles
    t0 = time.perf_counter()
    OC.run(dyn, b, 0.0, steps, 0.25)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    return {"value": n ** 3 * steps / dt, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": "C/OpenMP oracle, 3D %d^3 glide+climb Lambda=0.9 uniaxial load, %d TVD-RK3 step "
                      "(%.1f s) on %d threads" % (n, steps, dt, cores)}


def bench_loading(args):
    """3D n^3 (default 256) glide+climb, climb mobility 0.1, under the uniaxial
    strain-rate control, 1000 TVD-RK3 steps: the forward run with its per-step
    stress-strain record, then the checkpointed adjoint of the final mean
    stress sigma_zz with respect to the beta^P that starts the last 100 steps
    (back through those 100 steps, checkpoints every 10).  The steps follow
    the CFL rule (CalculateOneAdaptiveTimeStep, as RunLoadingMPI's mover
    does), so the strain 1000 steps reach at a given rate is only known after
    the run: two pilot runs of 1000 steps, at the rate r0 = 0.01 / (1000 dt0)
    (dt0 the initial CFL step) and at r1 = 0.01 / T0 (T0 the time the first
    covered), fix the rate by the secant of log strain against log rate, and
    the measured run ramps at that rate; the strain it reaches and its CFL
    numbers are reported."""
    import numpy as np
    import torch
    from Plasticity.amd import lib as L
    from Plasticity.amd.autograd import mean_stress_window
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    n, total, window, ck = args.n, 1000, 100, 10
    cells = n ** 3
    dev = torch.device("cuda", 0)
    beta = SpectralGaussianField((n,) * 3, 6.0, 0.01, seed=0)
    init_strain = float(beta[8].mean())
    beta_t = torch.from_numpy(beta).to(dev)
    del beta
    probe = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25)
    _, speed = probe.flux(beta_t)
    dt0 = 0.25 / (n * float(speed.max()))
    probe.close()
    del speed
    def pilot(r):
        pl = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25,
                    load_kind=L.LOAD_UNIAXIAL, load_init_strain=init_strain, load_uniaxial_rate=r)
        pl.set_state(beta_t)
        pl.set_time(0.0)
        pl.step(total)
        tt = pl.clock()[0]
        pl.close()
        print("pilot: rate %.4g, %d CFL steps cover t = %.6g, strain %.4g" % (r, total, tt, r * tt),
              file=sys.stderr, flush=True)
        return r * tt
    r0 = 0.01 / (total * dt0)
    s0 = pilot(r0)
    r1 = r0 * 0.01 / s0
    s1 = pilot(r1)
    rate = r1 * (0.01 / s1) ** (np.log(r0 / r1) / np.log(s0 / s1))
    pilots = [[r0, s0], [r1, s1]]
    load = dict(load_kind=L.LOAD_UNIAXIAL, load_init_strain=init_strain, load_uniaxial_rate=rate)
    plan = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25, **load)
    plan.set_state(beta_t)
    # CFL steps; every step's CFL speed is logged on the device and the run's
    # CFL number checked afterwards
    plan.set_time(0.0)
    slog = torch.zeros((total, 8), dtype=torch.float64, device=dev)
    plan.stress_record(slog)
    vlog = torch.zeros(total, dtype=torch.float64, device=dev)
    plan.speed_record(vlog)
    dtl = torch.zeros(total, dtype=torch.float64, device=dev)
    plan.dt_record(dtl)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.step(total - window)
    torch.cuda.synchronize()
    t_mid = plan.clock()[0]
    beta_mid = plan.get_state()  # the state that starts the last 100 steps
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plan.step(window)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    plan.stress_record(None)
    plan.speed_record(None)
    plan.dt_record(None)
    t_end = plan.clock()[0]
    fwd_s = (t1 - t0) + (t2 - t1)
    cfl_num = n * (vlog * dtl).cpu().numpy()  # N max speed dt per step
    rec = slog.cpu().numpy()
    print("forward %d steps: %.2f s" % (total, fwd_s), file=sys.stderr, flush=True)
    curve = [[float(rate * rec[k, 0]), float(rec[k, 3])] for k in range(0, total, 100)]
    final_state = plan.get_state()
    plan.close()
    # adjoint over the last window steps
    b = beta_mid.clone().requires_grad_(True)
    lt = torch.tensor(0.9, dtype=torch.float64, device=dev, requires_grad=True)
    torch.cuda.synchronize()
    ta = time.perf_counter()
    J = mean_stress_window(b, lt, window, 1, "zz", t0=t_mid, checkpoint_every=ck, rk_order=3, load=load)
    torch.cuda.synchronize()
    tb = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    tc = time.perf_counter()
    print("adjoint: forward %.2f s, backward %.2f s" % (tb - ta, tc - tb), file=sys.stderr, flush=True)
    g = b.grad
    out = {
        "metric": "config 5: 3D %d^3 loading ramp, %d RK3 steps fp64 + adjoint of the final stress over the last %d"
                  % (n, total, window),
        "value": cells * total / fwd_s,
        "unit": "cell-updates/s (forward)",
        "n_gpus": 1, "steps": total, "warmup": 0,
        "ms_per_step": 1e3 * fwd_s / total,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded Gaussian random beta^P (L=6 cells, RMS 0.01)",
        "config": {"workload": "cdd-3d-%d-loading" % n, "grid": [n, n, n], "dynamics": "glide+climb",
                   "Lambda": 0.9, "rk_order": 3, "cfl": 0.25,
                   "load": "uniaxial strain-rate control at rate %.4g, set so that %d CFL steps reach strain 0.01 "
                           "(two pilot runs; dt0 = %.4g the initial CFL step)" % (rate, total, dt0),
                   "pilot_runs_rate_strain": pilots,
                   "adjoint_window": window, "checkpoint_every": ck},
        "stress_strain_every_100_steps": curve,
        "final_strain": rate * t_end,
        "cfl_check": {"cfl": 0.25, "max_cfl_number": float(cfl_num.max()), "at_step": int(cfl_num.argmax()),
                      "first": float(cfl_num[0]), "last": float(cfl_num[-1]),
                      "within_cfl": bool(cfl_num.max() <= 0.25 * (1 + 1e-12))},
        "final_mean_sigma_zz": float(J.item()),
        "adjoint": {
            "objective": "mean sigma_zz (load included) of the final state",
            "wrt": "beta^P at the start of the last %d steps, and Lambda" % window,
            "forward_s": tb - ta, "backward_s": tc - tb,
            "cost_vs_forward": (tc - ta) / (tb - ta),
            "grad_l2": float(torch.linalg.norm(g)), "dJ_dLambda": float(lt.grad.item()),
            "grad_mean_zz": float(g[8].mean()),
        },
        "state_finite": bool(torch.isfinite(final_state).all()),
    }
    print(json.dumps(out))


def batched_vs_serial(K, W, count=48):
    """48 independent 2D realisations, sides drawn from {256,384,512,768,1024},
    per-instance seeds, L ~ U[3, 8] cells, RMS 0.01, glide-only ('lvp'),
    SSP-RK2, CFL 0.25: the batched hipGraph step (tiles autotuned) against a
    serial loop of plain single-instance plans over the same instances, K
    timed steps after W warm-up steps.  Returns a dict (ratio, times, the
    batched step's share of the 2D step roofline)."""
    import numpy as np
    import torch
    from Plasticity.amd import lib as L
    from Plasticity.amd.autotune import autotune_batch
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    rng = np.random.default_rng(2024)
    sizes = [int(s) for s in rng.choice([256, 384, 512, 768, 1024], size=count)]
    lens = rng.uniform(3.0, 8.0, size=count)
    betas = [torch.from_numpy(SpectralGaussianField((n, n), float(lens[i]), 0.01, seed=i)).cuda()
             for i, n in enumerate(sizes)]
    cells = sum(n * n for n in sizes)
    kw = dict(dynamics=L.GLIDE_ONLY, rk_order=2, cfl=0.25)
    b = L.Batch(sizes, **kw)
    for i in range(count):
        b.set_state(i, betas[i])
    b.set_time(-1, 0.0)
    choices, _ = autotune_batch(b, reps=3)
    b.step(W)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(K)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    steps_b = [b.clock(i)[3] for i in range(count)]
    # the dominant kernel: the persistent multi-instance flux launch (first
    # stage: beta and G in, rhs out over every instance), HIP events on the
    # batch's stream
    tile, per_cu = choices["multi"]
    flux_ms = b.time_flux_multi(tile, per_cu if per_cu else 2, reps=10)
    b.close()
    plans = []
    for i, n in enumerate(sizes):
        p = L.Plan(2, n, **kw)
        p.set_state(betas[i])
        p.set_time(0.0)
        plans.append(p)
    for p in plans:
        p.step(W)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p in plans:
        p.step(K)
    torch.cuda.synchronize()
    ts = time.perf_counter() - t0
    for p in plans:
        p.close()
    del betas
    torch.cuda.empty_cache()
    # the 2D RK2 step's canonical bytes, summed over the instances
    step_bytes = sum(n * n * canonical_bytes_per_cell_update(n, d=2, order=2) for n in sizes)
    flux_bytes = cells * spectral_bytes(512, d=2)["flux_stage1"]
    flux_gbs = flux_bytes / (flux_ms * 1e-3) / 1e9
    # calibrated counter bytes of the same launch (rocprofv3 FETCH_SIZE /
    # WRITE_SIZE of this workload), when the row march is the chosen form
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_batched.json")
    if tile == 3 and count == 48 and os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("k_flux_rows_stage1")
        except Exception:
            traffic = None
    return {"value": ts / tb, "unit": "x (serial time / batched time)", "steps": K, "warmup": W,
            "ms_per_step": 1e3 * tb / K, "serial_ms_per_step": 1e3 * ts / K,
            "cell_updates_per_s_batched": cells * K / tb, "cell_updates_per_s_serial": cells * K / ts,
            "step_roofline": {"bytes_per_step": step_bytes, "frac_batched": step_bytes * K / tb / (PEAK_HBM_GBS * 1e9),
                              "frac_serial": step_bytes * K / ts / (PEAK_HBM_GBS * 1e9)},
            "roofline": {"bound": "hbm", "kernel": "%s (first stage, all 48 instances in one launch)" % (
                "k_flux_rows" if tile == 3 else "k_flux_multi"),
                         "achieved": flux_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": flux_gbs / PEAK_HBM_GBS,
                         "traffic": traffic, "launch_ms": flux_ms, "algorithmic_bytes": flux_bytes},
            "sizes": sizes, "cells": cells, "flux_tiles": {str(k): v for k, v in choices.items()},
            "all_steps_ran": all(s == K + W for s in steps_b)}


def bench_batched(args):
    """The batched workload as its own line (batched_vs_serial)."""
    r = batched_vs_serial(args.steps, args.warmup)
    out = {
        "metric": "batched vs serial (48 uneven 2D realisations, SSP-RK2)",
        "value": r["value"], "unit": r["unit"],
        "n_gpus": 1, "steps": r["steps"], "warmup": r["warmup"],
        "ms_per_step": r["ms_per_step"], "serial_ms_per_step": r["serial_ms_per_step"],
        "cell_updates_per_s_batched": r["cell_updates_per_s_batched"],
        "cell_updates_per_s_serial": r["cell_updates_per_s_serial"],
        "step_roofline": r["step_roofline"],
        "roofline": r["roofline"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded Gaussian random beta^P per instance (L ~ U[3,8] cells, RMS 0.01)",
        "config": {"workload": "cdd-2d-batched-48", "sizes": r["sizes"], "cells": r["cells"], "dynamics": "lvp",
                   "rk_order": 2, "cfl": 0.25, "flux_tiles": r["flux_tiles"]},
        "all_steps_ran": r["all_steps_ran"],
    }
    print(json.dumps(out))


def adjoint_cost_2d(nsteps=400):
    """2D 512^2, nsteps SSP-RK2 steps, glide + climb with climb mobility 0.1
    (Lambda = 0.9), uniform shear sigma_xy = 0.005: the hand-written adjoint
    of the final total elastic energy w.r.t. the initial beta^P and Lambda
    (segments of 20 steps whose stage states are kept in device memory; the
    checkpoint-recomputation cost beside it) against the plain forward run.
    Returns (summary dict, context for the finite-difference report)."""
    import torch
    from Plasticity.amd import lib as L
    from Plasticity.amd.autograd import run_with_grad, elastic_energy
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    n, Lam = 512, 0.9
    shear = dict(load_kind=L.LOAD_TENSOR, load_init=(0, 0.005, 0, 0.005, 0, 0, 0, 0, 0))
    beta0 = torch.from_numpy(SpectralGaussianField((n, n), 4.0, 0.01, seed=4)).cuda()
    plan = L.Plan(2, n, dynamics=L.GLIDE_CLIMB, rk_order=2, Lambda=Lam, **shear)
    log = torch.zeros(nsteps, dtype=torch.float64, device="cuda")

    def forward(record):
        plan.set_state(beta0)
        plan.set_time(0.0)
        if record:
            plan.dt_record(log)
        plan.step(nsteps)
        e = plan.energy(plan.get_state(), 0.0)
        if record:
            plan.dt_record(None)
        return e
    forward(True)  # warm-up, records the schedule
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e_fwd = forward(False)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0

    def adjoint(store):
        b = beta0.clone().requires_grad_(True)
        lt = torch.tensor(Lam, dtype=torch.float64, device="cuda", requires_grad=True)
        E = elastic_energy(run_with_grad(b, lt, nsteps, checkpoint_every=20, rk_order=2, load=shear, store=store),
                           load=shear)
        E.backward()
        return b, lt
    # stored trajectories (every segment fits: 400 x 63 MB) versus checkpoint
    # recomputation; one untimed call first (plan set-up, device allocations)
    adjoint("auto")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b, lt = adjoint("auto")
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    adjoint(False)
    torch.cuda.synchronize()
    t_adj_rc = time.perf_counter() - t0
    torch.cuda.empty_cache()
    summary = {"value": t_adj / t_fwd, "unit": "x (forward+backward time / forward time)",
               "forward_ms_per_step": 1e3 * t_fwd / nsteps, "adjoint_ms_per_step": 1e3 * t_adj / nsteps,
               "cost_vs_forward_recompute": t_adj_rc / t_fwd, "steps": nsteps, "energy": e_fwd,
               "dE_dmobility": -lt.grad.item(),
               "config": {"workload": "cdd-2d-512-adjoint", "grid": [n, n], "dynamics": "glide+climb",
                          "Lambda": Lam, "rk_order": 2, "cfl": 0.25, "load": "sigma_xy = sigma_yx = 0.005",
                          "checkpoint_every": 20, "trajectory": "stored"}}
    ctx = dict(plan=plan, beta0=beta0, log=log, grad=b.grad.detach(), n=n, Lam=Lam, shear=shear, t_fwd=t_fwd,
               t_adj=t_adj, t_adj_rc=t_adj_rc, e_fwd=e_fwd, mob_grad=-lt.grad.item())
    return summary, ctx


def adjoint_cost_3d_window(n=256, window=100, ck=10):
    """3D n^3 glide + climb (Lambda = 0.9) under the uniaxial strain-rate
    control of the loading ramp (CFL steps): the adjoint of the final mean
    sigma_zz over a window of `window` RK3 steps (checkpoints every ck, stored
    while memory allows) against the plain forward run of the same steps,
    from the ramp's start state."""
    import torch
    from Plasticity.amd import lib as L
    from Plasticity.amd.autograd import mean_stress_window
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    beta = SpectralGaussianField((n,) * 3, 6.0, 0.01, seed=0)
    init_strain = float(beta[8].mean())
    beta0 = torch.from_numpy(beta).cuda()
    del beta
    probe = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25)
    _, speed = probe.flux(beta0)
    dt0 = 0.25 / (n * float(speed.max()))
    probe.close()
    del speed
    load = dict(load_kind=L.LOAD_UNIAXIAL, load_init_strain=init_strain, load_uniaxial_rate=0.01 / (1000 * dt0))
    plan = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25, **load)
    for rep in range(2):  # the first run warms up
        plan.set_state(beta0)
        plan.set_time(0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.step(window)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
    plan.close()
    b = beta0.clone().requires_grad_(True)
    lt = torch.tensor(0.9, dtype=torch.float64, device="cuda", requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = mean_stress_window(b, lt, window, 1, "zz", t0=0.0, checkpoint_every=ck, rk_order=3, load=load)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out = {"value": (t2 - t0) / t_fwd, "unit": "x (forward+backward time / forward time)",
           "forward_s": t_fwd, "adjoint_forward_s": t1 - t0, "adjoint_backward_s": t2 - t1, "steps": window,
           "objective": "final mean sigma_zz (load included)", "J": J.item(), "dJ_dLambda": lt.grad.item(),
           "config": {"workload": "cdd-3d-%d-loading-window" % n, "grid": [n, n, n], "rk_order": 3,
                      "checkpoint_every": ck, "steps": "CFL", "initial_dt": dt0}}
    del b, lt, J, beta0
    torch.cuda.empty_cache()
    return out


def bench_adjoint(args):
    """The 2D 512^2 adjoint workload (adjoint_cost_2d) as its own line, with
    central finite differences on random unit directions (eps = 1e-6, 1e-8)
    replaying the recorded time steps, reported as data: the scheme is only
    piecewise smooth, so these do not converge (the exactness checks are the
    tangent-linear dot products of the GPU tests)."""
    import numpy as np
    import torch
    nsteps = args.steps if args.steps != 20 else 400
    summary, c = adjoint_cost_2d(nsteps)
    plan, beta0, log, grad, n, Lam, shear = (c[k] for k in ("plan", "beta0", "log", "grad", "n", "Lam", "shear"))
    t_fwd, t_adj, t_adj_rc, e_fwd, mob_grad = (c[k] for k in ("t_fwd", "t_adj", "t_adj_rc", "e_fwd", "mob_grad"))
    from Plasticity.amd import lib as L
    store_bytes = L.Plan(2, n, dynamics=L.GLIDE_CLIMB, rk_order=2, Lambda=Lam, **shear).segment_store_bytes(20) * (
        nsteps // 20)

    def replay(b0):
        plan.set_state(b0)
        plan.set_time(0.0)
        plan.dt_replay(log)
        plan.step(nsteps)
        plan.dt_replay(None)
        return plan.energy(plan.get_state(), 0.0)
    e_base = replay(beta0)
    rng = np.random.default_rng(32)
    ndir = args.fd_dirs
    dirs = []
    for _ in range(ndir):
        d = torch.from_numpy(rng.standard_normal((9, n, n))).cuda()
        dirs.append(d / torch.linalg.norm(d))
    fd_stats = {}
    for eps in ((1e-6, 1e-8) if ndir else ()):
        errs, bracketed = [], 0
        for d in dirs:
            ep, em = replay(beta0 + eps * d), replay(beta0 - eps * d)
            cen, fwd, bwd = (ep - em) / (2 * eps), (ep - e_base) / eps, (e_base - em) / eps
            ad = float((grad * d).sum())
            errs.append(abs(cen - ad) / abs(ad))
            # a minmod / max-speed kink within +-eps puts the central difference
            # off by half the gap between the one-sided differences
            bracketed += abs(cen - ad) <= 1e-6 * abs(ad) + 0.5 * abs(fwd - bwd)
        fd_stats["%.0e" % eps] = {"rel_err_max": max(errs), "rel_err_median": float(np.median(errs)),
                                  "within_1e-6": int(sum(e <= 1e-6 for e in errs)),
                                  "within_1e-6_or_kink_bracket": int(bracketed)}
        print("fd eps %.0e: %s" % (eps, fd_stats["%.0e" % eps]), file=sys.stderr, flush=True)
    out = {
        "metric": "adjoint/forward cost (2D 512^2, 400 RK2 steps, energy gradient)",
        "value": t_adj / t_fwd,
        "unit": "x (forward+backward time / forward time)",
        "n_gpus": 1, "steps": nsteps, "warmup": 1,
        "ms_per_step": 1e3 * t_fwd / nsteps,
        "adjoint_ms_per_step": 1e3 * t_adj / nsteps,
        "stored_trajectory_bytes": store_bytes,
        "cost_vs_forward_recompute": t_adj_rc / t_fwd,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded Gaussian random beta^P (L=4 cells, RMS 0.01)",
        "config": {"workload": "cdd-2d-512-adjoint", "grid": [n, n], "dynamics": "glide+climb", "Lambda": Lam,
                   "rk_order": 2, "cfl": 0.25, "load": "sigma_xy = sigma_yx = 0.005", "checkpoint_every": 20,
                   "trajectory": "stored"},
        "energy": e_fwd,
        "dE_dmobility": mob_grad,
        "fd_directions": ndir,
        "fd_vs_adjoint": fd_stats,
    }
    print(json.dumps(out))


def bench_flux_vjp(args):
    """The hand-written flux adjoint on its own: cdd_flux_vjp (stress solve
    of beta, cell VJP, transposed stress solve, gather) on a seeded n^3
    field, `--steps` calls each for glide + climb (Lambda = 0.9) and glide
    only, then `--steps` forward glide-only fluxes; HIP events around each
    group.  One JSON line (ms per call)."""
    import torch
    from Plasticity.amd import lib as L
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField
    n = args.n
    beta = torch.from_numpy(SpectralGaussianField((n,) * 3, 6.0, 0.01, seed=0)).cuda()
    rbar = torch.from_numpy(SpectralGaussianField((n,) * 3, 4.0, 1.0, seed=1)).cuda()
    res = {}

    def timed(name, fn, reps):
        fn()  # warm-up
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        res[name] = e0.elapsed_time(e1) / reps
        print("%s: %.3f ms per call" % (name, res[name]), file=sys.stderr, flush=True)

    for label, dyn, lam in (("vjp_glide_climb", L.GLIDE_CLIMB, 0.9), ("vjp_glide_only", L.GLIDE_ONLY, 0.0)):
        p = L.Plan(3, n, dynamics=dyn, rk_order=3, Lambda=lam)
        timed(label, lambda: p.flux_vjp(beta, 0.0, rbar), args.steps)
        if dyn == L.GLIDE_ONLY:
            timed("flux_glide_only", lambda: p.flux(beta, 0.0), args.steps)
        p.close()
    out = {"metric": "flux adjoint (cdd_flux_vjp) time per call", "value": res["vjp_glide_climb"], "unit": "ms",
           "n_gpus": 1, "steps": args.steps, "warmup": 1, "higher_is_better": False, "dtype": "f64",
           "data": "synthetic: seeded Gaussian random beta^P (L=6 cells, RMS 0.01) and cotangent (L=4, RMS 1)",
           "config": {"workload": "cdd-3d-flux-vjp", "grid": [n] * 3}, "per_call_ms": res}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sub", action="store_true",
                    help="3d workload: skip the adjoint/forward and batched/serial sub-results")
    ap.add_argument("--fd-dirs", type=int, default=32, help="adjoint workload: finite-difference directions")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="3d workload: arithmetic of the plan (fp32 also reports its error against fp64)")
    ap.add_argument("--workload", default="3d", choices=["3d", "batched", "adjoint", "loading", "flux_vjp"],
                    help="3d: the headline 256^3 stepper (default); batched: 48 uneven 2D realisations, "
                         "batched hipGraph step vs a serial loop of plain plans")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(args.copy_calibration * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (independent realisations "
              "would shard one per GPU with no communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "flux_vjp":
        return bench_flux_vjp(args)
    if args.workload == "loading":
        return bench_loading(args)

    import numpy as np
    import torch
    from Plasticity.amd import lib as L
    from Plasticity.FieldInitializers.FieldInitializer import SpectralGaussianField

    n = args.n
    cells = n ** 3
    beta = SpectralGaussianField((n,) * 3, 6.0, 0.01, seed=0)
    init_strain = float(beta[8].mean())
    dev = torch.device("cuda", 0)
    beta_t = torch.from_numpy(beta).to(dev)
    del beta
    # load ramp: strain 0.01 reached after ~1000 steps of the initial time step
    probe = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25)
    _, speed = probe.flux(beta_t)
    dt0 = 0.25 / (n * float(speed.max()))
    probe.close()
    del speed
    rate = 0.01 / (1000.0 * dt0)
    loadkw = dict(load_kind=L.LOAD_UNIAXIAL, load_init_strain=init_strain, load_uniaxial_rate=rate)
    plan = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25,
                  precision=args.precision, **loadkw)
    plan.set_state(beta_t)
    plan.set_time(0.0)
    plan.step(args.warmup)
    torch.cuda.synchronize()
    plan.timing_enable(args.steps * 24 + 8)
    # stress-strain record of every timed step (written by the stress solve itself)
    slog = torch.zeros((args.steps, 8), dtype=torch.float64, device=dev)
    plan.stress_record(slog)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.step(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kt = plan.timing_read()
    plan.timing_enable(0)
    plan.stress_record(None)
    t, dt, done, nsteps = plan.clock()
    state = plan.get_state()
    finite = bool(torch.isfinite(state).all())
    elapsed = t1 - t0
    ms_step = 1e3 * elapsed / args.steps
    value = cells * args.steps / elapsed

    fp32_err = None
    if args.precision == "fp32":
        # the same run in float64: the fp32 variant's error is stated against it
        ref = L.Plan(3, n, dynamics=L.GLIDE_CLIMB, rk_order=3, Lambda=0.9, mu=1.0, nu=0.3, cfl=0.25, **loadkw)
        ref.set_state(beta_t)
        ref.set_time(0.0)
        ref.step(args.warmup + args.steps)
        s64 = ref.get_state()
        fp32_err = float(torch.linalg.norm(state.double() - s64) / torch.linalg.norm(s64))
        ref.close()
        del s64
    vb = 0.5 if args.precision == "fp32" else 1.0  # bytes per value relative to fp64
    per_cell = {k: v * vb for k, v in spectral_bytes(n).items()}
    kernels = {}
    for k, (ms, cnt) in kt.items():
        if cnt == 0:
            continue
        entry = {"total_ms": round(ms, 4), "launches": cnt, "avg_ms": round(ms / cnt, 5)}
        if k in per_cell:
            b = per_cell[k] * cells
            entry["gbs"] = round(b / (ms / cnt * 1e-3) / 1e9, 1)
        kernels[k] = entry
    # the dominant kernel by total time; the three flux stages are launches of
    # one kernel and count together (quoted on stage 2, the steady RK stage)
    fam = {}
    for k in kernels:
        if k in per_cell:
            f = "flux_stage2" if k.startswith("flux_stage") else k
            fam[f] = fam.get(f, 0.0) + kernels[k]["total_ms"]
    dom = max(fam, key=fam.get)
    dom_bytes = per_cell[dom] * cells
    achieved = dom_bytes / (kernels[dom]["avg_ms"] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_3d256.json")
    if os.path.exists(tpath) and args.precision == "fp64" and n == 256:
        try:
            traffic = json.load(open(tpath)).get(dom)
        except Exception:
            traffic = None
    # the dominant kernel's FP64 rate from its measured FP64 operations per cell
    fp64 = None
    fpath = os.path.join(ROOT, "profiles", "fp64_ops_3d256.json")
    if os.path.exists(fpath) and args.precision == "fp64" and n == 256:
        try:
            fpc = json.load(open(fpath)).get("flux" if dom.startswith("flux") else dom)
        except Exception:
            fpc = None
        if fpc:
            tf = fpc * cells / (kernels[dom]["avg_ms"] * 1e-3) / 1e12
            fp64 = {"bound": "fp64", "kernel": dom, "flops_per_cell": fpc, "achieved": round(tf, 2),
                    "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(tf / PEAK_FP64_TFLOPS, 4)}
    bpc = canonical_bytes_per_cell_update(n) * vb
    rec = slog.cpu().numpy()
    out = {
        "metric": "cell-updates/s (3D %d^3 %s glide+climb TVD-RK3)" % (n, args.precision),
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic: seeded Gaussian random beta^P (k-filter exp(-k^2 L^2/2), L=6 cells, RMS 0.01)",
        "config": {"workload": "cdd-3d-%d-glide-climb-rk3%s" % (n, "-fp32" if args.precision == "fp32" else ""),
                   "grid": [n, n, n], "dynamics": "glide+climb",
                   "Lambda": 0.9, "rk_order": 3, "cfl": 0.25, "mu": 1.0, "nu": 0.3,
                   "load": "uniaxial strain rate %.4g (0.01 after ~1000 steps)" % rate},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": dom_bytes},
        "roofline_fp64": fp64,
        "step_roofline": {"bytes_per_cell_update": bpc, "frac": round(bpc * value / (PEAK_HBM_GBS * 1e9), 4)},
        "kernels": kernels,
        "sim_time": t,
        "last_dt": dt,
        "state_finite": finite,
        "stress_strain": {"note": "per-step record of the timed steps: (strain = rate t, mean sigma_zz with load)",
                          "first": [float(rate * rec[0, 0]), float(rec[0, 3])],
                          "last": [float(rate * rec[-1, 0]), float(rec[-1, 3])]},
    }
    if fp32_err is not None:
        out["fp32_rel_l2_vs_fp64"] = fp32_err
    plan.close()
    del state, beta_t, slog
    torch.cuda.empty_cache()
    if not args.no_sub:
        # the two other metrics of the benchmark, compact: adjoint/forward cost
        # (2D 512^2 x 400 RK2 energy gradient; 3D 256^3 100-step stress window)
        # and the batched step against a serial loop of plain plans
        sub = {}
        for key, fn in (("2d_512_400_rk2_energy", lambda: adjoint_cost_2d(400)[0]),
                        ("3d_256_window_100_rk3_stress", lambda: adjoint_cost_3d_window(n))):
            try:
                sub[key] = fn()
            except Exception as e:  # reported, never fatal to the headline line
                sub[key] = {"error": repr(e)}
            print("adjoint %s: %s" % (key, sub[key].get("value", sub[key])), file=sys.stderr, flush=True)
            torch.cuda.empty_cache()
        out["adjoint_vs_forward"] = sub
        try:
            out["batched_vs_serial"] = batched_vs_serial(10, 3)
        except Exception as e:
            out["batched_vs_serial"] = {"error": repr(e)}
        print("batched: %s" % out["batched_vs_serial"].get("value"), file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        c = cpu_baseline(n)
        if c is not None:
            out["cpu_baseline"] = c
        out["cpu_baseline_numpy"] = cpu_baseline_numpy()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
