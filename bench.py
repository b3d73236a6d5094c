"""Benchmark of the stretch-sampler path on one MI355X.

Workload (one "step" = one whole job, inputs resident in HBM):
  4-D correlated Gaussian target, mu ~ N(0,1)^4, Sigma = A A' + I with A 4x4
  i.i.d. N(0,1) (numpy seed 1001); 1024 walkers initialised U(-5,5);
  burn-in 512 steps + run-sampler 4096 steps at a=2; then sample! 4096 x
  1024 draws and the dataset engine's mean, sd and 256-bin histogram of
  them.  ESS_d = Var_d(draws) / sigma_d^2 with sigma from the engine's acor
  of the run-sampler means; value = min_d ESS_d per second of the whole job.

Roofline: the plain stretch move (move-bare! kernel) on the same model at
2^24 walkers (256 MiB of positions, above the 256 MB Infinity Cache), timed
with HIP events on the sampler's stream; algorithmic bytes per walker-step
8*DIM + 4 + alpha*(4*DIM + 4).

CPU baseline: the same job on the host cores through the CPU restatement of
the reference kernels (oracle/), OpenMP over walkers.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("ESS/sec & walker-steps/sec, MFMA vs plain FP32 at posterior-mean err "
          "≤0.1 MC-SE")
DIM = 4
WALKERS = 1024
BURN_IN = 512
RUN_STEPS = 4096
SAMPLE_COLS = 4096 * 1024
A = 2.0
HBM_PEAK_GBS = 8000.0
ROOF_WALKERS = 1 << 24
ROOF_STEPS = 20


def c1_problem():
    rng = np.random.default_rng(1001)
    mu = rng.standard_normal(DIM)
    A_ = rng.standard_normal((DIM, DIM))
    sigma = A_ @ A_.T + np.eye(DIM)
    P = np.linalg.inv(sigma)
    params = np.concatenate([mu, P.reshape(-1)]).astype(np.float32)
    limits = np.tile(np.array([-5.0, 5.0], np.float32), DIM)
    return mu, sigma, params, limits


def ess_from(draws_var, sigma):
    return draws_var.astype(np.float64) / np.square(sigma.astype(np.float64))


def gpu_job(fac, sampler, sample_out, limits, seed):
    """One whole job on the GPU; returns per-phase timings and results."""
    import torch
    t0 = time.perf_counter()
    sampler.init(seed)
    sampler.init_position(seed + 1, limits)
    sampler.burn_in(BURN_IN, A)
    t1 = time.perf_counter()
    res = sampler.run_sampler(RUN_STEPS, A)
    t2 = time.perf_counter()
    sampler.sample(SAMPLE_COLS, out=sample_out)
    de = fac.dataset_engine()
    mean = de.data_mean(sample_out)
    var = de.data_variance(sample_out)
    hist = de.histogram(sample_out)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    return {"t_total": t3 - t0, "t_burn": t1 - t0, "t_run": t2 - t1, "t_sample": t3 - t2,
            "acor": res["autocorrelation"], "acc": res["acceptance-rate"], "mean": mean,
            "var": var, "hist": hist}


def roofline(fac, model, params):
    """HBM roofline of the move-bare! kernel at 2^24 walkers."""
    from bayadera_amd import _lib
    sf = fac.mcmc_factory(model)
    s = sf.create_sampler(77, ROOF_WALKERS, params)
    s.set_exec(_lib.BH_EXEC_STEPWISE).set_partner(_lib.BH_PARTNER_WALKER)
    s.init(77)
    s.init_position(78, np.tile(np.array([-5.0, 5.0], np.float32), DIM))
    s.burn_in(4, A)  # warm-up
    s.set_timing(True)
    s.burn_in(ROOF_STEPS, A)
    ms = s.last_gpu_ms()
    s.set_timing(False)
    alpha = s.acc_rate(A)
    launches = 2 * ROOF_STEPS
    avg_ms = ms / launches
    walkers_per_launch = ROOF_WALKERS // 2
    bytes_per_walker = 8 * DIM + 4 + alpha * (4 * DIM + 4)
    achieved = walkers_per_launch * bytes_per_walker / (avg_ms * 1e-3) / 1e9
    s.release()
    sf.release()
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "stretch_move_bare_2p24_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    return {"kernel": "bh_stretch_move_bare (DIM=4, K=2^23 walkers per launch)",
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "avg_kernel_ms": round(avg_ms, 4), "alpha": round(alpha, 4),
            "algorithmic_bytes_per_launch": int(walkers_per_launch * bytes_per_walker)}


def cpu_baseline(params, limits, budget_s=25.0):
    """The whole job on the host through the oracle (OpenMP over walkers)."""
    import oracle as O
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS"):
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    runs = []
    t_start = time.perf_counter()
    while not runs or (time.perf_counter() - t_start < budget_s / 2 and len(runs) < 3):
        t0 = time.perf_counter()
        s = O.Stretch(O.MVN_LOGPDF, DIM, WALKERS, params, 123, params_len=DIM + DIM * DIM,
                      partner=O.PARTNER_WALKER, threads=threads)
        s.init(123)
        s.init_position(124, limits.reshape(DIM, 2))
        s.burn_in(BURN_IN, A)
        res = s.run_sampler(RUN_STEPS, A)
        draws = s.sample(SAMPLE_COLS)
        O.mean_var_f64(draws)
        O.histogram(draws)
        t = time.perf_counter() - t0
        ess = ess_from(draws.astype(np.float64).var(0), res["autocorrelation"].sigma).min()
        runs.append((t, ess))
    t = float(np.median([r[0] for r in runs]))
    ess = float(np.median([r[1] for r in runs]))
    return {"value": round(ess / t, 1), "unit": "ESS/s", "cores": threads, "kind": "port",
            "sample": (f"whole job ({BURN_IN}+{RUN_STEPS} steps x {WALKERS} walkers, "
                       f"sample! {SAMPLE_COLS} draws, mean/var/histogram), median of "
                       f"{len(runs)} runs, {t:.2f} s each"),
            "walker_steps_per_s": round(WALKERS * (BURN_IN + RUN_STEPS + SAMPLE_COLS // WALKERS)
                                        / t, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding across GPUs is designed in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    from bayadera_amd import _lib, engine, models

    mu, sigma, params, limits = c1_problem()
    fac = engine.HipBayaderaFactory(0, 256)
    model = models.mvn_model(DIM)
    sf = fac.mcmc_factory(model)
    sampler = sf.create_sampler(123, WALKERS, params).set_partner(_lib.BH_PARTNER_WALKER)
    sample_out = torch.empty((SAMPLE_COLS, DIM), dtype=torch.float32, device="cuda")

    for i in range(args.warmup):
        gpu_job(fac, sampler, sample_out, limits, 1000 + i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    jobs = [gpu_job(fac, sampler, sample_out, limits, 123 + 17 * i) for i in range(args.steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3

    ess = np.array([ess_from(j["var"], j["acor"].sigma).min() for j in jobs])
    ess_per_s = float(ess.sum() / wall)
    run_time = sum(j["t_run"] for j in jobs)
    walker_steps = WALKERS * (BURN_IN + RUN_STEPS + SAMPLE_COLS // WALKERS) * args.steps
    last = jobs[-1]
    mcse = np.sqrt(np.diag(sigma) / ess_from(last["var"], last["acor"].sigma))
    line = {
        "metric": METRIC, "value": round(ess_per_s, 1), "unit": "ESS/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 4-D correlated Gaussian target)",
        "config": {"workload": "4-D correlated Gaussian, 1024 walkers, 512 burn-in + 4096 "
                               "run-sampler steps (a=2), sample! 4096x1024, dataset mean/sd/"
                               "256-bin histogram", "walkers": WALKERS, "dim": DIM,
                   "partner": "walker", "exec": "resident (one work-group, LDS)"},
        "walker_steps_per_s": round(walker_steps / wall, 1),
        "ess_per_s_run_phase": round(float(ess.sum() / run_time), 1),
        "ess_min_per_job": round(float(ess.mean()), 1),
        "tau": [round(float(t), 3) for t in last["acor"].tau],
        "acceptance_rate": round(float(last["acc"]), 4),
        "posterior_mean_err_over_mcse": [round(float(x), 3) for x in
                                         np.abs(last["mean"] - mu) / mcse],
        "phase_ms": {k: round(float(np.mean([j[k] for j in jobs])) * 1e3, 3)
                     for k in ("t_burn", "t_run", "t_sample")},
    }
    if not args.no_roofline:
        line["roofline"] = roofline(fac, model, params)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(params, limits)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
