"""Benchmark of the MI355X Bayadera engine on one GPU.

Default workload (the metric's comparison, matrix cores vs plain fp32):
  Bayesian logistic regression, precision form.  X is 1,000,000 x 64 with
  rows N(0,1)/8, stored bf16 (numpy seed 1003); beta* ~ N(0,1); y ~
  Bernoulli(sigmoid(X beta*)); prior beta ~ N(0, 5^2).  16384 walkers (8192
  per half), whole-walker partners, a = 2; walkers start uniform in a box of
  half-width sqrt(3) sd around the posterior mode per coordinate (Newton
  steps and the Laplace sd on the host, not timed).  Legs, from the same
  seeds:
    fp16    likelihood contraction on the matrix cores, fp16 copy of X with
            the proposals as one fp16 term and the log posterior corrected
            to second order for that rounding (the softplus sum's gradient
            and Hessian at a reference point of the ensemble; the headline,
            --c3-mfma),
    fp16x2  the same with the proposals split in two fp16 terms,
    bf16x3  the same on bf16 X with three bf16 terms (both with the
            first-order correction),
            all with fp32 accumulation, fp64 row sums and the guarded
            accept test (decisions near the threshold taken on fp64 values),
    fp32    the plain fp32 VALU likelihood kernel,
    fp64    the fp64 likelihood kernel (coupled windows; a whole job with
            --c3-fp64-full).
  Timed region of each leg: a fresh sampler, --warmup untimed move-bare!
  steps, then exactly --steps accounted move! steps between device
  synchronisations -> ms_per_step and walker-steps/s.  Efficiency:
  ESS_d = Var_d(final ensemble) / sigma_d^2 with sigma_d from the engine's
  acor of the run-sampler means (Sokal's windowed estimate where acor gives
  no value), ESS per walker-step = min_d ESS_d / (W x run steps), measured
  on guarded runs of 70000 steps (about 40 tau_max) from twelve seeds
  (--c3-ess-calibration, kept under profiles/); one job of the headline leg
  (500 burn-in + 12000 run-sampler steps) in every run reports its own
  (shorter, noisier) estimate beside it.  value = ESS per walker-step x the
  headline leg's walker-steps/s.  Every leg runs the same Markov kernel
  (the guarded legs' trajectories equal the fp64 sampler's step for step,
  checked in the coupled windows; the plain fp32 leg differs from it by
  rounding only), so each leg's ESS/s is the same ESS per walker-step times
  its own walker-steps/s.
  Roofline: the MFMA likelihood kernel, 2 N d K = 1.05e12 algorithmic flops
  per launch, every fourth of its launches in the timed region timed with
  HIP events on the sampler's stream.

--workload c1: the reference's CPU-runnable case.  4-D correlated Gaussian
  target, mu ~ N(0,1)^4, Sigma = A A' + I with A 4x4 i.i.d. N(0,1) (numpy
  seed 1001); 1024 walkers initialised U(-5,5); burn-in 512 + run-sampler
  4096 steps at a=2; then sample! 4096 x 1024 draws and the dataset engine's
  mean, sd and 256-bin histogram of them.  One "step" is one whole job.

Plain stretch move roofline (both workloads): the move-bare! kernel of the
correlated Gaussian model at 2^24 walkers and DIM 4 and 16, and of
independent Gaussian coordinates at DIM 32 (256 MiB, 1 GiB and 2 GiB of
positions, above the 256 MB Infinity Cache), its launches timed with HIP
events; algorithmic bytes per walker-step 8*DIM + 4 + alpha*(4*DIM + 4).

CPU baseline: the CPU restatement of the reference kernels (oracle/) on
the host cores, OpenMP over walkers, on a bounded sample of the workload;
for the logistic workload the reference's own arithmetic (the posterior as
an fp32 user model, its row loop inside LOGFN), a whole 8192-walker
half-step at -O3.
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
A = 2.0
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK = 2500.0  # TFLOP/s, dense
MFMA_FP8_PEAK = 5000.0  # TFLOP/s, dense

# correlated Gaussian workload
DIM = 4
WALKERS = 1024
BURN_IN = 512
RUN_STEPS = 4096
SAMPLE_COLS = 4096 * 1024
ROOF_WALKERS = 1 << 24
ROOF_STEPS = 20
# (model, DIM): the correlated Gaussian at DIM 4 and 16, independent
# coordinates at DIM 32 (a density of 2 DIM flops; the correlated one costs
# DIM^2 fma, 1024 at DIM 32, and would time the model rather than the move)
ROOF_DIMS = (("mvn", 4), ("mvn", 16), ("diag", 32))

# logistic-regression workload
C3_N, C3_D, C3_W = 1_000_000, 64, 16384
# the job that measures ESS per walker-step: tau_max is about 600 steps, so
# the run is about 20 tau
C3_BURN, C3_RUN = 500, 12000
C3_PRIOR_SD = 5.0
# half-width of the init-position! box per coordinate, in sd of the Laplace
# approximation (a uniform box of half-width sqrt(3) sd has variance sd^2)
C3_BOX_SD = 3.0 ** 0.5
# ESS calibration (--c3-ess-calibration): seeds and run length (about 40
# tau_max)
C3_CAL_SEEDS = tuple(range(123, 135))
C3_CAL_RUN = 70000
C3_CAL_FILE = "glm_logistic_1Mx64_ess_calibration.json"
# guard of the reduced-precision accept test per precision, in log units:
# walkers whose accept margin is smaller are decided on fp64 log posteriors.
# The measured error of lp(Y) - lp(X) at this size over 8192 pairs of a
# sampled ensemble: fp16 (one term, second-order correction of its
# rounding) sd 4.8e-4, max 2.1e-3; fp16x2 and bf16x3 (first-order
# correction) sd 1.9e-4, max 7.5e-4 (fp16x2 without the correction 1.9e-3,
# 8.2e-3).  Each guard is ten or more standard deviations (fp16's the most,
# since its remainder is of third order and grows fastest away from the
# reference point; the engine also widens it per walker by that distance).
C3_GUARDS = {"fp16": 6e-3, "fp16x2": 2e-3, "bf16x3": 2e-3}
C3_GUARD = C3_GUARDS["fp16x2"]
# the likelihood kernel's launches timed with HIP events: every fourth
C3_KTIME_EVERY = 4
# matrix-core precisions of the measured leg: (BH_GLM_*, products per
# matrix-core block, kernel)
C3_MFMA = {"bf16x3": (3, 3, "glm_logistic_mfma_kernel<64,3,bf16>"),
           "fp16x2": (4, 2, "glm_logistic_mfma_kernel<64,2,fp16>"),
           "fp16": (5, 1, "glm_logistic_mfma_kernel<64,1,fp16>")}
C3_MFMA_DEFAULT = "fp16"
C3_MFMA_DESC = {
    "fp16": "MFMA on the fp16 copy of X, proposals as one fp16 term, log posterior corrected "
            "to second order for that rounding (gradient and Hessian at a reference point)",
    "fp16x2": "MFMA split-fp16x2, first-order correction of the split's rounding",
    "bf16x3": "MFMA split-bf16x3, first-order correction of the split's rounding"}
# the batched GLMs' matrix-core precisions
C5_MFMA = {k: C3_MFMA[k] for k in ("bf16x3", "fp16x2")}
C5_MFMA_DEFAULT = "fp16x2"
# steps of each coupled window against the fp64 sampler
C3_WINDOW = 60


def _threads():
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS"):
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    return threads


def ess_from(draws_var, sigma):
    return draws_var.astype(np.float64) / np.square(sigma.astype(np.float64))


def ess_acor_or_sokal(ess_acor, var, series):
    """Per-dimension ESS: the reference's acor where it gives a value,
    Sokal's windowed estimate (var / (tau_sokal var(series) / n)) where it
    does not.  Returns (min over all dimensions, dims without an acor value,
    Sokal's per-dimension ESS)."""
    ess_acor = np.asarray(ess_acor, np.float64)
    m = np.asarray(series, np.float64)
    ess_s = np.asarray(var, np.float64) / (sokal_tau(m) * m.var(0) / m.shape[0])
    ok = np.isfinite(ess_acor)
    comb = np.where(ok, ess_acor, ess_s)
    return float(np.min(comb)), int((~ok).sum()), ess_s


def mcse_of(r):
    """Monte-Carlo standard error of each coordinate's posterior mean: acor's
    sigma where it gives one, else sqrt(tau_sokal var(series) / n)."""
    m = np.asarray(r["means"], np.float64)
    sig = np.asarray(r["sigma"], np.float64)
    sok = np.sqrt(sokal_tau(m) * m.var(0) / m.shape[0])
    return np.where(np.isfinite(sig), sig, sok)


def sokal_tau(series, c=5.0):
    """Integrated autocorrelation time of each column of an (n, dim) series:
    FFT autocorrelation, Sokal's self-consistent window (smallest M with
    M >= c * tau(M)).  A second estimator beside the reference's acor."""
    x = np.asarray(series, np.float64)
    x = x - x.mean(0)
    n = x.shape[0]
    f = np.fft.rfft(x, 2 * n, axis=0)
    ac = np.fft.irfft(f * np.conj(f), axis=0)[:n]
    ac /= ac[0]
    taus = 2.0 * np.cumsum(ac, axis=0) - 1.0
    out = np.empty(x.shape[1])
    for d in range(x.shape[1]):
        m = np.arange(n) < c * taus[:, d]
        out[d] = taus[np.argmin(m) if not m.all() else n - 1, d]
    return out


def mvn_problem(dim, seed):
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal(dim)
    A_ = rng.standard_normal((dim, dim))
    sigma = A_ @ A_.T + np.eye(dim)
    P = np.linalg.inv(sigma)
    params = np.concatenate([mu, P.reshape(-1)]).astype(np.float32)
    limits = np.tile(np.array([-5.0, 5.0], np.float32), dim)
    return mu, sigma, params, limits


def c1_problem():
    return mvn_problem(DIM, 1001)


# ---------------------------------------------------------------------------
# plain stretch move roofline
# ---------------------------------------------------------------------------
def traffic_of(name):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE,
    calibrated against a copy of known size), committed under profiles/."""
    tpath = os.path.join(ROOT, "profiles", name)
    if os.path.exists(tpath):
        with open(tpath) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def hbm_copy_gbps(mib=1024, reps=10):
    """HBM rate of a plain device copy (torch's copy kernel, mib MiB each
    way), read + write bytes over the copy time: the measured denominator
    beside the published 8 TB/s."""
    import torch
    n = mib << 18
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    return 2.0 * n * 4 / (ms * 1e-3) / 1e9


def device_rates(fac):
    """The device's measured rates; the HBM denominator is the engine's own
    streaming copy of 1 GiB each way (the better of plain and non-temporal
    accesses), torch's copy kernel beside it."""
    rates = fac.device_rates()
    cr = fac.device_copy_rate(1 << 30)
    rates["hbm_copy_gbps"] = max(cr["copy_gbps"], cr["copy_nt_gbps"])
    rates["hbm_copy_plain_gbps"] = cr["copy_gbps"]
    rates["hbm_copy_nt_gbps"] = cr["copy_nt_gbps"]
    rates["torch_copy_gbps"] = hbm_copy_gbps()
    return rates


def diag_problem(dim, seed):
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal(dim)
    w = 1.0 / rng.uniform(0.5, 2.0, dim) ** 2
    params = np.concatenate([mu, w]).astype(np.float32)
    limits = np.tile(np.array([-5.0, 5.0], np.float32), dim)
    return params, limits


def move_roofline(fac, dim, rates=None, model="mvn"):
    """HBM roofline of the move-bare! kernel at 2^24 walkers of a dim-D
    correlated Gaussian (model "mvn") or independent Gaussian coordinates
    ("diag"), and its bound from random line gathers: every walker reads one
    partner record from a random line of the other half, so a launch needs K
    line gathers (at the device's measured random-gather rate) plus its
    streamed bytes at the HBM peak."""
    from bayadera_amd import _lib, models
    if model == "mvn":
        _, _, params, limits = mvn_problem(dim, 1001 + dim)
        sf = fac.mcmc_factory(models.mvn_model(dim))
    else:
        params, limits = diag_problem(dim, 2001 + dim)
        sf = fac.mcmc_factory(models.diag_gaussian_model(dim))
    s = sf.create_sampler(77, ROOF_WALKERS, params)
    s.set_exec(_lib.BH_EXEC_STEPWISE).set_partner(_lib.BH_PARTNER_WALKER)
    s.init(77)
    s.init_position(78, limits)
    s.burn_in(4, A)  # warm-up
    s.set_kernel_timing(True)
    s.burn_in(ROOF_STEPS, A)
    kt = s.kernel_timing()
    s.set_kernel_timing(False)
    alpha = s.acc_rate(A)
    avg_ms = kt["avg_ms"]
    walkers_per_launch = ROOF_WALKERS // 2
    bytes_per_walker = 8 * dim + 4 + alpha * (4 * dim + 4)
    achieved = walkers_per_launch * bytes_per_walker / (avg_ms * 1e-3) / 1e9
    s.release()
    sf.release()
    gather = None
    if rates and rates.get("gathers_per_s"):
        streamed = walkers_per_launch * (bytes_per_walker - 4 * dim)
        bound_ms = (walkers_per_launch / rates["gathers_per_s"] + streamed / (HBM_PEAK_GBS * 1e9)) * 1e3
        gather = {"partner_line_gathers_per_launch": walkers_per_launch,
                  "gathers_per_s_measured": round(rates["gathers_per_s"], 0),
                  "streamed_bytes_per_launch": int(streamed),
                  "bound_ms": round(bound_ms, 4), "frac_of_bound": round(bound_ms / avg_ms, 4)}
    mdesc = "correlated Gaussian" if model == "mvn" else "independent Gaussian coordinates"
    tname = (f"stretch_move_bare_dim{dim}_2p24_traffic.json" if model == "mvn"
             else f"stretch_move_bare_{model}_dim{dim}_2p24_traffic.json")
    return {"kernel": f"bh_stretch_move_bare (DIM={dim}, {mdesc}, K=2^23 walkers per launch)",
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_of(tname),
            "avg_kernel_ms": round(avg_ms, 4), "launches_timed": kt["launches"],
            "alpha": round(alpha, 4),
            "algorithmic_bytes_per_launch": int(walkers_per_launch * bytes_per_walker),
            "hbm_copy_measured_gbps": round(rates["hbm_copy_gbps"], 1) if rates else None,
            "torch_copy_gbps": round(rates["torch_copy_gbps"], 1) if rates else None,
            "frac_of_measured_copy": (round(achieved / rates["hbm_copy_gbps"], 4)
                                      if rates else None),
            "gather_bound": gather}


# ---------------------------------------------------------------------------
# correlated Gaussian job (the reference's CPU-runnable case)
# ---------------------------------------------------------------------------
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


def cpu_baseline(params, limits, budget_s=25.0):
    """The whole job on the host through the oracle (OpenMP over walkers)."""
    import oracle as O
    threads = _threads()
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


def main_c1(args):
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
        rates = device_rates(fac)
        line["roofline"] = move_roofline(fac, DIM, rates)
        line["plain_move_roofline_more"] = [move_roofline(fac, d, rates, m)
                                            for m, d in ROOF_DIMS[1:]]
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(params, limits)
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# linear regression, plain fp32 path through the run-time compiled model
# ---------------------------------------------------------------------------
C2_N, C2_D, C2_W = 100_000, 16, 8192
C2_BURN, C2_RUN = 1000, 5000


def c2_problem(seed=1002):
    """X 100,000 x 16 i.i.d. N(0,1), beta* ~ N(0,1), y = X beta* + eps with
    eps ~ N(0, 0.5^2); params = rows [x_i, y_i] then hyperparameters [10, 1]
    (beta ~ N(0, 10^2), log sigma ~ N(0, 1)); DIM = 17."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((C2_N, C2_D))
    beta = rng.standard_normal(C2_D)
    y = X @ beta + 0.5 * rng.standard_normal(C2_N)
    rows = np.concatenate([X, y[:, None]], 1).astype(np.float32)
    params = np.concatenate([rows.reshape(-1), np.array([10.0, 1.0], np.float32)])
    # start box: least squares +- 4 standard errors (setup, not timed)
    Xd, yd = rows[:, :C2_D].astype(np.float64), rows[:, C2_D].astype(np.float64)
    bh, *_ = np.linalg.lstsq(Xd, yd, rcond=None)
    res = yd - Xd @ bh
    s2 = res @ res / (C2_N - C2_D)
    se = np.sqrt(np.diag(np.linalg.inv(Xd.T @ Xd)) * s2)
    ls = 0.5 * np.log(s2)
    lo = np.concatenate([bh - 4 * se, [ls - 0.02]])
    hi = np.concatenate([bh + 4 * se, [ls + 0.02]])
    limits = np.stack([lo, hi], 1).astype(np.float32).reshape(-1)
    return rows, params, bh, limits


def c2_loglik64(rows, xs):
    """fp64 log posterior of the walkers xs (n, 17): the model's formula in
    double (the oracle for the kernel's fp32 values)."""
    Xd = rows[:, :C2_D].astype(np.float64)
    yd = rows[:, C2_D].astype(np.float64)
    b = xs[:, :C2_D].astype(np.float64)
    ls = xs[:, C2_D].astype(np.float64)
    r = (yd[None, :] - b @ Xd.T) * np.exp(-ls)[:, None]
    ll = -0.5 * np.sum(r * r, 1) - C2_N * (ls + 0.9189385332046727)
    lp = -0.5 * (np.sum((b / 10.0) ** 2, 1) + ls ** 2)
    return ll + lp


def c2_job(sf, params, limits, burn, run, prec=None):
    from bayadera_amd import _lib
    s = sf.create_sampler(123, C2_W, params).set_partner(_lib.BH_PARTNER_WALKER)
    if prec is not None:
        s.set_glm_precision(prec)
    s.init(123)
    s.init_position(124, limits)
    t0 = time.perf_counter()
    s.burn_in(burn, A)
    s.set_timing(True)
    s.set_kernel_timing(True)
    res = s.run_sampler(run, A)
    run_ms = s.last_gpu_ms()
    var = s.variance()
    wall = time.perf_counter() - t0
    kt = s.kernel_timing()
    xs, lpv, _ = s.state()
    s.release()
    return res, var, wall, run_ms, kt, xs, lpv


def c2_cpu_baseline(params, limits, budget_s=12.0):
    """Walker-steps/s of the oracle's sampler with the linear-regression
    posterior as a user model (one walker per OpenMP iteration looping over
    the rows, float as the reference's kernel), whole steps of 1024 walkers
    within a bounded time."""
    import oracle as O
    threads = _threads()
    W = 1024
    s = O.Stretch(O.LINREG_POST, C2_D + 1, W, params, 123, data_len=C2_N * (C2_D + 1),
                  params_len=2, partner=O.PARTNER_WALKER, threads=threads)
    s.init(123)
    s.init_position(124, np.asarray(limits, np.float32).reshape(C2_D + 1, 2))
    t0 = time.perf_counter()
    n = 0
    while n == 0 or (time.perf_counter() - t0 < budget_s and n < 200):
        s.burn_in(1, A)
        n += 1
    t = time.perf_counter() - t0
    return {"walker_steps_per_s": W * n / t, "threads": threads,
            "sample": f"{n} steps of {W} walkers, the posterior as a user model, in {t:.1f} s"}


def main_c2(args):
    from bayadera_amd import _lib, engine, models
    rows, params, bh, limits = c2_problem()
    fac = engine.HipBayaderaFactory(0, 256)
    X, y = rows[:, :C2_D], rows[:, C2_D]
    # the posterior declared by form: one data-parallel fp32 pass per half
    sf = fac.mcmc_factory(models.glm_gaussian_model(X, y, 10.0, 1.0))
    burn, run = args.c2_burn, args.c2_run
    print("bench: linear-regression GLM run", file=sys.stderr, flush=True)
    c2_job(sf, params, limits, 2, 50)  # warm-up
    res, var, wall, run_ms, kt, xs, lpv = c2_job(sf, params, limits, burn, run, _lib.BH_GLM_FP32)
    ac = res["autocorrelation"]
    ess = ess_from(var, ac.sigma)
    ess_lo, nan_dims, _ = ess_acor_or_sokal(ess, var, res["means"])
    xs = xs.reshape(-1, C2_D + 1)
    lp64 = c2_loglik64(rows, xs[:2048])
    lp_err = np.abs(lpv[:2048].astype(np.float64) - lp64) / np.abs(lp64)
    m = res["means"].astype(np.float64)
    gm = m.mean(0)
    # MC-SE per dim from the reference's acor, Sokal's estimate where it
    # gives none
    sig = ac.sigma.astype(np.float64)
    sok = np.sqrt(sokal_tau(m) * m.var(0) / m.shape[0])
    sig = np.where(np.isfinite(sig), sig, sok)
    err = np.abs(gm[:C2_D] - bh) / sig[:C2_D]
    flops_ws = C2_N * (2 * C2_D + 3)
    achieved = (C2_W // 2) * flops_ws / (kt["avg_ms"] * 1e-3) / 1e12
    line = {
        "metric": METRIC, "value": round(ess_lo / wall, 2), "unit": "ESS/s", "n_gpus": 1,
        "steps": burn + run, "warmup": 52, "ms_per_step": round(wall / (burn + run) * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded linear-regression data)",
        "config": {"workload": f"linear regression {C2_N}x{C2_D} (DIM 17), {C2_W} walkers, "
                               f"{burn} burn-in + {run} run-sampler steps (a=2), fp32",
                   "likelihood": "declared linear-regression GLM, plain fp32 VALU pass "
                                 "(fp32 residuals, fp64 row sums)", "partner": "walker"},
        "walker_steps_per_s": round(C2_W * (burn + run) / wall, 1),
        "ess_per_s_run_phase": round(ess_lo / (run_ms * 1e-3), 2),
        "ess_min": round(ess_lo, 1), "acor_nan_dims": nan_dims,
        "tau_max": round(float(np.nanmax(ac.tau)), 2),
        "acceptance_rate": round(float(res["acceptance-rate"]), 4),
        "lp_rel_err_vs_fp64_max": float(lp_err.max()),
        "beta_mean_err_over_mcse_max_vs_least_squares": round(float(err.max()), 3),
        "roofline": {"kernel": "glm_gauss_fp32_kernel<16> (100,000 rows x 4096 points)",
                     "bound": "valu", "achieved": round(achieved, 2), "peak": 157.3,
                     "unit": "TFLOP/s", "frac": round(achieved / 157.3, 4), "traffic": None,
                     "avg_kernel_ms": round(kt["avg_ms"], 4), "launches_timed": kt["launches"],
                     "algorithmic_flops_per_walker_step": flops_ws},
    }
    if args.c2_user_steps:
        # the reference's plugin path: the LINREG_SOURCE user model through
        # hipRTC, one thread per walker looping over the rows (timed on a
        # short run)
        lik = models.likelihood_model(models.LINREG_SOURCE, name="linreg")
        prior = models.distribution_model(models.LINREG_SOURCE, name="linreg_prior",
                                          dimension=C2_D + 1, params_size=2)
        usf = fac.mcmc_factory(models.posterior_model(prior, "linreg_post", lik))
        n = args.c2_user_steps
        print("bench: user-model run", file=sys.stderr, flush=True)
        _, _, wall_u, run_u, kt_u, _, _ = c2_job(usf, params, limits, 0, n)
        line["user_model_path"] = {"steps": n, "ms_per_step": round(run_u / n, 3),
                                   "walker_steps_per_s": round(C2_W * n / (run_u * 1e-3), 1),
                                   "kernel_avg_ms": round(kt_u["avg_ms"], 4)}
    if not args.no_cpu_baseline:
        cb = c2_cpu_baseline(params, limits)
        ess_per_ws = ess_lo / (C2_W * (burn + run))
        line["cpu_baseline"] = {"value": round(cb["walker_steps_per_s"] * ess_per_ws, 2),
                                "unit": "ESS/s", "cores": cb["threads"], "kind": "port",
                                "sample": cb["sample"] + "; ESS per walker-step of the GPU run",
                                "walker_steps_per_s": round(cb["walker_steps_per_s"], 1)}
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# batched hierarchical beta-binomial posteriors
# ---------------------------------------------------------------------------
C4_B, C4_W, C4_GMAX = 4096, 256, 50
C4_BURN, C4_RUN = 1000, 2000
C4_HYPER_SD = 10.0


def c4_problems(B=C4_B, seed=1004):
    """B problems: G_i ~ U{5..50} groups, trials n_ig = round(log-uniform
    [10, 1e5]), theta_ig ~ Beta(2, 5), y_ig ~ Binomial(n_ig, theta_ig);
    params [n, y, s0] with log alpha, log beta ~ N(0, 10^2); DIM_i = G_i + 2.
    The start box: log alpha in [0, 1.5], log beta in [0.5, 2], each logit
    theta_g within 0.05 of the empirical logit."""
    rng = np.random.default_rng(seed)
    dims, params = [], []
    lim = np.zeros((B, C4_GMAX + 2, 2), np.float32)
    for b in range(B):
        G = int(rng.integers(5, C4_GMAX + 1))
        n = np.round(np.exp(rng.uniform(np.log(10.0), np.log(1e5), G)))
        th = rng.beta(2.0, 5.0, G)
        y = rng.binomial(n.astype(np.int64), th).astype(np.float64)
        dims.append(G + 2)
        params.append(np.concatenate([n, y, [C4_HYPER_SD]]).astype(np.float32))
        t = np.log((y + 0.5) / (n - y + 0.5))
        lim[b, 0] = [0.0, 1.5]
        lim[b, 1] = [0.5, 2.0]
        lim[b, 2:G + 2, 0] = t - 0.05
        lim[b, 2:G + 2, 1] = t + 0.05
    return dims, params, lim


# MI355X: 256 CUs x 4 SIMDs, a wave64 VALU instruction every 4 cycles per
# SIMD at the 2.4 GHz peak engine clock
VALU_WAVE_INSTS_PEAK = 1024 * 2.4e9 / 4


def c4_roofline(walker_steps_per_s, burn_ms_per_step):
    """VALU-issue roofline of bh_batch_run_g (the batch's burn-in launch):
    VALU instructions per walker-step from the committed counter profile
    (SQ_INSTS_VALU of the 100-step burn-in launch) times the burn-in's
    walker-steps/s, against one wave64 instruction per 4 cycles per SIMD."""
    path = os.path.join(ROOT, "profiles", "batch_hbb_4096_valu_pmc.json")
    if not os.path.exists(path):
        return None
    prof = json.load(open(path))
    per_ws = prof["valu_wave_insts_per_walker_step"]
    ws_burn = C4_B * C4_W / (burn_ms_per_step * 1e-3)
    achieved = per_ws * ws_burn
    return {"kernel": "bh_batch_run_g (4096 problems, one work-group each, burn-in launch)",
            "bound": "valu", "achieved": round(achieved * 64 / 1e12, 2),
            "peak": round(VALU_WAVE_INSTS_PEAK * 64 / 1e12, 2), "unit": "T VALU lane-instructions/s",
            "frac": round(achieved / VALU_WAVE_INSTS_PEAK, 4), "traffic": None,
            "valu_wave_insts_per_walker_step": per_ws,
            "burn_in_ms_per_step": round(burn_ms_per_step, 4),
            "source": "profiles/batch_hbb_4096_valu_pmc.json"}


def c4_cpu_baseline(dims, params, lim, budget_s=12.0):
    """Walker-steps/s of the oracle's single sampler (OpenMP over walkers)
    on the first problems of the batch, one after another, burn-in steps
    from the same start as the GPU run, within a bounded time."""
    import oracle as O
    threads = _threads()
    steps, done, b = 50, 0, 0
    t0 = time.perf_counter()
    while b < len(dims) and (b == 0 or time.perf_counter() - t0 < budget_s):
        G = dims[b] - 2
        s = O.Stretch(O.HBB_LOGPDF, dims[b], C4_W, params[b], 5 + 2 * b, wgs=128,
                      data_len=2 * G, params_len=1, threads=threads)
        s.init(5 + 2 * b)
        s.init_position(3 + b, lim[b, :dims[b]])
        s.burn_in(steps, A)
        done += C4_W * steps
        b += 1
    t = time.perf_counter() - t0
    return {"walker_steps_per_s": done / t, "threads": threads,
            "sample": f"{b} problems x {steps} steps x {C4_W} walkers (init + burn-in) in {t:.1f} s"}


def main_c4(args):
    import torch
    from bayadera_amd import engine, models
    dims, params, lim = c4_problems()
    fac = engine.HipBayaderaFactory(0, 256)
    bt = fac.batch(models.hbb_model(C4_GMAX), params, dims, [1] * len(dims), C4_W)
    burn, run = args.c4_burn, args.c4_run
    series = torch.empty(run * int(np.sum(dims)), dtype=torch.float32, device="cuda")
    bt.init_position(3, lim)
    bt.burn_in(1000003, 2, A)  # warm-up
    bt.init_position(3, lim)
    # launch shape (threads per problem, dimension classes) chosen by timing
    # the candidates from this state; the state is restored (setup, untimed)
    tuned = bt.autotune(10)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bt.burn_in(5, burn, A)
    burn_ms = bt.last_gpu_ms()
    res = bt.run(6, run, A, series=series)
    run_ms = bt.last_gpu_ms()
    wall = time.perf_counter() - t0
    # ESS per problem from the engine's acor of each recorded mean series
    # (after the timed region); the walkers' posterior variance per dim is
    # taken from the run's accounting (sd over all steps and walkers)
    acor = fac.acor_engine()
    blocks = bt.series_blocks(series, run)
    ess = np.zeros(len(dims))
    tau_max = np.zeros(len(dims))
    print("bench: acor of every problem", file=sys.stderr, flush=True)
    acor_ok = np.zeros(len(dims), bool)
    acs = acor.acor_batch(series, dims, run)  # one launch for all problems
    for b, blk in enumerate(blocks):
        ac = acs[b]
        var = res["sd"][b, :dims[b]] ** 2
        sig2 = np.square(ac.sigma.astype(np.float64))
        tau = ac.tau.astype(np.float64)
        acor_ok[b] = bool(np.all(np.isfinite(sig2)))
        if not acor_ok[b]:
            # no value from the reference's acor (tau long against the
            # series): Sokal's windowed estimate of the same quantities
            x = np.asarray(blk.cpu() if hasattr(blk, "cpu") else blk, np.float64)
            tau = sokal_tau(x)
            sig2 = tau * x.var(0) / x.shape[0]
        ess[b] = float(np.min(var / sig2))
        tau_max[b] = float(np.max(tau))
    walker_steps = len(dims) * C4_W * (burn + run)
    ok = np.isfinite(ess)
    line = {
        "metric": METRIC, "value": round(float(ess[ok].sum()) / wall, 1), "unit": "ESS/s",
        "n_gpus": 1, "steps": burn + run, "warmup": 2,
        "ms_per_step": round(wall / (burn + run) * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded hierarchical beta-binomial problems)",
        "config": {"workload": f"{len(dims)} hierarchical beta-binomial posteriors, G_i in "
                               f"[5, 50] (DIM 7-52), {C4_W} walkers each, {burn} burn-in + "
                               f"{run} run steps (a=2), one work-group per problem",
                   "threads_per_problem": tuned["threads_per_problem"],
                   "launch_classes": tuned["launch_classes"], "partner": "flat"},
        "walker_steps_per_s": round(walker_steps / wall, 1),
        "burn_in_gpu_ms": round(burn_ms, 2), "run_gpu_ms": round(run_ms, 2),
        "ess_sum": round(float(ess[ok].sum()), 1),
        "ess_per_problem_median": round(float(np.median(ess[ok])), 2),
        "problems_with_finite_ess": int(ok.sum()),
        "ess_estimator": {"acor": int(acor_ok.sum()), "sokal": int((~acor_ok).sum())},
        "tau_max_median": round(float(np.median(tau_max[ok])), 1),
        "acceptance_rate_median": round(float(np.median(res["acceptance-rate"])), 4),
        "acceptance_rate_min": round(float(np.min(res["acceptance-rate"])), 4),
    }
    ro = c4_roofline(walker_steps / wall, burn_ms / burn)
    if ro:
        line["roofline"] = ro
    if not args.no_cpu_baseline:
        cb = c4_cpu_baseline(dims, params, lim)
        ess_per_ws = float(ess[ok].sum()) / walker_steps
        line["cpu_baseline"] = {"value": round(cb["walker_steps_per_s"] * ess_per_ws, 1),
                                "unit": "ESS/s", "cores": cb["threads"], "kind": "port",
                                "sample": cb["sample"] + "; ESS per walker-step of the GPU run",
                                "walker_steps_per_s": round(cb["walker_steps_per_s"], 1)}
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# batched uneven logistic regressions, fp8 data
# ---------------------------------------------------------------------------
C5_P, C5_W = 512, 2048
C5_BURN, C5_RUN = 300, 1000


def c5_problems(P=C5_P, seed=1005):
    """P problems: N_p = round(log-uniform[1e3, 2e5]), d_p uniform in {8, 16,
    32, 64}, X ~ N(0,1)/sqrt(d) quantised to fp8 e4m3 per column, beta* ~
    N(0,1), y ~ Bernoulli(sigmoid(X beta*)), prior N(0, 5^2).  Start box:
    +-4 sd around each posterior mode (Newton on the dequantised data)."""
    from bayadera_amd import models
    rng = np.random.default_rng(seed)
    Xs, ys = [], []
    for _ in range(P):
        N = int(round(np.exp(rng.uniform(np.log(1e3), np.log(2e5)))))
        d = int(rng.choice([8, 16, 32, 64]))
        X = rng.standard_normal((N, d)) / np.sqrt(d)
        b = rng.standard_normal(d)
        Xs.append(X)
        ys.append((rng.random(N) < 1.0 / (1.0 + np.exp(-(X @ b)))).astype(np.float32))
    m = models.glm_logistic_batch_model(Xs, ys, prior_sd=C3_PRIOR_SD)
    lim = np.zeros((P, 64, 2), np.float32)
    for p in range(P):
        if p % 128 == 0:
            print(f"bench: start boxes {p}/{P}", file=sys.stderr, flush=True)
        # float32 Newton steps (setup only; the box need not be exact)
        Xd, y = m.dequantized(p).astype(np.float32), m.y[p]
        d = Xd.shape[1]
        b = np.zeros(d, np.float32)
        for _ in range(5):
            pr = 1.0 / (1.0 + np.exp(-(Xd @ b)))
            g = Xd.T @ (y - pr) - b / C3_PRIOR_SD ** 2
            H = (Xd * (pr * (1 - pr))[:, None]).T @ Xd + np.eye(d, dtype=np.float32) / C3_PRIOR_SD ** 2
            b = b + np.linalg.solve(H.astype(np.float64), g.astype(np.float64)).astype(np.float32)
        sd = np.sqrt(np.diag(np.linalg.inv(H.astype(np.float64))))
        lim[p, :d, 0] = b - 4 * sd
        lim[p, :d, 1] = b + 4 * sd
    return m, lim


def c5_cpu_baseline(m, lim, budget_s=12.0):
    """Walker-steps/s of the oracle's fp64 GLM sampler on the first problems'
    data (dequantized rows), half-steps of 1024 walkers in turn, within a
    bounded time (positions only: the initial log posteriors are not
    computed, the work per half-step is the same)."""
    import oracle as O
    threads = _threads()
    K = C5_W // 2
    work, p, t = 0.0, 0, 0.0
    while p < m.problems and (p == 0 or t < budget_s):
        Xd = np.ascontiguousarray(m.dequantized(p), np.float32)
        d = Xd.shape[1]
        g = O.GlmStretch(Xd, m.y[p], C3_PRIOR_SD, C5_W, 9 + 2 * p, partner=O.PARTNER_WALKER,
                         threads=threads)
        g.xs[:] = O.init_walkers(9 + p, lim[p, :d], d, C5_W)
        t0 = time.perf_counter()
        g._half(9 + 2 * p, 3333, 0, True, A, 1.0)
        t += time.perf_counter() - t0
        work += float(K) * Xd.shape[0] * d
        p += 1
    # a walker-step of problem p costs ~N_p d_p: the sampled rate in
    # (row x coordinate) walker-steps per second over the batch's mean N_p d_p
    mean_cost = float(np.mean(m.rows.astype(np.float64) * m.dims.astype(np.float64)))
    return {"walker_steps_per_s": work / t / mean_cost, "threads": threads,
            "sample": f"one half-step of {K} walkers for each of the first {p} problems "
                      f"({t:.1f} s), scaled by the batch's mean N_p d_p"}


def c5_run(fac, m, lim, prec, burn, run, series=None, tune=False):
    bt = fac.glm_batch(m, C5_W).set_precision(prec)
    bt.init_position(9, lim)
    # the contraction's tile shapes, chosen by the batch's autotuner on these
    # positions (setup, before the timed region)
    bt.tuned = bt.autotune(3) if tune else None
    t0 = time.perf_counter()
    bt.burn_in(100, burn, A)
    burn_ms = bt.last_gpu_ms()
    res = bt.run(200 + 2 * burn, run, A, series=series)
    wall = time.perf_counter() - t0
    return bt, res, wall, burn_ms, bt.last_gpu_ms()


def main_c5(args):
    import torch
    from bayadera_amd import _lib, engine
    t_setup = time.perf_counter()
    m, lim = c5_problems()
    t_setup = time.perf_counter() - t_setup
    fac = engine.HipBayaderaFactory(0, 256)
    burn, run = args.c5_burn, args.c5_run
    dims = m.dims.astype(np.int64)
    series = torch.empty(run * int(dims.sum()), dtype=torch.float32, device="cuda")
    mname = args.c5_mfma
    mprec, msplit, _ = C5_MFMA[mname]
    bt, _, _, _, _ = c5_run(fac, m, lim, mprec, 2, 50)  # warm-up
    bt.release()
    torch.cuda.synchronize()
    print(f"bench: {mname} batched run", file=sys.stderr, flush=True)
    bt, res, wall, burn_ms, run_ms = c5_run(fac, m, lim, mprec, burn, run, series, tune=True)
    tuned = bt.tuned
    # after the timed region: ESS per problem from the engine's acor of its
    # recorded means; the precision bias of the matrix-core chain's
    # stationary distribution, estimated per problem by reweighting its final
    # walkers with exp(lp_fp64 - lp_mfma) (the fp64 likelihood kernel at the
    # same positions), in MC-SE units
    acor = fac.acor_engine()
    xs, lp_b = bt.state()
    lp_64 = bt.log_density(_lib.BH_GLM_FP64)
    ess = np.zeros(m.problems)
    bias = np.zeros(m.problems)
    sd_bias = np.zeros(m.problems)
    lp_err = np.zeros(m.problems)
    acor_ok = np.zeros(m.problems, bool)
    acs = acor.acor_batch(series, dims, run)  # one launch for all problems
    for p, blk in enumerate(bt.series_blocks(series, run)):
        d = int(dims[p])
        ac = acs[p]
        var = res["sd"][p, :d] ** 2
        sig2 = np.square(ac.sigma.astype(np.float64))
        acor_ok[p] = bool(np.all(np.isfinite(sig2)))
        if not acor_ok[p]:
            # the reference's acor gives no value when tau is long against
            # the series: Sokal's windowed estimate of the same MC-SE^2
            b = np.asarray(blk.cpu() if hasattr(blk, "cpu") else blk, np.float64)
            sig2 = sokal_tau(b) * b.var(0) / b.shape[0]
        e = var / sig2
        ess[p] = float(np.min(e))
        mcse = np.sqrt(var / e)
        x = xs[p].astype(np.float64)
        lw = lp_64[p] - lp_b[p]
        lp_err[p] = float(np.max(np.abs(lw - lw.mean())))
        w = np.exp(lw - lw.max())
        w /= w.sum()
        mu_w = w @ x
        bias[p] = float(np.max(np.abs(mu_w - x.mean(0)) / mcse))
        sd_w = np.sqrt(np.maximum(w @ (x - mu_w) ** 2, 0.0))
        sd_bias[p] = float(np.max(np.abs(sd_w / x.std(0) - 1.0)))
    bt.release()
    flops = float(np.sum(2.0 * m.rows.astype(np.float64) * dims * (C5_W // 2)))
    half_ms = run_ms / (2 * run)
    achieved = flops / (half_ms * 1e-3) / 1e12
    ok = np.isfinite(ess)
    line = {
        "metric": METRIC, "value": round(float(ess[ok].sum()) / wall, 1), "unit": "ESS/s",
        "n_gpus": 1, "steps": burn + run, "warmup": 52,
        "ms_per_step": round(wall / (burn + run) * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": mname,
        "data": "synthetic (seeded logistic-regression problems, X stored fp8 e4m3 + "
                "per-column fp32 scales)",
        "config": {"workload": f"{m.problems} logistic regressions, N in [1e3, 2e5] "
                               f"(log-uniform), d in {{8,16,32,64}}, {C5_W} walkers each, "
                               f"{burn} burn-in + {run} run steps (a=2), prior N(0,5^2)",
                   "likelihood": f"grouped MFMA split-{mname} over fp8-decoded rows, fp32 "
                                 "accumulate, fp64 row sums", "partner": "walker"},
        "walker_steps_per_s": round(m.problems * C5_W * (burn + run) / wall, 1),
        "burn_in_gpu_ms": round(burn_ms, 1), "run_gpu_ms": round(run_ms, 1),
        "ess_sum": round(float(ess[ok].sum()), 1),
        "ess_estimator": {"acor": int(acor_ok.sum()), "sokal": int((~acor_ok).sum())},
        "ess_per_problem_median": round(float(np.median(ess[ok])), 1),
        "acceptance_rate_median": round(float(np.median(res["acceptance-rate"])), 4),
        "precision_bias_mean_over_mcse": {"median": round(float(np.nanmedian(bias)), 4),
                                          "max": round(float(np.nanmax(bias)), 4),
                                          "problems_over_0.1": int((bias > 0.1).sum()),
                                          "problems_without_mcse": int(np.isnan(bias).sum())},
        "precision_bias_sd_rel": {"median": float(np.median(sd_bias)),
                                  "max": float(sd_bias.max())},
        "lp_error_spread_max": float(lp_err.max()),
        "roofline": {"kernel": f"half-step (propose + grouped gb_mfma_kernel<D,{msplit}> + "
                               "accept), "
                               f"{m.problems} problems x 1024 points", "bound": "mfma",
                     "achieved": round(achieved, 1), "peak": MFMA_BF16_PEAK, "unit": "TFLOP/s",
                     "frac": round(achieved / MFMA_BF16_PEAK, 4), "traffic": None,
                     "frac_of_fp8_peak": round(achieved / MFMA_FP8_PEAK, 4),
                     "avg_half_step_ms": round(half_ms, 4),
                     "algorithmic_flops_per_half_step": flops,
                     "tiles_autotuned": tuned},
        "setup_s": round(t_setup, 1),
    }
    if not args.no_cpu_baseline:
        cb = c5_cpu_baseline(m, lim)
        ess_per_ws = float(ess[ok].sum()) / (m.problems * C5_W * (burn + run))
        line["cpu_baseline"] = {"value": round(cb["walker_steps_per_s"] * ess_per_ws, 2),
                                "unit": "ESS/s", "cores": cb["threads"], "kind": "port",
                                "sample": cb["sample"] + "; ESS per walker-step of the GPU run",
                                "walker_steps_per_s": round(cb["walker_steps_per_s"], 1)}
    if args.c5_fp32:
        bt32, res32, wall32, _, run32 = c5_run(fac, m, lim, _lib.BH_GLM_FP32, burn, run)
        bt32.release()
        line["fp32_walker_steps_per_s"] = round(m.problems * C5_W * (burn + run) / wall32, 1)
        line["mfma_vs_fp32_walker_steps_per_s"] = round(wall32 / wall, 2)
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# logistic regression (default)
# ---------------------------------------------------------------------------
def c3_problem(N=C3_N, d=C3_D, seed=1003):
    from bayadera_amd import models
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((N, d), dtype=np.float32) / np.float32(np.sqrt(d)))
    Xb = models.to_bf16(X)
    Xd = models.from_bf16(Xb).reshape(N, d).astype(np.float64)
    beta_true = rng.standard_normal(d)
    eta = Xd @ beta_true
    y = (rng.random(N) < 1.0 / (1.0 + np.exp(-eta))).astype(np.float32)
    # the init-position! box: Newton steps to the posterior mode (setup, not
    # timed), then per coordinate a uniform box of half-width sqrt(3) sd of
    # the Laplace approximation, whose variance is the Laplace variance
    b = np.zeros(d)
    for _ in range(8):
        p = 1.0 / (1.0 + np.exp(-(Xd @ b)))
        g = Xd.T @ (y - p) - b / C3_PRIOR_SD ** 2
        H = (Xd * (p * (1 - p))[:, None]).T @ Xd + np.eye(d) / C3_PRIOR_SD ** 2
        b = b + np.linalg.solve(H, g)
    sd = np.sqrt(np.diag(np.linalg.inv(H)))
    half = C3_BOX_SD * sd
    limits = np.stack([b - half, b + half], 1).astype(np.float32).reshape(-1)
    return Xb, Xd, y, b, sd, limits


def c3_sampler(sf, prec, guard=0.0, seed=123):
    from bayadera_amd import _lib
    s = sf.create_sampler(seed, C3_W, np.zeros(1, np.float32))
    s.set_glm_precision(prec).set_partner(_lib.BH_PARTNER_WALKER)
    if guard:
        s.set_glm_guard(guard)
    return s


def c3_timed(sf, prec, limits, warmup, steps, guard=0.0, seed=123):
    """The timed region: a fresh sampler, init-position!, `warmup` untimed
    move-bare! steps, then exactly `steps` accounted move! steps bracketed by
    device synchronisations; every C3_KTIME_EVERY-th launch of the likelihood
    kernel in them timed with HIP events on the sampler's stream (a pair of
    events costs the stream ~10 us, so timing every launch would slow the
    region it measures)."""
    import torch
    s = c3_sampler(sf, prec, guard, seed)
    s.init(seed)
    s.init_position(seed + 1, limits)
    if warmup:
        s.burn_in(warmup, A)
    s.init_move(A, steps)
    s.set_kernel_timing(True, C3_KTIME_EVERY)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.move()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    kt = s.kernel_timing()
    gs = s.glm_guard_stats()
    s.release()
    return {"wall_s": t, "ms_per_step": t / steps * 1e3,
            "walker_steps_per_s": C3_W * steps / t, "kernel": kt, "guard": gs}


def c3_run(sf, prec, limits, burn, run, guard=0.0, seed=123):
    """One job: init-position!, burn-in!, run-sampler!, variance; wall time
    of burn-in! plus run-sampler! and the variance read back (the state
    after burn-in! is copied out between the two, untimed)."""
    s = c3_sampler(sf, prec, guard, seed)
    s.init(seed)
    s.init_position(seed + 1, limits)
    t0 = time.perf_counter()
    s.burn_in(burn, A)
    after_burn = s.state()
    t1 = time.perf_counter()
    s.set_timing(True)
    s.set_kernel_timing(True, C3_KTIME_EVERY)
    res = s.run_sampler(run, A)
    run_gpu_ms = s.last_gpu_ms()
    s.set_timing(False)
    var = s.variance()
    t2 = time.perf_counter()
    kt = s.kernel_timing()
    end = s.state()
    gs = s.glm_guard_stats()
    s.release()
    ac = res["autocorrelation"]
    ess = ess_from(var, ac.sigma)
    return {"wall_s": (t1 - t0) + (t2 - t1), "run_gpu_ms": run_gpu_ms, "ess": ess,
            "means": res["means"], "grand_mean": res["means"].astype(np.float64).mean(0),
            "sigma": ac.sigma, "sd": np.sqrt(var.astype(np.float64)), "var": var,
            "tau": ac.tau, "acc": res["acceptance-rate"], "kernel": kt, "xs": end[0],
            "guard": gs, "after_burn": after_burn, "end": end, "burn": burn, "run": run}


def c3_ess_summary(r):
    """ESS of a job: min over dims of acor's ESS (Sokal's where acor gives
    none), per walker-step of the run phase and per second of the job."""
    m = np.asarray(r["means"], np.float64)
    ess, nan_dims, ess_sd = ess_acor_or_sokal(r["ess"], np.asarray(r["var"], np.float64), m)
    tau_s = sokal_tau(m)
    return {"ess_min": round(ess, 1), "acor_nan_dims": nan_dims,
            "ess_per_walker_step": ess / (C3_W * r["run"]),
            "ess_per_s_job": round(ess / r["wall_s"], 2),
            "ess_sokal_min": round(float(np.min(ess_sd)), 1),
            "tau_max": round(float(np.nanmax(r["tau"])), 2),
            "tau_sokal_max": round(float(tau_s.max()), 1),
            "run_over_tau_max": round(r["run"] / float(np.nanmax(r["tau"])), 1),
            "acceptance_rate": round(r["acc"], 4), "job_wall_s": round(r["wall_s"], 3),
            "job_steps": r["burn"] + r["run"],
            "job_ms_per_step_run": round(r["run_gpu_ms"] / r["run"], 3)}


def c3_window(sf, state, prec, guard, steps, accounted):
    """From a saved sampler state: `steps` run-sampler! (accounted) or
    burn-in! steps; returns the positions and, if accounted, the means."""
    s = c3_sampler(sf, prec, guard)
    s.set_state(state[0], None, state[2])
    means = None
    if accounted:
        means = s.run_sampler(steps, A)["means"]
    else:
        s.burn_in(steps, A)
    xs = s.state()[0]
    gs = s.glm_guard_stats()
    s.release()
    return xs, means, gs


def c3_cpu_baseline(Xd, y, limits, budget_s=30.0):
    """Walker-steps/s of the reference's arithmetic on the host: the logistic
    posterior as the reference's fp32 user model (posterior.cl:1-6 with the
    row loop of gaussian.cl:29-42 inside LOGFN), one stretch half-step of
    the whole 8192-walker half over the full 1M x 64 data, OpenMP over
    blocks of walkers (the oracle's -O3 AVX build)."""
    import oracle as O
    threads = _threads()
    L = O.fast_lib()
    N, d = Xd.shape
    params = np.concatenate([np.concatenate([Xd.astype(np.float32), y[:, None]], 1).reshape(-1),
                             [C3_PRIOR_SD]]).astype(np.float32)
    K = C3_W // 2
    xs = O.init_walkers(6, limits.reshape(-1, 2), d, 2 * K).astype(np.float32).reshape(-1)
    # the current half's log posteriors enter only the accept test; the
    # half-step's cost is the K proposals' log posteriors
    lf = np.zeros(2 * K, np.float32)
    acc = np.zeros(K, np.uint8)
    # a probe on 128 walkers sizes the sample to the budget (at least one
    # 128-walker block, at most the whole half)
    kk = 128
    t0 = time.perf_counter()
    L.bho_logistic_half_blocked(5, 3333, 0, kk, d, xs.ctypes.data + 4 * K * d, xs.ctypes.data,
                                lf.ctypes.data, A, 1.0, O.PARTNER_WALKER, N * (d + 1), params,
                                acc.ctypes.data, threads)
    t_probe = time.perf_counter() - t0
    blocks = int(min(K // 128, max(1, budget_s / max(t_probe, 1e-9))))
    kk = 128 * blocks
    t0 = time.perf_counter()
    L.bho_logistic_half_blocked(5, 3333, 1, kk, d, xs.ctypes.data + 4 * K * d, xs.ctypes.data,
                                lf.ctypes.data, A, 1.0, O.PARTNER_WALKER, N * (d + 1), params,
                                acc.ctypes.data, threads)
    t = time.perf_counter() - t0
    return {"walker_steps_per_s": kk / t, "threads": threads, "build": O.fast_lib_name(),
            "sample": (f"{kk} of the {K} walkers of one half-step over the full {N}x{d} data "
                       f"(fp32 user-model arithmetic, {O.fast_lib_name()}) in {t:.1f} s"
                       + ("" if kk == K else "; the rate scales linearly in walkers"))}


def c3_long_run(sf, prec, limits, burn, run, guard, seed, acor, chunk=10000):
    """A job of `run` run-sampler! steps as consecutive run-sampler! calls of
    `chunk` steps (the sampler's state and counters carry over), one
    progress line per call; acor of the whole concatenated means series."""
    s = c3_sampler(sf, prec, guard, seed)
    s.init(seed)
    s.init_position(seed + 1, limits)
    t0 = time.perf_counter()
    for b0 in range(0, burn, chunk):
        s.burn_in(min(chunk, burn - b0), A)
    parts, accs = [], []
    done = 0
    while done < run:
        n = min(chunk, run - done)
        res = s.run_sampler(n, A)
        parts.append(res["means"].copy())
        accs.append(res["acceptance-rate"] * n)
        done += n
        print(f"  {done}/{run} steps, {time.perf_counter() - t0:.0f} s", file=sys.stderr,
              flush=True)
    var = s.variance()
    wall = time.perf_counter() - t0
    gs = s.glm_guard_stats()
    xs = s.state()[0]
    s.release()
    means = np.ascontiguousarray(np.concatenate(parts))
    ac = acor.acor(means)
    return {"wall_s": wall, "run_gpu_ms": wall * 1e3, "ess": ess_from(var, ac.sigma),
            "means": means, "sigma": ac.sigma, "var": var, "tau": ac.tau,
            "grand_mean": means.astype(np.float64).mean(0), "sd": np.sqrt(var.astype(np.float64)),
            "xs": xs, "acc": float(np.sum(accs) / run), "guard": gs, "burn": burn, "run": run}


def c3_calibration(seeds, mname):
    """ESS per walker-step of the logistic workload, measured on long runs:
    for each seed, a guarded job of 500 burn-in + C3_CAL_RUN run-sampler
    steps (about 40 tau_max) on the matrix-core precision `mname` at its
    guard, and, from the first C3_RUN steps of the same series, the estimate
    a job of the bench's length gives.  The guarded chain is the fp64 chain
    step for step at every precision, so runs of different precisions
    measure one quantity; runs of other seeds already in the calibration
    file (same run length) are kept, and the summary covers all of them.
    (The plain fp32 chain is the same Markov kernel up to rounding, so its
    ESS per walker-step is the same quantity.)"""
    from bayadera_amd import _lib, engine, models
    Xb, Xd, y, _, _, limits = c3_problem()
    fac = engine.HipBayaderaFactory(0, 256)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y, prior_sd=C3_PRIOR_SD))
    acor = fac.acor_engine()
    prec, guard = C3_MFMA[mname][0], C3_GUARDS[mname]
    cmd = ("python bench.py --c3-ess-calibration --c3-cal-seeds "
           + ",".join(str(x) for x in seeds) + f" --c3-mfma {mname}")
    kept = []
    path = os.path.join(ROOT, "profiles", C3_CAL_FILE)
    if os.path.exists(path):
        with open(path) as f:
            old = json.load(f)
        if old.get("ess_per_walker_step", {}).get("run_steps") == C3_CAL_RUN:
            kept = [r for r in old["runs"] if r["seed"] not in seeds]
    out = {"workload": "logistic regression 1000000x64, 16384 walkers, a=2, prior N(0,5^2), "
                       "guarded matrix-core likelihood (the fp64 chain)",
           "runs": kept}
    for seed in seeds:
        print(f"calibration: seed {seed}, {mname}, {C3_BURN} + {C3_CAL_RUN} steps",
              file=sys.stderr, flush=True)
        r = c3_long_run(sf, prec, limits, C3_BURN, C3_CAL_RUN, guard, seed, acor)
        full = c3_ess_summary(r)
        full.pop("ess_per_s_job")
        full.pop("job_ms_per_step_run")
        pre = np.ascontiguousarray(r["means"][:C3_RUN])
        ac = acor.acor(pre)
        r_pre = dict(r, means=pre, ess=ess_from(r["var"], ac.sigma), tau=ac.tau, run=C3_RUN)
        prefix = c3_ess_summary(r_pre)
        for k in ("ess_per_s_job", "job_ms_per_step_run", "job_wall_s", "job_steps"):
            prefix.pop(k)
        out["runs"].append({"seed": seed, "precision": mname, "guard": guard, "command": cmd,
                            "full": full, f"first_{C3_RUN}_steps": prefix,
                            "guard_overflow": r["guard"]["overflow"]})
        print(json.dumps(out["runs"][-1]), file=sys.stderr, flush=True)
    out["runs"].sort(key=lambda x: x["seed"])
    e = np.array([x["full"]["ess_per_walker_step"] for x in out["runs"]])
    ep = np.array([x[f"first_{C3_RUN}_steps"]["ess_per_walker_step"] for x in out["runs"]])
    out["ess_per_walker_step"] = {"mean": float(e.mean()), "sd": float(e.std(ddof=1)),
                                  "sem": float(e.std(ddof=1) / np.sqrt(len(e))),
                                  "run_steps": C3_CAL_RUN, "seeds": len(e)}
    out[f"ess_per_walker_step_first_{C3_RUN}"] = {"mean": float(ep.mean()),
                                                  "sd": float(ep.std(ddof=1))}
    print(json.dumps(out), flush=True)


def c3_calibrated():
    path = os.path.join(ROOT, "profiles", C3_CAL_FILE)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main_c3(args):
    import torch
    from bayadera_amd import _lib, engine, models
    K, W = args.steps, args.warmup
    if K < 1 or W < 0:
        raise SystemExit("bench.py: --steps must be >= 1 and --warmup >= 0")
    t_setup = time.perf_counter()
    Xb, Xd, y, b_map, sd_map, limits = c3_problem()
    t_setup = time.perf_counter() - t_setup
    fac = engine.HipBayaderaFactory(0, 256)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y, prior_sd=C3_PRIOR_SD))
    burn, run = args.c3_burn, args.c3_run
    mname = args.c3_mfma
    mprec, msplit, mkern = C3_MFMA[mname]
    mguard = args.c3_guard if args.c3_guard is not None else C3_GUARDS[mname]
    others = [n for n in ("fp16", "fp16x2", "bf16x3") if n != mname]
    legs = ([(mname, mprec, mguard)] + [(n, C3_MFMA[n][0], C3_GUARDS[n]) for n in others] +
            [("fp32", _lib.BH_GLM_FP32, 0.0)])
    # first use of every kernel of the guarded and plain paths (fp16 copy of
    # X, hipRTC modules), outside every timed region
    for _n, prec, g in legs:
        c3_timed(sf, prec, limits, 1, 2, guard=g)
    if args.c3_bf16:
        legs.append(("bf16", _lib.BH_GLM_BF16, 0.0))
    timed = {}
    for name, prec, g in legs:
        print(f"bench: {name} timed region ({W} warm-up + {K} steps)", file=sys.stderr, flush=True)
        timed[name] = c3_timed(sf, prec, limits, W, K, guard=g)
    # the efficiency: ESS per walker-step of the MFMA leg's whole job
    print(f"bench: {mname} job ({burn} burn-in + {run} run-sampler steps)", file=sys.stderr,
          flush=True)
    job = c3_run(sf, mprec, limits, burn, run, guard=mguard)
    js = c3_ess_summary(job)
    ess_ws = js["ess_per_walker_step"]
    ref = r32 = None
    if args.c3_fp64_full:
        # run-sampler! in pieces with a progress line each (the trajectory
        # is that of one call: the step counter carries over)
        # (init-move! draws fresh seeds per call, so every leg here runs the
        # same pieces)
        acor = fac.acor_engine()
        print(f"bench: {mname} job in pieces", file=sys.stderr, flush=True)
        jc = c3_long_run(sf, mprec, limits, burn, run, mguard, 123, acor, chunk=500)
        print("bench: fp64 job", file=sys.stderr, flush=True)
        ref = c3_long_run(sf, _lib.BH_GLM_FP64, limits, burn, run, 0.0, 123, acor, chunk=500)
        print("bench: plain fp32 job", file=sys.stderr, flush=True)
        r32 = c3_long_run(sf, _lib.BH_GLM_FP32, limits, burn, run, 0.0, 123, acor, chunk=500)
    # log-posterior error of each precision against the fp64 kernel on the
    # job's final ensemble (16384 points), and of the Metropolis log ratio
    # lp(x_k) - lp(x_{k+K}) over pairs of them
    print("bench: log-density errors", file=sys.stderr, flush=True)
    pts = np.ascontiguousarray(job["end"][0])
    lp64 = sf.log_density(pts, _lib.BH_GLM_FP64)
    d64 = lp64[: C3_W // 2] - lp64[C3_W // 2:]
    lp_err = {}
    for name, prec, _g in legs:
        lp = sf.log_density(pts, prec)
        e = lp - lp64
        de = (lp[: C3_W // 2] - lp[C3_W // 2:]) - d64
        lp_err[name] = {"lp_abs_err_max": float(np.max(np.abs(e))),
                        "lp_rel_err_max": float(np.max(np.abs(e) / np.abs(lp64))),
                        "log_ratio_err_sd": float(np.std(de)),
                        "log_ratio_err_max": float(np.max(np.abs(de)))}
    windows = []
    if not args.no_fp64:
        # coupled windows against the fp64 sampler from the MFMA job's own
        # states: after burn-in (accounted steps) and at the end (bare steps)
        print("bench: fp64 windows", file=sys.stderr, flush=True)
        for where, st, accounted in (("after burn-in", job["after_burn"], True),
                                     ("after run", job["end"], False)):
            x64, m64, _ = c3_window(sf, st, _lib.BH_GLM_FP64, 0.0, args.c3_window, accounted)
            w = {"from": where, "steps": args.c3_window,
                 "kind": "run-sampler!" if accounted else "burn-in!"}
            for name, prec, g in legs:
                xg, mg, gs = c3_window(sf, st, prec, g, args.c3_window, accounted)
                key = f"{name}_guarded" if g else f"{name}_plain"
                w[f"{key}_positions_equal_fp64"] = bool(np.array_equal(xg, x64))
                w[f"{key}_walkers_differing"] = int(np.any(xg != x64, axis=1).sum())
                if accounted:
                    w[f"{key}_means_equal_fp64"] = bool(np.array_equal(mg, m64))
                if g:
                    w[f"{key}_listed"] = gs["listed"]
                    w[f"{key}_overflow"] = gs["overflow"]
            windows.append(w)
    cal = c3_calibrated()
    # ESS per walker-step: the long-run calibration where it exists (the
    # job's own estimate, from about 6 tau_max, is reported beside it)
    ess_ws = cal["ess_per_walker_step"]["mean"] if cal is not None else js["ess_per_walker_step"]
    out = {}
    for name, prec, g in legs:
        t = timed[name]
        o = {"walker_steps_per_s": round(t["walker_steps_per_s"], 1),
             "ms_per_step": round(t["ms_per_step"], 3),
             "likelihood_kernel_ms": round(t["kernel"]["avg_ms"], 4),
             # the same Markov kernel in every leg: the guarded matrix-core
             # chains are the fp64 chain step for step (coupled windows), the
             # plain ones differ from it by rounding only
             "ess_per_s": round(ess_ws * t["walker_steps_per_s"], 2)}
        if g:
            o["guard"] = g
            o["guard_listed_per_half_step"] = round(t["guard"]["listed"] / (2 * K), 3)
            o["guard_overflow"] = t["guard"]["overflow"]
        out[name] = o
    full = None
    if ref is not None:
        # the same job (seeds, steps) on the fp64 kernel: posterior means and
        # SDs of the guarded and plain legs against it, in MC standard errors
        full = {"fp64": c3_ess_summary(ref)}
        mcse = mcse_of(ref)
        for name, r in ((mname, jc), ("fp32", r32)):
            err = np.abs(r["grand_mean"] - ref["grand_mean"]) / mcse
            full[name] = {"mean_err_over_mcse_max": round(float(err.max()), 4),
                          "mean_err_over_mcse_rms": round(float(np.sqrt(np.mean(err ** 2))), 4),
                          "sd_rel_err_max": round(float(np.max(np.abs(r["sd"] / ref["sd"] - 1.0))),
                                                  6),
                          "trajectory_equals_fp64": bool(np.array_equal(r["xs"], ref["xs"])),
                          "acceptance_rate": round(r["acc"], 4)}
    mt, ft = timed[mname], timed["fp32"]
    ws_ratio = mt["walker_steps_per_s"] / ft["walker_steps_per_s"]
    kt = mt["kernel"]
    flops = 2.0 * C3_N * C3_D * (C3_W // 2)
    achieved = flops / (kt["avg_ms"] * 1e-3) / 1e12
    # the practical rates of the device, measured: the product issued
    # msplit times, and one v_exp_f32 per contraction value plus one
    # v_log_f32 per sixteen in the epilogue
    rates = device_rates(fac)
    issued = flops * msplit / (kt["avg_ms"] * 1e-3)
    trans = C3_N * (C3_W // 2) * (1.0 + 1.0 / 16.0)
    trans_rate = trans / (kt["avg_ms"] * 1e-3)
    line = {
        "metric": METRIC, "value": out[mname]["ess_per_s"], "unit": "ESS/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": out[mname]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": mname,
        "data": "synthetic (seeded logistic-regression data, X stored bf16)",
        "config": {"workload": f"logistic regression {C3_N}x{C3_D}, {C3_W} walkers (a=2), "
                               f"prior N(0,5^2); value = ESS per walker-step (long-run "
                               f"calibration) x walker-steps/s of {K} timed move! steps after "
                               f"{W} warm-up steps; a {burn} burn-in + {run} run-sampler job "
                               f"in the same run",
                   "likelihood": f"{C3_MFMA_DESC[mname]}; fp32 accumulate, fp64 row sums, "
                                 f"guarded accept test ({mguard})",
                   "partner": "walker"},
        "walker_steps_per_s": out[mname]["walker_steps_per_s"],
        "mfma_vs_fp32_walker_steps_per_s": round(ws_ratio, 2),
        "mfma_vs_fp32_ess_per_s": round(ws_ratio, 2),
        "ess_job": js,
        "precisions": out,
        "log_density_err_vs_fp64": lp_err,
        "fp64_coupled_windows": windows,
        "roofline": {"kernel": f"{mkern} (1M rows x 8192 points)",
                     "bound": "mfma", "achieved": round(achieved, 1), "peak": MFMA_BF16_PEAK,
                     "unit": "TFLOP/s", "frac": round(achieved / MFMA_BF16_PEAK, 4),
                     "traffic": traffic_of(f"glm_logistic_mfma_{mname}_1Mx64_traffic.json"),
                     "avg_kernel_ms": round(kt["avg_ms"], 4), "launches_timed": kt["launches"],
                     "algorithmic_flops_per_launch": flops, "mfma_issue_factor": msplit,
                     "measured_rates": {
                         "mfma_issued_tflops": round(issued / 1e12, 1),
                         "mfma_f16_peak_measured_tflops": round(rates["mfma_f16_flops_per_s"] / 1e12, 1),
                         "mfma_issued_frac_of_measured": round(issued / rates["mfma_f16_flops_per_s"], 4),
                         "transcendentals_per_launch": trans,
                         "exp_rate_achieved_t": round(trans_rate / 1e12, 3),
                         "exp_rate_measured_t": round(rates["exp2_per_s"] / 1e12, 3),
                         "fma_rate_measured_t": round(rates["fma_per_s"] / 1e12, 3)}},
        "setup_s": round(t_setup, 1),
    }
    fj = os.path.join(ROOT, "profiles", "glm_logistic_1Mx64_fp64_full_job.json")
    if full is not None:
        # this run's whole-job comparison (--c3-fp64-full)
        line["fp64_full_job"] = dict(full, steps=f"{burn} burn-in + {run} run-sampler")
    elif os.path.exists(fj):
        with open(fj) as f:
            fjd = json.load(f)
        g = fjd["fp64_full_job"].get(mname, {})
        line["fp64_full_job"] = {"profile": "profiles/glm_logistic_1Mx64_fp64_full_job.json",
                                 "steps": "500 burn-in + 2000 run-sampler",
                                 f"{mname}_trajectory_equals_fp64": g.get("trajectory_equals_fp64"),
                                 f"{mname}_mean_err_over_mcse_max": g.get("mean_err_over_mcse_max"),
                                 f"{mname}_sd_rel_err_max": g.get("sd_rel_err_max")}
    line["ess_per_walker_step"] = {
        "value": ess_ws,
        "source": (f"profiles/{C3_CAL_FILE} (mean over {cal['ess_per_walker_step']['seeds']} "
                   f"seeds of {cal['ess_per_walker_step']['run_steps']}-step runs, sd "
                   f"{cal['ess_per_walker_step']['sd']:.3g})") if cal is not None
        else "this run's job",
        "this_job": js["ess_per_walker_step"],
        "this_job_ess_per_s": js["ess_per_s_job"]}
    if not args.no_roofline:
        line["plain_move_roofline"] = [move_roofline(fac, d, rates, m) for m, d in ROOF_DIMS]
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cb = c3_cpu_baseline(Xd, y, limits)
        line["cpu_baseline"] = {
            "value": round(cb["walker_steps_per_s"] * ess_ws, 4), "unit": "ESS/s",
            "cores": cb["threads"], "kind": "port",
            "sample": cb["sample"] + "; ESS per walker-step of the GPU job (the same chain)",
            "walker_steps_per_s": round(cb["walker_steps_per_s"], 2)}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps: sampler steps of the logistic workload (default 20), "
                         "whole jobs of c1 (default 3)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps before them (default 5; c1: jobs, default 1)")
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4", "c5"], default="c3")
    ap.add_argument("--c2-burn", type=int, default=C2_BURN)
    ap.add_argument("--c2-run", type=int, default=C2_RUN)
    ap.add_argument("--c2-user-steps", type=int, default=50,
                    help="steps of the user-model comparison run (0 to skip)")
    ap.add_argument("--c5-burn", type=int, default=C5_BURN)
    ap.add_argument("--c5-run", type=int, default=C5_RUN)
    ap.add_argument("--c5-fp32", action="store_true", help="also the plain fp32 batched run")
    ap.add_argument("--c5-mfma", choices=sorted(C5_MFMA), default=C5_MFMA_DEFAULT,
                    help="matrix-core precision of the batched run")
    ap.add_argument("--c4-burn", type=int, default=C4_BURN)
    ap.add_argument("--c4-run", type=int, default=C4_RUN)
    ap.add_argument("--c3-burn", type=int, default=C3_BURN)
    ap.add_argument("--c3-run", type=int, default=C3_RUN)
    ap.add_argument("--c3-guard", type=float, default=None,
                    help="guard of the headline leg (default: its precision's, C3_GUARDS)")
    ap.add_argument("--c3-mfma", choices=sorted(C3_MFMA), default=C3_MFMA_DEFAULT,
                    help="matrix-core precision of the measured leg")
    ap.add_argument("--c3-ess-calibration", action="store_true",
                    help="only the ESS-per-walker-step calibration of the logistic workload "
                         "(about 4 minutes per seed on fp16); prints the calibration file's "
                         "runs of other seeds merged with these")
    ap.add_argument("--c3-cal-seeds", default=",".join(str(x) for x in C3_CAL_SEEDS),
                    help="comma-separated seeds of the calibration runs")
    ap.add_argument("--c3-bf16", action="store_true", help="also the single-bf16 likelihood")
    ap.add_argument("--no-fp64", action="store_true", help="skip the coupled fp64 windows")
    ap.add_argument("--c3-fp64-full", action="store_true",
                    help="also the whole job on the fp64 likelihood (about 5 minutes)")
    ap.add_argument("--c3-window", type=int, default=C3_WINDOW)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the move-bare! roofline runs (correlated Gaussian at DIM 4 "
                         "and 16, independent coordinates at DIM 32), plus a 256 MiB device "
                         "copy whose FETCH_SIZE/WRITE_SIZE calibrate the counters")
    args = ap.parse_args()
    c1 = args.workload == "c1"
    if args.steps is None:
        args.steps = 3 if c1 else 20
    if args.warmup is None:
        args.warmup = 1 if c1 else 5
    if 0 < args.c2_user_steps < 50:
        ap.error("--c2-user-steps: run-sampler! needs at least 50 steps (0 skips the run)")
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding across GPUs is designed, not built)", file=sys.stderr)
        sys.exit(2)
    if args.roofline_only:
        import torch
        from bayadera_amd import engine
        fac = engine.HipBayaderaFactory(0, 256)
        rates = device_rates(fac)
        r = [move_roofline(fac, d, rates, m) for m, d in ROOF_DIMS]
        # random 16-byte gathers from 1 GiB, plain and non-temporal: under
        # rocprofv3 --pmc TCC_EA0_RDREQ_*B_sum their request sizes show
        # whether a load policy fetches less than a line
        gathers = {"plain": fac.device_gather_rate(1 << 30, False),
                   "nontemporal": fac.device_gather_rate(1 << 30, True)}
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"roofline": r, "gathers_per_s_1GiB": gathers,
                          "calibration_copy_bytes": src.numel() * 4}), flush=True)
        return
    if args.c3_ess_calibration:
        seeds = tuple(int(x) for x in args.c3_cal_seeds.split(","))
        return c3_calibration(seeds, args.c3_mfma)
    if args.workload == "c3":
        return main_c3(args)
    if args.workload == "c2":
        return main_c2(args)
    if args.workload == "c4":
        return main_c4(args)
    if args.workload == "c5":
        return main_c5(args)
    return main_c1(args)


if __name__ == "__main__":
    main()
