"""The GLM (logistic regression) path on the GPU against the fp64 oracle."""
import numpy as np
import pytest

import oracle as O
from bayadera_amd import _lib, engine, models

pytestmark = pytest.mark.gpu


def _problem(N, d, seed):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((N, d)) / np.sqrt(d)).astype(np.float32)
    Xb = models.to_bf16(X)
    Xd = models.from_bf16(Xb).reshape(N, d)
    beta = rng.standard_normal(d)
    p = 1.0 / (1.0 + np.exp(-(Xd.astype(np.float64) @ beta)))
    y = (rng.random(N) < p).astype(np.float32)
    return Xb, Xd, y


@pytest.fixture(scope="module")
def fac():
    f = engine.HipBayaderaFactory(0, 256)
    yield f
    f.release()


@pytest.mark.parametrize("d", [32, 64])
@pytest.mark.parametrize("N", [1, 1000, 20011])
def test_glm_log_density_precisions(fac, N, d):
    Xb, Xd, y = _problem(N, d, 7 + N + d)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y, prior_sd=5.0))
    rng = np.random.default_rng(3)
    beta = (rng.standard_normal((700, d)) * 0.7).astype(np.float32)
    ref = O.glm_logpost(Xd, y, 5.0, beta)
    scale = np.abs(ref) + 1.0
    f64 = sf.log_density(beta, _lib.BH_GLM_FP64)
    assert np.max(np.abs(f64 - ref) / scale) < 1e-12
    f32 = sf.log_density(beta, _lib.BH_GLM_FP32)
    x3 = sf.log_density(beta, _lib.BH_GLM_BF16X3)
    h2 = sf.log_density(beta, _lib.BH_GLM_FP16X2)
    b1 = sf.log_density(beta, _lib.BH_GLM_BF16)
    # fp32-accurate contractions: error of a few 1e-7 of |ll| per point
    assert np.max(np.abs(f32 - ref) / scale) < 2e-6
    assert np.max(np.abs(x3 - ref) / scale) < 2e-6
    assert np.max(np.abs(h2 - ref) / scale) < 2e-6
    # one bf16 product: the proposals lose 16 mantissa bits
    assert np.max(np.abs(b1 - ref) / scale) < 2e-2


def test_glm_sampler_fp64_matches_oracle_trajectory(fac):
    N, d, W = 3000, 32, 1024
    Xb, Xd, y = _problem(N, d, 11)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y, prior_sd=5.0))
    s = sf.create_sampler(5, W, np.zeros(1, np.float32)).set_glm_precision(_lib.BH_GLM_FP64)
    s.set_partner(_lib.BH_PARTNER_WALKER)
    s.init(5)
    lim = np.tile(np.array([-0.5, 0.5], np.float32), d)
    s.init_position(6, lim)
    ref = O.GlmStretch(Xd, y, 5.0, W, 5, partner=O.PARTNER_WALKER, threads=8)
    ref.init(5)
    ref.init_position(6, lim.reshape(d, 2))
    xs, lp, _ = s.state()
    assert np.array_equal(xs.reshape(-1), ref.xs)
    assert np.allclose(lp, ref.lp.astype(np.float32), rtol=1e-6)
    s.burn_in(3, 2.0)
    ref.burn_in(3, 2.0)
    same = np.all(s.state()[0].reshape(-1, d) == ref.xs.reshape(-1, d), axis=1).mean()
    assert same > 0.99, same


@pytest.mark.parametrize("prec", [_lib.BH_GLM_FP32, _lib.BH_GLM_BF16X3, _lib.BH_GLM_FP16X2,
                                  _lib.BH_GLM_BF16])
def test_glm_sampler_runs_and_accounts(fac, prec):
    N, d, W = 4096, 64, 2048
    Xb, Xd, y = _problem(N, d, 13)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y, prior_sd=5.0))
    s = sf.create_sampler(8, W, np.zeros(1, np.float32)).set_glm_precision(prec)
    s.set_partner(_lib.BH_PARTNER_WALKER)
    s.init_position(9, np.tile(np.array([-0.1, 0.1], np.float32), d))
    s.burn_in(20, 2.0)
    res = s.run_sampler(60, 2.0)
    assert 0.05 < res["acceptance-rate"] < 0.9
    assert np.all(np.isfinite(res["means"]))
    xs, lp, _ = s.state()
    ref = O.glm_logpost(Xd, y, 5.0, xs[:64])
    # a single bf16 product rounds the coefficients to 8 bits (errors ~0.1
    # in a log posterior of ~-2700); the fp32-accurate modes agree to 1e-5
    tol = 5e-4 if prec == _lib.BH_GLM_BF16 else 1e-5
    assert np.allclose(lp[:64], ref, rtol=tol, atol=1e-2)


def test_glm_walker_rule(fac):
    Xb, Xd, y = _problem(100, 32, 1)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y))
    with pytest.raises(_lib.ArgumentError, match="multiple of 512"):
        sf.create_sampler(1, 768, np.zeros(1, np.float32))


def _coupled_run(sf, prec, guard, d, W=2048, burn=25, run=60):
    s = sf.create_sampler(21, W, np.zeros(1, np.float32)).set_glm_precision(prec)
    s.set_partner(_lib.BH_PARTNER_WALKER)
    if guard:
        s.set_glm_guard(guard)
    s.init(21)
    s.init_position(22, np.tile(np.array([-0.3, 0.3], np.float32), d))
    s.burn_in(burn, 2.0)
    res = s.run_sampler(run, 2.0)
    xs = s.state()[0].copy()
    stats = s.glm_guard_stats()
    s.release()
    return xs, res, stats


@pytest.mark.parametrize("prec", [_lib.BH_GLM_BF16X3, _lib.BH_GLM_FP16X2, _lib.BH_GLM_FP16,
                                  _lib.BH_GLM_FP32])
def test_glm_guard_makes_the_fp64_decisions(fac, prec):
    """Same seeds, same initial walkers: with the guarded accept test the
    reduced-precision sampler's trajectory is the fp64 sampler's, position
    for position, and its accounting matches."""
    # at this size the reduced precisions' log posteriors sit ~0.005 above
    # the fp64 ones (a common offset that cancels only between values of the
    # same precision) with ~5e-4 of scatter
    N, d = 262144, 64
    Xb, Xd, y = _problem(N, d, 17)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y, prior_sd=5.0))
    x64, r64, s64 = _coupled_run(sf, _lib.BH_GLM_FP64, 0.0, d)
    assert s64 == {"listed": 0, "overflow": 0}
    xg, rg, sg = _coupled_run(sf, prec, 5e-3, d)
    print("guard stats", sg)
    assert sg["overflow"] == 0
    assert np.array_equal(xg, x64)
    assert rg["acceptance-rate"] == r64["acceptance-rate"]
    assert np.array_equal(rg["means"], r64["means"])


def test_glm_guard_lists_near_threshold_walkers(fac):
    """A guard wide enough to list every walker whose margin is small still
    leaves the trajectory on the fp64 one, and the count is visible."""
    N, d = 4096, 32
    Xb, Xd, y = _problem(N, d, 19)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y, prior_sd=5.0))
    x64, r64, _ = _coupled_run(sf, _lib.BH_GLM_FP64, 0.0, d, W=1024)
    xg, rg, sg = _coupled_run(sf, _lib.BH_GLM_BF16X3, 0.5, d, W=1024)
    assert sg["listed"] > 0 and sg["overflow"] == 0
    assert np.array_equal(xg, x64)


def test_glm_fp16_refuses_values_outside_its_range(fac):
    """X with a value beyond the fp16 range (65504) cannot take the fp16
    matrix-core path; every other precision still runs."""
    Xb, Xd, y = _problem(256, 32, 3)
    X = models.from_bf16(Xb).reshape(256, 32).copy()
    X[100, 7] = 70000.0
    sf = fac.mcmc_factory(models.glm_logistic_model(models.to_bf16(X), y))
    beta = np.zeros((4, 32), np.float32)
    with pytest.raises(_lib.ArgumentError, match="fp16 range"):
        sf.log_density(beta, _lib.BH_GLM_FP16X2)
    s = sf.create_sampler(1, 1024, np.zeros(1, np.float32))
    with pytest.raises(_lib.ArgumentError, match="fp16 range"):
        s.set_glm_precision(_lib.BH_GLM_FP16X2)
    s.set_glm_precision(_lib.BH_GLM_BF16X3)
    assert np.all(np.isfinite(sf.log_density(beta, _lib.BH_GLM_BF16X3)))
    s.release()


def test_glm_guard_rejects_bad_values(fac):
    Xb, Xd, y = _problem(100, 32, 1)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y))
    s = sf.create_sampler(1, 1024, np.zeros(1, np.float32))
    with pytest.raises(_lib.ArgumentError):
        s.set_glm_guard(-1.0)
    with pytest.raises(_lib.ArgumentError):
        s.set_glm_guard(float("nan"))
    s.release()


def test_kernel_timing_counts_contractions(fac):
    Xb, Xd, y = _problem(20000, 64, 23)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y))
    s = sf.create_sampler(2, 1024, np.zeros(1, np.float32))
    s.init_position(3, np.tile(np.array([-0.1, 0.1], np.float32), 64))
    s.set_kernel_timing(True)
    s.burn_in(5, 2.0)
    t = s.kernel_timing()
    assert t["launches"] == 10 and t["avg_ms"] > 0.0
    # every fourth launch: 0, 4, 8, 12, 16 of 20
    s.set_kernel_timing(True, 4)
    s.burn_in(10, 2.0)
    t4 = s.kernel_timing()
    assert t4["launches"] == 5 and t4["avg_ms"] > 0.0
    s.set_kernel_timing(False)
    s.burn_in(1, 2.0)
    assert s.kernel_timing()["launches"] == 0
    s.release()


def _linreg(N, d, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, d)).astype(np.float32)
    beta = rng.standard_normal(d)
    y = (X.astype(np.float64) @ beta + 0.5 * rng.standard_normal(N)).astype(np.float32)
    pts = np.concatenate([beta + 0.01 * rng.standard_normal((300, d)),
                          np.log(0.5) + 0.05 * rng.standard_normal((300, 1))], 1)
    return X, y, pts.astype(np.float32)


def _linreg_lp64(X, y, pts, sb=10.0, ss=1.0):
    """The linear-regression log posterior in float64 (numpy)."""
    Xd, yd = X.astype(np.float64), y.astype(np.float64)
    b, ls = pts[:, :-1].astype(np.float64), pts[:, -1].astype(np.float64)
    r = yd[None, :] - b @ Xd.T
    N = X.shape[0]
    ll = -N * ls - 0.5 * N * np.log(2 * np.pi) - 0.5 * np.exp(-2 * ls) * np.sum(r * r, 1)
    return ll - 0.5 * (np.sum(b * b, 1) / sb ** 2 + ls ** 2 / ss ** 2)


@pytest.mark.parametrize("N,d", [(5000, 16), (777, 5), (20011, 33)])
def test_glm_gaussian_log_density_vs_fp64_and_user_model(fac, N, d):
    X, y, pts = _linreg(N, d, 29 + d)
    sf = fac.mcmc_factory(models.glm_gaussian_model(X, y, 10.0, 1.0))
    ref = _linreg_lp64(X, y, pts)
    scale = np.abs(ref) + 1.0
    f64 = sf.log_density(pts, _lib.BH_GLM_FP64)
    f32 = sf.log_density(pts, _lib.BH_GLM_FP32)
    assert np.max(np.abs(f64 - ref) / scale) < 1e-12
    assert np.max(np.abs(f32 - ref) / scale) < 2e-6
    with pytest.raises(_lib.ArgumentError):
        sf.log_density(pts, _lib.BH_GLM_BF16X3)
    # the same posterior through the reference's plugin path (user C99
    # source compiled by hipRTC, one walker per thread looping over rows)
    lik = models.likelihood_model(models.LINREG_SOURCE, name="linreg")
    prior = models.distribution_model(models.LINREG_SOURCE, name="linreg_prior",
                                      dimension=d + 1, params_size=2)
    post = models.posterior_model(prior, "linreg_post", lik)
    de = fac.distribution_engine(post)
    params = np.concatenate([np.concatenate([X, y[:, None]], 1).reshape(-1),
                             np.array([10.0, 1.0], np.float32)]).astype(np.float32)
    user = de.log_density(params, pts[:64]).astype(np.float64)
    assert np.max(np.abs(user - ref[:64]) / scale[:64]) < 2e-5
    de.release()


@pytest.mark.parametrize("d", [5, 16, 32])
def test_glm_gaussian_two_point_kernel_equals_one_point(fac, d):
    """The fp32 linear-regression kernel takes two points per thread when
    the point count is a multiple of 512 (each LDS read of a row feeds both
    points) and one otherwise, with the same arithmetic per point: at 3000
    rows both calls below take the capped range count, so the log densities
    of the same 768 points from a 1024-point call (two per thread) and a
    768-point call (one per thread) are bit-identical, and both agree with
    the float64 formula."""
    N = 3000
    X, y, _ = _linreg(N, d, 41 + d)
    bhat = np.linalg.lstsq(X.astype(np.float64), y.astype(np.float64), rcond=None)[0]
    rng = np.random.default_rng(5 + d)
    pts = np.concatenate([bhat + 0.02 * rng.standard_normal((1024, d)),
                          np.log(0.5) + 0.05 * rng.standard_normal((1024, 1))],
                         1).astype(np.float32)
    sf = fac.mcmc_factory(models.glm_gaussian_model(X, y, 10.0, 1.0))
    two = sf.log_density(pts, _lib.BH_GLM_FP32)
    one = sf.log_density(np.ascontiguousarray(pts[:768]), _lib.BH_GLM_FP32)
    assert np.array_equal(two[:768], one)
    ref = _linreg_lp64(X, y, pts)
    assert np.max(np.abs(two - ref) / (np.abs(ref) + 1.0)) < 2e-6


def test_glm_gaussian_sampler_moves_and_rejects_guard(fac):
    X, y, pts = _linreg(4000, 16, 41)
    sf = fac.mcmc_factory(models.glm_gaussian_model(X, y))
    s = sf.create_sampler(3, 1024, np.zeros(1, np.float32)).set_partner(_lib.BH_PARTNER_WALKER)
    with pytest.raises(_lib.ArgumentError):
        s.set_glm_guard(1e-2)
    with pytest.raises(_lib.ArgumentError):
        s.set_glm_precision(_lib.BH_GLM_BF16X3)
    lo = pts.min(0) - 0.01
    hi = pts.max(0) + 0.01
    s.init_position(4, np.stack([lo, hi], 1).reshape(-1))
    s.burn_in(50, 2.0)
    res = s.run_sampler(100, 2.0)
    assert 0.05 < res["acceptance-rate"] < 0.9
    xs, lp, _ = s.state()
    ref = _linreg_lp64(X, y, xs[:64])
    assert np.allclose(lp[:64], ref, rtol=2e-6)
    s.release()


def test_device_rates_are_measured(fac):
    """The issue rates that bound the softplus epilogue: both positive, and
    the transcendental slower than the fma."""
    r = fac.device_rates()
    assert r["exp2_per_s"] > 1e11 and r["fma_per_s"] > r["exp2_per_s"]
    assert 1e14 < r["mfma_f16_flops_per_s"] < 3e15
    assert 1e9 < r["gathers_per_s"] < 1e12


def test_glm_gaussian_sampler_state_at_8192_walkers(fac):
    """The linear regression's walkers carry d + 1 coordinates (log sigma):
    its work buffers are sized by the walker dimension, so at a large
    ensemble the stored log posteriors still equal the fp64 formula's."""
    X, y, pts = _linreg(20000, 16, 43)
    sf = fac.mcmc_factory(models.glm_gaussian_model(X, y, 10.0, 1.0))
    s = sf.create_sampler(5, 8192, np.zeros(1, np.float32)).set_partner(_lib.BH_PARTNER_WALKER)
    lo = pts.min(0) - 0.01
    hi = pts.max(0) + 0.01
    s.init(5)
    s.init_position(6, np.stack([lo, hi], 1).reshape(-1))
    s.burn_in(30, 2.0)
    xs, lp, _ = s.state()
    idx = np.r_[0:64, 4096:4160, 8128:8192]
    ref = _linreg_lp64(X, y, xs[idx])
    assert np.allclose(lp[idx], ref, rtol=2e-6)
    s.release()


def test_glm_guard_full_shape_equals_fp64(fac):
    """The logistic workload at its full shape (1,000,000 x 64 rows, 16384
    walkers): ten guarded fp16x2 steps take exactly the fp64 sampler's
    positions, nothing overflows the guard's list, and the fp16x2 and fp64
    log posteriors of a few walkers agree with the oracle's."""
    N, d, W = 1_000_000, 64, 16384
    print("full-shape guard: 1M x 64, 16384 walkers, 10 steps", flush=True)
    Xb, Xd, y = _problem(N, d, 1003)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y, prior_sd=5.0))
    out = {}
    for prec, guard in ((_lib.BH_GLM_FP64, 0.0), (_lib.BH_GLM_FP16X2, 1e-2),
                        (_lib.BH_GLM_FP16, 5e-3)):
        s = sf.create_sampler(31, W, np.zeros(1, np.float32)).set_glm_precision(prec)
        s.set_partner(_lib.BH_PARTNER_WALKER)
        if guard:
            s.set_glm_guard(guard)
        s.init(31)
        s.init_position(32, np.tile(np.array([-0.3, 0.3], np.float32), d))
        s.burn_in(10, 2.0)
        out[prec] = (s.state()[0].copy(), s.glm_guard_stats())
        s.release()
    x64, _ = out[_lib.BH_GLM_FP64]
    for prec in (_lib.BH_GLM_FP16X2, _lib.BH_GLM_FP16):
        # fp16 from this box, far from the posterior: its guard is widened
        # per walker by the distance from the reference point
        xg, st = out[prec]
        print("precision", prec, "guard stats", st, flush=True)
        assert st["overflow"] == 0 and st["listed"] > 0
        assert np.array_equal(xg, x64)
    pts = np.ascontiguousarray(x64[:4])
    ref = O.glm_logpost(Xd, y, 5.0, pts)
    assert np.max(np.abs(sf.log_density(pts, _lib.BH_GLM_FP64) - ref) / np.abs(ref)) < 1e-12
    assert np.max(np.abs(sf.log_density(pts, _lib.BH_GLM_FP16X2) - ref) / np.abs(ref)) < 2e-6


def test_glm_split_correction_keeps_log_ratios_near_fp64(fac):
    """The split precisions' log posteriors carry the first-order
    correction of the split's rounding (the softplus sum's gradient at the
    mean of 256 of the points): over 8192 pairs of points of a
    posterior-sized cloud at 1M x 64 rows, the error of the Metropolis log
    ratio lp(a) - lp(b) against the fp64 kernel has sd below 5e-4 and max
    below 2.5e-3 for fp16x2 and bf16x3 (fp16x2 without the correction: sd
    ~2e-3, max ~8e-3), and, with the second-order correction of its
    one-term rounding, sd below 1.5e-3 and max below 6e-3 for fp16 (without
    it: sd ~8); the fp64 kernel agrees with the oracle."""
    N, d = 1_000_000, 64
    print("split correction: 1M x 64, 16384 points", flush=True)
    Xb, Xd, y = _problem(N, d, 1003)
    sub = Xd[:100_000].astype(np.float64)
    ys = y[:100_000].astype(np.float64)
    b = np.zeros(d)
    for _ in range(8):  # the posterior mode of a tenth of the rows
        p = 1.0 / (1.0 + np.exp(-(sub @ b)))
        H = (sub * (p * (1 - p))[:, None]).T @ sub + np.eye(d) / 25.0
        b = b + np.linalg.solve(H, sub.T @ (ys - p) - b / 25.0)
    rng = np.random.default_rng(5)
    pts = (b + 0.02 * rng.standard_normal((16384, d))).astype(np.float32)
    sf = fac.mcmc_factory(models.glm_logistic_model(Xb, y, prior_sd=5.0))
    lp64 = sf.log_density(pts, _lib.BH_GLM_FP64)
    ref = O.glm_logpost(Xd, y, 5.0, pts[:4])
    assert np.max(np.abs(lp64[:4] - ref) / np.abs(ref)) < 1e-12
    d64 = lp64[:8192] - lp64[8192:]
    for prec, sd_max, err_max in ((_lib.BH_GLM_FP16X2, 5e-4, 2.5e-3),
                                  (_lib.BH_GLM_BF16X3, 5e-4, 2.5e-3),
                                  (_lib.BH_GLM_FP16, 1.5e-3, 6e-3)):
        lp = sf.log_density(pts, prec)
        de = (lp[:8192] - lp[8192:]) - d64
        print("precision", prec, "log-ratio error sd", np.std(de), "max", np.max(np.abs(de)),
              flush=True)
        assert np.std(de) < sd_max and np.max(np.abs(de)) < err_max
