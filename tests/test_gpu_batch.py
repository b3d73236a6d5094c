"""Batched posteriors: each problem of one launch equals the oracle's
single sampler with the problem's seeds."""
import numpy as np
import pytest

import oracle as O
from bayadera_amd import _lib, engine, models

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fac():
    f = engine.HipBayaderaFactory(0, 256)
    yield f
    f.release()


def _mvn_problems(B, seed):
    rng = np.random.default_rng(seed)
    dims, params = [], []
    for _ in range(B):
        d = int(rng.integers(1, 7))
        mu = rng.standard_normal(d)
        A = rng.standard_normal((d, d))
        P = np.linalg.inv(A @ A.T + np.eye(d))
        dims.append(d)
        params.append(np.concatenate([mu, P.reshape(-1)]).astype(np.float32))
    return dims, params


@pytest.mark.parametrize("where", [_lib.BH_BATCH_STATE_LDS, _lib.BH_BATCH_STATE_GLOBAL])
@pytest.mark.parametrize("partner", [_lib.BH_PARTNER_FLAT, _lib.BH_PARTNER_WALKER])
def test_batch_problems_equal_single_oracle_samplers(fac, partner, where):
    B, W, MAXD = 12, 256, 6
    dims, params = _mvn_problems(B, 4)
    bt = fac.batch(models.mvn_model(MAXD), params, dims, [0] * B, W).set_partner(partner)
    bt.set_state(where)
    lim = np.zeros((B, MAXD, 2), np.float32)
    lim[:, :, 0], lim[:, :, 1] = -4.0, 4.0
    bt.init_position(100, lim)
    bt.burn_in(50, 6, 2.0)
    xs, lf = bt.state()
    for b in range(B):
        ref = O.Stretch(O.MVN_LOGPDF, dims[b], W, params[b], 50 + 2 * b, wgs=128,
                        params_len=len(params[b]), partner=partner)
        ref.init(50 + 2 * b)
        ref.init_position(100 + b, lim[b, :dims[b]])
        ref.burn_in(6, 2.0)
        same = np.all(xs[b] == ref.xs.reshape(W, dims[b]), axis=1).mean()
        assert same > 0.99, (b, same)


def test_batch_moments_and_acceptance(fac):
    # 1-D Gaussians N(mu_b, s_b): posterior means and sds recovered
    B, W = 64, 512
    rng = np.random.default_rng(9)
    mus = rng.uniform(-3, 3, B).astype(np.float32)
    sds = rng.uniform(0.5, 2, B).astype(np.float32)
    params = [np.array([m, s], np.float32) for m, s in zip(mus, sds)]
    bt = fac.batch(models.gaussian_model(), params, [1] * B, [2] * B, W)
    lim = np.stack([mus - 3 * sds, mus + 3 * sds], 1).reshape(B, 1, 2)
    bt.init_position(7, lim)
    bt.burn_in(11, 200, 2.0)
    res = bt.run(13, 400, 2.0)
    assert np.all(np.abs(res["mean"][:, 0] - mus) < 0.08 * sds)
    assert np.all(np.abs(res["sd"][:, 0] / sds - 1) < 0.08)
    assert np.all((res["acceptance-rate"] > 0.5) & (res["acceptance-rate"] < 0.95))


def test_batch_hierarchical_beta_binomial_runs(fac):
    rng = np.random.default_rng(1004)
    B, W, GMAX = 32, 256, 50
    dims, params = [], []
    for _ in range(B):
        G = int(rng.integers(5, GMAX + 1))
        n = np.round(np.exp(rng.uniform(np.log(10), np.log(1e5), G)))
        th = rng.beta(2, 5, G)
        y = rng.binomial(n.astype(np.int64), th).astype(np.float64)
        dims.append(G + 2)
        params.append(np.concatenate([n, y, [10.0]]).astype(np.float32))
    bt = fac.batch(models.hbb_model(GMAX), params, dims, [1] * B, W)
    lim = np.zeros((B, GMAX + 2, 2), np.float32)
    for b in range(B):
        G = dims[b] - 2
        n, y = params[b][:G], params[b][G:2 * G]
        t = np.log((y + 0.5) / (n - y + 0.5))
        lim[b, 0] = [0.0, 1.5]
        lim[b, 1] = [0.5, 2.0]
        lim[b, 2:dims[b], 0] = t - 0.05
        lim[b, 2:dims[b], 1] = t + 0.05
    bt.init_position(3, lim)
    bt.burn_in(5, 100, 2.0)
    res = bt.run(6, 100, 2.0)
    assert np.all(np.isfinite(res["mean"]))
    # stretch moves in 7 to 52 dimensions after 100 burn-in steps: low but
    # nonzero acceptance in every problem (1-8% here)
    assert np.all(res["acceptance-rate"] > 0.005)
    _, lf = bt.state()
    assert np.all(np.isfinite(lf))


def test_batch_autotune_keeps_trajectories_and_restores_state(fac):
    """The chosen launch shape (threads per problem, dimension classes)
    changes no position; the tuning runs leave the state as it was."""
    B, W, MAXD = 40, 512, 6
    dims, params = _mvn_problems(B, 8)
    lim = np.zeros((B, MAXD, 2), np.float32)
    lim[:, :, 0], lim[:, :, 1] = -4.0, 4.0
    runs = []
    for tune in (False, True):
        bt = fac.batch(models.mvn_model(MAXD), params, dims, [0] * B, W)
        bt.init_position(3, lim)
        before = bt.state()[0]
        if tune:
            t = bt.autotune(5)
            assert t["threads_per_problem"] in (64, 128, 256) and t["launch_classes"] >= 1
            after = bt.state()[0]
            assert all(np.array_equal(x, y) for x, y in zip(before, after))
        bt.burn_in(21, 12, 2.0)
        res = bt.run(22, 60, 2.0)
        runs.append((bt.state()[0], res["acceptance-rate"]))
        bt.release()
    assert all(np.array_equal(x, y) for x, y in zip(runs[0][0], runs[1][0]))
    assert np.array_equal(runs[0][1], runs[1][1])


def _hbb_problems(B, seed, GMAX=50):
    rng = np.random.default_rng(seed)
    dims, params = [], []
    lim = np.zeros((B, GMAX + 2, 2), np.float32)
    for b in range(B):
        G = int(rng.integers(5, GMAX + 1))
        n = np.round(np.exp(rng.uniform(np.log(10), np.log(1e5), G)))
        th = rng.beta(2, 5, G)
        y = rng.binomial(n.astype(np.int64), th).astype(np.float64)
        dims.append(G + 2)
        params.append(np.concatenate([n, y, [10.0]]).astype(np.float32))
        t = np.log((y + 0.5) / (n - y + 0.5))
        lim[b, 0] = [0.0, 1.5]
        lim[b, 1] = [0.5, 2.0]
        lim[b, 2:G + 2, 0] = t - 0.05
        lim[b, 2:G + 2, 1] = t + 0.05
    return dims, params, lim


def test_batch_state_placements_agree_and_series_is_the_ensemble_mean(fac):
    """Walkers in LDS and in place in global memory: identical positions,
    acceptance, accounting and series (the same arithmetic in the same
    order); each step's recorded mean is the ensemble mean of that step's
    positions (checked at the last step)."""
    import torch
    B, W, GMAX = 24, 256, 50
    dims, params, lim = _hbb_problems(B, 77, GMAX)
    out = []
    for where in (_lib.BH_BATCH_STATE_LDS, _lib.BH_BATCH_STATE_GLOBAL):
        bt = fac.batch(models.hbb_model(GMAX), params, dims, [1] * B, W).set_state(where)
        assert bt.class_states() == [where]
        bt.init_position(3, lim)
        bt.burn_in(5, 20, 2.0)
        n = 30
        series = torch.zeros(n * int(np.sum(dims)), dtype=torch.float32, device="cuda")
        res = bt.run(6, n, 2.0, series=series)
        xs, lf = bt.state()
        blocks = [np.asarray(blk.cpu()) for blk in bt.series_blocks(series, n)]
        out.append((xs, lf, res, blocks))
        bt.release()
    (x0, l0, r0, s0), (x1, l1, r1, s1) = out
    assert all(np.array_equal(a, b) for a, b in zip(x0, x1))
    assert np.array_equal(l0, l1)
    for k in ("acceptance-rate", "mean", "sd"):
        assert np.array_equal(np.asarray(r0[k]), np.asarray(r1[k])), k
    assert all(np.array_equal(a, b) for a, b in zip(s0, s1))
    for b in range(B):
        want = x0[b].astype(np.float64).mean(0)
        assert np.allclose(s0[b][-1], want, rtol=1e-5, atol=1e-5), b


@pytest.mark.parametrize("where", [_lib.BH_BATCH_STATE_LDS, _lib.BH_BATCH_STATE_GLOBAL])
def test_batch_hbb_problems_equal_single_oracle_samplers(fac, where):
    """Hierarchical beta-binomials (DIM 7-52) in one batch: each problem's
    walkers after burn-in equal the oracle's single sampler with the same
    seeds (model evaluations differ in the last bits between the devices'
    log, exp and lgamma, so a few accept tests may go the other way)."""
    B, W, GMAX = 6, 256, 50
    dims, params, lim = _hbb_problems(B, 5, GMAX)
    bt = fac.batch(models.hbb_model(GMAX), params, dims, [1] * B, W).set_state(where)
    bt.init_position(40, lim)
    xs0, lf0 = bt.state()
    bt.burn_in(60, 4, 2.0)
    xs, _ = bt.state()
    for b in range(B):
        ref = O.Stretch(O.HBB_LOGPDF, dims[b], W, params[b], 60 + 2 * b, wgs=128,
                        data_len=2 * (dims[b] - 2), params_len=1, threads=4)
        ref.init(60 + 2 * b)
        ref.init_position(40 + b, lim[b, :dims[b]])
        # identical starting walkers; the model's log density within a few
        # float32 ulps of the oracle's
        assert np.array_equal(xs0[b], ref.xs.reshape(W, dims[b]))
        assert np.allclose(lf0[b], ref.logfn, rtol=2e-6, atol=0), b
        ref.burn_in(4, 2.0)
        x = xs[b]
        r = ref.xs.reshape(W, dims[b])
        same = np.all(x == r, axis=1).mean()
        assert same > 0.97, (b, dims[b], same)
        close = np.abs(x - r) <= 1e-5 * np.maximum(1.0, np.abs(r))
        assert np.all(close[np.all(x == r, axis=1)])


def test_acor_batch_equals_acor_per_series(fac):
    """acor of every problem's recorded mean series in one launch equals the
    engine's acor of each series alone, bit for bit (DIM 7-52, and 1)."""
    import torch
    B, W, GMAX = 8, 256, 50
    dims, params, lim = _hbb_problems(B, 21, GMAX)
    bt = fac.batch(models.hbb_model(GMAX), params, dims, [1] * B, W)
    bt.init_position(3, lim)
    bt.burn_in(5, 50, 2.0)
    n = 300
    series = torch.zeros(n * int(np.sum(dims)), dtype=torch.float32, device="cuda")
    bt.run(6, n, 2.0, series=series)
    acor = fac.acor_engine()
    batch = acor.acor_batch(series, dims, n)
    for b, blk in enumerate(bt.series_blocks(series, n)):
        one = acor.acor(blk.contiguous())
        for f in ("tau", "mean", "sigma"):
            assert np.array_equal(getattr(one, f), getattr(batch[b], f), equal_nan=True), (b, f)
        assert one.lag == batch[b].lag
    bt.release()
    # one-dimensional series too
    x = np.cumsum(np.random.default_rng(4).standard_normal((400, 3)), axis=0).astype(np.float32)
    joined = np.concatenate([x[:, :1].reshape(-1), x[:, 1:].reshape(-1)])
    res = acor.acor_batch(joined, [1, 2], 400)
    for got, cols in zip(res, (x[:, :1], x[:, 1:])):
        want = acor.acor(np.ascontiguousarray(cols))
        assert np.array_equal(want.tau, got.tau, equal_nan=True)
        assert np.array_equal(want.sigma, got.sigma, equal_nan=True)
