"""The batched workloads at their full sizes: 4096 hierarchical
beta-binomial posteriors (256 walkers each) and 512 uneven fp8 logistic
regressions (2048 walkers each, up to 200,000 rows), checked on sampled
problems against the oracle, which costs seconds per problem where the
whole batch would cost hours on the host."""
import numpy as np
import pytest

import oracle as O
from bayadera_amd import _lib, engine, models
from test_gpu_batch import _hbb_problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fac():
    f = engine.HipBayaderaFactory(0, 256)
    yield f
    f.release()


def test_hbb_batch_of_4096_problems_equals_single_oracle_samplers(fac):
    """4096 problems of DIM 7-52 in one batch, with the launch shape the
    batch's autotuner picks: eight sampled problems' walkers after burn-in
    equal the oracle's single samplers with the same seeds (up to the last
    bits of the devices' log, exp and lgamma, as for small batches)."""
    B, W, GMAX = 4096, 256, 50
    print("hbb full shape: 4096 problems x 256 walkers", flush=True)
    dims, params, lim = _hbb_problems(B, 1004, GMAX)
    bt = fac.batch(models.hbb_model(GMAX), params, dims, [1] * B, W)
    bt.init_position(40, lim)
    print("autotune", bt.autotune(5), flush=True)
    bt.burn_in(60, 4, 2.0)
    xs, lf = bt.state()
    assert all(np.all(np.isfinite(l)) for l in lf)
    rng = np.random.default_rng(7)
    for b in [0, 1, 2047, B - 1, *rng.integers(2, B - 1, 4)]:
        b = int(b)
        ref = O.Stretch(O.HBB_LOGPDF, dims[b], W, params[b], 60 + 2 * b, wgs=128,
                        data_len=2 * (dims[b] - 2), params_len=1, threads=4)
        ref.init(60 + 2 * b)
        ref.init_position(40 + b, lim[b, :dims[b]])
        ref.burn_in(4, 2.0)
        same = np.all(xs[b] == ref.xs.reshape(W, dims[b]), axis=1).mean()
        assert same > 0.97, (b, dims[b], same)
    bt.release()


def _c5_problems(P, seed):
    rng = np.random.default_rng(seed)
    Xs, ys = [], []
    for _ in range(P):
        N = int(round(np.exp(rng.uniform(np.log(1e3), np.log(2e5)))))
        d = int(rng.choice([8, 16, 32, 64]))
        X = (rng.standard_normal((N, d)) / np.sqrt(d)).astype(np.float32)
        b = rng.standard_normal(d)
        Xs.append(X)
        ys.append((rng.random(N) < 1.0 / (1.0 + np.exp(-(X @ b)))).astype(np.float32))
    return models.glm_logistic_batch_model(Xs, ys, prior_sd=5.0)


def test_glm_batch_of_512_problems(fac):
    """512 logistic regressions, N log-uniform in [1e3, 2e5], d in {8, 16,
    32, 64}, 2048 walkers, fp16x2 on the matrix cores with autotuned tiles:
    the initial log posteriors of sampled problems (the largest included)
    agree with the oracle's fp64 posterior to 2e-6, and accounted steps
    keep every problem's acceptance in (0, 1) and its log posteriors
    finite."""
    P, W = 512, 2048
    print("glm batch full shape: 512 problems x 2048 walkers", flush=True)
    m = _c5_problems(P, 1005)
    bt = fac.glm_batch(m, W).set_precision(_lib.BH_GLM_FP16X2)
    lim = np.tile(np.array([-0.5, 0.5], np.float32), (P, 64, 1))
    bt.init_position(9, lim)
    tuned = bt.autotune(2)
    print("autotune", tuned, flush=True)
    xs, lp = bt.state()
    rng = np.random.default_rng(11)
    for p in [int(np.argmax(m.rows)), 0, *rng.integers(1, P, 5)]:
        p = int(p)
        ref = O.glm_logpost(m.dequantized(p), m.y[p], 5.0, xs[p][:32])
        err = np.abs(lp[p, :32] - ref) / (np.abs(ref) + 1.0)
        assert err.max() < 2e-6, (p, int(m.rows[p]), int(m.dims[p]), err.max())
    bt.burn_in(100, 3, 2.0)
    res = bt.run(200, 3, 2.0)
    acc = res["acceptance-rate"]
    assert np.all((acc > 0.0) & (acc < 1.0))
    _, lp = bt.state()
    assert np.all(np.isfinite(lp))
    bt.release()
