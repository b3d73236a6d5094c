"""Work-group sizes other than 256 and the resident kernel's multi-step
sample!, through the C-ABI on the GPU against the oracle."""
import numpy as np
import pytest

import oracle as O
from bayadera_amd import _lib, engine, models

pytestmark = pytest.mark.gpu

STEPWISE, RESIDENT = _lib.BH_EXEC_STEPWISE, _lib.BH_EXEC_RESIDENT


def _mvn_params(dim, seed=1001):
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal(dim).astype(np.float32)
    A = rng.standard_normal((dim, dim))
    P = np.linalg.inv(A @ A.T + np.eye(dim)).astype(np.float32)
    return np.concatenate([mu, P.reshape(-1)]).astype(np.float32)


@pytest.mark.parametrize("dim,walkers", [(1, 4096), (4, 1024)])
def test_resident_sample_many_steps_equals_stepwise(dim, walkers):
    """sample! of many ensembles in one resident launch (each step's
    ensemble copied out while the next step starts) equals the stepwise
    path bit for bit; for the uniform model (exact decisions) also the
    oracle."""
    print(f"resident sample!: dim {dim}, {walkers} walkers, 60 steps", flush=True)
    fac = engine.HipBayaderaFactory(0, 256)
    if dim == 1:
        sf = fac.mcmc_factory(models.uniform_model())
        params, lim = np.float32([-1, 2]), np.float32([-1, 2])
    else:
        sf = fac.mcmc_factory(models.mvn_model(dim))
        params, lim = _mvn_params(dim), np.tile(np.float32([-5, 5]), dim)
    n = 60 * walkers
    out = []
    for mode in (STEPWISE, RESIDENT):
        s = sf.create_sampler(31, walkers, params).set_exec(mode)
        s.init(31)
        s.init_position(32, lim)
        s.burn_in(8, 2.0)
        out.append((s.sample(n), s.state()[0]))
        s.release()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    if dim == 1:
        ref = O.Stretch(O.UNIFORM_LOGPDF, 1, walkers, params, 31, params_len=2, threads=4)
        ref.init(31)
        ref.init_position(32, lim)
        ref.burn_in(8, 2.0)
        assert np.array_equal(out[1][0], ref.sample(n))
    sf.release()
    fac.release()


@pytest.mark.parametrize("wgs", [512, 1024])
def test_acor_on_wider_work_groups(wgs, goldens):
    """acor with a WGS of 512 or 1024 (the reduction trees and the lag rule
    follow WGS) equals the oracle's at that WGS bit for bit."""
    print(f"acor at WGS {wgs}", flush=True)
    fac = engine.HipBayaderaFactory(0, wgs)
    rng = np.random.default_rng(wgs)
    x = np.zeros((5000, 3), np.float32)
    for t in range(1, 5000):  # AR(1) series, tau ~ 19
        x[t] = 0.9 * x[t - 1] + rng.standard_normal(3).astype(np.float32)
    got = fac.acor_engine().acor(x)
    want = O.acor(x, wgs)
    assert got.lag == want.lag
    assert np.array_equal(got.tau, want.tau) and np.array_equal(got.sigma, want.sigma)
    assert np.array_equal(got.mean, want.mean)
    fac.release()


def test_run_sampler_wgs_1024_many_groups():
    """WGS 1024 with 1024 walkers per accounting group and 512 groups per
    half: the per-step means reduce 512 group sums per tree (more than the
    256-wide default), and the uniform model's accounting equals the
    oracle's bit for bit."""
    W, wgs = 2 * 1024 * 512, 1024
    print(f"run-sampler at WGS {wgs}, {W} walkers", flush=True)
    fac = engine.HipBayaderaFactory(0, wgs)
    sf = fac.mcmc_factory(models.uniform_model())
    s = sf.create_sampler(3, W, [-1, 2]).set_exec(STEPWISE)
    s.init(3)
    s.init_position(4, [-1, 2])
    res = s.run_sampler(60, 2.0)
    ref = O.Stretch(O.UNIFORM_LOGPDF, 1, W, [-1, 2], 3, params_len=2, wgs=wgs, threads=8)
    ref.init(3)
    ref.init_position(4, [-1, 2])
    rr = ref.run_sampler(60, 2.0)
    assert np.array_equal(res["means"], rr["means"])
    assert res["acceptance-rate"] == rr["acceptance_rate"]
    a, b = res["autocorrelation"], rr["autocorrelation"]
    assert np.array_equal(a.tau, b.tau) and np.array_equal(a.sigma, b.sigma)
    assert np.array_equal(s.state()[0].reshape(-1), ref.xs)
    s.release()
    sf.release()
    fac.release()
