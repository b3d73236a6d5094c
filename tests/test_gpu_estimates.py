"""Dataset engine, autocorrelation and density/likelihood engines on the GPU."""
import math

import numpy as np
import pytest

import oracle as O
from bayadera_amd import _lib, engine, models
from conftest import load_acor_fixture

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fac():
    f = engine.HipBayaderaFactory(0, 256)
    yield f
    f.release()


@pytest.mark.parametrize("m,n", [(1, 1), (1, 1000), (3, 4097), (22, 65536), (70, 3001)])
def test_dataset_histogram_exact_vs_oracle(fac, m, n):
    rng = np.random.default_rng(m * 1000 + n)
    a = rng.standard_normal((n, m)).astype(np.float32) * np.arange(1, m + 1, dtype=np.float32)
    h = fac.dataset_engine().histogram(a)
    hr = O.histogram(a)
    assert np.array_equal(h.limits, hr.limits)
    # n = 1 gives hi == lo: pdf = count * alpha / 0 (inf and nan), as the reference
    assert np.array_equal(h.pdf, hr.pdf, equal_nan=True)
    assert np.array_equal(h.bin_ranks, hr.bin_ranks)


@pytest.mark.parametrize("m,n", [(1, 1), (4, 1000), (22, 200000), (70, 3001)])
def test_dataset_moments_vs_oracle(fac, m, n):
    rng = np.random.default_rng(m + n)
    a = (rng.standard_normal((n, m)) * 3 + 100).astype(np.float32)
    mean, var = O.mean_var_f64(a)
    de = fac.dataset_engine()
    assert np.allclose(de.data_mean(a), mean.astype(np.float32), rtol=2e-7, atol=0)
    assert np.allclose(de.data_variance(a), var.astype(np.float32), rtol=1e-6, atol=0)


def test_dataset_like_core_test(fac):
    # core_test.clj:18-30: 22 x 2,031,616 uniform(0,1); histogram pdf
    # averages 1, mean and variance within 0.003 of a host computation
    rng = np.random.default_rng(22)
    n = 31 * 2 ** 16
    a = rng.random((n, 22), dtype=np.float32)
    de = fac.dataset_engine()
    h = de.histogram(a)
    assert float(h.pdf[4].mean()) == pytest.approx(1.0, abs=0.01)
    assert abs(float(np.sum(de.data_variance(a) - a.var(0, dtype=np.float64)))) < 0.003
    assert abs(float(np.sum(de.data_mean(a) - a.mean(0, dtype=np.float64)))) < 0.003
    counts = np.rint(h.pdf * (h.limits[:, 1:] - h.limits[:, :1]) * n / 256).astype(np.int64)
    assert np.all(counts.sum(1) == n)


def test_constant_data_histogram(fac):
    a = np.full((1000, 2), 3.0, np.float32)
    h = fac.dataset_engine().histogram(a)
    hr = O.histogram(a)
    assert np.array_equal(h.limits, hr.limits)
    assert np.array_equal(np.isinf(h.pdf), np.isinf(hr.pdf))


@pytest.mark.parametrize("n", [67, 367, 112640])
def test_acor_fixtures_exact_vs_oracle(fac, goldens, n):
    series = load_acor_fixture(n)
    a = fac.acor_engine().acor(series)
    r = O.acor(series)
    assert np.array_equal(a.tau, r.tau)
    assert np.array_equal(a.sigma, r.sigma)
    assert np.array_equal(a.mean, r.mean)
    g = goldens["G9_acor"][str(n)]
    if "tau" in g:
        assert a.tau.tolist() == g["tau"]


def test_acor_random_exact_vs_oracle(fac):
    rng = np.random.default_rng(9)
    x = np.cumsum(rng.standard_normal((5000, 5)), axis=0).astype(np.float32)
    a = fac.acor_engine().acor(x)
    r = O.acor(x)
    assert np.array_equal(a.tau, r.tau) and np.array_equal(a.sigma, r.sigma)


def test_acor_too_short(fac):
    with pytest.raises(_lib.ArgumentError, match="must not be less than 50"):
        fac.acor_engine().acor(np.ones((49, 1), np.float32))


def _gaussian_log_pdf(mu, sigma, x):
    return (x - mu) ** 2 / (-2.0 * sigma * sigma) - math.log(sigma) - 0.5 * math.log(2 * math.pi)


def test_gaussian_density_golden(fac):
    # amd_gcn_test.clj:107-129
    e = fac.distribution_engine(models.gaussian_model())
    x = np.arange(-10, 10, 0.1)[:200].astype(np.float32).reshape(200, 1)
    ref = np.array([_gaussian_log_pdf(0.0, 1.0, float(v)) for v in x[:, 0]])
    assert np.linalg.norm(e.log_density([0.0, 1.0], x) - ref) < 1e-4
    assert np.linalg.norm(e.density([0.0, 1.0], x) - np.exp(ref)) < 1e-5


def _lbeta(a, b):
    return math.lgamma(a) + math.lgamma(b) - math.lgamma(a + b)


def test_beta_binomial_posterior_and_evidence(fac, goldens):
    # amd_gcn_test.clj:131-157
    g = goldens["G10_density"]["beta_binomial"]
    n, z, a, b = g["n"], g["z"], g["a"], g["b"]
    beta = models.beta_model()
    lik = models.likelihood_models()["binomial"]
    post = fac.distribution_engine(beta.posterior_model("beta_binomial", lik))
    beng = fac.distribution_engine(beta)
    leng = fac.likelihood_engine(lik)
    x = (np.arange(1, 1000) * 0.001)[:200].astype(np.float32).reshape(200, 1)
    params = np.array([n, z, a, b, -_lbeta(a, b)], np.float32)
    bparams = np.array([a + z, b + n - z, -_lbeta(a + z, b + n - z)], np.float32)
    diff = post.log_density(params, x).astype(np.float64) - beng.log_density(bparams, x)
    assert np.linalg.norm(diff - g["posterior_minus_beta_logpdf"]) < 2e-4
    ev = leng.evidence(np.array([n, z], np.float32), x)
    assert ev == pytest.approx(g["evidence"], rel=1e-4)
    ll = leng.log_density(np.array([n, z], np.float32), x)
    assert np.allclose(ll, O.logfn_many(O.BINOMIAL_LOGLIK, [n, z], x, 1), rtol=1e-5, atol=1e-4)


def test_density_ragged_sizes(fac):
    e = fac.distribution_engine(models.gaussian_model())
    for n in (1, 255, 257, 100001):
        x = np.linspace(-3, 3, n, dtype=np.float32).reshape(n, 1)
        lp = e.log_density([0.5, 2.0], x)
        ref = O.logfn_many(O.GAUSSIAN_LOGPDF, [0.5, 2.0], x, 1)
        assert lp.shape == (n,)
        assert np.allclose(lp, ref, rtol=1e-5, atol=1e-5)


def test_rtc_error_surfaces(fac):
    bad = models.distribution_model("inline REAL nope(const uint a) { return q; }", name="nope")
    with pytest.raises(_lib.BayaderaError, match="failed to compile"):
        fac.distribution_engine(bad)


def test_device_copy_rate(fac):
    """The engine's streaming copy measures a plausible HBM rate both ways
    (read + written bytes per second of a 256 MiB copy)."""
    r = fac.device_copy_rate(256 << 20)
    assert 500.0 < r["copy_gbps"] < 20000.0
    assert 500.0 < r["copy_nt_gbps"] < 20000.0
    with pytest.raises(_lib.ArgumentError):
        fac.device_copy_rate(1000)
