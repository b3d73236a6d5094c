"""Pins the oracle's model log-densities and the evidence estimate."""
import math

import numpy as np
import pytest

import oracle as O


def _gaussian_log_pdf(mu, sigma, x):
    # distributions.clj:416-419 (host fp64 closed form)
    return (x - mu) ** 2 / (-2.0 * sigma * sigma) - math.log(sigma) - 0.5 * math.log(2 * math.pi)


def test_gaussian_logpdf_closed_form():
    x = np.arange(-10, 10, 0.1)[:200].astype(np.float32)
    lp = O.logfn_many(O.GAUSSIAN_LOGPDF, [0.0, 1.0], x, 1)
    ref = np.array([_gaussian_log_pdf(0.0, 1.0, float(v)) for v in x])
    assert np.linalg.norm(lp - ref) < 1e-4
    assert np.linalg.norm(np.exp(lp.astype(np.float64)) - np.exp(ref)) < 1e-5


def _lbeta(a, b):
    return math.lgamma(a) + math.lgamma(b) - math.lgamma(a + b)


def test_beta_binomial_posterior_constant(goldens):
    g = goldens["G10_density"]["beta_binomial"]
    n, z, a, b = g["n"], g["z"], g["a"], g["b"]
    x = (np.arange(1, 1000) * 0.001)[:200].astype(np.float32)
    params = np.array([n, z, a, b, -_lbeta(a, b)], np.float32)
    post = O.logfn_many(O.BETA_BINOMIAL_POST, params, x, 1, data_len=2, params_len=3)
    beta_params = np.array([a + z, b + n - z, -_lbeta(a + z, b + n - z)], np.float32)
    beta = O.logfn_many(O.BETA_LOGPDF, beta_params, x, 1, params_len=3)
    diff = post.astype(np.float64) - beta
    assert np.linalg.norm(diff - g["posterior_minus_beta_logpdf"]) < 2e-4
    # the constant itself is lbeta(a+z, b+n-z) - lbeta(a, b)
    assert _lbeta(a + z, b + n - z) - _lbeta(a, b) == pytest.approx(
        g["posterior_minus_beta_logpdf"], abs=1e-5)


def test_evidence_brute_force(goldens):
    g = goldens["G10_density"]["beta_binomial"]
    x = (np.arange(1, 1000) * 0.001)[:200].astype(np.float32)
    ll = O.logfn_many(O.BINOMIAL_LOGLIK, [g["n"], g["z"]], x, 1)
    ev = np.exp(ll.astype(np.float32)).astype(np.float64).mean()
    assert ev == pytest.approx(g["evidence"], rel=1e-5)
    exact = np.mean([v ** 15 * (1 - v) ** 35 for v in x.astype(np.float64)])
    assert ev == pytest.approx(exact, rel=1e-4)
