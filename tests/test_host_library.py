"""The library's parameter vectors and closed-form moments against scipy."""
import math

import numpy as np
import pytest
from scipy import stats

from bayadera_amd import library


@pytest.fixture(scope="module")
def lib():
    return library.Library(factory=None)


def test_log_normalisers_match_scipy(lib):
    t = lib.student_t(30.0, 0.5, 2.0)
    x = 1.3
    unscaled = -0.5 * (t.nu + 1) * math.log1p(((x - t.mu) / t.sigma) ** 2 / t.nu)
    assert abs(unscaled + t.params[3] - stats.t.logpdf(x, 30.0, 0.5, 2.0)) < 1e-5
    b = lib.beta(2.0, 5.0)
    assert abs(math.log(0.3) * 1 + math.log(0.7) * 4 + b.params[2]
               - stats.beta.logpdf(0.3, 2.0, 5.0)) < 1e-5
    g = lib.gamma(2.0, 5.0)
    assert abs(4 * math.log(3.0) - 3.0 / 2.0 + g.params[2]
               - stats.gamma.logpdf(3.0, 5.0, scale=2.0)) < 1e-5
    e = lib.erlang(3.5, 9)
    assert abs(8 * math.log(2.0) - 3.5 * 2.0 + e.params[2]
               - stats.gamma.logpdf(2.0, 9, scale=1 / 3.5)) < 1e-5
    ex = lib.exponential(3.5)
    assert abs(-3.5 * 0.4 + ex.params[1] - stats.expon.logpdf(0.4, scale=1 / 3.5)) < 1e-6


def test_moments_match_scipy(lib):
    cases = [(lib.uniform(100.0, 133.4), stats.uniform(100.0, 33.4)),
             (lib.gaussian(200.0, 10.0), stats.norm(200.0, 10.0)),
             (lib.student_t(30.0, 0.0, 1.0), stats.t(30.0)),
             (lib.beta(2.0, 5.0), stats.beta(2.0, 5.0)),
             (lib.gamma(2.0, 5.0), stats.gamma(5.0, scale=2.0)),
             (lib.exponential(3.5), stats.expon(scale=1 / 3.5)),
             (lib.erlang(3.5, 9), stats.gamma(9, scale=1 / 3.5))]
    for d, ref in cases:
        assert np.isclose(d.mean(), ref.mean()) and np.isclose(d.variance(), ref.var())


def test_beta_bernoulli_evidence_brute_force():
    a, b, n, z = 8.5, 3.5, 9, 6
    p = np.linspace(0, 1, 200001)
    f = p ** z * (1 - p) ** (n - z) * stats.beta.pdf(p, a, b)
    assert abs(np.trapezoid(f, p) / library.beta_bernoulli_evidence(a, b, n, z) - 1) < 1e-6
