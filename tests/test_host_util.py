"""HDI utilities (util.clj) and the MCMC API layer (mcmc.clj) on the host."""
import numpy as np
import pytest

from bayadera_amd import mcmc, util
from bayadera_amd.engine import Histogram

# util_test.clj:21-38
TPDF = np.array([0.02, 0.01, 0.04, 0.05, 0.07, 0.01, 0.2, 0.1, 0.3, 0.1, 0.02] + [0.01] * 68)
TRANK = np.array([8, 6, 7, 9, 4, 3, 2, 0, 10, 1, 5] + list(range(11, 79)), np.float64)


def test_hdi_rank_count_known_answers():
    asum = TPDF.sum()
    assert util.hdi_rank_count(TRANK, TPDF, 0.1) == 1
    assert util.hdi_rank_count(TRANK, TPDF, 0.3 / asum) == 1
    assert util.hdi_rank_count(TRANK, TPDF, 0.3) == 2
    assert util.hdi_rank_count(TRANK, TPDF, 0.7 / asum) == 4
    assert util.hdi_rank_count(TRANK, TPDF, 0.86 / asum) == 7
    assert util.hdi_rank_count(TRANK, TPDF, 0.9 / asum) == 9
    assert util.hdi_rank_count(TRANK, TPDF, 0.95 / asum) == 14


def test_hdi_bins_known_answers():
    assert util.hdi_bins(TRANK, 1) == [8.0, 8.0]
    assert util.hdi_bins(TRANK, 2) == [6.0, 6.0, 8.0, 8.0]
    assert util.hdi_bins(TRANK, 5) == [4.0, 4.0, 6.0, 9.0]
    assert util.hdi_bins(TRANK, 6) == [3.0, 4.0, 6.0, 9.0]
    assert util.hdi_bins(TRANK, 12) == [0.0, 11.0]
    assert util.hdi_bins(TRANK, 15) == [0.0, 14.0]


def test_hdi_regions_known_answers():
    # util_test.clj:50-58 with limits [1, 7] over 79 bins
    lim = np.array([1.0, 7.0])
    cases = [(1, [[1.61], [1.68]], 0.005), (2, [[1.46, 1.61], [1.53, 1.68]], 0.007),
             (5, [[1.30, 1.46], [1.38, 1.76]], 0.006), (6, [[1.23, 1.46], [1.38, 1.76]], 0.006),
             (12, [[1.00], [1.91]], 0.002)]
    for cnt, expect, tol in cases:
        got = util.hdi_regions(lim, TRANK, cnt)
        assert np.linalg.norm(got - np.array(expect)) < tol


def test_hdi_on_histogram():
    pdf = np.zeros((1, 8))
    pdf[0] = [0, 1, 4, 6, 5, 1, 0, 0]
    ranks = np.argsort(-pdf[0], kind="stable").astype(np.float32)[None]
    h = Histogram(np.array([[0.0, 8.0]]), pdf, ranks)
    # 0.9 of the mass 17 needs bins 3, 4, 2, 1 (6+5+4+1 = 16 >= 15.3): [1, 5)
    r = util.hdi(h, 0, 0.9)
    assert r.shape == (2, 1) and r[0, 0] == 1.0 and r[1, 0] == 5.0


class _FakeSampler:
    """Records the calls mcmc.py makes (host control flow only)."""

    def __init__(self, rates):
        self.calls, self.rates = [], list(rates)

    def model(self):
        return type("M", (), {"dimension": 4})()

    def anneal(self, schedule, n, a):
        self.calls.append(("anneal", n, a, schedule(0), schedule(n - 1)))

    def burn_in(self, n, a):
        self.calls.append(("burn_in", n, a))

    def run_sampler(self, n, a):
        self.calls.append(("run_sampler", n, a))

    def acc_rate(self, a):
        self.calls.append(("acc_rate", a))
        return self.rates.pop(0) if self.rates else 0.35


def test_defaults_and_schedules():
    s = _FakeSampler([])
    mcmc.burn_in(s)
    mcmc.run_sampler(s)
    mcmc.anneal(s)
    assert s.calls[0] == ("burn_in", 1024, 2.0)
    assert s.calls[1] == ("run_sampler", 256, 2.0)
    assert s.calls[2] == ("anneal", 1024, 2.0, 1024.0, 1.0)
    assert mcmc.sqrt_n(9.0)(5) == 2.0 and mcmc.pow_n(2.0)(5.0)(2) == 9.0


def test_mix_adjusts_a_like_the_reference():
    # acc 0.1 < 0.2: a <- 1 + (a-1)*0.1/0.35; then 0.6 > 0.5: a <- a*0.6/0.35; then in range
    s = _FakeSampler([0.1, 0.6, 0.3])
    res = mcmc.mix(s)
    a1 = 1 + (2.0 - 1) * (0.1 / 0.35)
    a2 = a1 * (0.6 / 0.35)
    assert res["a"] == pytest.approx(a2)
    n = int(64 * 4 ** 0.8)
    assert s.calls[0][0:3] == ("anneal", n, 2.0)
    assert ("burn_in", n, pytest.approx(a2)) in s.calls and ("burn_in", 64, 2.0) in s.calls
