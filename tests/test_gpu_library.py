"""The distribution library end to end on the GPU: the reference's
library_test.clj:24-161 (same distributions, parameters and tolerances:
`roughly100` is 1% relative), through the library's model sources, the
direct samplers, the stretch sampler, the distribution engine, the dataset
engine and the likelihood engine's evidence."""
import math

import numpy as np
import pytest

import oracle as O
from bayadera_amd import engine, library, mcmc, models

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    f = engine.HipBayaderaFactory(0, 256)
    L = library.Library(f)
    yield L
    L.release()
    f.release()


def _moments(lib, sample):
    de = lib.factory.dataset_engine()
    return float(de.data_mean(sample)[0]), float(de.data_variance(sample)[0])


def roughly100(x, ref):
    return abs(x - ref) <= abs(ref) / 100.0


def test_uniform(lib):
    d = lib.uniform(100.0, 133.4)
    s = d.sampler(7).sample()
    m, v = _moments(lib, s)
    assert roughly100(m, d.mean()) and roughly100(math.sqrt(v), d.sd())
    pdf = d.density(s)
    assert abs(pdf.mean() - 1.0 / 33.4) < 1e-3 / 33.4


def test_gaussian(lib):
    d = lib.gaussian(200.0, 10.0)
    s = d.sampler(8).sample()
    m, v = _moments(lib, s)
    assert roughly100(m, 200.0) and abs(v - 100.0) <= 2.0 and roughly100(math.sqrt(v), 10.0)


def test_student_t(lib):
    d = lib.student_t(30, 0.0, 1.0)
    s = d.sampler(9)
    x = s.sample()
    m, v = _moments(lib, x)
    assert abs(m - d.mean()) <= 0.01 and roughly100(math.sqrt(v), d.sd())
    assert abs(float(s.mean()[0]) - d.mean()) <= 0.01
    s.release()


def test_beta(lib):
    d = lib.beta(2.0, 5.0)
    s = d.sampler(10)
    x = s.sample()
    m, v = _moments(lib, x)
    assert roughly100(m, d.mean()) and roughly100(math.sqrt(v), d.sd())
    assert roughly100(float(s.mean()[0]), d.mean())
    # the engine's density against the closed form on the samples
    xs = x[:20000, 0].astype(np.float64)
    ref = np.exp((d.a - 1) * np.log(xs) + (d.b - 1) * np.log1p(-xs) - library._lbeta(d.a, d.b))
    assert abs(d.density(x[:20000]).sum() / ref.sum() - 1.0) < 1e-3
    s.release()


def test_gamma(lib):
    d = lib.gamma(2.0, 5.0)
    s = d.sampler(11)
    x = s.sample()
    m, v = _moments(lib, x)
    assert roughly100(m, d.mean()) and roughly100(math.sqrt(v), d.sd())
    s.release()


def test_exponential(lib):
    d = lib.exponential(3.5)
    m, v = _moments(lib, d.sampler(12).sample())
    assert abs(m - 1 / 3.5) <= 0.05 and abs(v - d.variance()) <= 0.05
    assert abs(math.sqrt(v) - d.sd()) <= 0.03


def test_erlang(lib):
    d = lib.erlang(3.5, 9)
    m, v = _moments(lib, d.sampler(13).sample())
    assert roughly100(m, 9 / 3.5) and abs(v - 0.75) <= 0.04 and abs(math.sqrt(v) - 0.866) <= 0.04


def test_beta_bernoulli(lib):
    a, b, z, N = 8.5, 3.5, 6, 9
    prior = lib.beta(a, b)
    ps = prior.sampler(14)
    prior_sample = ps.sample()
    lik = lib.likelihood("binomial")
    data = np.array([N, z], np.float32)
    le = lib.factory.likelihood_engine(lik)
    ev = le.evidence(data, prior_sample)
    assert abs(float(np.mean(prior_sample)) - prior.mean()) < 1e-3
    assert abs(ev / library.beta_bernoulli_evidence(a, b, N, z) - 1.0) < 1e-3
    post = lib.posterior(lik, prior)(data)
    s = post.sampler(15)
    # (doto (sampler post-dist) (mix!)), library_test.clj:135
    mixed = mcmc.mix(s)
    assert 0.2 <= mixed["acc-rate"] <= 0.5 or mixed["a"] != 2.0
    x = s.sample()
    real = lib.beta(a + z, b + N - z)
    m, v = _moments(lib, x)
    assert roughly100(m, real.mean()) and roughly100(math.sqrt(v), real.sd())
    # the normalised posterior density equals the conjugate beta's
    pdf = np.asarray(post.density(x[:20000]), np.float64) / ev
    rsamp = real.sampler(16)
    rs = rsamp.sample()
    rpdf = np.asarray(real.density(rs[:20000]), np.float64)
    assert roughly100(pdf.mean(), rpdf.mean())
    for h in (s, ps, rsamp, le):
        h.release()


class _OracleSampler:
    """The oracle's stretch sampler with the model() the mix! host loop reads."""

    def __init__(self, stretch, dimension):
        self.s, self.dimension = stretch, dimension

    def model(self):
        return self

    def anneal(self, schedule, n, a):
        return self.s.anneal(schedule, n, a)

    def acc_rate(self, a):
        return self.s.acc_rate(a)

    def burn_in(self, n, a):
        return self.s.burn_in(n, a)


def test_mix_uniform_hip_equals_oracle(lib):
    """mix! (mcmc.clj:66-101) on the uniform model through the HIP sampler
    and through the oracle: the uniform model's accept decisions are exact,
    so every acceptance rate, the tuned a and the final ensemble agree."""
    W, seed = 4096, 21
    sf = lib.factory.mcmc_factory(models.uniform_model())
    s = sf.create_sampler(seed, W, [-1, 2])
    s.init(seed)
    s.init_position(seed + 1, [-1, 2])
    got = mcmc.mix(s)
    ref = O.Stretch(O.UNIFORM_LOGPDF, 1, W, [-1, 2], seed, params_len=2, threads=4)
    ref.init(seed)
    ref.init_position(seed + 1, [-1, 2])
    want = mcmc.mix(_OracleSampler(ref, 1))
    assert got == want
    assert got["a"] != 2.0  # the loop moved a
    assert np.array_equal(s.state()[0].reshape(-1), ref.xs)
    s.release()
    sf.release()
