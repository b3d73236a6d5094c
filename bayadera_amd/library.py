"""The distribution library over the engine: library.clj and
internal/distribution_library.clj.

Each distribution carries its parameter vector in the layout the device
models read (distributions.clj: *-params, with the log normalisers the
`*_logpdf` functions add), its closed-form moments, a density through the
distribution engine, and a sampler: the direct samplers for uniform,
gaussian, exponential and erlang (GCNDirectSamplerEngine), the stretch
sampler for the others and for posteriors, created as the reference's
library does (walkers = processing elements x 256, start box, 128 warm-up
steps at a = 8; distribution_library.clj:95-105).
"""
from __future__ import annotations

import math

import numpy as np

from . import mcmc, models


def _lbeta(a: float, b: float) -> float:
    return math.lgamma(a) + math.lgamma(b) - math.lgamma(a + b)


class Library:
    """A DeviceLibrary over one engine factory (internal/device/models.clj
    ModelFactory/Library): model sources, cached engines and factories."""

    def __init__(self, factory):
        self.factory = factory
        self.models = models.distribution_models()
        self.lik_models = models.likelihood_models()
        self._mcmc, self._dist = {}, {}

    def mcmc_factory(self, model):
        key = model.name
        if key not in self._mcmc:
            self._mcmc[key] = self.factory.mcmc_factory(model)
        return self._mcmc[key]

    def distribution_engine(self, model):
        key = model.name
        if key not in self._dist:
            self._dist[key] = self.factory.distribution_engine(model)
        return self._dist[key]

    def release(self):
        for h in list(self._mcmc.values()) + list(self._dist.values()):
            h.release()
        self._mcmc.clear()
        self._dist.clear()

    # distributions (library.clj)
    def uniform(self, a, b):
        return Uniform(self, a, b)

    def gaussian(self, mu, sigma):
        return Gaussian(self, mu, sigma)

    def student_t(self, nu, mu=0.0, sigma=1.0):
        return StudentT(self, nu, mu, sigma)

    def beta(self, a, b):
        return Beta(self, a, b)

    def gamma(self, theta, k):
        return Gamma(self, theta, k)

    def exponential(self, lam):
        return Exponential(self, lam)

    def erlang(self, lam, k):
        return Erlang(self, lam, k)

    def likelihood(self, name: str):
        return self.lik_models[name]

    def posterior(self, lik, prior: "Distribution", name: str = "posterior"):
        """distribution (core.clj) of a likelihood and a prior: a function of
        the data returning the posterior distribution."""
        model = models.posterior_model(prior.model, name, lik)
        return lambda data: Posterior(self, model, np.asarray(data, np.float32), prior)


class Distribution:
    """A library distribution: model, params, moments, density, sampler."""
    model_name = ""
    direct = False

    def __init__(self, lib: Library):
        self.lib = lib
        self.model = lib.models[self.model_name]

    # Location / Spread
    def sd(self) -> float:
        return math.sqrt(self.variance())

    def log_density(self, x) -> np.ndarray:
        xx = np.asarray(x, np.float32).reshape(-1, self.model.dimension)
        return self.lib.distribution_engine(self.model).log_density(self.params, xx)

    def density(self, x) -> np.ndarray:
        xx = np.asarray(x, np.float32).reshape(-1, self.model.dimension)
        return self.lib.distribution_engine(self.model).density(self.params, xx)

    def start_box(self):
        m, s = self.mean(), self.sd()
        return [m - 10.0 * s, m + 10.0 * s]

    def sampler(self, seed: int = 42, walkers: int | None = None, warm_up: int = 128):
        if self.direct:
            return DirectSampler(self, seed)
        w = walkers or self.lib.factory.processing_elements() * 256
        s = self.lib.mcmc_factory(self.model).create_sampler(seed, w, self.params)
        s.init_position(seed + 1, np.asarray(self.start_box(), np.float32))
        s.burn_in(max(0, int(warm_up)), 8.0)
        return s


class DirectSampler:
    """RandomSampler of a distribution with a direct sampler
    (distribution_library.clj, GCNDirectSamplerEngine)."""

    def __init__(self, dist: Distribution, seed: int):
        self.dist, self.seed = dist, seed
        self.engine = dist.lib.factory.direct_sampler_engine(dist.model)

    def sample(self, n: int | None = None):
        n = n or self.dist.lib.factory.processing_elements() * 256
        self.seed += 1
        return self.engine.sample(self.seed, self.dist.direct_params, n).reshape(n, 1)


class Uniform(Distribution):
    model_name, direct = "uniform", True

    def __init__(self, lib, a, b):
        super().__init__(lib)
        self.a, self.b = float(a), float(b)
        self.params = np.array([a, b], np.float32)
        self.direct_params = self.params

    def mean(self):
        return 0.5 * (self.a + self.b)

    def variance(self):
        return (self.b - self.a) ** 2 / 12.0


class Gaussian(Distribution):
    model_name, direct = "gaussian", True

    def __init__(self, lib, mu, sigma):
        super().__init__(lib)
        self.mu, self.sigma = float(mu), float(sigma)
        self.params = np.array([mu, sigma], np.float32)
        self.direct_params = self.params

    def mean(self):
        return self.mu

    def variance(self):
        return self.sigma ** 2


class StudentT(Distribution):
    model_name = "student-t"

    def __init__(self, lib, nu, mu, sigma):
        super().__init__(lib)
        self.nu, self.mu, self.sigma = float(nu), float(mu), float(sigma)
        # distributions.clj:450-470: log scale of the density with sigma
        ls = (math.lgamma(0.5 * (nu + 1)) - 0.5 * math.log(nu) - 0.5 * math.log(math.pi)
              - math.lgamma(0.5 * nu) - math.log(sigma))
        self.params = np.array([nu, mu, sigma, ls], np.float32)

    def mean(self):
        return self.mu

    def variance(self):
        return self.sigma ** 2 * self.nu / (self.nu - 2.0)


class Beta(Distribution):
    model_name = "beta"

    def __init__(self, lib, a, b):
        super().__init__(lib)
        self.a, self.b = float(a), float(b)
        self.params = np.array([a, b, -_lbeta(a, b)], np.float32)

    def mean(self):
        return self.a / (self.a + self.b)

    def variance(self):
        s = self.a + self.b
        return self.a * self.b / (s * s * (s + 1.0))

    def start_box(self):
        return [0.0, 1.0]


class Gamma(Distribution):
    model_name = "gamma"

    def __init__(self, lib, theta, k):
        super().__init__(lib)
        self.theta, self.k = float(theta), float(k)
        self.params = np.array([theta, k, -(math.lgamma(k) + k * math.log(theta))], np.float32)

    def mean(self):
        return self.k * self.theta

    def variance(self):
        return self.k * self.theta ** 2

    def start_box(self):
        return [0.0, self.mean() + 10.0 * self.sd()]


class Exponential(Distribution):
    model_name, direct = "exponential", True

    def __init__(self, lib, lam):
        super().__init__(lib)
        self.lam = float(lam)
        self.params = np.array([lam, math.log(lam)], np.float32)
        self.direct_params = np.array([lam], np.float32)

    def mean(self):
        return 1.0 / self.lam

    def variance(self):
        return 1.0 / self.lam ** 2


class Erlang(Distribution):
    model_name, direct = "erlang", True

    def __init__(self, lib, lam, k):
        super().__init__(lib)
        self.lam, self.k = float(lam), int(k)
        self.params = np.array([lam, k, k * math.log(lam) - math.lgamma(k)], np.float32)
        self.direct_params = np.array([lam, k], np.float32)

    def mean(self):
        return self.k / self.lam

    def variance(self):
        return self.k / self.lam ** 2


class Posterior(Distribution):
    """The posterior of a likelihood and a prior for given data (params =
    [data, prior params]); sampled by the stretch sampler after mix!."""

    def __init__(self, lib, model, data, prior):
        self.lib, self.model, self.prior = lib, model, prior
        self.params = np.concatenate([data, prior.params]).astype(np.float32)

    def start_box(self):
        return self.prior.start_box()

    def sampler(self, seed: int = 42, walkers: int | None = None, warm_up: int = 128):
        w = walkers or self.lib.factory.processing_elements() * 256
        s = self.lib.mcmc_factory(self.model).create_sampler(seed, w, self.params)
        s.init_position(seed + 1, np.asarray(self.start_box(), np.float32))
        mcmc.mix(s)
        return s


def beta_bernoulli_evidence(a: float, b: float, n: int, z: int) -> float:
    """exp(log B(a + z, b + n - z) - log B(a, b)): the evidence of z
    successes in n trials under a Beta(a, b) prior (the unnormalised
    binomial likelihood, as binomial_loglik)."""
    return math.exp(_lbeta(a + z, b + n - z) - _lbeta(a, b))
