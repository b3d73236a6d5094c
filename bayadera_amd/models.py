"""Device models: C99 model sources and their metadata.

Mirrors ``src/clojure/uncomplicate/bayadera/internal/device/models.clj``:
a distribution model names its log-density functions, dimension, parameter
count and default limits (``DeviceDistributionModel``, models.clj:53-100); a
likelihood model names its log-likelihood (models.clj:25-51); a posterior
model concatenates both sources and adds two functions that sum the
likelihood and the prior (``device-posterior-model``, models.clj:102-115,
template ``distributions/posterior.cl``).

Sources keep the reference's plugin dialect so user models written for the
reference compile unchanged:
  logpdf  REAL f(const uint data_len, const uint params_len, const REAL* params,
                 const uint dim, const REAL* x)
  loglik  REAL f(const uint data_len, const REAL* data, const uint dim, const REAL* x)
The library sources below are this engine's own; the function names and
signatures are the reference's plugin interface (distributions/*.cl).
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

FLT_MAX = float(np.finfo(np.float32).max)
FLT_MIN = float(np.finfo(np.float32).tiny)

_gensym = itertools.count(1)


def gensym(prefix: str) -> str:
    return f"{prefix}{next(_gensym)}"


@dataclass
class LikelihoodModel:
    """DeviceLikelihoodModel (models.clj:25-44)."""
    name: str
    source: list[str]
    loglik: str
    dialect: str = "c99"

    def model(self):
        return self


@dataclass
class DistributionModel:
    """DeviceDistributionModel (models.clj:53-89)."""
    name: str
    source: list[str]
    logpdf: str
    mcmc_logpdf: str
    dimension: int = 1
    params_size: int = 1
    limits: np.ndarray | None = None
    sampler_source: list[str] = field(default_factory=list)
    dialect: str = "c99"

    def model(self):
        return self

    def posterior_model(self, name: str, likelihood: LikelihoodModel) -> "DistributionModel":
        return posterior_model(self, name, likelihood)


def likelihood_model(source, name: str | None = None, loglik: str | None = None,
                     dialect: str = "c99") -> LikelihoodModel:
    """device-likelihood-model (models.clj:46-51): defaults name=gensym,
    loglik=<name>_loglik."""
    name = name or gensym("likelihood")
    src = list(source) if isinstance(source, (list, tuple)) else [source]
    return LikelihoodModel(name, src, loglik or f"{name}_loglik", dialect)


def distribution_model(source, name: str | None = None, logpdf: str | None = None,
                       mcmc_logpdf: str | None = None, dimension: int = 1, params_size: int = 1,
                       limits=None, sampler_source=None, dialect: str = "c99") -> DistributionModel:
    """device-distribution-model (models.clj:91-100): defaults name=gensym,
    logpdf=<name>_logpdf, mcmc-logpdf=logpdf, dimension 1, params-size 1."""
    name = name or gensym("distribution")
    logpdf = logpdf or f"{name}_logpdf"
    src = list(source) if isinstance(source, (list, tuple)) else [source]
    lim = None if limits is None else np.asarray(limits, np.float32).reshape(-1, 2)
    ss = [] if sampler_source is None else (
        list(sampler_source) if isinstance(sampler_source, (list, tuple)) else [sampler_source])
    return DistributionModel(name, src, logpdf, mcmc_logpdf or logpdf, dimension, params_size,
                             lim, ss, dialect)


POSTERIOR_TEMPLATE = """
inline REAL {name}(const uint data_len, const uint hyperparams_len, const REAL* params,
                   const uint dim, const REAL* x) {{
    const REAL log_likelihood = {loglik}(data_len, params, dim, x);
    const REAL log_prior = {logpdf}(data_len, hyperparams_len, params + data_len, dim, x);
    return log_likelihood + log_prior;
}}
"""


def posterior_model(prior: DistributionModel, name: str, lik: LikelihoodModel) -> DistributionModel:
    """device-posterior-model (models.clj:102-115): the posterior's logpdf and
    mcmc-logpdf are both loglik(data) + prior logpdf(hyperparams); params are
    [data, hyperparams] with data_len = count(params) - params-size."""
    if prior.dialect != lik.dialect:
        raise ValueError("Incompatible types of likelihood and prior models.")
    post_name = gensym(name)
    post_logpdf = f"{post_name}_logpdf"
    post_mcmc = f"{post_name}_mcmc_logpdf"
    src = []
    for s in prior.source + lik.source:
        if s not in src:
            src.append(s)
    src.append(POSTERIOR_TEMPLATE.format(name=post_logpdf, loglik=lik.loglik, logpdf=prior.logpdf)
               + POSTERIOR_TEMPLATE.format(name=post_mcmc, loglik=lik.loglik, logpdf=prior.logpdf))
    return DistributionModel(post_name, src, post_logpdf, post_mcmc, prior.dimension,
                             prior.params_size,
                             None if prior.limits is None else prior.limits.copy(), [],
                             prior.dialect)


# ---------------------------------------------------------------------------
# Library sources (the reference's library: distributions/*.cl)
# ---------------------------------------------------------------------------
UNIFORM_SOURCE = """
inline REAL uniform_pdf(const REAL lower, const REAL upper, const REAL x) {
    const bool inside = (lower <= x) && (x <= upper);
    return inside ? (1.0f / (upper - lower)) : 0.0f;
}

inline REAL uniform_log(const REAL lower, const REAL upper, const REAL x) {
    return native_log(uniform_pdf(lower, upper, x));
}

inline REAL uniform_logpdf(const uint data_len, const uint params_len, const REAL* params,
                           const uint dim, const REAL* x) {
    return uniform_log(params[0], params[1], x[0]);
}
"""

GAUSSIAN_SOURCE = """
#ifndef BH_LOG_SQRT_2PI
#define BH_LOG_SQRT_2PI 0.9189385332046727f
#endif

inline REAL gaussian_log_unscaled(const REAL mu, const REAL sigma, const REAL x) {
    const REAL d = x - mu;
    return d * d / (-2.0f * sigma * sigma);
}

inline REAL gaussian_log_scale(const REAL sigma) {
    return -native_log(sigma) - BH_LOG_SQRT_2PI;
}

inline REAL gaussian_log(const REAL mu, const REAL sigma, const REAL x) {
    return gaussian_log_unscaled(mu, sigma, x) + gaussian_log_scale(sigma);
}

inline REAL gaussian_mcmc_logpdf(const uint data_len, const uint params_len, const REAL* params,
                                 const uint dim, const REAL* x) {
    return gaussian_log_unscaled(params[0], params[1], x[0]);
}

inline REAL gaussian_logpdf(const uint data_len, const uint params_len, const REAL* params,
                            const uint dim, const REAL* x) {
    return gaussian_log(params[0], params[1], x[0]);
}

inline REAL gaussian_loglik(const uint data_len, const REAL* data, const uint dim,
                            const REAL* mu_sigma) {
    const REAL mu = mu_sigma[0];
    const REAL sigma = mu_sigma[1];
    if (!(0.0f < sigma)) {
        return NAN;
    }
    REAL acc = gaussian_log_unscaled(mu, sigma, data[0]) + data_len * gaussian_log_scale(sigma);
    for (uint i = 1; i < data_len; i++) {
        acc += gaussian_log_unscaled(mu, sigma, data[i]);
    }
    return acc;
}
"""

STUDENT_T_SOURCE = """
#ifndef BH_LOG_SQRT_PI
#define BH_LOG_SQRT_PI 0.5723649429247001f
#endif

inline REAL student_t_log_unscaled(const REAL nu, const REAL mu, const REAL sigma, const REAL x) {
    const REAL z = (x - mu) / sigma;
    return -0.5f * (nu + 1.0f) * native_log(1.0f + z * z / nu);
}

inline REAL student_t_log_scale(const REAL nu, const REAL sigma) {
    return lgamma(0.5f * (nu + 1.0f)) - lgamma(0.5f * nu) - BH_LOG_SQRT_PI
        - 0.5f * native_log(nu) - native_log(sigma);
}

inline REAL student_t_log(const REAL nu, const REAL mu, const REAL sigma, const REAL x) {
    return student_t_log_unscaled(nu, mu, sigma, x) + student_t_log_scale(nu, sigma);
}

inline REAL student_t_mcmc_logpdf(const uint data_len, const uint params_len, const REAL* params,
                                  const uint dim, const REAL* x) {
    return student_t_log_unscaled(params[0], params[1], params[2], x[0]);
}

inline REAL student_t_logpdf(const uint data_len, const uint params_len, const REAL* params,
                             const uint dim, const REAL* x) {
    return student_t_log_unscaled(params[0], params[1], params[2], x[0]) + params[3];
}

inline REAL student_t_loglik(const uint data_len, const REAL* data, const uint dim,
                             const REAL* nu_mu_sigma) {
    const REAL nu = nu_mu_sigma[0];
    const REAL mu = nu_mu_sigma[1];
    const REAL sigma = nu_mu_sigma[2];
    if (!((0.0f < nu) && (0.0f < sigma))) {
        return NAN;
    }
    const REAL scale = student_t_log_scale(nu, sigma);
    REAL acc = 0.0f;
    for (uint i = 0; i < data_len; i++) {
        acc += student_t_log_unscaled(nu, mu, sigma, data[i]) + scale;
    }
    return acc;
}
"""

BETA_SOURCE = """
inline REAL lbeta(const REAL a, const REAL b) {
    return lgamma(a) + lgamma(b) - lgamma(a + b);
}

inline REAL beta_log_unscaled(const REAL a, const REAL b, const REAL x) {
    return (a - 1.0f) * native_log(x) + (b - 1.0f) * native_log(1.0f - x);
}

inline REAL beta_log(const REAL a, const REAL b, const REAL x) {
    return beta_log_unscaled(a, b, x) - lbeta(a, b);
}

inline REAL beta_mcmc_logpdf(const uint data_len, const uint params_len, const REAL* params,
                             const uint dim, const REAL* x) {
    return beta_log_unscaled(params[0], params[1], x[0]);
}

inline REAL beta_logpdf(const uint data_len, const uint params_len, const REAL* params,
                        const uint dim, const REAL* x) {
    return beta_log_unscaled(params[0], params[1], x[0]) + params[2];
}
"""

BINOMIAL_SOURCE = """
inline REAL lbinco(const REAL n, const REAL k) {
    return lgamma(n + 1.0f) - lgamma(k + 1.0f) - lgamma(n - k + 1.0f);
}

inline REAL binomial_log_unscaled(const REAL n, const REAL p, const REAL k) {
    return k * native_log(p) + (n - k) * native_log(1.0f - p);
}

inline REAL binomial_log(const REAL n, const REAL p, const REAL k) {
    return binomial_log_unscaled(n, p, k) + lbinco(n, k);
}

inline REAL binomial_mcmc_logpdf(const uint data_len, const uint params_len, const REAL* params,
                                 const uint dim, const REAL* k) {
    return binomial_log_unscaled(params[0], params[1], k[0]);
}

inline REAL binomial_logpdf(const uint data_len, const uint params_len, const REAL* params,
                            const uint dim, const REAL* k) {
    return binomial_log(params[0], params[1], k[0]);
}

inline REAL binomial_loglik(const uint data_len, const REAL* data, const uint dim, const REAL* p) {
    return binomial_log_unscaled(data[0], p[0], data[1]);
}
"""

EXPONENTIAL_SOURCE = """
inline REAL exponential_log_unscaled(const REAL lambda, const REAL x) {
    return -lambda * x;
}

inline REAL exponential_log(const REAL lambda, const REAL x) {
    return native_log(lambda) - lambda * x;
}

inline REAL exponential_mcmc_logpdf(const uint data_len, const uint params_len,
                                    const REAL* params, const uint dim, const REAL* x) {
    return (0.0f < x[0]) ? exponential_log_unscaled(params[0], x[0]) : NAN;
}

inline REAL exponential_logpdf(const uint data_len, const uint params_len, const REAL* params,
                               const uint dim, const REAL* x) {
    return (0.0f < x[0]) ? exponential_log_unscaled(params[0], x[0]) + params[1] : NAN;
}
"""

ERLANG_SOURCE = """
inline REAL erlang_log_unscaled(const REAL lambda, const REAL k, const REAL x) {
    return (k - 1.0f) * native_log(x) - lambda * x;
}

inline REAL erlang_log_scale(const REAL lambda, const REAL k) {
    return k * native_log(lambda) - lgamma(k);
}

inline REAL erlang_log(const REAL lambda, const REAL k, const REAL x) {
    return erlang_log_unscaled(lambda, k, x) + erlang_log_scale(lambda, k);
}

inline REAL erlang_mcmc_logpdf(const uint data_len, const uint params_len, const REAL* params,
                               const uint dim, const REAL* x) {
    return erlang_log_unscaled(params[0], params[1], x[0]);
}

inline REAL erlang_logpdf(const uint data_len, const uint params_len, const REAL* params,
                          const uint dim, const REAL* x) {
    return erlang_log_unscaled(params[0], params[1], x[0]) + params[2];
}
"""

GAMMA_SOURCE = """
inline REAL gamma_log_unscaled(const REAL theta, const REAL k, const REAL x) {
    return (k - 1.0f) * native_log(x) - x / theta;
}

inline REAL gamma_log_scale(const REAL theta, const REAL k) {
    return -lgamma(k) - k * native_log(theta);
}

inline REAL gamma_log(const REAL theta, const REAL k, const REAL x) {
    return gamma_log_unscaled(theta, k, x) + gamma_log_scale(theta, k);
}

inline REAL gamma_mcmc_logpdf(const uint data_len, const uint params_len, const REAL* params,
                              const uint dim, const REAL* x) {
    return gamma_log_unscaled(params[0], params[1], x[0]);
}

inline REAL gamma_logpdf(const uint data_len, const uint params_len, const REAL* params,
                         const uint dim, const REAL* x) {
    return gamma_log_unscaled(params[0], params[1], x[0]) + params[2];
}
"""

# Models of the benchmark workloads (not in the reference's library).
MVN_SOURCE = """
// Correlated Gaussian log-density up to a constant: -(x-mu)' P (x-mu) / 2,
// params = [mu (dim), P (dim x dim, row-major)].
inline REAL mvn_logpdf(const uint data_len, const uint params_len, const REAL* params,
                       const uint dim, const REAL* x) {
    const REAL* mu = params;
    const REAL* P = params + dim;
    REAL q = 0.0f;
    for (uint r = 0; r < dim; r++) {
        REAL row = 0.0f;
        for (uint c = 0; c < dim; c++) {
            row = fma(P[r * dim + c], x[c] - mu[c], row);
        }
        q = fma(x[r] - mu[r], row, q);
    }
    return -0.5f * q;
}
"""

DIAG_GAUSS_SOURCE = """
// Independent Gaussian coordinates up to a constant:
// -sum_i w_i (x_i - mu_i)^2 / 2, params = [mu (dim), w (dim)], w_i = 1/sigma_i^2.
inline REAL diag_gaussian_logpdf(const uint data_len, const uint params_len, const REAL* params,
                                 const uint dim, const REAL* x) {
    const REAL* mu = params;
    const REAL* w = params + dim;
    REAL q = 0.0f;
    for (uint i = 0; i < dim; i++) {
        const REAL d = x[i] - mu[i];
        q = fma(w[i] * d, d, q);
    }
    return -0.5f * q;
}
"""

LINREG_SOURCE = """
// Gaussian linear regression, x = [beta (dim-1), log sigma];
// data rows [x_i (dim-1), y_i], so rows = data_len / dim.
inline REAL linreg_loglik(const uint data_len, const REAL* data, const uint dim, const REAL* x) {
    const uint d = dim - 1;
    const uint rows = data_len / dim;
    const REAL log_sigma = x[d];
    const REAL inv_sigma = native_exp(-log_sigma);
    REAL ss = 0.0f;
    for (uint i = 0; i < rows; i++) {
        const REAL* row = data + i * dim;
        REAL eta = 0.0f;
        for (uint j = 0; j < d; j++) {
            eta = fma(row[j], x[j], eta);
        }
        const REAL r = (row[d] - eta) * inv_sigma;
        ss = fma(r, r, ss);
    }
    return -0.5f * ss - rows * (log_sigma + 0.9189385332046727f);
}

// beta ~ N(0, s_beta^2) iid, log sigma ~ N(0, s_sigma^2); hyperparams [s_beta, s_sigma]
inline REAL linreg_prior_logpdf(const uint data_len, const uint params_len, const REAL* params,
                                const uint dim, const REAL* x) {
    const REAL ib = 1.0f / params[0];
    const REAL is = 1.0f / params[1];
    REAL q = 0.0f;
    for (uint j = 0; j + 1 < dim; j++) {
        const REAL z = x[j] * ib;
        q = fma(z, z, q);
    }
    const REAL zs = x[dim - 1] * is;
    return -0.5f * fma(zs, zs, q);
}
"""

LOGISTIC_SOURCE = """
// Bernoulli-logit regression, x = beta (dim); data rows [x_i (dim), y_i],
// rows = data_len / (dim + 1).  log p(y|eta) = -softplus(-(2y-1) eta).
inline REAL logistic_loglik(const uint data_len, const REAL* data, const uint dim, const REAL* x) {
    const uint rows = data_len / (dim + 1);
    REAL acc = 0.0f;
    for (uint i = 0; i < rows; i++) {
        const REAL* row = data + i * (dim + 1);
        REAL eta = 0.0f;
        for (uint j = 0; j < dim; j++) {
            eta = fma(row[j], x[j], eta);
        }
        const REAL t = (row[dim] > 0.5f) ? -eta : eta;
        // softplus(t) = max(t, 0) + log(1 + exp(-|t|))
        acc -= fmax(t, 0.0f) + native_log(1.0f + native_exp(-fabs(t)));
    }
    return acc;
}

// beta ~ N(0, s^2) iid; hyperparams [s]
inline REAL gaussian_iid_logpdf(const uint data_len, const uint params_len, const REAL* params,
                                const uint dim, const REAL* x) {
    const REAL is = 1.0f / params[0];
    REAL q = 0.0f;
    for (uint j = 0; j < dim; j++) {
        const REAL z = x[j] * is;
        q = fma(z, z, q);
    }
    return -0.5f * q;
}
"""

HBB_SOURCE = """
// Hierarchical beta-binomial: x = [log alpha, log beta, t_1 .. t_G] with
// theta_g = 1 / (1 + exp(-t_g)); data [n_1 .. n_G, y_1 .. y_G], hyperparameter
// [s0]: log alpha, log beta ~ N(0, s0^2).  The logit transform's Jacobian
// theta (1 - theta) is folded into the Beta(alpha, beta) prior's exponents.
inline REAL hbb_softplus(const REAL t) {
    return fmax(t, 0.0f) + native_log(1.0f + native_exp(-fabs(t)));
}

inline REAL hbb_logpdf(const uint data_len, const uint params_len, const REAL* params,
                       const uint dim, const REAL* x) {
    const uint G = dim - 2;
    const REAL* n = params;
    const REAL* y = params + G;
    const REAL s0 = params[data_len];
    const REAL la = x[0];
    const REAL lb = x[1];
    const REAL a = native_exp(la);
    const REAL b = native_exp(lb);
    REAL acc = -(REAL)G * (lgamma(a) + lgamma(b) - lgamma(a + b));
    for (uint g = 0; g < G; g++) {
        const REAL t = x[2 + g];
        acc -= (y[g] + a) * hbb_softplus(-t) + (n[g] - y[g] + b) * hbb_softplus(t);
    }
    return acc - 0.5f * (la * la + lb * lb) / (s0 * s0);
}
"""

_LIM_ALL = [[-FLT_MAX, FLT_MAX]]


def uniform_model() -> DistributionModel:
    # models.clj:146-150
    return distribution_model(UNIFORM_SOURCE, name="uniform", params_size=2, limits=_LIM_ALL)


def gaussian_model() -> DistributionModel:
    # models.clj:152-156
    return distribution_model(GAUSSIAN_SOURCE, name="gaussian", mcmc_logpdf="gaussian_mcmc_logpdf",
                              params_size=2, limits=_LIM_ALL)


def student_t_model() -> DistributionModel:
    # models.clj:158-161
    return distribution_model(STUDENT_T_SOURCE, name="student_t",
                              mcmc_logpdf="student_t_mcmc_logpdf", params_size=4, limits=_LIM_ALL)


def beta_model() -> DistributionModel:
    # models.clj:163-166
    return distribution_model(BETA_SOURCE, name="beta", mcmc_logpdf="beta_mcmc_logpdf",
                              params_size=3, limits=[[0.0, 1.0]])


def exponential_model() -> DistributionModel:
    # models.clj:168-172
    return distribution_model(EXPONENTIAL_SOURCE, name="exponential",
                              mcmc_logpdf="exponential_mcmc_logpdf", params_size=2,
                              limits=[[FLT_MIN, FLT_MAX]])


def erlang_model() -> DistributionModel:
    # models.clj:174-178
    return distribution_model(ERLANG_SOURCE, name="erlang", mcmc_logpdf="erlang_mcmc_logpdf",
                              params_size=3, limits=[[0.0, FLT_MAX]])


def gamma_model() -> DistributionModel:
    # models.clj:180-183
    return distribution_model(GAMMA_SOURCE, name="gamma", mcmc_logpdf="gamma_mcmc_logpdf",
                              params_size=2, limits=[[0.0, FLT_MAX]])


def binomial_model() -> DistributionModel:
    # models.clj:185-188
    return distribution_model(BINOMIAL_SOURCE, name="binomial",
                              mcmc_logpdf="binomial_mcmc_logpdf", params_size=3,
                              limits=[[0.0, FLT_MAX]])


def distribution_models() -> dict:
    """distribution-models (models.clj:190-203)."""
    return {"uniform": uniform_model(), "gaussian": gaussian_model(),
            "student-t": student_t_model(), "beta": beta_model(),
            "exponential": exponential_model(), "erlang": erlang_model(),
            "gamma": gamma_model(), "binomial": binomial_model()}


def likelihood_models() -> dict:
    """likelihood-models (models.clj:205-208)."""
    return {"gaussian": likelihood_model(GAUSSIAN_SOURCE, name="gaussian"),
            "student-t": likelihood_model(STUDENT_T_SOURCE, name="student_t"),
            "binomial": likelihood_model(BINOMIAL_SOURCE, name="binomial")}


@dataclass
class GlmModel:
    """A GLM posterior declared by form, not by source: the engine evaluates
    its likelihood for a whole half-ensemble as one contraction X.B on the
    matrix cores (bh_glm_stretch_factory_create).  ``X`` holds bf16 bit
    patterns (uint16, N x d row-major), ``y`` labels in {0, 1}; the prior is
    beta ~ N(0, prior_sd^2 I)."""
    kind: str
    X: np.ndarray
    y: np.ndarray
    prior_sd: float
    name: str = "glm"
    log_sigma_sd: float = 1.0

    @property
    def dimension(self) -> int:
        # linear regression carries log sigma after the coefficients
        return int(self.X.shape[1]) + (1 if self.kind == "gaussian" else 0)

    @property
    def params_size(self) -> int:
        return 0

    def model(self):
        return self


def glm_logistic_model(X_bf16, y, prior_sd: float = 5.0) -> GlmModel:
    X = np.ascontiguousarray(X_bf16)
    if X.dtype != np.uint16 or X.ndim != 2:
        raise ValueError("X must be an (N, d) array of bf16 bit patterns (uint16)")
    yv = np.ascontiguousarray(y, np.float32).reshape(-1)
    if yv.size != X.shape[0]:
        raise ValueError("y must have one label per row of X")
    return GlmModel("logistic", X, yv, float(prior_sd))


def glm_gaussian_model(X, y, beta_sd: float = 10.0, log_sigma_sd: float = 1.0) -> GlmModel:
    """Linear regression y_i ~ N(x_i . beta, sigma^2) declared by form:
    walkers [beta (d), log sigma], priors beta ~ N(0, beta_sd^2 I) and
    log sigma ~ N(0, log_sigma_sd^2); X fp32 (N, d).  The same posterior as
    the LINREG_SOURCE user model, evaluated as one data-parallel pass per
    half-ensemble."""
    Xf = np.ascontiguousarray(X, np.float32)
    if Xf.ndim != 2 or not 1 <= Xf.shape[1] <= 64:
        raise ValueError("X must be an (N, d) float array with 1 <= d <= 64")
    yv = np.ascontiguousarray(y, np.float32).reshape(-1)
    if yv.size != Xf.shape[0]:
        raise ValueError("y must have one value per row of X")
    return GlmModel("gaussian", Xf, yv, float(beta_sd), log_sigma_sd=float(log_sigma_sd))


def to_bf16(a) -> np.ndarray:
    """float32 -> bf16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32)
    r = ((u >> 16) & 1) + 0x7FFF
    return ((u + r) >> 16).astype(np.uint16)


def from_bf16(b) -> np.ndarray:
    return (np.ascontiguousarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def _fp8_e4m3_table() -> np.ndarray:
    """The 256 e4m3 values (OCP: bias 7, subnormals, no infinities,
    S.1111.111 is NaN)."""
    c = np.arange(256)
    s = np.where(c >= 128, -1.0, 1.0)
    e = (c >> 3) & 15
    m = c & 7
    v = np.where(e == 0, m * 2.0 ** -9, (1.0 + m / 8.0) * 2.0 ** (e - 7.0))
    v = np.where((e == 15) & (m == 7), np.nan, v)
    return (s * v).astype(np.float32)


FP8_E4M3_VALUES = _fp8_e4m3_table()
FP8_E4M3_MAX = 448.0


def to_fp8_e4m3(a) -> np.ndarray:
    """float -> e4m3 codes (uint8), round to nearest with ties to the even
    code, saturating at +-448.  With a = |x| in binade 2^e (e clamped to
    [-6, 8]) the e4m3 grid step is 2^(e-3), so the rounded magnitude is
    rint(a / 2^(e-3)) 2^(e-3) (rint: ties to even, which is the even code)."""
    x = np.asarray(a, np.float64)
    mag = np.minimum(np.abs(x), FP8_E4M3_MAX)
    _, E = np.frexp(mag)                      # mag = f 2^E, f in [0.5, 1)
    e = np.clip(E - 1, -6, 8)
    step = np.ldexp(1.0, e - 3)
    v = np.rint(mag / step) * step            # exact: division by a power of two
    _, E2 = np.frexp(v)
    e2 = np.clip(E2 - 1, -6, 8)
    normal = v >= 2.0 ** -6
    mant = np.where(normal, np.rint(np.ldexp(v, -e2) * 8.0) - 8.0, np.rint(v * 2.0 ** 9))
    code = np.where(normal, ((e2 + 7) << 3) + mant, mant).astype(np.uint8)
    return np.where(x < 0, code | 0x80, code).astype(np.uint8)


def from_fp8_e4m3(codes) -> np.ndarray:
    return FP8_E4M3_VALUES[np.asarray(codes, np.uint8)]


@dataclass
class GlmBatchModel:
    """Many independent logistic-regression posteriors (the batched GLM
    runner): per problem X_p quantised to fp8 e4m3 codes (N_p x d_p, uint8)
    with one fp32 scale per column, labels y_p in {0, 1}, prior
    beta ~ N(0, prior_sd^2 I)."""
    X: list
    scales: list
    y: list
    prior_sd: float
    kind: str = "logistic"

    @property
    def problems(self) -> int:
        return len(self.X)

    @property
    def dims(self) -> np.ndarray:
        return np.array([x.shape[1] for x in self.X], np.uint32)

    @property
    def rows(self) -> np.ndarray:
        return np.array([x.shape[0] for x in self.X], np.uint64)

    def dequantized(self, p: int) -> np.ndarray:
        """x_ij = scale_j * fp8(q_ij) in float64, the data the posterior of
        problem p is defined on."""
        return (from_fp8_e4m3(self.X[p]).astype(np.float64)
                * np.asarray(self.scales[p], np.float32).astype(np.float64)[None, :])


def quantize_fp8_columns(X) -> tuple[np.ndarray, np.ndarray]:
    """Per-column fp32 scales max|x_j| / 448 (1 for an all-zero column) and
    the e4m3 codes of x / scale."""
    X = np.asarray(X, np.float64)
    m = np.abs(X).max(axis=0)
    sc = np.where(m > 0, m / FP8_E4M3_MAX, 1.0).astype(np.float32)
    return to_fp8_e4m3(X / sc.astype(np.float64)[None, :]), sc


def glm_logistic_batch_model(Xs, ys, prior_sd: float = 5.0) -> GlmBatchModel:
    """From float data (each X_p N_p x d_p): per-column fp8 quantisation."""
    q, sc = zip(*(quantize_fp8_columns(x) for x in Xs))
    yv = [np.ascontiguousarray(y, np.float32).reshape(-1) for y in ys]
    for x, y in zip(q, yv):
        if x.shape[0] != y.size:
            raise ValueError("y must have one label per row of X")
        if not 1 <= x.shape[1] <= 64:
            raise ValueError("batched GLM problems have 1 to 64 coefficients")
    return GlmBatchModel(list(q), list(sc), yv, float(prior_sd))


def hbb_model(max_groups: int = 50) -> DistributionModel:
    """Hierarchical beta-binomial with up to max_groups groups (dimension
    max_groups + 2; each batched problem has its own G)."""
    return distribution_model(HBB_SOURCE, name="hbb", dimension=max_groups + 2, params_size=1)


def mvn_model(dim: int) -> DistributionModel:
    return distribution_model(MVN_SOURCE, name="mvn", dimension=dim, params_size=dim + dim * dim,
                              limits=[[-FLT_MAX, FLT_MAX]] * dim)


def diag_gaussian_model(dim: int) -> DistributionModel:
    """Independent Gaussian coordinates (params: dim means, dim precisions):
    a model whose density costs 2 DIM flops, for measuring the move itself."""
    return distribution_model(DIAG_GAUSS_SOURCE, name="diag_gaussian", dimension=dim,
                              params_size=2 * dim, limits=[[-FLT_MAX, FLT_MAX]] * dim)
