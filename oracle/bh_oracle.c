/*
 * bh_oracle.c -- CPU restatement of the reference's OpenCL hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for the HIP
 * engine in bayadera_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or, for
 * cpu_baseline, as the timed CPU restatement).  The product path never calls
 * into it.
 *
 * Every function cites the reference file:line it restates.  Path
 * abbreviations:
 *   stretch.cl  = src/device/.../opencl/engines/amd-gcn-mcmc-stretch.cl
 *   estimate.cl = src/device/.../opencl/engines/amd-gcn-estimate.cl
 *   acor.cl     = src/device/.../opencl/engines/amd-gcn-acor.cl
 *   uniform.cl  = src/device/.../opencl/rng/uniform-sampler.cl
 *   dist/NAME.cl = src/device/.../opencl/distributions/NAME.cl
 *   amd_gcn.clj = src/clojure/uncomplicate/bayadera/internal/device/amd_gcn.clj
 *
 * Compiled with -ffp-contract=off: every fused multiply-add the reference's
 * OpenCL compiler forms (FP_CONTRACT is on by default in OpenCL C) is written
 * here as an explicit fmaf(), and every other operation rounds separately.
 * Which expressions contract was fixed by the reference's golden vectors
 * (test/.../internal/amd_gcn_test.clj:40-54, 196-272).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BHO_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As   */
/* Easy as 1, 2, 3", SC'11.  Random123's philox4x32 with R=10 (the reference */
/* includes Random123/philox.h at stretch.cl:1 and uniform.cl:1).            */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    uint64_t p0 = (uint64_t)PHILOX_M0 * c[0];
    uint64_t p1 = (uint64_t)PHILOX_M1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

BHO_EXPORT void bho_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        philox_round(c, k);
        k[0] += PHILOX_W0;
        k[1] += PHILOX_W1;
    }
    memcpy(out, c, sizeof c);
}

/* u01fpt_oo_32_24, uniform.cl:8-10: (0.5 + (i >> 9)) * 2^-23, exact in fp32. */
BHO_EXPORT float bho_u01(uint32_t i) {
    return (0.5f + (float)(i >> 9)) * 1.1920928955078125E-7f;
}

static inline void u01x4(const uint32_t ctr[4], const uint32_t key[2], float u[4]) {
    uint32_t r[4];
    bho_philox4x32_10(ctr, key, r);
    for (int m = 0; m < 4; ++m) u[m] = bho_u01(r[m]);
}

/* Uniform direct sampler, uniform.cl:19-34: one Philox call per 4 outputs,
 * key {seed, 0xdecafaaa}, counter {gid, 0xf00dcafe, 0xdeadbeef, 0xbeeff00d};
 * x = u*(upper-lower) + lower, contracted to fma(u, upper-lower, lower). */
BHO_EXPORT void bho_uniform_sample(uint32_t seed, float lower, float upper,
                                   uint32_t n_quads, float* out) {
    const uint32_t key[2] = {seed, 0xdecafaaau};
    const float range = upper - lower;
    for (uint32_t g = 0; g < n_quads; ++g) {
        const uint32_t ctr[4] = {g, 0xf00dcafeu, 0xdeadbeefu, 0xbeeff00du};
        float u[4];
        u01x4(ctr, key, u);
        for (int m = 0; m < 4; ++m) out[4 * g + m] = fmaf(u[m], range, lower);
    }
}

/* Direct samplers of rng/NAME-sampler.cl (the uniform one above):
 *   kind 1 gaussian    (gaussian-sampler.cl:17-44): Box-Muller, sin/cos of
 *          2pi*u0 (2pi*u2) times sqrt(-2 log u1) (sqrt(-2 log u3)), then
 *          fma(., sigma, mu) with params {mu, sigma}
 *   kind 2 exponential (exponential-sampler.cl:16-29): (-1/lambda) log(1-u)
 *   kind 3 erlang      (erlang-sampler.cl:20-40): -(sum_{i<k} log u_i)/lambda,
 *          counter word 3 = i, params {lambda, k}
 * The reference uses native_sin/cos/log there; this uses the C library's. */
BHO_EXPORT void bho_direct_sample(int kind, uint32_t seed, float p0, float p1, uint32_t n_quads,
                                  float* out) {
    const uint32_t key[2] = {seed, 0xdecafaaau};
    const float two_pi = 6.2831855f;
    for (uint32_t g = 0; g < n_quads; ++g) {
        float r[4];
        if (kind == 3) {
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (uint32_t i = 0; (float)i < p1; ++i) {
                const uint32_t ctr[4] = {g, 0xf00dcafeu, 0xdeadbeefu, i};
                float u[4];
                u01x4(ctr, key, u);
                for (int m = 0; m < 4; ++m) acc[m] += logf(u[m]);
            }
            for (int m = 0; m < 4; ++m) r[m] = -(acc[m] / p0);
        } else {
            const uint32_t ctr[4] = {g, 0xf00dcafeu, 0xdeadbeefu, 0xbeeff00du};
            float u[4];
            u01x4(ctr, key, u);
            if (kind == 1) {
                const float ra = sqrtf(-2.0f * logf(u[1])), rb = sqrtf(-2.0f * logf(u[3]));
                const float bm[4] = {sinf(two_pi * u[0]) * ra, cosf(two_pi * u[0]) * ra,
                                     sinf(two_pi * u[2]) * rb, cosf(two_pi * u[2]) * rb};
                for (int m = 0; m < 4; ++m) r[m] = fmaf(bm[m], p1, p0);
            } else if (kind == 2) {
                const float c = -1.0f / p0;
                for (int m = 0; m < 4; ++m) r[m] = c * logf(1.0f - u[m]);
            } else {
                for (int m = 0; m < 4; ++m) r[m] = fmaf(u[m], p1 - p0, p0);
            }
        }
        for (int m = 0; m < 4; ++m) out[4 * g + m] = r[m];
    }
}

/* init_walkers, stretch.cl:138-159: work-item g writes xs[4g .. 4g+3];
 * coordinate i gets limits row (i % dim); x = u*hi + (1-u)*lo, contracted
 * to fma(u, hi, (1-u)*lo).  Global size is dim*walkers/4 (amd_gcn.clj:399). */
BHO_EXPORT void bho_init_walkers(uint32_t seed, const float* limits, uint32_t dim,
                                 uint32_t walkers, float* xs) {
    const uint32_t key[2] = {seed, 0xdecafaaau};
    const uint32_t items = dim * walkers / 4;
    for (uint32_t g = 0; g < items; ++g) {
        const uint32_t ctr[4] = {g, 0xf00dcafeu, 0xdeadbeefu, 0xbeeff00du};
        float u[4];
        u01x4(ctr, key, u);
        for (uint32_t m = 0; m < 4; ++m) {
            const uint32_t i = 4 * g + m;
            const float lo = limits[2 * (i % dim)], hi = limits[2 * (i % dim) + 1];
            xs[i] = fmaf(u[m], hi, (1.0f - u[m]) * lo);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Models.  The reference's model sources (dist/NAME.cl) restated in C; the        */
/* library's native_log/native_exp become logf/expf (the device versions     */
/* differ in the last bits, so model-dependent results are compared with a   */
/* tolerance).                                                               */
/* ------------------------------------------------------------------------ */
enum {
    BHO_UNIFORM_LOGPDF = 0,        /* dist/uniform.cl:12-15 */
    BHO_GAUSSIAN_MCMC_LOGPDF = 1,  /* dist/gaussian.cl:19-22 */
    BHO_GAUSSIAN_LOGPDF = 2,       /* dist/gaussian.cl:24-27 */
    BHO_BETA_MCMC_LOGPDF = 3,      /* dist/beta.cl:19-22 */
    BHO_BETA_LOGPDF = 4,           /* dist/beta.cl:24-27 */
    BHO_BINOMIAL_LOGLIK = 5,       /* dist/binomial.cl:25-27 (loglik signature) */
    BHO_BETA_BINOMIAL_POST = 6,    /* dist/posterior.cl:1-6 with binomial_loglik + beta_logpdf */
    BHO_MVN_LOGPDF = 7,            /* correlated Gaussian -(x-mu)'P(x-mu)/2 (4-D benchmark model) */
    BHO_GAUSSIAN_LOGLIK = 8,       /* dist/gaussian.cl:29-42 */
    BHO_DIAG_GAUSS_LOGPDF = 9,     /* independent Gaussian coordinates -sum w_i (x_i-mu_i)^2/2 (benchmark model) */
    BHO_HBB_LOGPDF = 10,           /* hierarchical beta-binomial on logits (batched benchmark model) */
    BHO_LINREG_POST = 11,          /* Gaussian linear regression with normal priors (benchmark model) */
    BHO_LOGISTIC_POST = 12,        /* dist/posterior.cl:1-6: logistic loglik + iid normal prior (benchmark model) */
};

static const float LOG_SQRT_2PI_F = 0.9189385332046727f;

static inline float gaussian_log_unscaled(float mu, float sigma, float x) {
    /* dist/gaussian.cl:5-7 */
    return (x - mu) * (x - mu) / (-2.0f * sigma * sigma);
}

static inline float beta_log_unscaled(float a, float b, float x) {
    /* dist/beta.cl:9-11: (a-1)log x + (b-1)log(1-x), the second product
     * contracts into the add. */
    return fmaf(b - 1.0f, logf(1.0f - x), (a - 1.0f) * logf(x));
}

static inline float binomial_loglik(const float* data, const float* p) {
    /* dist/binomial.cl:5-7, 25-27: k*log(p) + (n-k)*log(1-p) */
    const float n = data[0], k = data[1];
    return fmaf(n - k, logf(1.0f - p[0]), k * logf(p[0]));
}

/* One data row of the logistic loglik in fp32: softplus(t) with t = -eta
 * for y = 1 and eta for y = 0, as max(t, 0) + log(1 + exp(-|t|)). */
static inline float logistic_row_term(float eta, float yv) {
    const float t = (yv > 0.5f) ? -eta : eta;
    return fmaxf(t, 0.0f) + logf(1.0f + expf(-fabsf(t)));
}

/* The logistic posterior as a user model in the reference's plugin form:
 * posterior.cl:1-6 (LOGLIK + LOGPDF) with LOGLIK a loop over the data rows
 * [x_i (dim), y_i] (params[0 .. data_len), data_len / (dim + 1) rows) in
 * fp32, the accumulation shape of gaussian.cl:29-42, and LOGPDF the iid
 * N(0, s^2) prior (hyperparameter params[data_len] = s) without its
 * constant. */
static float logistic_post_f32(uint32_t data_len, const float* params, uint32_t dim,
                               const float* x) {
    const uint32_t rows = data_len / (dim + 1);
    float acc = 0.0f;
    for (uint32_t i = 0; i < rows; ++i) {
        const float* row = params + (size_t)i * (dim + 1);
        float eta = 0.0f;
        for (uint32_t j = 0; j < dim; ++j) eta = fmaf(row[j], x[j], eta);
        acc -= logistic_row_term(eta, row[dim]);
    }
    const float is = 1.0f / params[data_len];
    float q = 0.0f;
    for (uint32_t j = 0; j < dim; ++j) {
        const float z = x[j] * is;
        q = fmaf(z, z, q);
    }
    return acc + -0.5f * q;
}

BHO_EXPORT float bho_model_logfn(int model, uint32_t data_len, uint32_t params_len,
                                 const float* params, uint32_t dim, const float* x) {
    (void)params_len;
    switch (model) {
    case BHO_UNIFORM_LOGPDF: {
        const float lo = params[0], hi = params[1];
        const int in_range = (lo <= x[0]) && (x[0] <= hi);
        return logf(in_range ? (1.0f / (hi - lo)) : 0.0f);
    }
    case BHO_GAUSSIAN_MCMC_LOGPDF:
        return gaussian_log_unscaled(params[0], params[1], x[0]);
    case BHO_GAUSSIAN_LOGPDF:
        return gaussian_log_unscaled(params[0], params[1], x[0]) + (-logf(params[1]) - LOG_SQRT_2PI_F);
    case BHO_BETA_MCMC_LOGPDF:
        return beta_log_unscaled(params[0], params[1], x[0]);
    case BHO_BETA_LOGPDF:
        return beta_log_unscaled(params[0], params[1], x[0]) + params[2];
    case BHO_BINOMIAL_LOGLIK:
        return binomial_loglik(params, x);
    case BHO_BETA_BINOMIAL_POST: {
        const float* hyper = params + data_len;
        return binomial_loglik(params, x) + (beta_log_unscaled(hyper[0], hyper[1], x[0]) + hyper[2]);
    }
    case BHO_MVN_LOGPDF: {
        /* params = [mu (dim), P (dim*dim, row-major)] */
        const float* mu = params;
        const float* P = params + dim;
        float q = 0.0f;
        for (uint32_t r = 0; r < dim; ++r) {
            float row = 0.0f;
            for (uint32_t c = 0; c < dim; ++c) row = fmaf(P[r * dim + c], x[c] - mu[c], row);
            q = fmaf(x[r] - mu[r], row, q);
        }
        return -0.5f * q;
    }
    case BHO_DIAG_GAUSS_LOGPDF: {
        /* params = [mu (dim), w (dim)], w_i = 1 / sigma_i^2 */
        const float* mu = params;
        const float* w = params + dim;
        float q = 0.0f;
        for (uint32_t i = 0; i < dim; ++i) {
            const float d = x[i] - mu[i];
            q = fmaf(w[i] * d, d, q);
        }
        return -0.5f * q;
    }
    case BHO_HBB_LOGPDF: {
        /* x = [log alpha, log beta, t_1 .. t_G], theta_g = 1/(1+exp(-t_g));
         * params = [n_1 .. n_G, y_1 .. y_G, s0] (data_len = 2G):
         * -G lnB(a, b) - sum_g [(y_g + a) sp(-t_g) + (n_g - y_g + b) sp(t_g)]
         * - (la^2 + lb^2) / (2 s0^2), sp(t) = max(t, 0) + log(1 + exp(-|t|)) */
        const uint32_t G = dim - 2;
        const float* n = params;
        const float* y = params + G;
        const float s0 = params[data_len];
        const float la = x[0], lb = x[1];
        const float a = expf(la), b = expf(lb);
        float acc = -(float)G * (lgammaf(a) + lgammaf(b) - lgammaf(a + b));
        for (uint32_t g = 0; g < G; ++g) {
            const float t = x[2 + g];
            const float spm = fmaxf(-t, 0.0f) + logf(1.0f + expf(-fabsf(t)));
            const float spp = fmaxf(t, 0.0f) + logf(1.0f + expf(-fabsf(t)));
            acc -= (y[g] + a) * spm + (n[g] - y[g] + b) * spp;
        }
        return acc - 0.5f * (la * la + lb * lb) / (s0 * s0);
    }
    case BHO_LINREG_POST: {
        /* x = [beta (dim-1), log sigma]; data rows [x_i (dim-1), y_i]
         * (data_len / dim rows); hyperparameters [s_beta, s_sigma]:
         * -1/2 sum ((y_i - x_i.beta) / sigma)^2 - rows (log sigma + log sqrt(2 pi))
         * - 1/2 (|beta / s_beta|^2 + (log sigma / s_sigma)^2) */
        const uint32_t d = dim - 1;
        const uint32_t rows = data_len / dim;
        const float log_sigma = x[d];
        const float inv_sigma = expf(-log_sigma);
        float ss = 0.0f;
        for (uint32_t i = 0; i < rows; ++i) {
            const float* row = params + (size_t)i * dim;
            float eta = 0.0f;
            for (uint32_t j = 0; j < d; ++j) eta = fmaf(row[j], x[j], eta);
            const float r = (row[d] - eta) * inv_sigma;
            ss = fmaf(r, r, ss);
        }
        const float ll = -0.5f * ss - (float)rows * (log_sigma + LOG_SQRT_2PI_F);
        const float* hyper = params + data_len;
        const float ib = 1.0f / hyper[0];
        const float is = 1.0f / hyper[1];
        float q = 0.0f;
        for (uint32_t j = 0; j < d; ++j) {
            const float z = x[j] * ib;
            q = fmaf(z, z, q);
        }
        const float zs = x[d] * is;
        return ll + -0.5f * fmaf(zs, zs, q);
    }
    case BHO_LOGISTIC_POST:
        return logistic_post_f32(data_len, params, dim, x);
    case BHO_GAUSSIAN_LOGLIK: {
        /* data = params[0 .. data_len), x = (mu, sigma) */
        const float mu = x[0], sigma = x[1];
        if (0.0f < sigma) {
            float res = gaussian_log_unscaled(mu, sigma, params[0])
                        + (float)data_len * (-logf(sigma) - LOG_SQRT_2PI_F);
            for (uint32_t i = 1; i < data_len; ++i) res += gaussian_log_unscaled(mu, sigma, params[i]);
            return res;
        }
        return NAN;
    }
    default:
        return NAN;
    }
}

/* ------------------------------------------------------------------------ */
/* Stretch move                                                              */
/* ------------------------------------------------------------------------ */
enum { BHO_PARTNER_FLAT = 0, BHO_PARTNER_WALKER = 1 };

/* Planted-mistake switch for tests: bit 0 rounds z's product before the add
 * (no contraction), bit 1 does the same for Y.  Zero in every real use. */
static int g_variant = 0;
BHO_EXPORT void bho_set_variant(int flags) { g_variant = flags; }

static inline float pown_f(float z, int n) {
    /* pown(z, n) for the small non-negative n = DIM-1 used at stretch.cl:69 */
    float r = 1.0f;
    for (int i = 0; i < n; ++i) r *= z;
    return r;
}

/* One walker's stretch move, stretch.cl:31-81.
 *   key {seed, 0xdecafbad}, counter {k, step, odd_or_even, 0xbeeff00d};
 *   z = (a-2+1/a) u1^2 + 2(1-1/a) u1 + 1/a, contracted as
 *       fma((a-2+1/a)*u1, u1, (2(1-1/a))*u1) + 1/a             (stretch.cl:53-54)
 *   partner: FLAT   -> element offset j0 = (uint)((u0*K)*DIM), read
 *                      Scompl[(j0+i) mod K*DIM]                 (stretch.cl:57,62)
 *            WALKER -> walker j = min((uint)(u0*K), K-1), Scompl[j*DIM+i]
 *   Y_i = Xj_i + z (Xk_i - Xj_i), contracted to fma(z, Xk_i-Xj_i, Xj_i) (:64)
 *   q = isfinite(ly) ? z^(DIM-1) exp(beta (ly - lX)) : 0 ; accept iff u2 <= q (:68-71)
 * The wrap modulo K*DIM defines the reference's out-of-bounds read at
 * j0+i >= K*DIM (it only occurs for DIM > 1). */
static inline float propose_one(uint32_t k, uint32_t seed, uint32_t odd_or_even, uint32_t step,
                                uint32_t K, uint32_t dim, const float* Scompl, const float* X,
                                float a, int partner_mode, float* Y, float* z_out) {
    const uint32_t key[2] = {seed, 0xdecafbadu};
    const uint32_t ctr[4] = {k, step, odd_or_even, 0xbeeff00du};
    float u[4];
    u01x4(ctr, key, u);
    const float inv_a = 1.0f / a;
    const float ca = a - 2.0f + inv_a;
    const float cb = 2.0f * (1.0f - inv_a);
    const float z = (g_variant & 1) ? ((ca * u[1]) * u[1] + cb * u[1]) + inv_a
                                    : fmaf(ca * u[1], u[1], cb * u[1]) + inv_a;
    const uint32_t k0 = k * dim;
    if (partner_mode == BHO_PARTNER_FLAT) {
        const uint32_t total = K * dim;
        uint32_t j0 = (uint32_t)((u[0] * (float)K) * (float)dim);
        if (j0 >= total) j0 -= total;
        for (uint32_t i = 0; i < dim; ++i) {
            uint32_t e = j0 + i;
            if (e >= total) e -= total;
            const float xj = Scompl[e];
            Y[i] = (g_variant & 2) ? xj + z * (X[k0 + i] - xj) : fmaf(z, X[k0 + i] - xj, xj);
        }
    } else {
        uint32_t j = (uint32_t)(u[0] * (float)K);
        if (j >= K) j = K - 1;
        for (uint32_t i = 0; i < dim; ++i) {
            const float xj = Scompl[j * dim + i];
            Y[i] = (g_variant & 2) ? xj + z * (X[k0 + i] - xj) : fmaf(z, X[k0 + i] - xj, xj);
        }
    }
    *z_out = z;
    return u[2];
}

static inline int stretch_one(uint32_t k, uint32_t seed, uint32_t odd_or_even, uint32_t step,
                              uint32_t K, uint32_t dim, const float* Scompl, float* X,
                              float* logfn_X, float a, float beta, int partner_mode, int model,
                              uint32_t data_len, uint32_t params_len, const float* params,
                              float* Y) {
    float z;
    const float u2 = propose_one(k, seed, odd_or_even, step, K, dim, Scompl, X, a, partner_mode,
                                 Y, &z);
    const float ly = bho_model_logfn(model, data_len, params_len, params, dim, Y);
    const float q = isfinite(ly) ? pown_f(z, (int)dim - 1) * expf(beta * (ly - logfn_X[k])) : 0.0f;
    const int accepted = u2 <= q;
    if (accepted) {
        for (uint32_t i = 0; i < dim; ++i) X[k * dim + i] = Y[i];
        logfn_X[k] = ly;
    }
    return accepted;
}

/* One half-step over all K walkers of X (stretch_move_bare / the move part
 * of stretch_move_accu, stretch.cl:83-134).  Walkers are independent given
 * Scompl, so the loop is parallel.  Returns the number of accepted moves and
 * optionally the per-walker accept flags. */
BHO_EXPORT uint32_t bho_stretch_half(uint32_t seed, uint32_t odd_or_even, uint32_t step,
                                     uint32_t K, uint32_t dim, const float* Scompl, float* X,
                                     float* logfn_X, float a, float beta, int partner_mode,
                                     int model, uint32_t data_len, uint32_t params_len,
                                     const float* params, uint8_t* accepted, int threads) {
    uint32_t total = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = 1;
#pragma omp parallel num_threads(threads) reduction(+ : total)
#endif
    {
        float* Y = (float*)malloc(sizeof(float) * (dim ? dim : 1));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t kk = 0; kk < (int64_t)K; ++kk) {
            const uint32_t k = (uint32_t)kk;
            const int acc = stretch_one(k, seed, odd_or_even, step, K, dim, Scompl, X, logfn_X, a,
                                        beta, partner_mode, model, data_len, params_len, params, Y);
            if (accepted) accepted[k] = (uint8_t)acc;
            total += (uint32_t)acc;
        }
        free(Y);
    }
    (void)threads;
    return total;
}

/* bho_stretch_half for the logistic posterior (BHO_LOGISTIC_POST) with the
 * walkers taken BHO_WBLK at a time: every walker's proposal, log posterior
 * and accept test are the per-walker ones (the same operations in the same
 * order), but the row loop runs once per block, each row's values applied
 * to the block's walkers side by side -- the work-item vectorisation an
 * OpenCL CPU device applies to the reference's kernel.  This is the timed
 * CPU baseline of the logistic workload; with -ffp-contract=off it equals
 * bho_stretch_half bit for bit. */
#define BHO_WBLK 128
BHO_EXPORT uint32_t bho_logistic_half_blocked(uint32_t seed, uint32_t odd_or_even, uint32_t step,
                                              uint32_t K, uint32_t dim, const float* Scompl,
                                              float* X, float* logfn_X, float a, float beta,
                                              int partner_mode, uint32_t data_len,
                                              const float* params, uint8_t* accepted,
                                              int threads) {
    uint32_t total = 0;
    const uint32_t nblk = (K + BHO_WBLK - 1) / BHO_WBLK;
    const uint32_t rows = data_len / (dim + 1);
    const float is = 1.0f / params[data_len];
#ifdef _OPENMP
    if (threads <= 0) threads = 1;
#pragma omp parallel num_threads(threads) reduction(+ : total)
#endif
    {
        float* Yt = (float*)malloc(sizeof(float) * (size_t)dim * BHO_WBLK);
        float* Y = (float*)malloc(sizeof(float) * (dim ? dim : 1));
        float z[BHO_WBLK], u2[BHO_WBLK], acc[BHO_WBLK], eta[BHO_WBLK];
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t bb = 0; bb < (int64_t)nblk; ++bb) {
            const uint32_t k0 = (uint32_t)bb * BHO_WBLK;
            const uint32_t nw = K - k0 < BHO_WBLK ? K - k0 : BHO_WBLK;
            for (uint32_t w = 0; w < BHO_WBLK; ++w) {
                if (w < nw) {
                    u2[w] = propose_one(k0 + w, seed, odd_or_even, step, K, dim, Scompl, X, a,
                                        partner_mode, Y, &z[w]);
                    for (uint32_t j = 0; j < dim; ++j) Yt[(size_t)j * BHO_WBLK + w] = Y[j];
                } else {
                    for (uint32_t j = 0; j < dim; ++j) Yt[(size_t)j * BHO_WBLK + w] = 0.0f;
                }
                acc[w] = 0.0f;
            }
            for (uint32_t i = 0; i < rows; ++i) {
                const float* row = params + (size_t)i * (dim + 1);
                for (uint32_t w = 0; w < BHO_WBLK; ++w) eta[w] = 0.0f;
                for (uint32_t j = 0; j < dim; ++j) {
                    const float xv = row[j];
                    const float* yj = Yt + (size_t)j * BHO_WBLK;
#ifdef _OPENMP
#pragma omp simd
#endif
                    for (uint32_t w = 0; w < BHO_WBLK; ++w) eta[w] = fmaf(xv, yj[w], eta[w]);
                }
                const float yv = row[dim];
#ifdef _OPENMP
#pragma omp simd
#endif
                for (uint32_t w = 0; w < BHO_WBLK; ++w) acc[w] -= logistic_row_term(eta[w], yv);
            }
            for (uint32_t w = 0; w < nw; ++w) {
                const uint32_t k = k0 + w;
                float q = 0.0f;
                for (uint32_t j = 0; j < dim; ++j) {
                    const float zz = Yt[(size_t)j * BHO_WBLK + w] * is;
                    q = fmaf(zz, zz, q);
                }
                const float ly = acc[w] + -0.5f * q;
                const float qa = isfinite(ly)
                                     ? pown_f(z[w], (int)dim - 1) * expf(beta * (ly - logfn_X[k]))
                                     : 0.0f;
                const int acc_k = u2[w] <= qa;
                if (acc_k) {
                    for (uint32_t j = 0; j < dim; ++j) X[(size_t)k * dim + j] = Yt[(size_t)j * BHO_WBLK + w];
                    logfn_X[k] = ly;
                }
                if (accepted) accepted[k] = (uint8_t)acc_k;
                total += (uint32_t)acc_k;
            }
        }
        free(Yt);
        free(Y);
    }
    (void)threads;
    return total;
}

/* logfn kernel, stretch.cl:161-167: res[w] = LOGFN(x + DIM*w). */
BHO_EXPORT void bho_logfn(int model, uint32_t data_len, uint32_t params_len, const float* params,
                          uint32_t dim, uint32_t n, const float* x, float* res) {
    for (uint32_t w = 0; w < n; ++w)
        res[w] = bho_model_logfn(model, data_len, params_len, params, dim, x + (size_t)dim * w);
}

/* ------------------------------------------------------------------------ */
/* Work-group reductions                                                     */
/* ------------------------------------------------------------------------ */

/* The work-group sum tree of stretch.cl:5-29 (uint twin of clojurecl's
 * work_group_reduction_sum; the 2-D variant estimate.cl:38-74 has the same
 * shape along dimension 1).  For a group of n items: while i > 0, the item
 * (i/2 - 1) first adds lacc[(i/2 - 1) + i/2 + 1] when i is odd, then every
 * item l < i/2 adds lacc[l + i/2].  Restated for float. */
BHO_EXPORT float bho_tree_sum_f32(const float* v, uint32_t n) {
    if (n == 0) return 0.0f;
    float* lacc = (float*)malloc(sizeof(float) * n);
    float* pacc = (float*)malloc(sizeof(float) * n);
    memcpy(lacc, v, sizeof(float) * n);
    memcpy(pacc, v, sizeof(float) * n);
    uint32_t i = n;
    while (i > 0) {
        const int odd = (i & 1u) != 0;
        const uint32_t odd_item = (i >> 1) - 1; /* only used when odd and i > 1 */
        i >>= 1;
        /* every read below indexes >= i, every write < i: no intra-level race */
        if (odd && i > 0) pacc[odd_item] += lacc[odd_item + i + 1];
        for (uint32_t l = 0; l < i; ++l) {
            pacc[l] += lacc[l + i];
            lacc[l] = pacc[l];
        }
    }
    float r = lacc[0];
    free(lacc);
    free(pacc);
    return r;
}

BHO_EXPORT double bho_tree_sum_f64(const double* v, uint32_t n) {
    if (n == 0) return 0.0;
    double* lacc = (double*)malloc(sizeof(double) * n);
    double* pacc = (double*)malloc(sizeof(double) * n);
    memcpy(lacc, v, sizeof(double) * n);
    memcpy(pacc, v, sizeof(double) * n);
    uint32_t i = n;
    while (i > 0) {
        const int odd = (i & 1u) != 0;
        const uint32_t odd_item = (i >> 1) - 1;
        i >>= 1;
        if (odd && i > 0) pacc[odd_item] += lacc[odd_item + i + 1];
        for (uint32_t l = 0; l < i; ++l) {
            pacc[l] += lacc[l + i];
            lacc[l] = pacc[l];
        }
    }
    double r = lacc[0];
    free(lacc);
    free(pacc);
    return r;
}

/* The accounting half of stretch_move_accu, stretch.cl:98-114, applied after
 * a half-step: accept[g] += (number accepted in group g); for each dim i,
 * means[groups*DIM*step + i*groups + g] += tree_sum(X[k*DIM+i], k in g). */
BHO_EXPORT void bho_stretch_accu(uint32_t K, uint32_t dim, uint32_t wgs, uint32_t step,
                                 const float* X, const uint8_t* accepted, uint32_t* accept,
                                 float* means) {
    const uint32_t groups = K / wgs;
    float* v = (float*)malloc(sizeof(float) * wgs);
    for (uint32_t g = 0; g < groups; ++g) {
        uint32_t cnt = 0;
        for (uint32_t l = 0; l < wgs; ++l) cnt += accepted[g * wgs + l];
        accept[g] += cnt;
        for (uint32_t i = 0; i < dim; ++i) {
            for (uint32_t l = 0; l < wgs; ++l) v[l] = X[(size_t)(g * wgs + l) * dim + i];
            means[(size_t)groups * dim * step + (size_t)i * groups + g] += bho_tree_sum_f32(v, wgs);
        }
    }
    free(v);
}

/* The 2-D reduce chain of clojurecl's enq-reduce! (amd_gcn.clj:437-438,
 * 155-159) over an m x n column-major block a[j*m + r] along n with local
 * sizes (local_m, local_n): the first pass reduces each run of local_n
 * columns with the tree above, and passes repeat over the partial sums until
 * one remains (the final group is smaller when local_n does not divide the
 * count).  out[r] = the float total for row r. */
BHO_EXPORT void bho_reduce_rows_f32(const float* a, uint32_t m, uint32_t n, uint32_t local_n,
                                    float* out) {
    float* cur = (float*)malloc(sizeof(float) * (size_t)m * (n ? n : 1));
    memcpy(cur, a, sizeof(float) * (size_t)m * n);
    float* v = (float*)malloc(sizeof(float) * local_n);
    uint32_t count = n;
    do {
        const uint32_t groups = (count + local_n - 1) / local_n;
        for (uint32_t r = 0; r < m; ++r) {
            for (uint32_t g = 0; g < groups; ++g) {
                const uint32_t start = g * local_n;
                const uint32_t len = (count - start < local_n) ? (count - start) : local_n;
                for (uint32_t l = 0; l < len; ++l) v[l] = cur[(size_t)(start + l) * m + r];
                /* all groups read before any group writes its slot */
                float s = bho_tree_sum_f32(v, len);
                v[0] = s;
                cur[(size_t)g * m + r] = s;
            }
        }
        count = groups;
    } while (count > 1);
    for (uint32_t r = 0; r < m; ++r) out[r] = cur[r];
    free(cur);
    free(v);
}

/* ------------------------------------------------------------------------ */
/* Dataset estimates, estimate.cl                                            */
/* ------------------------------------------------------------------------ */

/* min_max_reduce chain, estimate.cl:38-93: per-row [min, max] of an m x n
 * column-major matrix with leading dimension ld.  Order-independent. */
BHO_EXPORT void bho_min_max(const float* a, uint32_t m, uint32_t n, uint32_t ld, float* limits) {
    for (uint32_t r = 0; r < m; ++r) {
        float lo = a[r], hi = a[r];
        for (uint32_t j = 1; j < n; ++j) {
            const float x = a[(size_t)j * ld + r];
            lo = fminf(lo, x);
            hi = fmaxf(hi, x);
        }
        limits[2 * r] = lo;
        limits[2 * r + 1] = hi;
    }
}

/* histogram, estimate.cl:1-25: bin = (uint)((x-lo)/(hi-lo)*bins), clamped
 * to bins-1; the float->uint conversion saturates (negative and NaN -> 0). */
static inline uint32_t f2u_sat(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}

BHO_EXPORT uint32_t bho_hist_bin(float x, float lo, float hi, uint32_t bins) {
    uint32_t bin = f2u_sat((x - lo) / (hi - lo) * (float)bins);
    return bin < bins ? bin : bins - 1;
}

BHO_EXPORT void bho_histogram_add(const float* limits, const float* a, uint32_t m, uint32_t n,
                                  uint32_t ld, uint32_t bins, uint32_t* counts) {
    for (uint32_t r = 0; r < m; ++r) {
        const float lo = limits[2 * r], hi = limits[2 * r + 1];
        for (uint32_t j = 0; j < n; ++j) counts[r * bins + bho_hist_bin(a[(size_t)j * ld + r], lo, hi, bins)]++;
    }
}

/* uint_to_real, estimate.cl:27-34: pdf = alpha/(hi-lo) * count with
 * alpha = bins/n as float (amd_gcn.clj:206, 481). */
BHO_EXPORT void bho_uint_to_real(float alpha, const float* limits, const uint32_t* counts,
                                 uint32_t m, uint32_t bins, float* pdf) {
    for (uint32_t r = 0; r < m; ++r) {
        const float s = alpha / (limits[2 * r + 1] - limits[2 * r]);
        for (uint32_t b = 0; b < bins; ++b) pdf[r * bins + b] = s * (float)counts[r * bins + b];
    }
}

/* bitonic_local, estimate.cl:95-123, simulated step by step for one group
 * of n (a power of two) values: the output is the bin indices ordered by
 * descending value, ties to the lower index. */
BHO_EXPORT void bho_bitonic_ranks(const float* in, uint32_t n, float* out) {
    float* vx = (float*)malloc(sizeof(float) * n);
    float* vy = (float*)malloc(sizeof(float) * n);
    float* ax = (float*)malloc(sizeof(float) * n);
    float* ay = (float*)malloc(sizeof(float) * n);
    for (uint32_t l = 0; l < n; ++l) { vx[l] = (float)l; vy[l] = in[l]; ax[l] = vx[l]; ay[l] = vy[l]; }
    for (uint32_t length = 1; length < n; length <<= 1) {
        for (int inc = (int)length; inc > 0; inc >>= 1) {
            for (uint32_t l = 0; l < n; ++l) {
                const int direction = (l & (length << 1)) != 0;
                const uint32_t j = l ^ (uint32_t)inc;
                const float ox = ax[j], oy = ay[j];
                const int smaller = (vy[l] < oy) || (oy == vy[l] && j < l);
                const int swap = smaller ^ (j < l) ^ direction;
                if (swap) { vx[l] = ox; vy[l] = oy; }
            }
            for (uint32_t l = 0; l < n; ++l) { ax[l] = vx[l]; ay[l] = vy[l]; }
        }
    }
    for (uint32_t l = 0; l < n; ++l) out[l] = vx[l];
    free(vx); free(vy); free(ax); free(ay);
}

/* Two-pass mean and population variance per row of an m x n column-major
 * matrix (mean_reduce / variance_reduce, estimate.cl:135-157; divide by n,
 * amd_gcn.clj:159,180).  FP64 accumulation. */
BHO_EXPORT void bho_mean_var_f64(const float* a, uint32_t m, uint64_t n, uint32_t ld,
                                 double* mean, double* var) {
    for (uint32_t r = 0; r < m; ++r) {
        double s = 0.0;
        for (uint64_t j = 0; j < n; ++j) s += (double)a[j * ld + r];
        const double mu = s / (double)n;
        double v = 0.0;
        /* the reference subtracts the fp32-rounded mean (estimate.cl:152) */
        const float muf = (float)mu;
        for (uint64_t j = 0; j < n; ++j) {
            const double d = (double)(a[j * ld + r] - muf);
            v += d * d;
        }
        mean[r] = mu;
        if (var) var[r] = v / (double)n;
    }
}

/* ------------------------------------------------------------------------ */
/* Autocorrelation, acor.cl + amd_gcn.clj:216-267                           */
/* ------------------------------------------------------------------------ */

/* acor_1d, acor.cl:26-86, for one work-group [start, start+lsize) of a
 * series of length gsize stored at x[t*stride + dim_id]; writes the group's
 * tree sums of x^2 and of x(x + 2*sum_{s=1..lag} x_{t+s}). */
static void acor_1d_group(const float* x, uint32_t stride, uint32_t dim_id, uint32_t lag,
                          uint32_t gsize, uint32_t start, uint32_t lsize, int has_next,
                          float* c0_out, float* d_out, float* lm, float* va, float* vb) {
    for (uint32_t l = 0; l < 2 * lsize; ++l) lm[l] = 0.0f;
    for (uint32_t l = 0; l < lsize; ++l) {
        const uint32_t gid = start + l;
        const int compute = gid + lag < gsize;
        /* the reference reads past the series when the next group holds
         * fewer than lag points; that read is defined here as 0 */
        const int load_lag = (l < lag) && has_next && (gid + lsize < gsize);
        lm[l] = compute ? x[(size_t)gid * stride + dim_id] : 0.0f;
        lm[l + lsize] = load_lag ? x[(size_t)(gid + lsize) * stride + dim_id] : 0.0f;
    }
    for (uint32_t l = 0; l < lsize; ++l) {
        const float xv = lm[l];
        float xacc = 0.0f;
        for (uint32_t s = 0; s < lag; ++s) {
            const uint32_t idx = l + s + 1;
            /* the local buffer is 2*WGS floats; indexes past 2*lsize read
             * the zero-filled tail */
            xacc += (idx < 2 * lsize) ? lm[idx] : 0.0f;
        }
        xacc = xv * fmaf(2.0f, xacc, xv);
        va[l] = xv * xv;
        vb[l] = xacc;
    }
    *c0_out = bho_tree_sum_f32(va, lsize);
    *d_out = bho_tree_sum_f32(vb, lsize);
}

/* One level of acor_1d over a series of length gsize in groups of wgs. */
static uint32_t acor_level(const float* x, uint32_t stride, uint32_t dim_id, uint32_t lag,
                           uint32_t gsize, uint32_t wgs, float* c0acc, float* dacc) {
    const uint32_t groups = (gsize + wgs - 1) / wgs;
    float* lm = (float*)malloc(sizeof(float) * 2 * wgs + 256 * sizeof(float));
    float* va = (float*)malloc(sizeof(float) * wgs);
    float* vb = (float*)malloc(sizeof(float) * wgs);
    for (uint32_t g = 0; g < groups; ++g) {
        const uint32_t start = g * wgs;
        const uint32_t lsize = (gsize - start < wgs) ? gsize - start : wgs;
        acor_1d_group(x, stride, dim_id, lag, gsize, start, lsize, g + 1 < groups,
                      &c0acc[g], &dacc[g], lm, va, vb);
    }
    free(lm); free(va); free(vb);
    return groups;
}

/* sum_reduction_acor's reduction, acor.cl:88-130: a group of
 * min(WGS, groups) items, item l sums partials l, l+lsize, ... then the tree. */
static void acor_reduce(const float* c0acc, const float* dacc, uint32_t groups, uint32_t wgs,
                        float* c0, float* d) {
    const uint32_t lsize = groups < wgs ? groups : wgs;
    float* va = (float*)malloc(sizeof(float) * lsize);
    float* vb = (float*)malloc(sizeof(float) * lsize);
    for (uint32_t l = 0; l < lsize; ++l) {
        float a = 0.0f, b = 0.0f;
        for (uint32_t t = l; t < groups; t += lsize) { a += c0acc[t]; b += dacc[t]; }
        va[l] = a; vb[l] = b;
    }
    *c0 = bho_tree_sum_f32(va, lsize);
    *d = bho_tree_sum_f32(vb, lsize);
    free(va); free(vb);
}

/* GCNAcorEngine.acor, amd_gcn.clj:216-267, with the device-enqueued
 * recursion of acor.cl:132-214 run as a loop.  series is dim x n
 * column-major (series[t*dim + d]) and is overwritten (centred, then
 * pair-summed in place as the reference does).  Returns 0, or -1 when
 * n < 5*lag (amd_gcn.clj:264-267). */
BHO_EXPORT int bho_acor(float* series, uint32_t dim, uint32_t n, uint32_t wgs, float* tau,
                        float* mean, float* sigma, uint32_t* lag_out) {
    const uint32_t min_fac = 5, win_mult = 5, min_lag = 10, max_lag = 64;
    uint32_t lag = n / min_fac;
    if (wgs < lag) lag = wgs;
    if (max_lag < lag) lag = max_lag;
    if (lag < min_lag) lag = min_lag;
    if (lag_out) *lag_out = lag;
    if (lag * min_fac > n) return -1;
    /* mean: enq-reduce! with local (min(16,dim,WGS), WGS/that), then scal! 1/n */
    uint32_t wgsm = dim < 16 ? dim : 16;
    if (wgs < wgsm) wgsm = wgs;
    const uint32_t wgsn = wgs / wgsm;
    float* sums = (float*)malloc(sizeof(float) * dim);
    bho_reduce_rows_f32(series, dim, n, wgsn, sums);
    const float inv_n = (float)(1.0 / (double)n);
    for (uint32_t d = 0; d < dim; ++d) mean[d] = sums[d] * inv_n;
    free(sums);
    for (uint32_t t = 0; t < n; ++t)
        for (uint32_t d = 0; d < dim; ++d) series[(size_t)t * dim + d] -= mean[d];
    const uint32_t max_groups = (n + wgs - 1) / wgs;
    float* c0acc = (float*)malloc(sizeof(float) * (max_groups + 1));
    float* dacc = (float*)malloc(sizeof(float) * (max_groups + 1));
    for (uint32_t d = 0; d < dim; ++d) {
        /* level 1: acor_2d + sum_reduction_acor(stride 1) */
        uint32_t groups = acor_level(series, dim, d, lag, n, wgs, c0acc, dacc);
        float c0, dd;
        acor_reduce(c0acc, dacc, groups, wgs, &c0, &dd);
        uint32_t stride = 1, cur_lag = lag, prev_n = n;
        float orig_c0 = 0.0f, prev_c0 = 0.0f;
        for (;;) {
            const float start_c0 = (stride == 1) ? c0 : orig_c0;
            const float t = dd / ((stride == 1) ? c0 : prev_c0);
            const uint32_t lag2 = (cur_lag * win_mult < prev_n)
                                      ? cur_lag
                                      : ((prev_n / win_mult) > 10u ? prev_n / win_mult : 10u);
            if (min_lag < lag2 && (float)lag2 < t * (float)win_mult) {
                /* sum_pairwise, acor_1d and the next sum_reduction_acor (acor.cl:143-168) */
                const uint32_t n2 = prev_n / 2;
                const uint32_t w2 = wgs < n2 ? wgs : n2;
                const size_t step = (size_t)stride * dim;
                for (uint32_t g = 0; g < n2; ++g)
                    series[d + 2 * g * step] += series[d + (2 * (size_t)g + 1) * step];
                groups = acor_level(series, (uint32_t)(2 * step), d, lag2, n2, w2, c0acc, dacc);
                float c0n, dn;
                acor_reduce(c0acc, dacc, groups, wgs, &c0n, &dn);
                orig_c0 = start_c0;
                prev_c0 = c0;
                c0 = c0n;
                dd = dn;
                stride *= 2;
                cur_lag = lag2;
                prev_n = n2;
            } else {
                /* acor.cl:174-177 */
                const float scale = (float)stride * (float)(prev_n - lag2);
                tau[d] = dd * (float)(n - cur_lag) / (scale * start_c0);
                sigma[d] = sqrtf(dd / (scale * (float)n));
                break;
            }
        }
    }
    free(c0acc);
    free(dacc);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Logistic-regression posterior (GLM path of the engine).  Not in the       */
/* reference: its models loop over data rows inside LOGFN.  Everything in    */
/* double: eta_i = x_i . beta (the bf16 data times fp32 coefficients are     */
/* exact in double), ll = sum_i y_i eta_i - softplus(eta_i), prior          */
/* beta ~ N(0, 1/prior_prec I).                                             */
/* ------------------------------------------------------------------------ */
BHO_EXPORT double bho_glm_logpost(const float* Xd, const float* y, uint64_t N, uint32_t d,
                                  double prior_prec, const float* beta) {
    double ll = 0.0;
    for (uint64_t i = 0; i < N; ++i) {
        double eta = 0.0;
        for (uint32_t j = 0; j < d; ++j) eta = fma((double)Xd[i * d + j], (double)beta[j], eta);
        if (g_variant & 8) eta = -eta; /* planted mistake (tests only) */
        ll += (double)y[i] * eta - (fmax(eta, 0.0) + log1p(exp(-fabs(eta))));
    }
    double q = 0.0;
    for (uint32_t j = 0; j < d; ++j) q = fma((double)beta[j], (double)beta[j], q);
    if (g_variant & 4) q = 0.0; /* planted mistake (tests only) */
    return ll - 0.5 * prior_prec * q;
}

/* One GLM half-step: the proposals of stretch_one, log posteriors kept in
 * double, accept iff u2 <= z^(d-1) exp(beta * (float)(lp(Y) - lp(X))). */
BHO_EXPORT uint32_t bho_glm_stretch_half(uint32_t seed, uint32_t odd_or_even, uint32_t step,
                                         uint32_t K, uint32_t dim, const float* Scompl, float* X,
                                         double* lp, float a, float beta, int partner_mode,
                                         const float* Xd, const float* y, uint64_t N,
                                         double prior_prec, uint8_t* accepted, int threads) {
    uint32_t total = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = 1;
#pragma omp parallel num_threads(threads) reduction(+ : total)
#endif
    {
        float* Y = (float*)malloc(sizeof(float) * dim);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t kk = 0; kk < (int64_t)K; ++kk) {
            const uint32_t k = (uint32_t)kk;
            float z;
            const float u2 = propose_one(k, seed, odd_or_even, step, K, dim, Scompl, X, a,
                                         partner_mode, Y, &z);
            const double lpy = bho_glm_logpost(Xd, y, N, dim, prior_prec, Y);
            const float q = isfinite(lpy)
                                ? pown_f(z, (int)dim - 1) * expf(beta * (float)(lpy - lp[k]))
                                : 0.0f;
            const int acc = u2 <= q;
            if (acc) {
                for (uint32_t i = 0; i < dim; ++i) X[k * dim + i] = Y[i];
                lp[k] = lpy;
            }
            if (accepted) accepted[k] = (uint8_t)acc;
            total += (uint32_t)acc;
        }
        free(Y);
    }
    (void)threads;
    return total;
}
