/*
 * bayadera_hip.h -- C-ABI of the MI355X engine for Bayadera's stretch-sampler
 * path.
 *
 * This is the boundary a host binds through its FFI (JNA from Clojure,
 * ctypes from Python): plain C types, opaque handles, caller-owned arrays.
 * It replaces the OpenCL engine of the reference,
 * src/clojure/uncomplicate/bayadera/internal/device/amd_gcn.clj (cited below
 * as amd_gcn.clj), behind the protocols of
 * src/clojure/uncomplicate/bayadera/internal/protocols.clj (protocols.clj).
 *
 * Conventions
 *  - Every function returns BH_OK (0) or a negative BH_ERR_* code; the
 *    calling thread's last error message is bh_last_error().
 *  - Matrices are column-major DIM x n blocks, as neanderthal's `ge`
 *    (element (d, j) at a[j*ld + d]).
 *  - Array arguments may be host or device pointers (the engine asks the HIP
 *    runtime which); results land where the pointer points.
 *  - Each handle owns one HIP stream; calls on one handle are ordered and a
 *    call returns after its results are written.  Handles are not
 *    thread-safe (as the reference's engines are not, amd_gcn.clj:273-274).
 *  - Model sources are the reference's C99 plugin dialect (REAL, native_exp,
 *    native_log, pown, lgamma, NAN), compiled at run time with hipRTC, as
 *    the reference compiles OpenCL C at run time (amd_gcn.clj:573-631).
 */
#ifndef BAYADERA_HIP_H
#define BAYADERA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BH_OK 0
#define BH_ERR_ARG (-1)     /* invalid argument (dragan-says-ex / IllegalArgumentException) */
#define BH_ERR_HIP (-2)     /* HIP runtime error */
#define BH_ERR_RTC (-3)     /* model source failed to compile */
#define BH_ERR_RANGE (-4)   /* size rule violated (walker count, acor series too short) */
#define BH_ERR_STATE (-5)   /* call out of order (e.g. move before init-move) */

#define BH_PARTNER_FLAT 0   /* the reference's partner offset (amd-gcn-mcmc-stretch.cl:57) */
#define BH_PARTNER_WALKER 1 /* whole-walker partner (Goodman-Weare as published) */

#define BH_EXEC_AUTO 0      /* resident kernel when the ensemble fits in LDS */
#define BH_EXEC_STEPWISE 1  /* two kernels per step, as the reference */
#define BH_EXEC_RESIDENT 2  /* one work-group runs many steps from LDS */

/* GLM posteriors (a capability the reference does not have: its models
 * loop over data rows serially inside each walker's LOGFN, e.g.
 * distributions/gaussian.cl:29-42).  Precision of the contraction X.B: */
#define BH_GLM_LOGISTIC 1   /* y_i ~ Bernoulli(sigmoid(x_i . beta)), beta ~ N(0, s^2 I) */
#define BH_GLM_GAUSSIAN 2   /* y_i ~ N(x_i . beta, sigma^2), beta ~ N(0, s_b^2 I), log sigma ~ N(0, s_s^2) */
#define BH_GLM_FP64 0       /* fp64 VALU: the accuracy reference */
#define BH_GLM_FP32 1       /* plain fp32 VALU contraction and softplus */
#define BH_GLM_BF16 2       /* MFMA, bf16 x bf16 -> fp32, one product */
#define BH_GLM_BF16X3 3     /* MFMA, proposals split in three bf16 terms (fp32-accurate) */
#define BH_GLM_FP16X2 4     /* MFMA on an fp16 copy of X, proposals split in two fp16 terms
                               (fp32-accurate; refused when |x| > 65504 somewhere; values
                               below 2^-17 in magnitude keep an absolute error <= 2^-25) */
#define BH_GLM_FP16 5       /* MFMA on the fp16 copy of X, proposals as one fp16 term, the
                               log posterior corrected to second order for that rounding
                               (the softplus sum's gradient and Hessian at a reference point
                               of the ensemble); single samplers only */
/* The split precisions (BF16, BF16X3, FP16X2) carry the first-order part of
 * that correction (the gradient term). */

typedef struct bh_ctx bh_ctx;
typedef struct bh_dist_engine bh_dist_engine;
typedef struct bh_lik_engine bh_lik_engine;
typedef struct bh_stretch_factory bh_stretch_factory;
typedef struct bh_stretch bh_stretch;

/* Message of the last failed call on this thread ("" if none). */
const char* bh_last_error(void);
/* ABI version: major*10000 + minor*100 + patch. */
int bh_version(void);

/* Compile a model source with the engine kernels for `arch` (e.g. "gfx950")
 * without a device: the check a host runs before creating engines (the
 * reference surfaces the same errors from build-program!, amd_gcn.clj:585-627).
 * Any of the three function names may be NULL; dim applies to logfn. */
int bh_rtc_compile_check(const char* source, const char* logfn_name, const char* logpdf_name,
                         const char* loglik_name, uint32_t dim, uint32_t wgs, const char* arch,
                         char* log, size_t log_size);

/* ---------------------------------------------------------------------- */
/* Factory: GCNBayaderaFactory / gcn-bayadera-factory (amd_gcn.clj:635-678) */
/* ---------------------------------------------------------------------- */
/* wgs = the reference's WGS: work-group size, histogram bin count and
 * walker-count quantum (amd_gcn.clj:516, 678); a power of two in [64, 1024]. */
int bh_ctx_create(int device, uint32_t wgs, bh_ctx** out);
int bh_ctx_destroy(bh_ctx* ctx);
/* compute units and WGS; processing-elements = cu * wgs (amd_gcn.clj:663-664) */
int bh_ctx_info(const bh_ctx* ctx, int* compute_units, uint32_t* wgs);
/* The device's streaming copy rate (no reference counterpart): read plus
 * written bytes per second (GB/s) of a plain copy kernel of `bytes` each way
 * (four 16-byte pieces per thread in flight; 10 launches after a warm-up),
 * and the same with non-temporal loads and stores; either output may be
 * NULL.  The measured HBM denominator beside the published 8 TB/s. */
int bh_device_copy_rate(bh_ctx* ctx, uint64_t bytes, double* gbps, double* gbps_nt);
/* Random 16-byte gathers per second (no reference counterpart) from a
 * buffer of `bytes` (16 x a power of two, at most 4 GiB), one per thread, 5
 * launches after a warm-up; nontemporal != 0 issues them as non-temporal
 * loads.  The stretch move's partner reads at large ensembles are such
 * gathers. */
int bh_device_gather_rate(bh_ctx* ctx, uint64_t bytes, int nontemporal, double* gathers_per_s);
/* Measured rates of the device (no reference counterpart): lane operations
 * per second of v_exp_f32 (the transcendental of the GLM likelihood's
 * softplus epilogue) and of v_fma_f32, flops per second of
 * v_mfma_f32_16x16x32_f16 (the likelihood's product), each from one timed
 * launch of independent chains on all compute units, and random 16-byte
 * gathers per second from a 128 MiB buffer (the stretch move's partner
 * reads; each costs a memory line).  The practical rates against which the
 * likelihood and move kernels are read; a few ms in all.  Any output may be
 * NULL (the gathers are skipped then). */
int bh_device_rates(bh_ctx* ctx, double* exp2_per_s, double* fma_per_s,
                    double* mfma_f16_flops_per_s, double* gathers_per_s);

/* ---------------------------------------------------------------------- */
/* Direct sampler: GCNDirectSamplerEngine.sample for the uniform model      */
/* (amd_gcn.clj:49-62, uniform-sampler.cl:19-34).  n a multiple of 4.      */
/* ---------------------------------------------------------------------- */
int bh_uniform_sample(bh_ctx* ctx, uint32_t seed, float lower, float upper, float* out,
                      uint64_t n);
/* The other direct samplers of GCNDirectSamplerEngine (rng/NAME-sampler.cl):
 * kind BH_DIRECT_GAUSSIAN params {mu, sigma}, BH_DIRECT_EXPONENTIAL {lambda},
 * BH_DIRECT_ERLANG {lambda, k}, BH_DIRECT_UNIFORM {lower, upper}. */
#define BH_DIRECT_UNIFORM 0
#define BH_DIRECT_GAUSSIAN 1
#define BH_DIRECT_EXPONENTIAL 2
#define BH_DIRECT_ERLANG 3
int bh_direct_sample(bh_ctx* ctx, int kind, uint32_t seed, const float* params,
                     uint32_t params_count, float* out, uint64_t n);

/* ---------------------------------------------------------------------- */
/* Distribution engine: gcn-distribution-engine + GCNDistributionEngine      */
/* (amd_gcn.clj:599-605, 66-93).  logpdf_name has the signature             */
/* REAL f(uint data_len, uint params_len, const REAL* params, uint dim,     */
/*        const REAL* x).  data_len = params_count - params_len (:75-77).    */
/* ---------------------------------------------------------------------- */
int bh_dist_engine_create(bh_ctx* ctx, const char* source, const char* logpdf_name,
                          uint32_t params_len, bh_dist_engine** out, char* log, size_t log_size);
int bh_dist_engine_destroy(bh_dist_engine* e);
/* DensityEngine.log-density / density: res[j] for the n columns of x (dim x n) */
int bh_log_density(bh_dist_engine* e, const float* params, uint32_t params_count, const float* x,
                   uint32_t dim, uint64_t n, float* res);
int bh_density(bh_dist_engine* e, const float* params, uint32_t params_count, const float* x,
               uint32_t dim, uint64_t n, float* res);

/* ---------------------------------------------------------------------- */
/* Likelihood engine: gcn-likelihood-engine + GCNLikelihoodEngine           */
/* (amd_gcn.clj:607-613, 97-133).  loglik_name has the signature            */
/* REAL f(uint data_len, const REAL* data, uint dim, const REAL* x).        */
/* ---------------------------------------------------------------------- */
int bh_lik_engine_create(bh_ctx* ctx, const char* source, const char* loglik_name,
                         bh_lik_engine** out, char* log, size_t log_size);
int bh_lik_engine_destroy(bh_lik_engine* e);
int bh_loglik(bh_lik_engine* e, const float* data, uint32_t data_len, const float* x,
              uint32_t dim, uint64_t n, float* res);
int bh_lik(bh_lik_engine* e, const float* data, uint32_t data_len, const float* x, uint32_t dim,
           uint64_t n, float* res);
/* LikelihoodEngine.evidence: mean over the n columns of exp(loglik), FP64
 * accumulator (amd_gcn.clj:122-133, 580) */
int bh_evidence(bh_lik_engine* e, const float* data, uint32_t data_len, const float* x,
                uint32_t dim, uint64_t n, double* out);

/* ---------------------------------------------------------------------- */
/* Dataset engine: GCNDatasetEngine (amd_gcn.clj:137-214)                   */
/* ---------------------------------------------------------------------- */
/* DatasetEngine.data-mean / data-variance (population variance, divide by
 * n, :159, 180), FP64 accumulation; out has m floats. */
int bh_dataset_mean(bh_ctx* ctx, const float* data, uint32_t m, uint64_t n, uint32_t ld,
                    float* out);
int bh_dataset_variance(bh_ctx* ctx, const float* data, uint32_t m, uint64_t n, uint32_t ld,
                        float* out);
/* bh_dataset_variance without allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  data and
 * out must be device pointers (nothing is staged: BH_ERR_ARG otherwise) and
 * m <= 256 (BH_ERR_ARG beyond).  The partial sums and row means live in one
 * static device array per GPU (4 MiB, held until the process ends) instead
 * of allocated buffers; the result equals bh_dataset_variance's bit for bit.
 * Every call on a GPU shares that GPU's array, so calls that overlap on one
 * GPU (from several threads or contexts) race. */
int bh_dataset_variance_noalloc(bh_ctx* ctx, const float* data, uint32_t m, uint64_t n,
                                uint32_t ld, float* out);
/* EstimateEngine.histogram: Histogram{limits 2 x m, pdf WGS x m, bin-ranks
 * WGS x m} (protocols.clj:13-18); WGS bins per row (amd-gcn-estimate.cl:1-25). */
int bh_dataset_histogram(bh_ctx* ctx, const float* data, uint32_t m, uint64_t n, uint32_t ld,
                         float* limits, float* pdf, float* bin_ranks);

/* ---------------------------------------------------------------------- */
/* Autocorrelation: GCNAcorEngine.acor (amd_gcn.clj:216-267)                */
/* series dim x n column-major (the per-step ensemble means).  Returns      */
/* BH_ERR_RANGE when n < 5*lag (:264-267).  lag may be NULL.                */
/* ---------------------------------------------------------------------- */
int bh_acor(bh_ctx* ctx, const float* series, uint32_t dim, uint32_t n, float* tau, float* mean,
            float* sigma, uint32_t* lag);
/* acor of many series of n steps in one launch (no reference counterpart;
 * each series is what bh_acor computes for it alone, amd_gcn.clj:216-267):
 * series p is an n x dims[p] block (row-major, step by step) after those of
 * series 0 .. p-1 (the layout bh_batch_run_series records); tau, mean and
 * sigma receive dims[0] + .. + dims[problems-1] values in the same order.
 * BH_ERR_RANGE when n < 5 * lag, as bh_acor. */
int bh_acor_batch(bh_ctx* ctx, const float* series, uint32_t problems, const uint32_t* dims,
                  uint32_t n, float* tau, float* mean, float* sigma, uint32_t* lag);

/* ---------------------------------------------------------------------- */
/* Stretch sampler factory: gcn-stretch-factory (amd_gcn.clj:623-631) with  */
/* LOGFN = logfn_name (the model's mcmc-logpdf) and DIM = dim.              */
/* ---------------------------------------------------------------------- */
int bh_stretch_factory_create(bh_ctx* ctx, const char* source, const char* logfn_name,
                              uint32_t dim, uint32_t params_len, bh_stretch_factory** out,
                              char* log, size_t log_size);
int bh_stretch_factory_destroy(bh_stretch_factory* f);

/* A stretch-sampler factory whose LOGFN is a GLM posterior evaluated for a
 * whole half-ensemble at once (no reference counterpart; it sits behind the
 * same SamplerFactory.create-sampler, amd_gcn.clj:514-567).  X: N x d bf16
 * (IEEE bfloat16 bit patterns, row-major), y: N labels in {0, 1}; d is 32
 * or 64, N < 2^31; the data stay in HBM until the factory is destroyed. */
int bh_glm_stretch_factory_create(bh_ctx* ctx, int kind, const void* X_bf16, const float* y,
                                  uint64_t N, uint32_t d, float prior_sd,
                                  bh_stretch_factory** out, char* log, size_t log_size);
/* Log posteriors (likelihood + prior, FP64) of n points beta (d x n). */
/* Linear regression declared by form (no reference counterpart; the
 * reference runs it as a user model whose LOGFN loops over the rows,
 * posterior.cl:1-6): walkers x = [beta (d), log sigma], DIM = d + 1 <= 65;
 * X fp32 N x d row-major, y fp32 N.  The likelihood of a half-ensemble is
 * one data-parallel pass computing sum_i (y_i - x_i . beta)^2 per walker;
 * precisions BH_GLM_FP32 (the default) and BH_GLM_FP64. */
int bh_glm_gaussian_factory_create(bh_ctx* ctx, const float* X, const float* y, uint64_t N,
                                   uint32_t d, float beta_sd, float log_sigma_sd,
                                   bh_stretch_factory** out, char* log, size_t log_size);
int bh_glm_log_density(bh_stretch_factory* f, const float* beta, uint64_t n, int precision,
                       double* out);

/* SamplerFactory.create-sampler (amd_gcn.clj:514-567): walkers a multiple of
 * 2*WGS (BH_ERR_RANGE otherwise); params = [data, hyperparameters]; calls
 * init! with seed. */
int bh_stretch_create(bh_stretch_factory* f, uint32_t seed, uint32_t walkers,
                      const float* params, uint32_t params_count, bh_stretch** out);
int bh_stretch_destroy(bh_stretch* s);
int bh_stretch_set_partner(bh_stretch* s, int partner_mode);
int bh_stretch_set_exec(bh_stretch* s, int exec_mode);
/* GLM samplers: precision of the likelihood contraction (BH_GLM_*). */
int bh_stretch_set_glm_precision(bh_stretch* s, int precision);
/* GLM samplers, reduced precisions: guarded accept test (no reference
 * counterpart; the test it guards is amd-gcn-mcmc-stretch.cl:66-71).  A
 * walker whose log margin log u2 - (DIM-1) log z - beta (lp(Y) - lp(X)) is
 * within guard * max(1, beta) of zero is decided on fp64 log posteriors of
 * Y and X instead; while the reduced-precision error of lp(Y) - lp(X) stays
 * below the guard, every decision equals the fp64 sampler's.  guard = 0
 * (the default) turns it off.  Every walker inside the guard is
 * re-evaluated (the list holds a whole half-ensemble), so no decision inside
 * the guard falls back to the reduced precision.  At BH_GLM_FP16 the
 * guard is widened per walker by max(1, q / (2 DIM)), q the larger squared
 * H0-norm of Y's and X's distances from the correction's reference point
 * (its remainder grows with that distance). */
int bh_stretch_set_glm_guard(bh_stretch* s, double guard);
/* Walkers re-evaluated in fp64, summed over all guarded half-steps of the
 * sampler; overflow counts listed walkers left to the reduced precision,
 * which the whole-half list keeps at 0. */
int bh_stretch_glm_guard_stats(bh_stretch* s, uint64_t* listed, uint64_t* overflow);
/* Info.info: {walker-count, iteration-counter} (amd_gcn.clj:321-324) */
int bh_stretch_info(const bh_stretch* s, uint32_t* walkers, uint64_t* iteration_counter);

/* MCMC (protocols.clj:97-103) */
int bh_stretch_init(bh_stretch* s, uint32_t seed);                               /* init! :378-387 */
int bh_stretch_init_position(bh_stretch* s, uint32_t seed, const float* limits); /* :394-402, limits 2 x DIM */
int bh_stretch_init_position_from(bh_stretch* s, const bh_stretch* position);    /* :388-393 */
int bh_stretch_burn_in(bh_stretch* s, uint64_t n, float a);                      /* :403-411 */
/* anneal!: temperatures[i] is the schedule's value at step i (:412-420) */
int bh_stretch_anneal(bh_stretch* s, const double* temperatures, uint64_t n, float a);
/* run-sampler! (:421-443): acceptance rate, the per-step ensemble means
 * (dim x n, may be NULL) and their autocorrelation (tau, mean, sigma: dim
 * floats each, lag).  BH_ERR_RANGE when n is too short for acor. */
int bh_stretch_run(bh_stretch* s, uint64_t n, float a, double* acceptance_rate, float* means,
                   float* tau, float* mean, float* sigma, uint32_t* lag);
int bh_stretch_acc_rate(bh_stretch* s, float a, double* out);                    /* :444-453 */

/* MCMCStretch (protocols.clj:105-109) */
int bh_stretch_init_move(bh_stretch* s, float a, uint64_t max_steps);            /* :329-339 */
int bh_stretch_move(bh_stretch* s);                                              /* :340-346 */
int bh_stretch_move_bare(bh_stretch* s);                                         /* :347-353 */
int bh_stretch_set_temperature(bh_stretch* s, double t);                         /* :354-358 */
/* the accounting buffers of the last init-move: accept[K/WGS] and
 * means_acc[DIM*(K/WGS)*steps] (either may be NULL) */
int bh_stretch_read_accounting(bh_stretch* s, uint32_t* accept, float* means_acc,
                               uint64_t steps);

/* RandomSampler.sample! (:360-376): out is DIM x ncols. */
int bh_stretch_sample(bh_stretch* s, float* out, uint64_t ncols);
/* EstimateEngine.histogram! (:457-487) */
int bh_stretch_histogram(bh_stretch* s, uint32_t cycles, float* limits, float* pdf,
                         float* bin_ranks);
/* Location.mean / Spread.variance of the current walkers (:489-505) */
int bh_stretch_mean(bh_stretch* s, float* out);
int bh_stretch_variance(bh_stretch* s, float* out);

/* GPU time of the sampler's own stream: with timing on, burn-in!, anneal!,
 * run-sampler!, sample!, move! and move-bare! record HIP events around their
 * kernels, and bh_stretch_last_gpu_ms returns the elapsed milliseconds of
 * the last such call (run-sampler!: the steps only).  No reference
 * counterpart: the reference has no timers (amd_gcn.clj:321-324). */
int bh_stretch_set_timing(bh_stretch* s, int on);
int bh_stretch_last_gpu_ms(const bh_stretch* s, double* ms);
/* Kernel timing (no reference counterpart): with on = n >= 1, every n-th
 * launch of the sampler's dominant kernel (the stretch-move kernels, the
 * resident kernel, or a GLM sampler's likelihood contraction) is bracketed
 * by HIP events on the sampler's stream, the first 1024 such launches after
 * the call that turned it on (each pair of events costs the stream ~10 us);
 * on = 0 turns it off.  bh_stretch_kernel_timing waits for the stream and
 * returns their average duration and count.  Turning it on again restarts
 * the count. */
int bh_stretch_set_kernel_timing(bh_stretch* s, int on);
int bh_stretch_kernel_timing(bh_stretch* s, double* avg_ms, uint32_t* launches);

/* ---------------------------------------------------------------------- */
/* Batched posteriors (no reference counterpart: the reference runs one      */
/* posterior per sampler).  `problems` independent posteriors of one model   */
/* source, problem b with dimension dims[b] <= max_dim and params            */
/* params[offsets[b] .. offsets[b+1]) of which the last hyper_len[b] are     */
/* hyperparameters (data_len = the rest, as create-sampler splits them).     */
/* Every problem has `walkers` walkers (a multiple of 128) and runs the     */
/* single-sampler algorithm with bare seed seed0 + 2b; one launch runs all   */
/* problems, one work-group each, for all steps of a call, its walkers in   */
/* LDS or in place in global memory (bh_batch_set_state).                   */
/* ---------------------------------------------------------------------- */
typedef struct bh_batch bh_batch;
#define BH_BATCH_STATE_LDS 0    /* positions copied into LDS for the call */
#define BH_BATCH_STATE_GLOBAL 1 /* positions worked on in place (L2-resident) */
int bh_batch_create(bh_ctx* ctx, const char* source, const char* logfn_name, uint32_t max_dim,
                    uint32_t problems, const uint32_t* dims, const uint64_t* offsets,
                    const uint32_t* hyper_len, const float* params, uint32_t walkers,
                    bh_batch** out, char* log, size_t log_size);
int bh_batch_destroy(bh_batch* b);
int bh_batch_set_partner(bh_batch* b, int partner_mode);
/* init-position! of every problem: key {seed + b, 0xdecafaaa}; limits
 * problems x max_dim x 2 (rows beyond a problem's dim are ignored). */
int bh_batch_init_position(bh_batch* b, uint32_t seed, const float* limits);
int bh_batch_burn_in(bh_batch* b, uint32_t seed0, uint64_t n, float a);
/* Where every launch class keeps its walkers during a call
 * (BH_BATCH_STATE_LDS or BH_BATCH_STATE_GLOBAL); BH_ERR_RANGE when a class
 * does not fit the LDS.  Created batches keep them in LDS when that fits.
 * Positions do not depend on it.  No reference counterpart. */
int bh_batch_set_state(bh_batch* b, int where);
/* Chooses the launch shape of the batch by timing `steps` bare steps of each
 * candidate from the current state (restored afterwards): threads per
 * problem (walkers per work-group pass; 64 .. 1024), launch classes
 * (problems split by dimension into 1-4 launches, each with the LDS its
 * largest problem needs) and, per class, the state placement that runs it
 * faster.  Positions do not depend on the shape; the accounting sums'
 * rounding does (each thread sums its own walker residue).  No reference
 * counterpart. */
int bh_batch_autotune(bh_batch* b, uint32_t steps, uint32_t* threads_per_problem,
                      uint32_t* launch_classes, double* best_ms);
/* The state placement of each launch class, smallest dimensions first: on
 * entry *count is the capacity of states (may be 0 with states NULL), on
 * return the number of classes.  No reference counterpart. */
int bh_batch_class_states(const bh_batch* b, uint32_t* count, int* states);
/* n steps with accounting: per problem the acceptance rate and, over the n
 * steps and all walkers, the posterior mean and sd of each coordinate
 * (problems x max_dim; NULL to skip). */
int bh_batch_run(bh_batch* b, uint32_t seed0, uint64_t n, float a, double* acceptance_rate,
                 double* mean, double* sd);
/* bh_batch_run that also records each step's ensemble mean of every problem
 * (the series acor takes, amd_gcn.clj:437-439 for one sampler): problem p's
 * n x dims[p] block (row-major, step by step) starts at
 * series + n * (dims[0] + .. + dims[p-1]).  series is a device or host
 * pointer; n < 2^20. */
int bh_batch_run_series(bh_batch* b, uint32_t seed0, uint64_t n, float a,
                        double* acceptance_rate, double* mean, double* sd, float* series);
/* positions (problems x walkers x max_dim slots, each problem packed by its
 * own dim) and logfn values (problems x walkers) */
int bh_batch_state_get(const bh_batch* b, float* xs, float* logfn);
/* GPU time of the last burn-in or run call (HIP events on the batch's stream) */
int bh_batch_last_gpu_ms(const bh_batch* b, double* ms);

/* ---- Batched GLM posteriors (no reference counterpart: many independent
 * logistic regressions, each with its own rows N_p and dimension d_p <= 64,
 * advanced in lockstep; the likelihood of every problem is one grouped
 * contraction).  X is fp8 e4m3 (OCP: bias 7, max 448, S.1111.111 NaN) with
 * one fp32 scale per column: x_ij = scale_j * fp8(q_ij).  Arguments are
 * concatenated over problems: X_fp8 the N_p x d_p row-major bytes, col_scales
 * the d_p scales, y the N_p labels (0/1).  Prior beta ~ N(0, prior_sd^2).
 * Each problem runs the single GLM sampler's moves (move-bare! arithmetic,
 * amd-gcn-mcmc-stretch.cl:31-81) with whole-walker partners and seeds
 * seed0 + 2p / seed0 + 2p + 1 for its two halves.  walkers: a multiple of
 * 512.  Precision as bh_stretch_set_glm_precision (BH_GLM_* up to
 * BH_GLM_FP16X2; the batch does not take BH_GLM_FP16). */
typedef struct bh_glm_batch bh_glm_batch;
int bh_glm_batch_create(bh_ctx* ctx, int kind, uint32_t problems, const uint64_t* rows,
                        const uint32_t* dims, const void* X_fp8, const float* col_scales,
                        const float* y, float prior_sd, uint32_t walkers, bh_glm_batch** out);
int bh_glm_batch_destroy(bh_glm_batch* b);
int bh_glm_batch_set_precision(bh_glm_batch* b, int precision);
/* Tile shape of the matrix-core contraction for the problems of one row
 * stride (32: d_p <= 32; 64: d_p > 32), no reference counterpart: the rows
 * per unit (a multiple of 16 in [256, 65536]; each unit is one work-group's
 * range of a problem's rows) and the 16-point column blocks per wave (4, or
 * 8 for stride 32; a work-group takes 64 x col_blocks points; needs
 * walkers / 2 a multiple of 64 x col_blocks).  The stored log posteriors are
 * recomputed over the new units; contraction values are the same for every
 * shape, their sums over units differ only in rounding order. */
int bh_glm_batch_set_tiles(bh_glm_batch* b, uint32_t row_stride, uint32_t unit_rows,
                           uint32_t col_blocks);
/* The batch's own tile autotuner (no reference counterpart): for each row
 * stride, times `reps` launches of the half-step contraction (walkers / 2
 * points per problem, the batch's precision, its current positions) for
 * every candidate shape (unit rows 1024 .. 16384, column blocks 4 and, for
 * stride 32 at matrix-core precisions, 8) with HIP events, and keeps the
 * fastest.  unit_rows, col_blocks and best_ms (index 0: stride 32, 1: stride
 * 64; best_ms 0 for a stride with no problems) may be NULL.  Rebuilds the
 * units and recomputes the stored log posteriors. */
int bh_glm_batch_autotune(bh_glm_batch* b, uint32_t reps, uint32_t* unit_rows,
                          uint32_t* col_blocks, float* best_ms);
/* limits: problems x 64 x 2 (lo, hi per coordinate; rows past d_p unused);
 * problem p's walkers from init_walkers' stream with seed + p
 * (amd-gcn-mcmc-stretch.cl:138-159), then their log posteriors. */
int bh_glm_batch_init_position(bh_glm_batch* b, uint32_t seed, const float* limits);
int bh_glm_batch_burn_in(bh_glm_batch* b, uint32_t seed0, uint64_t n, float a);
/* n accounted steps: per problem the acceptance rate and, over the steps and
 * walkers, the mean and sd of every coordinate (problems x 64), and when
 * series is not NULL each step's ensemble means, problem p's n x d_p block
 * at series + n * (d_0 + .. + d_{p-1}). */
int bh_glm_batch_run(bh_glm_batch* b, uint32_t seed0, uint64_t n, float a,
                     double* acceptance_rate, double* mean, double* sd, float* series);
/* positions (problems x walkers x 64, padded) and log posteriors (double) */
int bh_glm_batch_state_get(const bh_glm_batch* b, float* xs, double* lp);
/* Log posteriors of the current positions in the given precision
 * (problems x walkers, double); the sampler's own values are kept. */
int bh_glm_batch_log_density(bh_glm_batch* b, int precision, double* lp);
int bh_glm_batch_last_gpu_ms(const bh_glm_batch* b, double* ms);
/* The device's fp8 e4m3 decoding of n bytes (n a multiple of 4). */
int bh_fp8_decode(bh_ctx* ctx, const void* in, float* out, uint64_t n);
/* The fp16 matrix-core path's decoding of the same bytes (fp16 values of
 * x * 2^-8 on the device, returned times 2^8 as floats). */
int bh_fp8_decode_f16(bh_ctx* ctx, const void* in, float* out, uint64_t n);

/* Sampler state for resume: positions (DIM x W), logfn (W), and the host
 * counters {bare seed, move seed, bare counter, move counter, iteration
 * counter low, high} (amd_gcn.clj:273-274, 330-334, 385-386). */
int bh_stretch_state_get(const bh_stretch* s, float* xs, float* logfn, uint32_t counters[6]);
int bh_stretch_state_set(bh_stretch* s, const float* xs, const float* logfn,
                         const uint32_t counters[6]);

#ifdef __cplusplus
}
#endif

#endif /* BAYADERA_HIP_H */
