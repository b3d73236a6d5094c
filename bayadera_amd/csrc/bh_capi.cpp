// C-ABI of the HIP engine (include/bayadera_hip.h): handle types, hipRTC
// compilation of model sources together with the engine kernels, and the
// host loops of the stretch sampler (GCNStretch, amd_gcn.clj:271-567).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <array>
#include <vector>

#include "../../include/bayadera_hip.h"
#include "bh_glm.hpp"
#include "bh_glm_batch.hpp"
#include "bh_kernels.hpp"

// Sources handed to hipRTC, embedded at build time as raw string literals.
#include "bh_rtc_sources.inc"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define BH_HIP(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(BH_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                             \
    } while (0)

#define BH_TRY(expr)              \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != BH_OK) return rc_; \
    } while (0)

#define BH_ARG(cond, ...)                                  \
    do {                                                   \
        if (!(cond)) return fail(BH_ERR_ARG, __VA_ARGS__); \
    } while (0)

bool is_device_ptr(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Device buffer owned for the duration of a call.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n) {
        if (n <= bytes && p) return BH_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        BH_HIP(hipMalloc(&p, n ? n : 1));
        bytes = n;
        return BH_OK;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// An input array as a device pointer: used in place when it already is one,
// else copied up on the stream.
struct DevIn {
    const void* dev = nullptr;
    DevBuf own;
    int init(const void* src, size_t bytes, hipStream_t s) {
        if (is_device_ptr(src)) {
            dev = src;
            return BH_OK;
        }
        BH_TRY(own.alloc(bytes));
        BH_HIP(hipMemcpyAsync(own.p, src, bytes, hipMemcpyHostToDevice, s));
        dev = own.p;
        return BH_OK;
    }
    template <class T>
    const T* as() const {
        return static_cast<const T*>(dev);
    }
};

// An output array: written in place when it is a device pointer, else
// staged and copied down by finish().
struct DevOut {
    void* dst = nullptr;
    void* dev = nullptr;
    size_t bytes = 0;
    DevBuf own;
    int init(void* d, size_t n) {
        dst = d;
        bytes = n;
        if (is_device_ptr(d)) {
            dev = d;
            return BH_OK;
        }
        BH_TRY(own.alloc(n));
        dev = own.p;
        return BH_OK;
    }
    int finish(hipStream_t s) {
        if (dev != dst) BH_HIP(hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToHost, s));
        return BH_OK;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(dev);
    }
};

// ------------------------------------------------------------------------
// hipRTC: model source + engine kernels -> code object (cached per process)
// ------------------------------------------------------------------------
std::mutex g_rtc_mu;
std::unordered_map<std::string, std::vector<char>> g_rtc_cache;

int rtc_compile(const std::string& arch, const std::string& model_source,
                const std::vector<std::string>& defines, std::vector<char>& code, char* log,
                size_t log_size) {
    std::string src;
    src += "#include \"bh_rtc_prelude.h\"\n";
    src += "#pragma clang force_cuda_host_device begin\n";
    src += model_source;
    src += "\n#pragma clang force_cuda_host_device end\n";
    src += kRtcEngineSource;
    std::string key = arch + "\x1f" + src;
    for (const auto& d : defines) key += "\x1f" + d;
    {
        std::lock_guard<std::mutex> lk(g_rtc_mu);
        auto it = g_rtc_cache.find(key);
        if (it != g_rtc_cache.end()) {
            code = it->second;
            if (log && log_size) log[0] = 0;
            return BH_OK;
        }
    }
    const char* hdr_src[] = {kRtcPhiloxSource, kRtcPreludeSource};
    const char* hdr_name[] = {"bh_philox.hpp", "bh_rtc_prelude.h"};
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "bh_model.hip", 2, hdr_src, hdr_name) !=
        HIPRTC_SUCCESS)
        return fail(BH_ERR_RTC, "hiprtcCreateProgram failed");
    std::vector<std::string> opts = {"--offload-arch=" + arch,
                                     "-O3",
                                     "-std=c++17",
                                     "-ffp-contract=on",
                                     "-fhip-fp32-correctly-rounded-divide-sqrt",
                                     "-Xclang",
                                     "-target-feature",
                                     "-Xclang",
                                     "-packed-fp32-ops"};
    for (const auto& d : defines) opts.push_back(d);
    std::vector<const char*> copts;
    for (const auto& o : opts) copts.push_back(o.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string plog(ls + 1, '\0');
    if (ls) hiprtcGetProgramLog(prog, &plog[0]);
    if (log && log_size) {
        snprintf(log, log_size, "%s", plog.c_str());
    }
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return fail(BH_ERR_RTC, "model source failed to compile: %.1500s", plog.c_str());
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    std::lock_guard<std::mutex> lk(g_rtc_mu);
    g_rtc_cache[key] = code;
    return BH_OK;
}

struct Module {
    hipModule_t mod = nullptr;
    ~Module() {
        if (mod) (void)hipModuleUnload(mod);
    }
    int load(const std::vector<char>& code) {
        BH_HIP(hipModuleLoadData(&mod, code.data()));
        return BH_OK;
    }
    int fn(const char* name, hipFunction_t* f) {
        BH_HIP(hipModuleGetFunction(f, mod, name));
        return BH_OK;
    }
};

int launch(hipFunction_t f, unsigned grid, unsigned block, unsigned lds, hipStream_t s,
           void** args) {
    BH_HIP(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, lds, s, args, nullptr));
    return BH_OK;
}

bool valid_name(const char* n) {
    if (!n || !*n) return false;
    for (const char* p = n; *p; ++p)
        if (!(isalnum((unsigned char)*p) || *p == '_')) return false;
    return true;
}

}  // namespace

// --------------------------------------------------------------------------
// Handle types
// --------------------------------------------------------------------------
struct bh_ctx {
    int device = 0;
    uint32_t wgs = 256;
    int cu = 0;
    std::string arch;
    hipStream_t stream = nullptr;
    DevBuf scratch;
};

struct bh_dist_engine {
    bh_ctx* ctx = nullptr;
    uint32_t params_len = 0;
    Module mod;
    hipFunction_t logpdf = nullptr, pdf = nullptr;
};

struct bh_lik_engine {
    bh_ctx* ctx = nullptr;
    Module mod;
    hipFunction_t loglik = nullptr, lik = nullptr, evidence = nullptr;
};

// A GLM posterior's data, resident in HBM for the factory's lifetime.
struct GlmData {
    int kind = 0;
    unsigned long long N = 0;
    uint32_t d = 0;
    uint32_t D = 0;           // linear regression: padded column count of Xf
    double prior_prec = 1.0;  // logistic: 1 / prior_sd^2
    DevBuf X;                 // logistic: bf16, N x d row-major
    DevBuf lin;               // logistic: double[d]: X'y; linear regression: the
                              // four constants of its log posterior
    DevBuf lin_half;          // logistic: double[d]: X'y - (1/2) X'1 (log2-unit kernels)
    DevBuf half_colsum;       // logistic: double[d]: (1/2) X'1
    DevBuf X16;               // logistic: fp16 copy of X in fragment order, made on first fp16 use
    mutable unsigned bpc[BH_GLM_FP16 + 1] = {};  // resident work-groups per CU, by precision
    DevBuf Xf, y;             // linear regression: fp32 N x D and N
};

// The linear coefficients of a precision's log-likelihood: the logistic
// fp64 kernel sums whole softplus values, the others leave out u/2 per row.
const double* glm_lin(const GlmData& g, int prec) {
    if (g.kind == BH_GLM_GAUSSIAN) return g.lin.as<double>();
    return (prec == BH_GLM_FP64 ? g.lin : g.lin_half).as<double>();
}

// The likelihood contraction of n points P (row stride = the walker
// dimension) into partial[nranges][n].
int glm_contract(const GlmData& g, int prec, hipStream_t st, const float* P, unsigned pstride,
                 const void* ys, unsigned n, unsigned nranges, unsigned long long rpr,
                 float yscale, double* partial) {
    if (g.kind == BH_GLM_GAUSSIAN) {
        BH_HIP(bh::launch_glm_gauss(prec, st, g.Xf.as<float>(), g.y.as<float>(), g.N, g.D, P,
                                    pstride, g.d, n, nranges, rpr, partial));
    } else {
        const void* X = (prec == BH_GLM_FP16X2 || prec == BH_GLM_FP16) ? g.X16.p : g.X.p;
        BH_HIP(bh::launch_glm_logistic(prec, st, X, g.N, g.d, P, ys, n, nranges, rpr, yscale,
                                       partial));
    }
    return BH_OK;
}

struct bh_stretch_factory {
    bh_ctx* ctx = nullptr;
    uint32_t dim = 0, params_len = 0;
    Module mod;
    hipFunction_t logfn = nullptr, bare = nullptr, accu = nullptr, resident = nullptr;
    std::unique_ptr<GlmData> glm;
    hipFunction_t glm_propose = nullptr, glm_split = nullptr, glm_logpost = nullptr,
                  glm_accept = nullptr, glm_screen = nullptr, glm_resolve = nullptr,
                  glm_accept_split = nullptr;
};

struct bh_batch {
    bh_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    Module mod;
    hipFunction_t run = nullptr, run_g = nullptr, init = nullptr, logfn = nullptr;
    uint32_t max_dim = 0, problems = 0, W = 0, nb = 0, dim_hi = 0;
    int lds_max = 0;
    // launch classes: problems order[begin .. begin + count) with dimensions
    // up to dim_hi, each launched with the LDS its largest problem needs
    struct Class {
        uint32_t begin, count, dim_hi;
        int gstate;  // BH_BATCH_STATE_LDS or BH_BATCH_STATE_GLOBAL
    };
    int gstate = BH_BATCH_STATE_LDS;  // placement of new classes
    std::vector<Class> classes;
    std::vector<uint32_t> dims_h;
    DevBuf order;
    uint32_t bare_counter = 0;
    int partner = BH_PARTNER_FLAT;
    DevBuf dims, poff, dlen, plen, params, xs, logfn_buf, accept, sums;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
};

struct bh_stretch {
    bh_stretch_factory* f = nullptr;
    bh_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    uint32_t dim = 0, W = 0, K = 0, wgs = 0, groups = 0;
    uint32_t params_count = 0, data_len = 0, params_len = 0;
    DevBuf xs, logfn, params, accept, means_acc, tmp;
    uint64_t means_cap = 0;
    uint32_t bare_seed = 0, move_seed = 0, bare_counter = 0, move_counter = 0;
    uint64_t iter = 0;
    float bare_a = 2.0f, beta = 1.0f, accu_a = 2.0f;
    int partner = BH_PARTNER_FLAT, exec = BH_EXEC_AUTO;
    bool move_ready = false;
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    // GLM samplers: log posteriors in double, proposal and partial-sum buffers
    int glm_prec = BH_GLM_BF16X3;
    DevBuf lp, gY, gz, gu2, gsplit, gpartial;
    // guarded accept test of the reduced precisions (guard 0: off)
    double glm_guard = 0.0;
    DevBuf gflag, glist, gcount, gtotals, gpart64, gpts64, glpy;
    // the split precisions' first-order correction (logistic): the softplus
    // sum's gradient at a reference point, its block partials, and each
    // point's correction
    DevBuf gg0, ggpart, gcorr;
    int g0_age = -1;  // steps since gg0 was computed; -1: not valid
    // the one-term fp16 precision's second-order correction: the Hessian at
    // the same reference point (d x d floats) and its work
    DevBuf gh0, ghpart;
    bool h0_valid = false;
    DevBuf gscale;  // per-walker widening of the guard (second order)
    // event pairs around the dominant kernel's launches (the stretch move
    // kernels, or the GLM likelihood contraction), up to the pool's size
    bool ktiming = false;
    std::vector<hipEvent_t> kev;
    size_t kused = 0;
    uint32_t kstride = 1, kcount = 0;  // every kstride-th launch is timed
};

constexpr size_t kKernelTimerPairs = 1024;

// rounds of work-groups per logistic likelihood launch
constexpr unsigned kGlmRounds = 4;

// Guarded accept test: the list holds every walker of a half (K), so no
// walker inside the guard is ever decided on reduced-precision values; the
// fp64 pass over the listed points uses up to kGuardRangesMax row ranges
// (one work-group each, ~4 rows per thread at 1M rows), fewer for very
// large halves so that its partial sums (ranges x 2K doubles) stay within
// kGuardPartialMax doubles.
constexpr unsigned kGuardRangesMax = 1024;
constexpr size_t kGuardPartialMax = (size_t)1 << 25;  // 256 MB of doubles
constexpr unsigned kGuardResolveBlocks = 16;

static unsigned guard_ranges(unsigned k_half) {
    size_t r = kGuardPartialMax / (2 * (size_t)k_half);
    r = std::min<size_t>(r, kGuardRangesMax) / 8 * 8;
    return (unsigned)std::max<size_t>(r, 8);
}

namespace {

int ctx_scratch(bh_ctx* c, size_t bytes, void** p) {
    BH_TRY(c->scratch.alloc(bytes));
    *p = c->scratch.p;
    return BH_OK;
}

int sync(hipStream_t s) {
    BH_HIP(hipStreamSynchronize(s));
    return BH_OK;
}

// Event pair around a call's kernels on the sampler's stream.
struct Timed {
    bh_stretch* s;
    explicit Timed(bh_stretch* st) : s(st) {
        if (s->timing) (void)hipEventRecord(s->ev0, s->stream);
    }
    int stop() {
        if (!s->timing) return BH_OK;
        BH_HIP(hipEventRecord(s->ev1, s->stream));
        BH_HIP(hipEventSynchronize(s->ev1));
        float ms = 0.0f;
        BH_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        s->last_ms = ms;
        return BH_OK;
    }
};

// Launches a dominant kernel between two events of the sampler's timing pool
// when kernel timing is on and the pool has room.
template <class F>
int ktimed(bh_stretch* s, F&& launch_fn) {
    const bool rec = s->ktiming && (s->kcount++ % s->kstride) == 0 &&
                     s->kused + 2 <= s->kev.size();
    if (rec) BH_HIP(hipEventRecord(s->kev[s->kused], s->stream));
    BH_TRY(launch_fn());
    if (rec) {
        BH_HIP(hipEventRecord(s->kev[s->kused + 1], s->stream));
        s->kused += 2;
    }
    return BH_OK;
}

// block size of the resident kernel: a multiple of WGS dividing K, <= 1024
unsigned resident_block(const bh_stretch* s) {
    unsigned nb = s->K <= 1024 ? s->K : 1024;
    while (nb > s->wgs && (s->K % nb) != 0) nb -= s->wgs;
    return nb;
}

size_t resident_lds(const bh_stretch* s, int mode) {
    const unsigned nb = resident_block(s);
    // positions, logfn values, and the accounting staging areas (the
    // kernel's BH_ACCU_CHUNK = min(DIM, 8) floats per thread, the odd half's
    // group sums and the groups' accept counts)
    const size_t chunk = s->dim < 8 ? s->dim : 8;
    const size_t groups = s->K / s->wgs;
    size_t floats = (size_t)s->W * s->dim + s->W +
                    (mode == 1 ? chunk * nb + groups * s->dim + groups : 0);
    return floats * 4 + (nb / 64) * 4;
}

constexpr size_t kResidentLdsMax = 64 * 1024;

bool use_resident(const bh_stretch* s, int mode) {
    if (s->exec == BH_EXEC_STEPWISE || s->f->glm) return false;
    const bool fits = resident_lds(s, mode) <= kResidentLdsMax;
    if (s->exec == BH_EXEC_RESIDENT) return fits;
    // a single work-group pays off while the ensemble is small
    return fits && s->W <= 4096;
}

int run_resident(bh_stretch* s, int mode, uint32_t seed_odd, uint32_t seed_even, uint32_t c_odd,
                 uint32_t c_even, float a, float beta, const float* d_betas, uint32_t step0,
                 uint32_t nsteps, float* sample_out, unsigned long long sample_floats) {
    unsigned data_len = s->data_len, params_len = s->params_len, K = s->K;
    const float* params = s->params.as<float>();
    float* xs = s->xs.as<float>();
    float* lf = s->logfn.as<float>();
    int partner = s->partner;
    unsigned* acc = mode == 1 ? s->accept.as<unsigned>() : nullptr;
    float* means = mode == 1 ? s->means_acc.as<float>() : nullptr;
    void* args[] = {&seed_odd, &seed_even, &c_odd,   &c_even,      &data_len, &params_len,
                    &params,   &xs,        &lf,      &K,           &a,        &beta,
                    &d_betas,  &step0,     &nsteps,  &partner,     &mode,     &acc,
                    &means,    &sample_out, &sample_floats};
    return ktimed(s, [&] {
        return launch(s->f->resident, 1, resident_block(s), (unsigned)resident_lds(s, mode),
                      s->stream, args);
    });
}

// ---------------------------------------------------------------- GLM
constexpr double kLn2 = 0.69314718055994530942;
constexpr float kLog2e = 1.44269504088896340736f;
// threads per walker of bh_glm_propose and bh_glm_screen (their
// BH_GLM_TPW, BH_GLM_STPW)
constexpr unsigned kGlmProposeThreads = 16;
constexpr unsigned kGlmScreenThreads = 8;
// steps between refreshes of the split precisions' correction gradient and
// Hessian (the ensemble's reference point moves little in 16 steps, and
// both are always taken at the same point; the correction needs g0 to
// ~1e-3 relative)
constexpr int kGlmGradRefresh = 16;

int glm_nsplit(int prec) {
    return prec == BH_GLM_BF16X3 ? 3
           : prec == BH_GLM_FP16X2 ? 2
           : (prec == BH_GLM_BF16 || prec == BH_GLM_FP16) ? 1
                                                           : 0;
}
int glm_f16(int prec) { return (prec == BH_GLM_FP16X2 || prec == BH_GLM_FP16) ? 1 : 0; }
// the one-term fp16 precision's correction goes to second order
bool glm_second_order(int prec) { return prec == BH_GLM_FP16; }

// For the split precisions' correction: g0, the gradient at the mean b of
// the first min(n, 256) points of P (device, row stride d) of the part of
// the logistic softplus sum the kernels evaluate at the split coefficients
// (everything but the linear half (1/2) X'1 . y, which they add at the
// exact y): sum_r sigma(x_r . b) x_r - (1/2) X'1; for the one-term fp16
// precision the sum's Hessian H0 at b too.  nullptr (no correction) for the
// other precisions and models.
int glm_grad_ref(bh_stretch* s, const float* P, unsigned n, const double** g0) {
    const GlmData& g = *s->f->glm;
    *g0 = nullptr;
    if (g.kind != BH_GLM_LOGISTIC || glm_nsplit(s->glm_prec) == 0 || n == 0) return BH_OK;
    BH_HIP(bh::launch_glm_grad(s->stream, g.X.p, g.N, g.d, P, std::min(n, 256u),
                               g.half_colsum.as<double>(), s->ggpart.as<double>(),
                               s->gg0.as<double>()));
    if (glm_second_order(s->glm_prec)) {
        BH_TRY(s->gh0.alloc((size_t)g.d * g.d * sizeof(float)));
        BH_TRY(s->ghpart.alloc(bh::glm_hess_work_doubles(g.d) * sizeof(double)));
        BH_HIP(bh::launch_glm_hess(s->stream, g.X.p, g.N, g.d,
                                   bh::glm_grad_bref(s->ggpart.as<double>(), g.d),
                                   s->ghpart.as<double>(), s->gh0.as<float>()));
        s->h0_valid = true;
    }
    *g0 = s->gg0.as<double>();
    return BH_OK;
}

// The second-order correction's H0 and reference point, or nullptrs.
void glm_hess_args(const bh_stretch* s, const float** H0, const float** bref) {
    const bool on = glm_second_order(s->glm_prec) && s->h0_valid;
    *H0 = on ? s->gh0.as<float>() : nullptr;
    *bref = on ? bh::glm_grad_bref(s->ggpart.as<double>(), s->f->glm->d) : nullptr;
}

// Checks a precision against the model and makes what it needs: the fp16
// copy of X for BH_GLM_FP16X2 and BH_GLM_FP16 (refused when a value of X is
// outside the fp16 range).
int glm_prepare(bh_stretch_factory* f, int prec) {
    BH_ARG(prec >= BH_GLM_FP64 && prec <= BH_GLM_FP16, "bad precision %d", prec);
    GlmData& g = *f->glm;
    BH_ARG(g.kind != BH_GLM_GAUSSIAN || prec <= BH_GLM_FP32,
           "linear-regression models run in fp64 or fp32");
    if (!glm_f16(prec) || g.X16.p) return BH_OK;
    bh_ctx* c = f->ctx;
    BH_HIP(hipSetDevice(c->device));
    DevBuf x16, over;
    BH_TRY(x16.alloc(bh::glm_x_f16_bytes(g.N, g.d)));
    BH_TRY(over.alloc(sizeof(unsigned)));
    BH_HIP(hipMemsetAsync(over.p, 0, sizeof(unsigned), c->stream));
    BH_HIP(bh::launch_glm_x_to_f16(c->stream, g.X.p, g.N, g.d, x16.p, over.as<unsigned>()));
    unsigned bad = 0;
    BH_HIP(hipMemcpyAsync(&bad, over.p, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    BH_TRY(sync(c->stream));
    BH_ARG(bad == 0, "%u values of X are outside the fp16 range (|x| > 65504)", bad);
    std::swap(g.X16.p, x16.p);
    std::swap(g.X16.bytes, x16.bytes);
    return BH_OK;
}

// row ranges of one likelihood launch over n points: a multiple of 8.  The
// logistic kernels get whole rounds of work-groups (the compute units times
// the work-groups one holds at once for the kernel's registers): a partial
// last round leaves compute units idle; a few rounds of smaller ranges even
// out the ranges' finishing times.
unsigned glm_ranges(const bh_ctx* c, const GlmData& g, int prec, unsigned n) {
    unsigned groups = std::max(n / bh::glm_point_quantum(prec), 1u);
    // the linear-regression kernel's work-groups can take twice 256 points:
    // twice the ranges keep the work-group count
    if (g.kind == BH_GLM_GAUSSIAN)
        groups = std::max(groups / bh::glm_gauss_points_per_thread(prec, g.D, n), 1u);
    unsigned blocks = 4u * (unsigned)c->cu;
    if (g.kind == BH_GLM_LOGISTIC) {
        if (!g.bpc[prec]) g.bpc[prec] = bh::glm_blocks_per_cu(prec, g.d);
        blocks = kGlmRounds * g.bpc[prec] * (unsigned)c->cu;
    }
    unsigned r = (blocks + groups - 1) / groups;
    r = (r + 7) / 8 * 8;
    const unsigned long long max_r = std::max(8ull, (g.N / 64 + 7) / 8 * 8);
    if (r > max_r) r = (unsigned)max_r;
    return std::max(r, 8u);
}

// log posteriors of n points P (device, n x d) into lp (double, device).
int glm_logpost_points(bh_stretch* s, const float* P, unsigned n, double* lp) {
    bh_stretch_factory* f = s->f;
    const GlmData& g = *f->glm;
    const int prec = s->glm_prec;
    const unsigned nranges = glm_ranges(s->ctx, g, prec, n);
    const unsigned long long rpr = ((g.N + nranges - 1) / nranges + 15) / 16 * 16;
    int nsplit = glm_nsplit(prec);
    int f16 = glm_f16(prec);
    float yscale = kLog2e;
    void* ys = s->gsplit.p;
    double* corr = nullptr;
    if (nsplit) {
        const double* g0 = nullptr;
        BH_TRY(glm_grad_ref(s, P, n, &g0));
        double lnscale = kLn2;
        corr = s->gcorr.as<double>();
        const float *H0, *bref;
        glm_hess_args(s, &H0, &bref);
        void* a[] = {&P, &n, &ys, &yscale, &nsplit, &f16, &g0, &lnscale, &corr, &H0, &bref};
        BH_TRY(launch(f->glm_split, (n + s->wgs - 1) / s->wgs, s->wgs, 0, s->stream, a));
    }
    double* partial = s->gpartial.as<double>();
    BH_TRY(glm_contract(g, prec, s->stream, P, s->dim, ys, n, nranges, rpr, yscale, partial));
    const double* lin = glm_lin(g, prec);
    double scale = prec == BH_GLM_FP64 ? 1.0 : kLn2;
    double pp = g.prior_prec;
    unsigned nr = nranges;
    const double* pc = partial;
    const double* cc = corr;
    void* a[] = {&P, &n, &pc, &nr, &lin, &scale, &pp, &lp, &cc};
    return launch(f->glm_logpost, (n + s->wgs - 1) / s->wgs, s->wgs, 0, s->stream, a);
}

// One GLM half-step: proposals, the likelihood contraction, the accept test.
int glm_half(bh_stretch* s, bool first, unsigned seed, unsigned oe, unsigned step, float a,
             float beta, int accu, unsigned mstep) {
    bh_stretch_factory* f = s->f;
    const GlmData& g = *f->glm;
    const int prec = s->glm_prec;
    unsigned K = s->K;
    float* X = s->xs.as<float>() + (first ? 0 : (size_t)K * s->dim);
    const float* S = s->xs.as<float>() + (first ? (size_t)K * s->dim : 0);
    double* lp = s->lp.as<double>() + (first ? 0 : K);
    float* Y = s->gY.as<float>();
    float* z = s->gz.as<float>();
    float* u2 = s->gu2.as<float>();
    void* ys = s->gsplit.p;
    int partner = s->partner;
    int nsplit = glm_nsplit(prec);
    int f16 = glm_f16(prec);
    float yscale = kLog2e;
    const float* Xc = X;
    const bool guarded = s->glm_guard > 0.0 && prec != BH_GLM_FP64;
    unsigned* gcount = guarded ? s->gcount.as<unsigned>() : nullptr;
    // the correction's gradient at the mean of 256 walkers of the ensemble,
    // every kGlmGradRefresh steps (both halves use it)
    const double* g0 = nullptr;
    if (nsplit && g.kind == BH_GLM_LOGISTIC) {
        if (first && (s->g0_age < 0 || ++s->g0_age >= kGlmGradRefresh)) {
            BH_TRY(glm_grad_ref(s, s->xs.as<float>(), s->W, &g0));
            s->g0_age = 0;
        }
        g0 = s->gg0.as<double>();
    }
    double lnscale = kLn2;
    double* corr = nsplit ? s->gcorr.as<double>() : nullptr;
    const float *H0, *bref;
    glm_hess_args(s, &H0, &bref);
    if (!g0) H0 = bref = nullptr;
    // the second-order precision's per-walker widening of the guard
    float* gscale = (H0 && guarded) ? s->gscale.as<float>() : nullptr;
    void* pa[] = {&seed, &oe, &S, &Xc, &a, &step, &K, &partner, &Y, &z, &u2, &ys, &yscale, &nsplit,
                  &f16, &gcount, &g0, &lnscale, &corr, &H0, &bref, &gscale};
    BH_TRY(launch(f->glm_propose, K * kGlmProposeThreads / s->wgs, s->wgs, 0, s->stream, pa));
    unsigned nranges = glm_ranges(s->ctx, g, prec, K);
    const unsigned long long rpr = ((g.N + nranges - 1) / nranges + 15) / 16 * 16;
    double* partial = s->gpartial.as<double>();
    BH_TRY(ktimed(s, [&]() -> int {
        BH_TRY(glm_contract(g, prec, s->stream, Y, s->dim, ys, K, nranges, rpr, yscale,
                            partial));
        return BH_OK;
    }));
    const double* lin = glm_lin(g, prec);
    double scale = prec == BH_GLM_FP64 ? 1.0 : kLn2;
    double pp = g.prior_prec;
    const float* Yc = Y;
    const float* zc = z;
    const float* u2c = u2;
    const double* pc = partial;
    unsigned* acc = accu ? s->accept.as<unsigned>() : nullptr;
    float* means = accu ? s->means_acc.as<float>() : nullptr;
    unsigned* flag = nullptr;
    if (guarded) {
        // screen, fp64 evaluation of the listed walkers, their decisions
        flag = s->gflag.as<unsigned>();
        unsigned* list = s->glist.as<unsigned>();
        unsigned lmax = K;
        double guard = s->glm_guard * std::max(1.0f, beta);
        const double* lpc = lp;
        double* pts64 = s->gpts64.as<double>();
        double* lpy = s->glpy.as<double>();
        const double* cc = corr;
        const float* gsc = gscale;
        void* sa[] = {&Yc, &zc, &u2c, &pc, &nranges, &lin, &scale, &pp, &lpc, &beta, &K, &guard,
                      &flag, &list, &gcount, &lmax, &Xc, &pts64, &lpy, &cc, &gsc};
        BH_TRY(launch(f->glm_screen, K * kGlmScreenThreads / s->wgs, s->wgs, 0, s->stream, sa));
        double* part64 = s->gpart64.as<double>();
        BH_HIP(bh::launch_glm_fp64_list(s->stream, g.X.p, g.N, g.d, pts64, gcount, lmax,
                                        guard_ranges(K), part64));
        const double* p64 = part64;
        unsigned nr64 = guard_ranges(K), stride = 2 * K;
        const double* lin64 = glm_lin(g, BH_GLM_FP64);
        const unsigned* listc = list;
        const unsigned* countc = gcount;
        unsigned long long* totals = s->gtotals.as<unsigned long long>();
        void* ra[] = {&Yc, &Xc, &zc, &u2c, &p64, &nr64, &stride, &lin64, &pp, &beta, &listc,
                      &countc, &lmax, &flag, &totals};
        BH_TRY(launch(f->glm_resolve, kGuardResolveBlocks, s->wgs, 0, s->stream, ra));
        // every walker decided: the accept step in (group, coordinate chunk)
        // work-groups
        const unsigned chunk = std::min(s->dim, 8u);
        const unsigned nchunks = (s->dim + chunk - 1) / chunk;
        const double* lpyc = lpy;
        const unsigned* flagc = flag;
        void* xa[] = {&Yc, &lpyc, &X, &lp, &accu, &acc, &means, &mstep, &flagc};
        BH_HIP(hipModuleLaunchKernel(f->glm_accept_split, K / s->wgs, nchunks, 1, s->wgs, 1, 1, 0,
                                     s->stream, xa, nullptr));
        return BH_OK;
    }
    const unsigned* flagc = flag;
    const double* cc = corr;
    void* aa[] = {&Yc, &zc, &u2c, &pc, &nranges, &lin, &scale, &pp, &X, &lp, &beta, &K, &accu,
                  &acc, &means, &mstep, &flagc, &cc};
    return launch(f->glm_accept, K / s->wgs, s->wgs, 0, s->stream, aa);
}

// Work buffers of a GLM sampler (or of a log-density call) for up to n
// points per launch.  Per-point buffers hold the walker's dimension
// (s->dim: d, or d + 1 for the linear regression's log sigma).
int glm_alloc(bh_stretch* s, unsigned n_points, unsigned k_half) {
    const GlmData& g = *s->f->glm;
    const size_t dim = s->dim;
    size_t maxr = 0;
    for (int prec = BH_GLM_FP64; prec <= BH_GLM_FP16; ++prec)
        for (unsigned n : {n_points, k_half})
            if (n && n % bh::glm_point_quantum(prec) == 0)
                maxr = std::max(maxr, (size_t)glm_ranges(s->ctx, g, prec, n) * n);
    BH_TRY(s->gpartial.alloc(std::max<size_t>(maxr, 1) * sizeof(double)));
    BH_TRY(s->gsplit.alloc((size_t)3 * n_points * dim * sizeof(uint16_t)));
    BH_TRY(s->gcorr.alloc((size_t)std::max(std::max(n_points, k_half), 1u) * sizeof(double)));
    BH_TRY(s->gscale.alloc((size_t)std::max(k_half, 1u) * sizeof(float)));
    BH_TRY(s->gg0.alloc((size_t)std::max<size_t>(dim, 64) * sizeof(double)));
    BH_TRY(s->ggpart.alloc(bh::glm_grad_work_doubles(std::max<unsigned>(s->dim, 64)) *
                           sizeof(double)));
    if (k_half) {
        BH_TRY(s->gY.alloc((size_t)k_half * dim * sizeof(float)));
        BH_TRY(s->gz.alloc((size_t)k_half * sizeof(float)));
        BH_TRY(s->gu2.alloc((size_t)k_half * sizeof(float)));
        BH_TRY(s->gflag.alloc((size_t)k_half * sizeof(unsigned)));
        BH_TRY(s->glpy.alloc((size_t)k_half * sizeof(double)));
        BH_TRY(s->glist.alloc((size_t)k_half * sizeof(unsigned)));
        BH_TRY(s->gcount.alloc(sizeof(unsigned)));
        BH_TRY(s->gtotals.alloc(2 * sizeof(unsigned long long)));
        BH_TRY(s->gpart64.alloc((size_t)guard_ranges(k_half) * 2 * k_half * sizeof(double)));
        BH_TRY(s->gpts64.alloc((size_t)2 * k_half * dim * sizeof(double)));
        BH_HIP(hipMemsetAsync(s->gtotals.p, 0, 2 * sizeof(unsigned long long), s->stream));
    }
    return BH_OK;
}

// One bare half-step pair (move-bare!, amd_gcn.clj:347-353): odd half with
// seed and constant 3333, even half with seed+1 and 4444 (:547-550).
int bare_pair(bh_stretch* s, float beta) {
    if (s->f->glm) {
        BH_TRY(glm_half(s, true, s->bare_seed, 3333, s->bare_counter, s->bare_a, beta, 0, 0));
        return glm_half(s, false, s->bare_seed + 1u, 4444, s->bare_counter, s->bare_a, beta, 0, 0);
    }
    unsigned data_len = s->data_len, params_len = s->params_len, K = s->K;
    const float* params = s->params.as<float>();
    float* x0 = s->xs.as<float>();
    float* x1 = x0 + (size_t)s->K * s->dim;
    float* l0 = s->logfn.as<float>();
    float* l1 = l0 + s->K;
    float a = s->bare_a;
    unsigned step = s->bare_counter;
    int partner = s->partner;
    unsigned seed0 = s->bare_seed, seed1 = s->bare_seed + 1u, c0 = 3333, c1 = 4444;
    const float* S1 = x1;
    const float* S0 = x0;
    void* a0[] = {&seed0, &c0, &data_len, &params_len, &params, &S1, &x0, &l0, &a, &beta, &step, &K, &partner};
    void* a1[] = {&seed1, &c1, &data_len, &params_len, &params, &S0, &x1, &l1, &a, &beta, &step, &K, &partner};
    const unsigned grid = s->K / s->wgs;
    BH_TRY(ktimed(s, [&] { return launch(s->f->bare, grid, s->wgs, 0, s->stream, a0); }));
    BH_TRY(ktimed(s, [&] { return launch(s->f->bare, grid, s->wgs, 0, s->stream, a1); }));
    return BH_OK;
}

// One accounting pair (move!, amd_gcn.clj:340-346): constants 1111/2222.
int accu_pair(bh_stretch* s) {
    if (s->f->glm) {
        BH_TRY(glm_half(s, true, s->move_seed, 1111, s->move_counter, s->accu_a, 1.0f, 1,
                        s->move_counter));
        return glm_half(s, false, s->move_seed + 1u, 2222, s->move_counter, s->accu_a, 1.0f, 1,
                        s->move_counter);
    }
    unsigned data_len = s->data_len, params_len = s->params_len, K = s->K;
    const float* params = s->params.as<float>();
    float* x0 = s->xs.as<float>();
    float* x1 = x0 + (size_t)s->K * s->dim;
    float* l0 = s->logfn.as<float>();
    float* l1 = l0 + s->K;
    unsigned* acc = s->accept.as<unsigned>();
    float* means = s->means_acc.as<float>();
    float a = s->accu_a;
    unsigned step = s->move_counter;
    int partner = s->partner;
    unsigned seed0 = s->move_seed, seed1 = s->move_seed + 1u, c0 = 1111, c1 = 2222;
    const float* S1 = x1;
    const float* S0 = x0;
    void* a0[] = {&seed0, &c0, &data_len, &params_len, &params, &S1, &x0, &l0, &acc, &means, &a, &step, &K, &partner};
    void* a1[] = {&seed1, &c1, &data_len, &params_len, &params, &S0, &x1, &l1, &acc, &means, &a, &step, &K, &partner};
    const unsigned grid = s->K / s->wgs;
    BH_TRY(ktimed(s, [&] { return launch(s->f->accu, grid, s->wgs, 0, s->stream, a0); }));
    BH_TRY(ktimed(s, [&] { return launch(s->f->accu, grid, s->wgs, 0, s->stream, a1); }));
    return BH_OK;
}

int bare_steps(bh_stretch* s, uint64_t n, const float* h_betas) {
    if (n == 0) return BH_OK;
    if (use_resident(s, 0)) {
        DevBuf d_betas;
        const float* db = nullptr;
        if (h_betas) {
            BH_TRY(d_betas.alloc(n * sizeof(float)));
            BH_HIP(hipMemcpyAsync(d_betas.p, h_betas, n * sizeof(float), hipMemcpyHostToDevice,
                                  s->stream));
            db = d_betas.as<float>();
        }
        // the kernel's step counter is 32-bit, as the reference's (int-array)
        uint64_t done = 0;
        while (done < n) {
            const uint32_t chunk = (uint32_t)std::min<uint64_t>(n - done, 1u << 20);
            BH_TRY(run_resident(s, 0, s->bare_seed, s->bare_seed + 1u, 3333, 4444, s->bare_a,
                                s->beta, db ? db + done : nullptr, s->bare_counter, chunk, nullptr,
                                0));
            s->bare_counter += chunk;
            done += chunk;
        }
        if (h_betas) s->beta = h_betas[n - 1];
        BH_TRY(sync(s->stream));
        return BH_OK;
    }
    for (uint64_t i = 0; i < n; ++i) {
        if (h_betas) s->beta = h_betas[i];
        BH_TRY(bare_pair(s, s->beta));
        s->bare_counter += 1;
    }
    return sync(s->stream);
}

int init_move(bh_stretch* s, float a, uint64_t steps) {
    s->move_seed += 2u;
    s->move_counter = 0;
    s->accu_a = a;
    BH_TRY(s->accept.alloc((size_t)s->groups * sizeof(unsigned)));
    BH_HIP(hipMemsetAsync(s->accept.p, 0, (size_t)s->groups * sizeof(unsigned), s->stream));
    const size_t mbytes = (size_t)s->dim * s->groups * (steps ? steps : 1) * sizeof(float);
    BH_TRY(s->means_acc.alloc(mbytes));
    BH_HIP(hipMemsetAsync(s->means_acc.p, 0, mbytes, s->stream));
    s->means_cap = steps ? steps : 1;
    s->move_ready = true;
    return BH_OK;
}

int accu_steps(bh_stretch* s, uint64_t n) {
    if (!s->move_ready) return fail(BH_ERR_STATE, "move! before init-move!");
    if (s->move_counter + n > s->means_cap)
        return fail(BH_ERR_STATE, "move! beyond the %llu steps reserved by init-move!",
                    (unsigned long long)s->means_cap);
    if (n == 0) return BH_OK;
    if (use_resident(s, 1)) {
        BH_TRY(run_resident(s, 1, s->move_seed, s->move_seed + 1u, 1111, 2222, s->accu_a, 1.0f,
                            nullptr, s->move_counter, (uint32_t)n, nullptr, 0));
        s->move_counter += (uint32_t)n;
        return sync(s->stream);
    }
    for (uint64_t i = 0; i < n; ++i) {
        BH_TRY(accu_pair(s));
        s->move_counter += 1;
    }
    return sync(s->stream);
}

uint32_t acor_lag(uint32_t n, uint32_t wgs) {
    // amd_gcn.clj:224-228
    uint32_t lag = n / 5;
    lag = std::min(lag, wgs);
    lag = std::min(lag, 64u);
    return std::max(lag, 10u);
}

// acor on a device series (dim x n), which it overwrites.
int acor_device(bh_ctx* c, hipStream_t st, float* d_series, uint32_t dim, uint32_t n, float* tau,
                float* mean, float* sigma, uint32_t* lag_out) {
    const uint32_t lag = acor_lag(n, c->wgs);
    if (lag_out) *lag_out = lag;
    if ((uint64_t)lag * 5 > n)
        return fail(BH_ERR_RANGE,
                    "The autocorrelation time is too long relative to the variance. Number of "
                    "steps (%u) must not be less than %u.",
                    n, lag * 5);
    const uint32_t wgsm = std::min(std::min(dim, 16u), c->wgs);
    const uint32_t local_n_mean = c->wgs / wgsm;
    const uint32_t half = (n + local_n_mean - 1) / local_n_mean + (n + c->wgs - 1) / c->wgs + 8;
    const uint32_t per_dim = 2 * half;
    DevBuf scratch, out;
    BH_TRY(scratch.alloc((size_t)per_dim * dim * sizeof(float)));
    BH_TRY(out.alloc((size_t)3 * dim * sizeof(float)));
    float* o = out.as<float>();
    BH_HIP(bh::launch_acor(st, d_series, dim, n, lag, c->wgs, scratch.as<float>(), per_dim, o,
                           o + dim, o + 2 * dim));
    std::vector<float> h(3 * dim);
    BH_HIP(hipMemcpyAsync(h.data(), o, h.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    BH_TRY(sync(st));
    for (uint32_t d = 0; d < dim; ++d) {
        if (tau) tau[d] = h[d];
        if (mean) mean[d] = h[dim + d];
        if (sigma) sigma[d] = h[2 * dim + d];
    }
    return BH_OK;
}

int histogram_device(bh_ctx* c, hipStream_t st, const float* d_a, uint32_t m, uint64_t n,
                     uint32_t ld, float* d_limits, unsigned* d_counts, bool compute_limits) {
    void* scratch = nullptr;
    BH_TRY(ctx_scratch(c, bh::dataset_scratch_bytes(m, n), &scratch));
    if (compute_limits) BH_HIP(bh::launch_min_max(st, d_a, m, n, ld, d_limits, scratch));
    BH_HIP(bh::launch_histogram(st, d_a, m, n, ld, d_limits, c->wgs, d_counts));
    return BH_OK;
}

int check_matrix(uint32_t m, uint64_t n, uint32_t ld) {
    BH_ARG(m > 0 && n > 0, "empty matrix (%u x %llu)", m, (unsigned long long)n);
    BH_ARG(ld >= m, "leading dimension %u < rows %u", ld, m);
    return BH_OK;
}

}  // namespace

// ==========================================================================
// C-ABI
// ==========================================================================
extern "C" {

const char* bh_last_error(void) { return g_err.c_str(); }

int bh_version(void) { return 100; }

int bh_rtc_compile_check(const char* source, const char* logfn_name, const char* logpdf_name,
                         const char* loglik_name, uint32_t dim, uint32_t wgs, const char* arch,
                         char* log, size_t log_size) {
    BH_ARG(source && arch, "NULL argument");
    BH_ARG(wgs >= 64 && wgs <= 1024 && (wgs & (wgs - 1)) == 0, "bad WGS %u", wgs);
    std::vector<std::string> defs = {"-DWGS=" + std::to_string(wgs)};
    if (logfn_name) {
        BH_ARG(valid_name(logfn_name), "invalid logfn function name");
        BH_ARG(dim >= 1 && dim <= 256, "dimension %u outside [1, 256]", dim);
        defs.push_back("-DDIM=" + std::to_string(dim));
        defs.push_back(std::string("-DLOGFN=") + logfn_name);
    }
    if (logpdf_name) {
        BH_ARG(valid_name(logpdf_name), "invalid logpdf function name");
        defs.push_back(std::string("-DLOGPDF=") + logpdf_name);
    }
    if (loglik_name) {
        BH_ARG(valid_name(loglik_name), "invalid loglik function name");
        defs.push_back(std::string("-DLOGLIK=") + loglik_name);
    }
    std::vector<char> code;
    return rtc_compile(arch, source, defs, code, log, log_size);
}

int bh_ctx_create(int device, uint32_t wgs, bh_ctx** out) {
    BH_ARG(out, "out is NULL");
    BH_ARG(wgs >= 64 && wgs <= 1024 && (wgs & (wgs - 1)) == 0,
           "WGS must be a power of two in [64, 1024], got %u", wgs);
    BH_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    BH_HIP(hipGetDeviceProperties(&prop, device));
    auto c = std::make_unique<bh_ctx>();
    c->device = device;
    c->wgs = wgs;
    c->cu = prop.multiProcessorCount;
    std::string arch = prop.gcnArchName;
    c->arch = arch.substr(0, arch.find(':'));
    // the address of this device's copy of bh_dataset_variance_noalloc's
    // working memory, looked up once per device with it current
    BH_HIP(bh::noalloc_scratch_init(device));
    BH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *out = c.release();
    return BH_OK;
}

int bh_ctx_destroy(bh_ctx* ctx) {
    if (!ctx) return BH_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BH_OK;
}

int bh_ctx_info(const bh_ctx* ctx, int* compute_units, uint32_t* wgs) {
    BH_ARG(ctx, "ctx is NULL");
    if (compute_units) *compute_units = ctx->cu;
    if (wgs) *wgs = ctx->wgs;
    return BH_OK;
}

int bh_device_copy_rate(bh_ctx* ctx, uint64_t bytes, double* gbps, double* gbps_nt) {
    BH_ARG(ctx, "ctx is NULL");
    BH_ARG(bytes >= (1u << 20) && bytes % 16 == 0 && bytes <= (16ull << 30),
           "copy size %llu: a multiple of 16 bytes in [1 MiB, 16 GiB]",
           (unsigned long long)bytes);
    BH_HIP(hipSetDevice(ctx->device));
    DevBuf src, dst;
    BH_TRY(src.alloc(bytes));
    BH_TRY(dst.alloc(bytes));
    BH_HIP(hipMemsetAsync(src.p, 0, bytes, ctx->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BH_HIP(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    double rate[2] = {0.0, 0.0};
    const int reps = 10;
    for (int nt = 0; nt < 2 && err == hipSuccess; ++nt) {
        err = bh::launch_copy16(ctx->stream, src.p, dst.p, bytes / 16, nt != 0);  // warm-up
        if (err == hipSuccess) err = hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < reps && err == hipSuccess; ++r)
            err = bh::launch_copy16(ctx->stream, src.p, dst.p, bytes / 16, nt != 0);
        if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        // bytes read plus bytes written, per second
        if (err == hipSuccess && ms > 0.0f) rate[nt] = 2.0 * reps * (double)bytes / (ms * 1e-3) / 1e9;
    }
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    BH_HIP(err);
    if (gbps) *gbps = rate[0];
    if (gbps_nt) *gbps_nt = rate[1];
    return BH_OK;
}

int bh_device_gather_rate(bh_ctx* ctx, uint64_t bytes, int nontemporal, double* gathers_per_s) {
    BH_ARG(ctx && gathers_per_s, "NULL argument");
    const uint64_t n16 = bytes / 16;
    BH_ARG(n16 >= 256 && (n16 & (n16 - 1)) == 0 && n16 <= (1ull << 28) && bytes % 16 == 0,
           "gather buffer %llu bytes: 16 x a power of two in [4 KiB, 4 GiB]",
           (unsigned long long)bytes);
    BH_HIP(hipSetDevice(ctx->device));
    DevBuf src, out;
    BH_TRY(src.alloc(bytes));
    BH_TRY(out.alloc(n16 * sizeof(float)));
    BH_HIP(hipMemsetAsync(src.p, 0, bytes, ctx->stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BH_HIP(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    const bool nt = nontemporal != 0;
    for (int rep = 0; rep < 2 && err == hipSuccess; ++rep)
        err = bh::launch_glm_gather(ctx->stream, src.p, out.as<float>(), (unsigned)n16, nt);
    if (err == hipSuccess) err = hipEventRecord(e0, ctx->stream);
    for (int rep = 0; rep < 5 && err == hipSuccess; ++rep)
        err = bh::launch_glm_gather(ctx->stream, src.p, out.as<float>(), (unsigned)n16, nt);
    if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    BH_HIP(err);
    *gathers_per_s = ms > 0.0f ? 5.0 * (double)n16 / ((double)ms * 1e-3) : 0.0;
    return BH_OK;
}

int bh_device_rates(bh_ctx* ctx, double* exp2_per_s, double* fma_per_s,
                    double* mfma_f16_flops_per_s, double* gathers_per_s) {
    BH_ARG(ctx, "ctx is NULL");
    BH_HIP(hipSetDevice(ctx->device));
    const unsigned blocks = 8u * (unsigned)ctx->cu;
    const int iters = 2048;
    const unsigned n16 = 1u << 23;  // gather records of 16 bytes: 128 MiB
    DevBuf out, src, gout;
    BH_TRY(out.alloc((size_t)blocks * 256 * sizeof(float)));
    if (gathers_per_s) {
        BH_TRY(src.alloc((size_t)n16 * 16));
        BH_TRY(gout.alloc((size_t)n16 * sizeof(float)));
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BH_HIP(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    double rate[3] = {0.0, 0.0, 0.0};
    for (int which = 0; which < 3 && err == hipSuccess; ++which) {
        err = bh::launch_glm_rate(ctx->stream, which, out.as<float>(), blocks, iters);  // warm-up
        if (err == hipSuccess) err = hipEventRecord(e0, ctx->stream);
        if (err == hipSuccess) err = bh::launch_glm_rate(ctx->stream, which, out.as<float>(), blocks, iters);
        if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        // lane operations; for the products, flops (8 per wave and
        // iteration, 16 x 16 x 32 x 2 each)
        const double ops = which < 2 ? (double)blocks * 256.0 * 8.0 * iters
                                     : (double)blocks * 4.0 * 8.0 * iters * 16384.0;
        if (err == hipSuccess && ms > 0.0f) rate[which] = ops / ((double)ms * 1e-3);
    }
    // random 16-byte gathers from 128 MiB (above the L2s, within the
    // Infinity Cache, as the stretch move's partner reads at 2^23 walkers)
    double grate = 0.0;
    if (err == hipSuccess && gathers_per_s) {
        err = hipMemsetAsync(src.p, 0, (size_t)n16 * 16, ctx->stream);
        for (int rep = 0; rep < 2 && err == hipSuccess; ++rep)
            err = bh::launch_glm_gather(ctx->stream, src.p, gout.as<float>(), n16);
        if (err == hipSuccess) err = hipEventRecord(e0, ctx->stream);
        for (int rep = 0; rep < 5 && err == hipSuccess; ++rep)
            err = bh::launch_glm_gather(ctx->stream, src.p, gout.as<float>(), n16);
        if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (err == hipSuccess && ms > 0.0f) grate = 5.0 * n16 / ((double)ms * 1e-3);
    }
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    BH_HIP(err);
    if (gathers_per_s) *gathers_per_s = grate;
    if (exp2_per_s) *exp2_per_s = rate[0];
    if (fma_per_s) *fma_per_s = rate[1];
    if (mfma_f16_flops_per_s) *mfma_f16_flops_per_s = rate[2];
    return BH_OK;
}

int bh_uniform_sample(bh_ctx* ctx, uint32_t seed, float lower, float upper, float* out,
                      uint64_t n) {
    BH_ARG(ctx && out, "NULL argument");
    BH_ARG(n % 4 == 0 && n / 4 <= 0xffffffffull,
           "the direct sampler needs a multiple of 4 samples (amd_gcn.clj:55-62), got %llu",
           (unsigned long long)n);
    BH_HIP(hipSetDevice(ctx->device));
    DevOut o;
    BH_TRY(o.init(out, n * sizeof(float)));
    BH_HIP(bh::launch_uniform_sample(ctx->stream, seed, lower, upper, (unsigned)(n / 4),
                                     o.as<float>()));
    BH_TRY(o.finish(ctx->stream));
    return sync(ctx->stream);
}

int bh_direct_sample(bh_ctx* ctx, int kind, uint32_t seed, const float* params,
                     uint32_t params_count, float* out, uint64_t n) {
    BH_ARG(ctx && params && out, "NULL argument");
    BH_ARG(kind >= BH_DIRECT_UNIFORM && kind <= BH_DIRECT_ERLANG, "unknown sampler %d", kind);
    const uint32_t need = (kind == BH_DIRECT_EXPONENTIAL) ? 1u : 2u;
    BH_ARG(params_count >= need, "sampler %d needs %u params", kind, need);
    BH_ARG(n % 4 == 0 && n / 4 <= 0xffffffffull,
           "the direct sampler needs a multiple of 4 samples (amd_gcn.clj:55-62), got %llu",
           (unsigned long long)n);
    BH_HIP(hipSetDevice(ctx->device));
    float p[2] = {0.0f, 0.0f};
    BH_HIP(hipMemcpy(p, params, need * sizeof(float), hipMemcpyDefault));
    DevOut o;
    BH_TRY(o.init(out, n * sizeof(float)));
    BH_HIP(bh::launch_direct_sample(ctx->stream, kind, seed, p[0], p[1], (unsigned)(n / 4),
                                    o.as<float>()));
    BH_TRY(o.finish(ctx->stream));
    return sync(ctx->stream);
}

// ---------------------------------------------------------------- density
int bh_dist_engine_create(bh_ctx* ctx, const char* source, const char* logpdf_name,
                          uint32_t params_len, bh_dist_engine** out, char* log,
                          size_t log_size) {
    BH_ARG(ctx && source && out, "NULL argument");
    BH_ARG(valid_name(logpdf_name), "invalid logpdf function name");
    BH_HIP(hipSetDevice(ctx->device));
    std::vector<char> code;
    BH_TRY(rtc_compile(ctx->arch, source,
                       {"-DWGS=" + std::to_string(ctx->wgs), std::string("-DLOGPDF=") + logpdf_name},
                       code, log, log_size));
    auto e = std::make_unique<bh_dist_engine>();
    e->ctx = ctx;
    e->params_len = params_len;
    BH_TRY(e->mod.load(code));
    BH_TRY(e->mod.fn("bh_logpdf", &e->logpdf));
    BH_TRY(e->mod.fn("bh_pdf", &e->pdf));
    *out = e.release();
    return BH_OK;
}

int bh_dist_engine_destroy(bh_dist_engine* e) {
    delete e;
    return BH_OK;
}

static int density_impl(bh_dist_engine* e, hipFunction_t fn, const float* params,
                        uint32_t params_count, const float* x, uint32_t dim, uint64_t n,
                        float* res) {
    BH_ARG(e && params && x && res, "NULL argument");
    BH_ARG(dim > 0 && n > 0, "empty input");
    bh_ctx* c = e->ctx;
    BH_HIP(hipSetDevice(c->device));
    // data_len = max(0, params_count - params_len) (amd_gcn.clj:75-77)
    unsigned data_len = params_count > e->params_len ? params_count - e->params_len : 0u;
    unsigned params_len = e->params_len;
    DevIn dp, dx;
    DevOut dr;
    BH_TRY(dp.init(params, (size_t)params_count * sizeof(float), c->stream));
    BH_TRY(dx.init(x, (size_t)dim * n * sizeof(float), c->stream));
    BH_TRY(dr.init(res, n * sizeof(float)));
    const float* pp = dp.as<float>();
    const float* xp = dx.as<float>();
    float* rp = dr.as<float>();
    unsigned long long nn = n;
    void* args[] = {&data_len, &params_len, &pp, &dim, &xp, &rp, &nn};
    BH_TRY(launch(fn, (unsigned)((n + c->wgs - 1) / c->wgs), c->wgs, 0, c->stream, args));
    BH_TRY(dr.finish(c->stream));
    return sync(c->stream);
}

int bh_log_density(bh_dist_engine* e, const float* params, uint32_t params_count, const float* x,
                   uint32_t dim, uint64_t n, float* res) {
    return density_impl(e, e ? e->logpdf : nullptr, params, params_count, x, dim, n, res);
}

int bh_density(bh_dist_engine* e, const float* params, uint32_t params_count, const float* x,
               uint32_t dim, uint64_t n, float* res) {
    return density_impl(e, e ? e->pdf : nullptr, params, params_count, x, dim, n, res);
}

// ------------------------------------------------------------- likelihood
int bh_lik_engine_create(bh_ctx* ctx, const char* source, const char* loglik_name,
                         bh_lik_engine** out, char* log, size_t log_size) {
    BH_ARG(ctx && source && out, "NULL argument");
    BH_ARG(valid_name(loglik_name), "invalid loglik function name");
    BH_HIP(hipSetDevice(ctx->device));
    std::vector<char> code;
    BH_TRY(rtc_compile(ctx->arch, source,
                       {"-DWGS=" + std::to_string(ctx->wgs), std::string("-DLOGLIK=") + loglik_name},
                       code, log, log_size));
    auto e = std::make_unique<bh_lik_engine>();
    e->ctx = ctx;
    BH_TRY(e->mod.load(code));
    BH_TRY(e->mod.fn("bh_loglik", &e->loglik));
    BH_TRY(e->mod.fn("bh_lik", &e->lik));
    BH_TRY(e->mod.fn("bh_evidence_partial", &e->evidence));
    *out = e.release();
    return BH_OK;
}

int bh_lik_engine_destroy(bh_lik_engine* e) {
    delete e;
    return BH_OK;
}

static int lik_impl(bh_lik_engine* e, hipFunction_t fn, const float* data, uint32_t data_len,
                    const float* x, uint32_t dim, uint64_t n, float* res, double* evidence) {
    BH_ARG(e && x && (res || evidence), "NULL argument");
    BH_ARG(data || data_len == 0, "NULL data");
    BH_ARG(dim > 0 && n > 0, "empty input");
    bh_ctx* c = e->ctx;
    BH_HIP(hipSetDevice(c->device));
    DevIn dd, dx;
    float zero = 0.0f;
    BH_TRY(dd.init(data ? (const void*)data : (const void*)&zero,
                   (size_t)std::max(data_len, 1u) * sizeof(float), c->stream));
    BH_TRY(dx.init(x, (size_t)dim * n * sizeof(float), c->stream));
    const float* dp = dd.as<float>();
    const float* xp = dx.as<float>();
    unsigned long long nn = n;
    const unsigned grid = (unsigned)((n + c->wgs - 1) / c->wgs);
    if (evidence) {
        DevBuf partial, total;
        BH_TRY(partial.alloc((size_t)grid * sizeof(double)));
        BH_TRY(total.alloc(sizeof(double)));
        double* pp = partial.as<double>();
        void* args[] = {&data_len, &dp, &dim, &xp, &nn, &pp};
        BH_TRY(launch(fn, grid, c->wgs, 0, c->stream, args));
        BH_HIP(bh::launch_sum_f64(c->stream, pp, grid, total.as<double>()));
        double h = 0.0;
        BH_HIP(hipMemcpyAsync(&h, total.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        BH_TRY(sync(c->stream));
        *evidence = h / (double)n;
        return BH_OK;
    }
    DevOut dr;
    BH_TRY(dr.init(res, n * sizeof(float)));
    float* rp = dr.as<float>();
    void* args[] = {&data_len, &dp, &dim, &xp, &rp, &nn};
    BH_TRY(launch(fn, grid, c->wgs, 0, c->stream, args));
    BH_TRY(dr.finish(c->stream));
    return sync(c->stream);
}

int bh_loglik(bh_lik_engine* e, const float* data, uint32_t data_len, const float* x,
              uint32_t dim, uint64_t n, float* res) {
    return lik_impl(e, e ? e->loglik : nullptr, data, data_len, x, dim, n, res, nullptr);
}

int bh_lik(bh_lik_engine* e, const float* data, uint32_t data_len, const float* x, uint32_t dim,
           uint64_t n, float* res) {
    return lik_impl(e, e ? e->lik : nullptr, data, data_len, x, dim, n, res, nullptr);
}

int bh_evidence(bh_lik_engine* e, const float* data, uint32_t data_len, const float* x,
                uint32_t dim, uint64_t n, double* out) {
    BH_ARG(out, "NULL argument");
    return lik_impl(e, e ? e->evidence : nullptr, data, data_len, x, dim, n, nullptr, out);
}

// ---------------------------------------------------------------- dataset
static int moments_impl(bh_ctx* ctx, const float* data, uint32_t m, uint64_t n, uint32_t ld,
                        float* out, bool variance) {
    BH_ARG(ctx && data && out, "NULL argument");
    BH_TRY(check_matrix(m, n, ld));
    BH_HIP(hipSetDevice(ctx->device));
    DevIn dx;
    BH_TRY(dx.init(data, ((size_t)(n - 1) * ld + m) * sizeof(float), ctx->stream));
    DevBuf mv;
    BH_TRY(mv.alloc((size_t)2 * m * sizeof(float)));
    void* scratch = nullptr;
    BH_TRY(ctx_scratch(ctx, bh::dataset_scratch_bytes(m, n), &scratch));
    float* mean = mv.as<float>();
    float* var = mean + m;
    BH_HIP(bh::launch_moments(ctx->stream, dx.as<float>(), m, n, ld, mean,
                              variance ? var : nullptr, scratch));
    DevOut o;
    BH_TRY(o.init(out, (size_t)m * sizeof(float)));
    BH_HIP(hipMemcpyAsync(o.dev, variance ? var : mean, (size_t)m * sizeof(float),
                          hipMemcpyDeviceToDevice, ctx->stream));
    BH_TRY(o.finish(ctx->stream));
    return sync(ctx->stream);
}

int bh_dataset_mean(bh_ctx* ctx, const float* data, uint32_t m, uint64_t n, uint32_t ld,
                    float* out) {
    return moments_impl(ctx, data, m, n, ld, out, false);
}

int bh_dataset_variance(bh_ctx* ctx, const float* data, uint32_t m, uint64_t n, uint32_t ld,
                        float* out) {
    return moments_impl(ctx, data, m, n, ld, out, true);
}

int bh_dataset_variance_noalloc(bh_ctx* ctx, const float* data, uint32_t m, uint64_t n,
                                uint32_t ld, float* out) {
    BH_ARG(ctx && data && out, "NULL argument");
    BH_TRY(check_matrix(m, n, ld));
    BH_ARG(m <= bh::kNoAllocMaxRows,
           "%u rows: bh_dataset_variance_noalloc takes at most %u (bh_dataset_variance has no "
           "limit)",
           m, bh::kNoAllocMaxRows);
    BH_ARG(is_device_ptr(data) && is_device_ptr(out),
           "bh_dataset_variance_noalloc needs data and out in device memory (it stages nothing)");
    BH_HIP(hipSetDevice(ctx->device));
    void* scr = nullptr;
    BH_HIP(bh::noalloc_scratch_init(ctx->device));
    BH_HIP(bh::noalloc_scratch(ctx->device, &scr));
    // the same launches as bh_dataset_variance, with the partial sums and the
    // row means in the device array instead of allocated buffers
    float* mean = reinterpret_cast<float*>(static_cast<char*>(scr) + bh::kNoAllocPartialBytes);
    BH_HIP(bh::launch_moments(ctx->stream, data, m, n, ld, mean, out, scr));
    return sync(ctx->stream);
}

int bh_dataset_histogram(bh_ctx* ctx, const float* data, uint32_t m, uint64_t n, uint32_t ld,
                         float* limits, float* pdf, float* bin_ranks) {
    BH_ARG(ctx && data && limits && pdf && bin_ranks, "NULL argument");
    BH_TRY(check_matrix(m, n, ld));
    BH_HIP(hipSetDevice(ctx->device));
    const uint32_t bins = ctx->wgs;
    DevIn dx;
    BH_TRY(dx.init(data, ((size_t)(n - 1) * ld + m) * sizeof(float), ctx->stream));
    DevBuf counts;
    BH_TRY(counts.alloc((size_t)m * bins * sizeof(unsigned)));
    BH_HIP(hipMemsetAsync(counts.p, 0, (size_t)m * bins * sizeof(unsigned), ctx->stream));
    DevOut dl, dp, dr;
    BH_TRY(dl.init(limits, (size_t)2 * m * sizeof(float)));
    BH_TRY(dp.init(pdf, (size_t)bins * m * sizeof(float)));
    BH_TRY(dr.init(bin_ranks, (size_t)bins * m * sizeof(float)));
    BH_TRY(histogram_device(ctx, ctx->stream, dx.as<float>(), m, n, ld, dl.as<float>(),
                            counts.as<unsigned>(), true));
    // alpha = WGS / n as the reference's float (amd_gcn.clj:206)
    const float alpha = (float)((double)bins / (double)n);
    BH_HIP(bh::launch_histogram_finalize(ctx->stream, counts.as<unsigned>(), dl.as<float>(), alpha,
                                         m, bins, dp.as<float>(), dr.as<float>()));
    BH_TRY(dl.finish(ctx->stream));
    BH_TRY(dp.finish(ctx->stream));
    BH_TRY(dr.finish(ctx->stream));
    return sync(ctx->stream);
}

int bh_acor_batch(bh_ctx* ctx, const float* series, uint32_t problems, const uint32_t* dims,
                  uint32_t n, float* tau, float* mean, float* sigma, uint32_t* lag) {
    BH_ARG(ctx && series && dims, "NULL argument");
    BH_ARG(problems > 0 && n > 0, "empty series");
    const uint32_t wgs = ctx->wgs;
    const uint32_t lg = acor_lag(n, wgs);
    if (lag) *lag = lg;
    if ((uint64_t)lg * 5 > n)
        return fail(BH_ERR_RANGE,
                    "The autocorrelation time is too long relative to the variance. Number of "
                    "steps (%u) must not be less than %u.",
                    n, lg * 5);
    // one column per (problem, coordinate): problem p's n x dims[p] block
    // starts after those of problems 0 .. p-1
    std::vector<unsigned long long> off;
    std::vector<unsigned> stride;
    uint64_t base = 0;
    uint32_t per_dim = 0;
    for (uint32_t p = 0; p < problems; ++p) {
        BH_ARG(dims[p] > 0, "problem %u: empty series", p);
        for (uint32_t d = 0; d < dims[p]; ++d) {
            off.push_back(base + d);
            stride.push_back(dims[p]);
        }
        base += (uint64_t)n * dims[p];
        const uint32_t wgsm = std::min(std::min(dims[p], 16u), wgs);
        const uint32_t lnm = wgs / wgsm;
        per_dim = std::max(per_dim, 2 * ((n + lnm - 1) / lnm + (n + wgs - 1) / wgs + 8));
    }
    const uint32_t cols = (uint32_t)off.size();
    BH_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevBuf copy, scratch, out, doff, dstride;
    BH_TRY(copy.alloc((size_t)base * sizeof(float)));
    BH_TRY(scratch.alloc((size_t)per_dim * cols * sizeof(float)));
    BH_TRY(out.alloc((size_t)3 * cols * sizeof(float)));
    BH_TRY(doff.alloc(off.size() * sizeof(unsigned long long)));
    BH_TRY(dstride.alloc(stride.size() * sizeof(unsigned)));
    BH_HIP(hipMemcpyAsync(copy.p, series, (size_t)base * sizeof(float), hipMemcpyDefault, st));
    BH_HIP(hipMemcpyAsync(doff.p, off.data(), off.size() * sizeof(unsigned long long),
                          hipMemcpyHostToDevice, st));
    BH_HIP(hipMemcpyAsync(dstride.p, stride.data(), stride.size() * sizeof(unsigned),
                          hipMemcpyHostToDevice, st));
    float* o = out.as<float>();
    BH_HIP(bh::launch_acor_batch(st, copy.as<float>(), doff.as<unsigned long long>(),
                                 dstride.as<unsigned>(), cols, n, lg, wgs, scratch.as<float>(),
                                 per_dim, o, o + cols, o + 2 * cols));
    std::vector<float> h(3 * (size_t)cols);
    BH_HIP(hipMemcpyAsync(h.data(), o, h.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    BH_TRY(sync(st));
    for (uint32_t b = 0; b < cols; ++b) {
        if (tau) tau[b] = h[b];
        if (mean) mean[b] = h[cols + b];
        if (sigma) sigma[b] = h[2 * (size_t)cols + b];
    }
    return BH_OK;
}

int bh_acor(bh_ctx* ctx, const float* series, uint32_t dim, uint32_t n, float* tau, float* mean,
            float* sigma, uint32_t* lag) {
    BH_ARG(ctx && series, "NULL argument");
    BH_ARG(dim > 0 && n > 0, "empty series");
    BH_HIP(hipSetDevice(ctx->device));
    DevBuf copy;
    BH_TRY(copy.alloc((size_t)dim * n * sizeof(float)));
    BH_HIP(hipMemcpyAsync(copy.p, series, (size_t)dim * n * sizeof(float), hipMemcpyDefault,
                          ctx->stream));
    return acor_device(ctx, ctx->stream, copy.as<float>(), dim, n, tau, mean, sigma, lag);
}

// ---------------------------------------------------------------- stretch
int bh_stretch_factory_create(bh_ctx* ctx, const char* source, const char* logfn_name,
                              uint32_t dim, uint32_t params_len, bh_stretch_factory** out,
                              char* log, size_t log_size) {
    BH_ARG(ctx && source && out, "NULL argument");
    BH_ARG(valid_name(logfn_name), "invalid logfn function name");
    BH_ARG(dim >= 1 && dim <= 256, "dimension %u outside [1, 256]", dim);
    BH_HIP(hipSetDevice(ctx->device));
    std::vector<char> code;
    BH_TRY(rtc_compile(ctx->arch, source,
                       {"-DWGS=" + std::to_string(ctx->wgs), "-DDIM=" + std::to_string(dim),
                        std::string("-DLOGFN=") + logfn_name},
                       code, log, log_size));
    auto f = std::make_unique<bh_stretch_factory>();
    f->ctx = ctx;
    f->dim = dim;
    f->params_len = params_len;
    BH_TRY(f->mod.load(code));
    BH_TRY(f->mod.fn("bh_logfn", &f->logfn));
    BH_TRY(f->mod.fn("bh_stretch_move_bare", &f->bare));
    BH_TRY(f->mod.fn("bh_stretch_move_accu", &f->accu));
    BH_TRY(f->mod.fn("bh_stretch_resident", &f->resident));
    *out = f.release();
    return BH_OK;
}

int bh_stretch_factory_destroy(bh_stretch_factory* f) {
    delete f;
    return BH_OK;
}

int bh_glm_stretch_factory_create(bh_ctx* ctx, int kind, const void* X_bf16, const float* y,
                                  uint64_t N, uint32_t d, float prior_sd,
                                  bh_stretch_factory** out, char* log, size_t log_size) {
    BH_ARG(ctx && X_bf16 && y && out, "NULL argument");
    BH_ARG(kind == BH_GLM_LOGISTIC, "unknown GLM kind %d", kind);
    BH_ARG(bh::glm_supported_dim(d), "GLM dimension %u not supported (32 or 64)", d);
    BH_ARG(N > 0 && N < (1ull << 31), "bad row count");
    BH_ARG(prior_sd > 0.0f, "prior sd must be positive");
    BH_HIP(hipSetDevice(ctx->device));
    std::vector<char> code;
    BH_TRY(rtc_compile(ctx->arch, "",
                       {"-DWGS=" + std::to_string(ctx->wgs), "-DDIM=" + std::to_string(d),
                        "-DBH_GLM=1"},
                       code, log, log_size));
    auto f = std::make_unique<bh_stretch_factory>();
    f->ctx = ctx;
    f->dim = d;
    f->params_len = 0;
    BH_TRY(f->mod.load(code));
    BH_TRY(f->mod.fn("bh_glm_propose", &f->glm_propose));
    BH_TRY(f->mod.fn("bh_glm_split_points", &f->glm_split));
    BH_TRY(f->mod.fn("bh_glm_logpost", &f->glm_logpost));
    BH_TRY(f->mod.fn("bh_glm_accept", &f->glm_accept));
    BH_TRY(f->mod.fn("bh_glm_screen", &f->glm_screen));
    BH_TRY(f->mod.fn("bh_glm_accept_split", &f->glm_accept_split));
    BH_TRY(f->mod.fn("bh_glm_resolve", &f->glm_resolve));
    auto g = std::make_unique<GlmData>();
    g->kind = kind;
    g->N = N;
    g->d = d;
    g->prior_prec = 1.0 / ((double)prior_sd * (double)prior_sd);
    BH_TRY(g->X.alloc((size_t)N * d * sizeof(uint16_t)));
    BH_HIP(hipMemcpyAsync(g->X.p, X_bf16, (size_t)N * d * sizeof(uint16_t), hipMemcpyDefault,
                          ctx->stream));
    DevIn dy;
    BH_TRY(dy.init(y, (size_t)N * sizeof(float), ctx->stream));
    BH_TRY(g->lin.alloc((size_t)d * sizeof(double)));
    const unsigned chunks = 1024;
    DevBuf part;
    BH_TRY(part.alloc((size_t)chunks * d * sizeof(double)));
    BH_HIP(bh::launch_glm_lin(ctx->stream, g->X.p, dy.as<float>(), N, d, part.as<double>(),
                              chunks, g->lin.as<double>()));
    BH_TRY(g->lin_half.alloc((size_t)d * sizeof(double)));
    BH_HIP(bh::launch_glm_lin(ctx->stream, g->X.p, nullptr, N, d, part.as<double>(), chunks,
                              g->lin_half.as<double>()));
    std::vector<double> lin(d), colsum(d);
    BH_HIP(hipMemcpyAsync(lin.data(), g->lin.p, d * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    BH_HIP(hipMemcpyAsync(colsum.data(), g->lin_half.p, d * sizeof(double),
                          hipMemcpyDeviceToHost, ctx->stream));
    BH_TRY(sync(ctx->stream));
    std::vector<double> half(d);
    for (uint32_t j = 0; j < d; ++j) {
        half[j] = 0.5 * colsum[j];
        colsum[j] = lin[j] - half[j];
    }
    BH_HIP(hipMemcpy(g->lin_half.p, colsum.data(), d * sizeof(double), hipMemcpyHostToDevice));
    BH_TRY(g->half_colsum.alloc((size_t)d * sizeof(double)));
    BH_HIP(hipMemcpy(g->half_colsum.p, half.data(), d * sizeof(double), hipMemcpyHostToDevice));
    f->glm = std::move(g);
    *out = f.release();
    return BH_OK;
}

int bh_glm_gaussian_factory_create(bh_ctx* ctx, const float* X, const float* y, uint64_t N,
                                   uint32_t d, float beta_sd, float log_sigma_sd,
                                   bh_stretch_factory** out, char* log, size_t log_size) {
    BH_ARG(ctx && X && y && out, "NULL argument");
    BH_ARG(d >= 1 && d <= 64, "regression dimension %u outside [1, 64]", d);
    BH_ARG(N > 0 && N < (1ull << 31), "bad row count");
    BH_ARG(beta_sd > 0.0f && log_sigma_sd > 0.0f, "prior sds must be positive");
    BH_HIP(hipSetDevice(ctx->device));
    std::vector<char> code;
    BH_TRY(rtc_compile(ctx->arch, "",
                       {"-DWGS=" + std::to_string(ctx->wgs), "-DDIM=" + std::to_string(d + 1),
                        "-DBH_GLM=1", "-DBH_GLM_KIND=2"},
                       code, log, log_size));
    auto f = std::make_unique<bh_stretch_factory>();
    f->ctx = ctx;
    f->dim = d + 1;
    f->params_len = 0;
    BH_TRY(f->mod.load(code));
    BH_TRY(f->mod.fn("bh_glm_propose", &f->glm_propose));
    BH_TRY(f->mod.fn("bh_glm_split_points", &f->glm_split));
    BH_TRY(f->mod.fn("bh_glm_logpost", &f->glm_logpost));
    BH_TRY(f->mod.fn("bh_glm_accept", &f->glm_accept));
    BH_TRY(f->mod.fn("bh_glm_screen", &f->glm_screen));
    BH_TRY(f->mod.fn("bh_glm_accept_split", &f->glm_accept_split));
    BH_TRY(f->mod.fn("bh_glm_resolve", &f->glm_resolve));
    auto g = std::make_unique<GlmData>();
    g->kind = BH_GLM_GAUSSIAN;
    g->N = N;
    g->d = d;
    g->D = d <= 16 ? 16 : (d <= 32 ? 32 : 64);
    std::vector<float> xp((size_t)N * g->D, 0.0f);
    const float* xh = X;
    std::vector<float> xcopy;
    if (is_device_ptr(X)) {
        xcopy.resize((size_t)N * d);
        BH_HIP(hipMemcpy(xcopy.data(), X, xcopy.size() * 4, hipMemcpyDeviceToHost));
        xh = xcopy.data();
    }
    for (uint64_t i = 0; i < N; ++i)
        std::memcpy(&xp[(size_t)i * g->D], xh + (size_t)i * d, (size_t)d * 4);
    BH_TRY(g->Xf.alloc(xp.size() * 4));
    BH_HIP(hipMemcpy(g->Xf.p, xp.data(), xp.size() * 4, hipMemcpyHostToDevice));
    BH_TRY(g->y.alloc((size_t)N * 4));
    BH_HIP(hipMemcpy(g->y.p, y, (size_t)N * 4, hipMemcpyDefault));
    const double c[4] = {(double)N, 0.5 * (double)N * std::log(2.0 * M_PI),
                         1.0 / ((double)beta_sd * beta_sd),
                         1.0 / ((double)log_sigma_sd * log_sigma_sd)};
    BH_TRY(g->lin.alloc(sizeof(c)));
    BH_HIP(hipMemcpy(g->lin.p, c, sizeof(c), hipMemcpyHostToDevice));
    f->glm = std::move(g);
    *out = f.release();
    return BH_OK;
}

int bh_glm_log_density(bh_stretch_factory* f, const float* beta, uint64_t n, int precision,
                       double* out) {
    BH_ARG(f && beta && out, "NULL argument");
    BH_ARG(f->glm, "not a GLM factory");
    BH_ARG(n > 0 && n < (1ull << 26), "bad point count");
    BH_TRY(glm_prepare(f, precision));
    bh_ctx* c = f->ctx;
    BH_HIP(hipSetDevice(c->device));
    const unsigned q = bh::glm_point_quantum(precision);
    const unsigned npad = (unsigned)((n + q - 1) / q * q);
    bh_stretch t;  // work buffers and stream of this call
    t.f = f;
    t.ctx = c;
    t.stream = c->stream;
    t.wgs = c->wgs;
    t.dim = f->dim;
    t.glm_prec = precision;
    BH_TRY(glm_alloc(&t, npad, 0));
    DevBuf P, lp;
    BH_TRY(P.alloc((size_t)npad * f->dim * sizeof(float)));
    BH_HIP(hipMemsetAsync(P.p, 0, (size_t)npad * f->dim * sizeof(float), c->stream));
    BH_HIP(hipMemcpyAsync(P.p, beta, (size_t)n * f->dim * sizeof(float), hipMemcpyDefault,
                          c->stream));
    BH_TRY(lp.alloc((size_t)npad * sizeof(double)));
    BH_TRY(glm_logpost_points(&t, P.as<float>(), npad, lp.as<double>()));
    BH_HIP(hipMemcpyAsync(out, lp.p, n * sizeof(double), hipMemcpyDefault, c->stream));
    t.stream = nullptr;  // owned by the context
    return sync(c->stream);
}

int bh_stretch_create(bh_stretch_factory* f, uint32_t seed, uint32_t walkers,
                      const float* params, uint32_t params_count, bh_stretch** out) {
    BH_ARG(f && out, "NULL argument");
    BH_ARG(params || params_count == 0, "NULL params");
    bh_ctx* c = f->ctx;
    if (!(walkers >= 2 * c->wgs && walkers % (2 * c->wgs) == 0))
        return fail(BH_ERR_RANGE, "Number of walkers (%u) must be a multiple of %u.", walkers,
                    2 * c->wgs);
    BH_ARG((uint64_t)walkers * f->dim < (1ull << 31), "ensemble too large");
    BH_HIP(hipSetDevice(c->device));
    auto s = std::make_unique<bh_stretch>();
    s->f = f;
    s->ctx = c;
    s->dim = f->dim;
    s->W = walkers;
    s->K = walkers / 2;
    s->wgs = c->wgs;
    s->groups = s->K / c->wgs;
    s->params_count = params_count;
    s->params_len = f->params_len;
    s->data_len = params_count > f->params_len ? params_count - f->params_len : 0u;
    BH_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    if (f->glm) {
        if (f->glm->kind == BH_GLM_GAUSSIAN) s->glm_prec = BH_GLM_FP32;
        BH_TRY(glm_alloc(s.get(), walkers, walkers / 2));
        BH_TRY(s->lp.alloc((size_t)walkers * sizeof(double)));
        BH_HIP(hipMemsetAsync(s->lp.p, 0, (size_t)walkers * sizeof(double), s->stream));
    }
    BH_TRY(s->xs.alloc((size_t)walkers * f->dim * sizeof(float)));
    BH_TRY(s->logfn.alloc((size_t)walkers * sizeof(float)));
    BH_TRY(s->params.alloc((size_t)std::max(params_count, 1u) * sizeof(float)));
    BH_HIP(hipMemsetAsync(s->xs.p, 0, (size_t)walkers * f->dim * sizeof(float), s->stream));
    BH_HIP(hipMemsetAsync(s->logfn.p, 0, (size_t)walkers * sizeof(float), s->stream));
    if (params_count)
        BH_HIP(hipMemcpyAsync(s->params.p, params, (size_t)params_count * sizeof(float),
                              hipMemcpyDefault, s->stream));
    BH_TRY(sync(s->stream));
    BH_TRY(bh_stretch_init(s.get(), seed));
    *out = s.release();
    return BH_OK;
}

int bh_stretch_destroy(bh_stretch* s) {
    if (!s) return BH_OK;
    if (s->stream) {
        (void)hipStreamSynchronize(s->stream);
        (void)hipStreamDestroy(s->stream);
    }
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    for (hipEvent_t e : s->kev) (void)hipEventDestroy(e);
    delete s;
    return BH_OK;
}

int bh_stretch_set_timing(bh_stretch* s, int on) {
    BH_ARG(s, "NULL sampler");
    BH_HIP(hipSetDevice(s->ctx->device));
    if (on && !s->ev0) {
        BH_HIP(hipEventCreate(&s->ev0));
        BH_HIP(hipEventCreate(&s->ev1));
    }
    s->timing = on != 0;
    return BH_OK;
}

int bh_stretch_set_kernel_timing(bh_stretch* s, int on) {
    BH_ARG(s, "NULL sampler");
    BH_HIP(hipSetDevice(s->ctx->device));
    if (on && s->kev.empty()) {
        s->kev.resize(2 * kKernelTimerPairs, nullptr);
        for (hipEvent_t& e : s->kev) BH_HIP(hipEventCreate(&e));
    }
    BH_HIP(hipStreamSynchronize(s->stream));
    s->ktiming = on > 0;
    s->kstride = on > 0 ? (uint32_t)on : 1u;
    s->kcount = 0;
    s->kused = 0;
    return BH_OK;
}

int bh_stretch_kernel_timing(bh_stretch* s, double* avg_ms, uint32_t* launches) {
    BH_ARG(s && avg_ms && launches, "NULL argument");
    BH_HIP(hipSetDevice(s->ctx->device));
    BH_HIP(hipStreamSynchronize(s->stream));
    double total = 0.0;
    for (size_t i = 0; i + 1 < s->kused; i += 2) {
        float ms = 0.0f;
        BH_HIP(hipEventElapsedTime(&ms, s->kev[i], s->kev[i + 1]));
        total += ms;
    }
    *launches = (uint32_t)(s->kused / 2);
    *avg_ms = s->kused ? total / (double)(s->kused / 2) : 0.0;
    return BH_OK;
}

int bh_stretch_last_gpu_ms(const bh_stretch* s, double* ms) {
    BH_ARG(s && ms, "NULL argument");
    *ms = s->last_ms;
    return BH_OK;
}

int bh_stretch_set_partner(bh_stretch* s, int mode) {
    BH_ARG(s, "NULL sampler");
    BH_ARG(mode == BH_PARTNER_FLAT || mode == BH_PARTNER_WALKER, "unknown partner mode %d", mode);
    s->partner = mode;
    return BH_OK;
}

int bh_stretch_set_exec(bh_stretch* s, int mode) {
    BH_ARG(s, "NULL sampler");
    BH_ARG(mode == BH_EXEC_AUTO || mode == BH_EXEC_STEPWISE || mode == BH_EXEC_RESIDENT,
           "unknown execution mode %d", mode);
    if (mode == BH_EXEC_RESIDENT && resident_lds(s, 1) > kResidentLdsMax)
        return fail(BH_ERR_RANGE, "ensemble (%u walkers x %u) does not fit in one work-group's LDS",
                    s->W, s->dim);
    s->exec = mode;
    return BH_OK;
}

int bh_stretch_set_glm_precision(bh_stretch* s, int precision) {
    BH_ARG(s, "NULL sampler");
    BH_ARG(s->f->glm, "not a GLM sampler");
    BH_TRY(glm_prepare(s->f, precision));
    if (precision != s->glm_prec) {
        s->glm_prec = precision;
        // the next step computes the correction's gradient (and Hessian) afresh
        s->g0_age = -1;
        s->h0_valid = false;
    }
    return BH_OK;
}

int bh_stretch_set_glm_guard(bh_stretch* s, double guard) {
    BH_ARG(s, "NULL sampler");
    BH_ARG(s->f->glm, "not a GLM sampler");
    BH_ARG(s->f->glm->kind == BH_GLM_LOGISTIC, "the guarded accept test is for logistic GLMs");
    BH_ARG(guard >= 0.0 && std::isfinite(guard), "guard must be finite and >= 0");
    s->glm_guard = guard;
    return BH_OK;
}

int bh_stretch_glm_guard_stats(bh_stretch* s, uint64_t* listed, uint64_t* overflow) {
    BH_ARG(s && listed && overflow, "NULL argument");
    BH_ARG(s->f->glm, "not a GLM sampler");
    BH_HIP(hipSetDevice(s->ctx->device));
    unsigned long long t[2];
    BH_HIP(hipMemcpyAsync(t, s->gtotals.p, sizeof(t), hipMemcpyDeviceToHost, s->stream));
    BH_TRY(sync(s->stream));
    *listed = t[0];
    *overflow = t[1];
    return BH_OK;
}

// ------------------------------------------------------------------ batch
// LDS of one work-group of bh_batch_run: the accounting (2 max_dim doubles,
// 2 nb floats), the proposals (nb x dim_hi floats) and, with the state in
// LDS, the positions and log densities (W x dim_hi + W floats).
static size_t batch_lds(const bh_batch* b, uint32_t dim_hi, uint32_t nb, int gstate) {
    size_t floats = (size_t)2 * nb + (size_t)nb * dim_hi;
    if (gstate == BH_BATCH_STATE_LDS) floats += (size_t)b->W * dim_hi + b->W;
    return (size_t)2 * b->max_dim * sizeof(double) + floats * sizeof(float);
}

// Launch classes from dimension bounds (ascending; the last >= every dim):
// problems sorted by dimension, one class per bound that holds any.
static int batch_set_classes(bh_batch* b, const std::vector<uint32_t>& bounds) {
    std::vector<uint32_t> ord(b->problems);
    for (uint32_t i = 0; i < b->problems; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(),
                     [&](uint32_t x, uint32_t y) { return b->dims_h[x] < b->dims_h[y]; });
    b->classes.clear();
    uint32_t i = 0;
    for (uint32_t bound : bounds) {
        const uint32_t begin = i;
        uint32_t hi = 0;
        while (i < b->problems && b->dims_h[ord[i]] <= bound) hi = std::max(hi, b->dims_h[ord[i++]]);
        if (i > begin) b->classes.push_back({begin, i - begin, hi, b->gstate});
    }
    if (i != b->problems) return fail(BH_ERR_ARG, "class bounds do not cover every dimension");
    BH_HIP(hipMemcpyAsync(b->order.p, ord.data(), ord.size() * 4, hipMemcpyHostToDevice,
                          b->stream));
    return sync(b->stream);
}

int bh_batch_create(bh_ctx* ctx, const char* source, const char* logfn_name, uint32_t max_dim,
                    uint32_t problems, const uint32_t* dims, const uint64_t* offsets,
                    const uint32_t* hyper_len, const float* params, uint32_t walkers,
                    bh_batch** out, char* log, size_t log_size) {
    BH_ARG(ctx && source && dims && offsets && hyper_len && params && out, "NULL argument");
    BH_ARG(valid_name(logfn_name), "invalid logfn function name");
    BH_ARG(max_dim >= 1 && max_dim <= 256, "max_dim %u outside [1, 256]", max_dim);
    BH_ARG(problems >= 1 && problems <= (1u << 20), "bad problem count %u", problems);
    if (!(walkers >= 128 && walkers % 128 == 0 && walkers <= 8192))
        return fail(BH_ERR_RANGE, "batched samplers need a multiple of 128 walkers (<= 8192), got %u",
                    walkers);
    BH_HIP(hipSetDevice(ctx->device));
    auto b = std::make_unique<bh_batch>();
    b->ctx = ctx;
    b->max_dim = max_dim;
    b->problems = problems;
    b->W = walkers;
    const uint32_t K = walkers / 2;
    uint32_t nb = 1;
    while (nb * 2 <= std::min(K, 1024u) && K % (nb * 2) == 0) nb *= 2;
    b->nb = std::max(nb, 64u);
    std::vector<uint32_t> dl(problems), pl(problems);
    for (uint32_t i = 0; i < problems; ++i) {
        BH_ARG(dims[i] >= 1 && dims[i] <= max_dim, "problem %u: dim %u outside [1, %u]", i,
               dims[i], max_dim);
        BH_ARG((uint64_t)dims[i] * walkers % 4 == 0, "problem %u: dim*walkers not a multiple of 4", i);
        BH_ARG(offsets[i + 1] >= offsets[i] && offsets[i + 1] - offsets[i] >= hyper_len[i],
               "problem %u: bad params range", i);
        pl[i] = hyper_len[i];
        dl[i] = (uint32_t)(offsets[i + 1] - offsets[i] - hyper_len[i]);
        b->dim_hi = std::max(b->dim_hi, dims[i]);
    }
    int lds_max = 0;
    BH_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
    // the state goes to LDS when it fits there, else it stays in global
    // memory (the proposals alone must fit)
    if (batch_lds(b.get(), b->dim_hi, b->nb, BH_BATCH_STATE_LDS) > (size_t)lds_max)
        b->gstate = BH_BATCH_STATE_GLOBAL;
    if (batch_lds(b.get(), b->dim_hi, b->nb, BH_BATCH_STATE_GLOBAL) > (size_t)lds_max)
        return fail(BH_ERR_RANGE, "%u walkers x %u dims need %zu bytes of LDS per problem (%d available)",
                    walkers, b->dim_hi,
                    batch_lds(b.get(), b->dim_hi, b->nb, BH_BATCH_STATE_GLOBAL), lds_max);
    b->lds_max = lds_max;
    std::vector<char> code;
    BH_TRY(rtc_compile(ctx->arch, source,
                       {"-DWGS=" + std::to_string(ctx->wgs), "-DDIM=" + std::to_string(max_dim),
                        std::string("-DLOGFN=") + logfn_name, "-DBH_BATCH=1"},
                       code, log, log_size));
    BH_TRY(b->mod.load(code));
    BH_TRY(b->mod.fn("bh_batch_run", &b->run));
    BH_TRY(b->mod.fn("bh_batch_run_g", &b->run_g));
    BH_TRY(b->mod.fn("bh_batch_init", &b->init));
    BH_TRY(b->mod.fn("bh_batch_logfn", &b->logfn));
    BH_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BH_HIP(hipEventCreate(&b->ev0));
    BH_HIP(hipEventCreate(&b->ev1));
    const uint64_t total = offsets[problems];
    BH_TRY(b->dims.alloc((size_t)problems * 4));
    BH_TRY(b->dlen.alloc((size_t)problems * 4));
    BH_TRY(b->plen.alloc((size_t)problems * 4));
    BH_TRY(b->poff.alloc((size_t)problems * 8));
    BH_TRY(b->params.alloc(std::max<size_t>(total, 1) * sizeof(float)));
    BH_TRY(b->xs.alloc((size_t)problems * walkers * max_dim * sizeof(float)));
    BH_TRY(b->logfn_buf.alloc((size_t)problems * walkers * sizeof(float)));
    BH_TRY(b->accept.alloc((size_t)problems * 8));
    BH_TRY(b->sums.alloc((size_t)problems * max_dim * 2 * sizeof(double)));
    BH_TRY(b->order.alloc((size_t)problems * 4));
    BH_HIP(hipMemcpyAsync(b->dims.p, dims, (size_t)problems * 4, hipMemcpyDefault, b->stream));
    BH_HIP(hipMemcpyAsync(b->dlen.p, dl.data(), (size_t)problems * 4, hipMemcpyDefault, b->stream));
    BH_HIP(hipMemcpyAsync(b->plen.p, pl.data(), (size_t)problems * 4, hipMemcpyDefault, b->stream));
    BH_HIP(hipMemcpyAsync(b->poff.p, offsets, (size_t)problems * 8, hipMemcpyDefault, b->stream));
    if (total)
        BH_HIP(hipMemcpyAsync(b->params.p, params, total * sizeof(float), hipMemcpyDefault,
                              b->stream));
    BH_HIP(hipMemsetAsync(b->xs.p, 0, (size_t)problems * walkers * max_dim * sizeof(float),
                          b->stream));
    BH_HIP(hipMemsetAsync(b->logfn_buf.p, 0, (size_t)problems * walkers * sizeof(float), b->stream));
    b->dims_h.assign(dims, dims + problems);
    BH_TRY(batch_set_classes(b.get(), {b->dim_hi}));
    *out = b.release();
    return BH_OK;
}

int bh_batch_destroy(bh_batch* b) {
    if (!b) return BH_OK;
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipStreamDestroy(b->stream);
    }
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
    return BH_OK;
}

int bh_batch_set_partner(bh_batch* b, int mode) {
    BH_ARG(b, "NULL batch");
    BH_ARG(mode == BH_PARTNER_FLAT || mode == BH_PARTNER_WALKER, "unknown partner mode %d", mode);
    b->partner = mode;
    return BH_OK;
}

int bh_batch_init_position(bh_batch* b, uint32_t seed, const float* limits) {
    BH_ARG(b && limits, "NULL argument");
    BH_HIP(hipSetDevice(b->ctx->device));
    DevIn dl;
    BH_TRY(dl.init(limits, (size_t)b->problems * b->max_dim * 2 * sizeof(float), b->stream));
    const unsigned* dims = b->dims.as<unsigned>();
    const unsigned long long* poff = b->poff.as<unsigned long long>();
    const unsigned* dlen = b->dlen.as<unsigned>();
    const unsigned* plen = b->plen.as<unsigned>();
    const float* params = b->params.as<float>();
    const float* lim = dl.as<float>();
    float* xs = b->xs.as<float>();
    float* lf = b->logfn_buf.as<float>();
    unsigned W = b->W;
    void* ai[] = {&seed, &dims, &poff, &dlen, &plen, &params, &lim, &xs, &lf, &W};
    const unsigned gx = std::max(1u, (b->dim_hi * W / 4 + 255) / 256);
    BH_HIP(hipModuleLaunchKernel(b->init, gx, b->problems, 1, 256, 1, 1, 0, b->stream, ai, nullptr));
    void* al[] = {&dims, &poff, &dlen, &plen, &params, &xs, &lf, &W};
    BH_HIP(hipModuleLaunchKernel(b->logfn, (W + 255) / 256, b->problems, 1, 256, 1, 1, 0, b->stream,
                                 al, nullptr));
    b->bare_counter = 0;
    return sync(b->stream);
}

static int batch_steps(bh_batch* b, uint32_t seed0, uint64_t n, float a, int mode,
                       const unsigned long long* soff = nullptr, float* series = nullptr,
                       std::vector<hipEvent_t>* class_events = nullptr) {
    BH_HIP(hipSetDevice(b->ctx->device));
    const unsigned* dims = b->dims.as<unsigned>();
    const unsigned long long* poff = b->poff.as<unsigned long long>();
    const unsigned* dlen = b->dlen.as<unsigned>();
    const unsigned* plen = b->plen.as<unsigned>();
    const float* params = b->params.as<float>();
    float* xs = b->xs.as<float>();
    float* lf = b->logfn_buf.as<float>();
    unsigned W = b->W;
    int partner = b->partner;
    unsigned long long* acc = b->accept.as<unsigned long long>();
    double* sums = b->sums.as<double>();
    BH_HIP(hipEventRecord(b->ev0, b->stream));
    uint64_t done = 0;
    while (done < n) {
        unsigned step0 = b->bare_counter;
        unsigned chunk = (unsigned)std::min<uint64_t>(n - done, 1u << 20);
        float* ser = series;  // recorded runs are a single launch (n < 2^20)
        const unsigned* ord = b->order.as<unsigned>();
        for (size_t c = 0; c < b->classes.size(); ++c) {
            const auto& cl = b->classes[c];
            unsigned base = cl.begin;
            void* args[] = {&seed0, &dims, &poff, &dlen, &plen, &params, &xs, &lf, &W, &a,
                            &step0, &chunk, &partner, &mode, &acc, &sums, &soff, &ser, &ord,
                            &base};
            if (class_events) BH_HIP(hipEventRecord((*class_events)[c], b->stream));
            BH_HIP(hipModuleLaunchKernel(cl.gstate == BH_BATCH_STATE_GLOBAL ? b->run_g : b->run,
                                         cl.count, 1, 1, b->nb, 1, 1,
                                         (unsigned)batch_lds(b, cl.dim_hi, b->nb, cl.gstate),
                                         b->stream, args, nullptr));
        }
        if (class_events) BH_HIP(hipEventRecord(class_events->back(), b->stream));
        b->bare_counter += chunk;
        done += chunk;
    }
    BH_HIP(hipEventRecord(b->ev1, b->stream));
    BH_HIP(hipEventSynchronize(b->ev1));
    float ms = 0.0f;
    BH_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    return BH_OK;
}

int bh_batch_set_state(bh_batch* b, int where) {
    BH_ARG(b, "NULL batch");
    BH_ARG(where == BH_BATCH_STATE_LDS || where == BH_BATCH_STATE_GLOBAL,
           "unknown state placement %d", where);
    for (const auto& cl : b->classes)
        if (batch_lds(b, cl.dim_hi, b->nb, where) > (size_t)b->lds_max)
            return fail(BH_ERR_RANGE, "%u walkers x %u dims do not fit in %d bytes of LDS", b->W,
                        cl.dim_hi, b->lds_max);
    b->gstate = where;
    for (auto& cl : b->classes) cl.gstate = where;
    return BH_OK;
}

int bh_batch_autotune(bh_batch* b, uint32_t steps, uint32_t* block_out, uint32_t* classes_out,
                      double* ms_out) {
    BH_ARG(b, "NULL batch");
    BH_ARG(steps >= 1 && steps <= 10000, "bad step count");
    BH_HIP(hipSetDevice(b->ctx->device));
    // the state every candidate starts from, restored afterwards
    const size_t xbytes = (size_t)b->problems * b->W * b->max_dim * sizeof(float);
    const size_t lbytes = (size_t)b->problems * b->W * sizeof(float);
    DevBuf xs0, lf0;
    BH_TRY(xs0.alloc(xbytes));
    BH_TRY(lf0.alloc(lbytes));
    BH_HIP(hipMemcpyAsync(xs0.p, b->xs.p, xbytes, hipMemcpyDeviceToDevice, b->stream));
    BH_HIP(hipMemcpyAsync(lf0.p, b->logfn_buf.p, lbytes, hipMemcpyDeviceToDevice, b->stream));
    const uint32_t counter0 = b->bare_counter;
    std::vector<uint32_t> sorted(b->dims_h);
    std::sort(sorted.begin(), sorted.end());
    std::vector<std::vector<uint32_t>> bound_sets;
    for (uint32_t nc = 1; nc <= 4; ++nc) {
        std::vector<uint32_t> bs;
        for (uint32_t c = 1; c < nc; ++c) bs.push_back(sorted[(size_t)sorted.size() * c / nc]);
        bs.push_back(sorted.back());
        bs.erase(std::unique(bs.begin(), bs.end()), bs.end());
        bound_sets.push_back(bs);
    }
    // one event before each class launch and one after the last
    struct Events {
        std::vector<hipEvent_t> ev;
        ~Events() {
            for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        }
    } events;
    for (int i = 0; i < 5; ++i) {
        hipEvent_t e;
        BH_HIP(hipEventCreate(&e));
        events.ev.push_back(e);
    }
    const uint32_t K = b->W / 2;
    const int gstate0 = b->gstate;
    double best = -1.0;
    uint32_t best_nb = b->nb;
    std::vector<uint32_t> best_bounds = {b->dim_hi};
    std::vector<int> best_states;
    for (uint32_t nb = 64; nb <= 1024 && nb <= std::max(64u, K); nb *= 2) {
        for (const auto& bs : bound_sets) {
            b->nb = nb;
            BH_TRY(batch_set_classes(b, bs));
            const size_t nc = b->classes.size();
            // per class, the time of each placement that fits (-1: does not)
            std::vector<std::array<double, 2>> cms(nc, {-1.0, -1.0});
            for (int where = BH_BATCH_STATE_LDS; where <= BH_BATCH_STATE_GLOBAL; ++where) {
                bool any = false;
                for (auto& cl : b->classes) {
                    cl.gstate = batch_lds(b, cl.dim_hi, nb, where) <= (size_t)b->lds_max
                                    ? where
                                    : BH_BATCH_STATE_GLOBAL;
                    any = any || cl.gstate == where;
                }
                if (!any) continue;
                if (batch_lds(b, b->classes.back().dim_hi, nb, BH_BATCH_STATE_GLOBAL) >
                    (size_t)b->lds_max)
                    break;
                BH_HIP(hipMemcpyAsync(b->xs.p, xs0.p, xbytes, hipMemcpyDeviceToDevice,
                                      b->stream));
                BH_HIP(hipMemcpyAsync(b->logfn_buf.p, lf0.p, lbytes, hipMemcpyDeviceToDevice,
                                      b->stream));
                b->bare_counter = counter0;
                std::vector<hipEvent_t> ce(events.ev.begin(), events.ev.begin() + nc + 1);
                BH_TRY(batch_steps(b, 0x5eedu, steps, 2.0f, 0, nullptr, nullptr, &ce));
                for (size_t c = 0; c < nc; ++c) {
                    if (b->classes[c].gstate != where) continue;
                    float ms = 0.0f;
                    BH_HIP(hipEventElapsedTime(&ms, ce[c], ce[c + 1]));
                    cms[c][where] = ms;
                }
            }
            double total = 0.0;
            std::vector<int> states(nc);
            bool ok = true;
            for (size_t c = 0; c < nc; ++c) {
                const double l = cms[c][0], g = cms[c][1];
                if (l < 0.0 && g < 0.0) ok = false;
                states[c] = (g >= 0.0 && (l < 0.0 || g < l)) ? BH_BATCH_STATE_GLOBAL
                                                            : BH_BATCH_STATE_LDS;
                total += states[c] == BH_BATCH_STATE_GLOBAL ? g : l;
            }
            if (!ok) continue;
            if (best < 0.0 || total < best) {
                best = total;
                best_nb = nb;
                best_bounds = bs;
                best_states = states;
            }
        }
    }
    b->nb = best_nb;
    b->gstate = gstate0;
    BH_TRY(batch_set_classes(b, best_bounds));
    for (size_t c = 0; c < best_states.size() && c < b->classes.size(); ++c)
        b->classes[c].gstate = best_states[c];
    BH_HIP(hipMemcpyAsync(b->xs.p, xs0.p, xbytes, hipMemcpyDeviceToDevice, b->stream));
    BH_HIP(hipMemcpyAsync(b->logfn_buf.p, lf0.p, lbytes, hipMemcpyDeviceToDevice, b->stream));
    b->bare_counter = counter0;
    BH_TRY(sync(b->stream));
    if (block_out) *block_out = b->nb;
    if (classes_out) *classes_out = (uint32_t)b->classes.size();
    if (ms_out) *ms_out = best;
    return BH_OK;
}

int bh_batch_class_states(const bh_batch* b, uint32_t* count, int* states) {
    BH_ARG(b && count, "NULL argument");
    if (states)
        for (size_t c = 0; c < b->classes.size() && c < *count; ++c) states[c] = b->classes[c].gstate;
    *count = (uint32_t)b->classes.size();
    return BH_OK;
}

int bh_batch_burn_in(bh_batch* b, uint32_t seed0, uint64_t n, float a) {
    BH_ARG(b, "NULL batch");
    BH_HIP(hipMemsetAsync(b->accept.p, 0, (size_t)b->problems * 8, b->stream));
    return batch_steps(b, seed0, n, a, 0);
}

int bh_batch_run_series(bh_batch* b, uint32_t seed0, uint64_t n, float a,
                        double* acceptance_rate, double* mean, double* sd, float* series) {
    BH_ARG(b, "NULL batch");
    BH_ARG(n > 0, "no steps");
    BH_ARG(!series || n < (1u << 20), "a recorded run takes fewer than 2^20 steps");
    BH_HIP(hipSetDevice(b->ctx->device));
    BH_HIP(hipMemsetAsync(b->accept.p, 0, (size_t)b->problems * 8, b->stream));
    BH_HIP(hipMemsetAsync(b->sums.p, 0, (size_t)b->problems * b->max_dim * 2 * sizeof(double),
                          b->stream));
    DevOut ser;
    DevBuf soff;
    if (series) {
        // problem p's n x dims[p] block starts at n * (dims[0] + .. + dims[p-1])
        std::vector<uint32_t> dh(b->problems);
        BH_HIP(hipMemcpyAsync(dh.data(), b->dims.p, dh.size() * 4, hipMemcpyDeviceToHost,
                              b->stream));
        BH_TRY(sync(b->stream));
        std::vector<unsigned long long> off(b->problems);
        unsigned long long o = 0;
        for (uint32_t p = 0; p < b->problems; ++p) {
            off[p] = o;
            o += (unsigned long long)dh[p] * n;
        }
        BH_TRY(soff.alloc(off.size() * 8));
        BH_HIP(hipMemcpyAsync(soff.p, off.data(), off.size() * 8, hipMemcpyHostToDevice,
                              b->stream));
        BH_TRY(ser.init(series, (size_t)o * sizeof(float)));
    }
    BH_TRY(batch_steps(b, seed0, n, a, 1, series ? soff.as<unsigned long long>() : nullptr,
                       series ? ser.as<float>() : nullptr));
    if (series) BH_TRY(ser.finish(b->stream));
    std::vector<unsigned long long> acc(b->problems);
    std::vector<double> sums((size_t)b->problems * b->max_dim * 2);
    std::vector<uint32_t> dims(b->problems);
    BH_HIP(hipMemcpyAsync(acc.data(), b->accept.p, acc.size() * 8, hipMemcpyDeviceToHost, b->stream));
    BH_HIP(hipMemcpyAsync(sums.data(), b->sums.p, sums.size() * 8, hipMemcpyDeviceToHost, b->stream));
    BH_HIP(hipMemcpyAsync(dims.data(), b->dims.p, dims.size() * 4, hipMemcpyDeviceToHost, b->stream));
    BH_TRY(sync(b->stream));
    const double cnt = (double)b->W * (double)n;
    for (uint32_t p = 0; p < b->problems; ++p) {
        if (acceptance_rate) acceptance_rate[p] = (double)acc[p] / cnt;
        for (uint32_t i = 0; i < b->max_dim; ++i) {
            const size_t o = (size_t)p * b->max_dim + i;
            const bool valid = i < dims[p];
            const double m = valid ? sums[o * 2] / cnt : 0.0;
            const double v = valid ? std::max(0.0, sums[o * 2 + 1] / cnt - m * m) : 0.0;
            if (mean) mean[o] = m;
            if (sd) sd[o] = std::sqrt(v);
        }
    }
    return BH_OK;
}

int bh_batch_run(bh_batch* b, uint32_t seed0, uint64_t n, float a, double* acceptance_rate,
                 double* mean, double* sd) {
    return bh_batch_run_series(b, seed0, n, a, acceptance_rate, mean, sd, nullptr);
}

int bh_batch_state_get(const bh_batch* b, float* xs, float* logfn) {
    BH_ARG(b, "NULL batch");
    BH_HIP(hipSetDevice(b->ctx->device));
    if (xs)
        BH_HIP(hipMemcpyAsync(xs, b->xs.p, (size_t)b->problems * b->W * b->max_dim * sizeof(float),
                              hipMemcpyDefault, b->stream));
    if (logfn)
        BH_HIP(hipMemcpyAsync(logfn, b->logfn_buf.p, (size_t)b->problems * b->W * sizeof(float),
                              hipMemcpyDefault, b->stream));
    return sync(b->stream);
}

int bh_batch_last_gpu_ms(const bh_batch* b, double* ms) {
    BH_ARG(b && ms, "NULL argument");
    *ms = b->last_ms;
    return BH_OK;
}

// ---------------------------------------------------------------- GLM batch
}  // extern "C"

struct bh_glm_batch {
    bh_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    uint32_t P = 0, W = 0, K = 0, U = 0, n32 = 0, n64 = 0;
    int prec = BH_GLM_BF16X3;
    double prior_prec = 1.0;
    std::vector<uint32_t> dims_h;
    std::vector<uint64_t> rows_h;
    // tile shape of the matrix-core contraction per row stride (index 0:
    // stride 32, 1: stride 64): rows per unit, 16-point column blocks per wave
    uint32_t unit_rows[2] = {bh::kGbUnitRows, bh::kGbUnitRows};
    uint32_t col_blocks[2] = {4, 4};
    DevBuf dims, scales, tscale, X8, xoff, units32, units64, ubeg, lin, lin_half;
    DevBuf xs, lp, Y, z, u2, ysc, ysplit, partial, acc_cnt, msum, msum2, acc1, acc2;
    uint32_t bare_counter = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
};

namespace {

// e4m3 (OCP: bias 7, no infinities, S.1111.111 is NaN) to float
float fp8_e4m3_value(uint8_t c) {
    const int s = c >> 7, e = (c >> 3) & 15, m = c & 7;
    float v;
    if (e == 15 && m == 7) v = NAN;
    else if (e == 0) v = std::ldexp((float)m, -9);
    else v = std::ldexp(1.0f + (float)m / 8.0f, e - 7);
    return s ? -v : v;
}

bool gb_mfma(int prec) {
    return prec == BH_GLM_BF16 || prec == BH_GLM_BF16X3 || prec == BH_GLM_FP16X2;
}

// The contraction units of every problem: its rows in pieces of the tile's
// unit rows of its stride class, the units of one stride in problem order,
// global unit ids in problem order (ubeg[p] the first of problem p).
int gb_build_units(bh_glm_batch* b) {
    std::vector<uint32_t> u32, u64, ub(b->P + 1);
    uint32_t uid = 0;
    for (uint32_t p = 0; p < b->P; ++p) {
        const int cls = b->dims_h[p] <= 32 ? 0 : 1;
        const unsigned long long N = b->rows_h[p], ur = b->unit_rows[cls];
        auto& dst = cls == 0 ? u32 : u64;
        ub[p] = uid;
        for (unsigned long long r0 = 0; r0 < N; r0 += ur) {
            const unsigned long long r1 = std::min<unsigned long long>(N, r0 + ur);
            dst.insert(dst.end(), {p, (uint32_t)r0, (uint32_t)r1, uid++});
        }
    }
    ub[b->P] = uid;
    // the kernels running on the stream may still read the old units
    BH_TRY(sync(b->stream));
    b->U = uid;
    b->n32 = (uint32_t)(u32.size() / 4);
    b->n64 = (uint32_t)(u64.size() / 4);
    auto up = [&](DevBuf& d, const void* h, size_t bytes) -> int {
        BH_TRY(d.alloc(std::max<size_t>(bytes, 16)));
        if (bytes) BH_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, b->stream));
        return BH_OK;
    };
    BH_TRY(up(b->units32, u32.data(), u32.size() * 4));
    BH_TRY(up(b->units64, u64.data(), u64.size() * 4));
    BH_TRY(up(b->ubeg, ub.data(), ub.size() * 4));
    BH_TRY(b->partial.alloc((size_t)std::max(uid, 1u) * b->W * 8));
    return sync(b->stream);
}

// one stride class's contraction of n points (cls 0: stride 32, 1: 64)
int gb_contract_class(bh_glm_batch* b, int cls, const float* Ypts, unsigned n) {
    const int cb = gb_mfma(b->prec) ? (int)b->col_blocks[cls] : 4;
    BH_HIP(bh::launch_gb_contract(b->prec, cls == 0 ? 32 : 64, cb, b->stream,
                                  b->X8.as<unsigned char>(), b->xoff.as<unsigned long long>(),
                                  cls == 0 ? b->units32.p : b->units64.p,
                                  cls == 0 ? b->n32 : b->n64, b->P, b->scales.as<float>(), Ypts,
                                  b->ysc.as<float>(), b->ysplit.p, n, b->partial.as<double>()));
    return BH_OK;
}

// log posteriors of the walkers of every problem (n = W points each)
int gb_contract(bh_glm_batch* b, const float* Ypts, unsigned n) {
    BH_TRY(gb_contract_class(b, 0, Ypts, n));
    return gb_contract_class(b, 1, Ypts, n);
}

const double* gb_lin(const bh_glm_batch* b) {
    return (b->prec == BH_GLM_FP64 ? b->lin : b->lin_half).as<double>();
}

double gb_scale(const bh_glm_batch* b) { return b->prec == BH_GLM_FP64 ? 1.0 : kLn2; }

int gb_logpost_all(bh_glm_batch* b) {
    const int ns = glm_nsplit(b->prec);
    BH_HIP(bh::launch_gb_points(b->stream, b->xs.as<float>(), b->dims.as<unsigned>(),
                                b->tscale.as<float>(), b->P, b->W, b->ysc.as<float>(),
                                b->ysplit.p, ns, glm_f16(b->prec)));
    BH_TRY(gb_contract(b, b->xs.as<float>(), b->W));
    BH_HIP(bh::launch_gb_logpost(b->stream, b->xs.as<float>(), b->dims.as<unsigned>(),
                                 b->ubeg.as<unsigned>(), b->partial.as<double>(), b->P, b->W,
                                 gb_lin(b), gb_scale(b), b->prior_prec, b->lp.as<double>()));
    return BH_OK;
}

// n steps (both halves each) with seeds seed0 + 2p / + 1 and constants
// 3333/4444; with accu the accounting and, if series, the per-step means.
int gb_steps(bh_glm_batch* b, uint32_t seed0, uint64_t n, float a, int accu,
             const unsigned long long* soff, float* series) {
    const int ns = glm_nsplit(b->prec);
    BH_HIP(hipEventRecord(b->ev0, b->stream));
    for (uint64_t st = 0; st < n; ++st) {
        const unsigned step = b->bare_counter;
        for (int half = 0; half < 2; ++half) {
            BH_HIP(bh::launch_gb_propose(b->stream, seed0, half, half ? 4444u : 3333u, step, a,
                                         b->dims.as<unsigned>(), b->tscale.as<float>(),
                                         b->xs.as<float>(), b->P, b->K, b->Y.as<float>(),
                                         b->z.as<float>(), b->u2.as<float>(), b->ysc.as<float>(),
                                         b->ysplit.p, ns, glm_f16(b->prec)));
            BH_TRY(gb_contract(b, b->Y.as<float>(), b->K));
            BH_HIP(bh::launch_gb_accept(b->stream, half, b->dims.as<unsigned>(),
                                        b->ubeg.as<unsigned>(), b->Y.as<float>(),
                                        b->z.as<float>(), b->u2.as<float>(),
                                        b->partial.as<double>(), gb_lin(b), gb_scale(b),
                                        b->prior_prec, b->xs.as<float>(), b->lp.as<double>(),
                                        1.0f, b->P, b->K, accu,
                                        b->acc_cnt.as<unsigned long long>(), b->msum.as<float>(),
                                        b->msum2.as<float>()));
        }
        if (accu)
            BH_HIP(bh::launch_gb_step(b->stream, b->dims.as<unsigned>(), b->msum.as<float>(),
                                      b->msum2.as<float>(), b->P, b->W, b->acc1.as<double>(),
                                      b->acc2.as<double>(), soff, series, (unsigned)st));
        b->bare_counter += 1;
    }
    BH_HIP(hipEventRecord(b->ev1, b->stream));
    BH_HIP(hipEventSynchronize(b->ev1));
    float ms = 0.0f;
    BH_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    return BH_OK;
}

}  // namespace

extern "C" {

static int fp8_decode(bh_ctx* ctx, const void* in, float* out, uint64_t n, int f16) {
    BH_ARG(ctx && in && out, "NULL argument");
    BH_ARG(n > 0 && n % 4 == 0, "n must be a positive multiple of 4");
    BH_HIP(hipSetDevice(ctx->device));
    DevIn din;
    BH_TRY(din.init(in, n, ctx->stream));
    DevOut dout;
    BH_TRY(dout.init(out, n * sizeof(float)));
    BH_HIP(bh::launch_gb_fp8_decode(ctx->stream, din.dev, dout.as<float>(), n, f16));
    BH_TRY(dout.finish(ctx->stream));
    return sync(ctx->stream);
}

int bh_fp8_decode(bh_ctx* ctx, const void* in, float* out, uint64_t n) {
    return fp8_decode(ctx, in, out, n, 0);
}

int bh_fp8_decode_f16(bh_ctx* ctx, const void* in, float* out, uint64_t n) {
    return fp8_decode(ctx, in, out, n, 1);
}

int bh_glm_batch_create(bh_ctx* ctx, int kind, uint32_t problems, const uint64_t* rows,
                        const uint32_t* dims, const void* X_fp8, const float* col_scales,
                        const float* y, float prior_sd, uint32_t walkers, bh_glm_batch** out) {
    BH_ARG(ctx && rows && dims && X_fp8 && col_scales && y && out, "NULL argument");
    BH_ARG(kind == BH_GLM_LOGISTIC, "unknown GLM kind %d", kind);
    BH_ARG(problems >= 1 && problems <= 65535, "bad problem count %u", problems);
    BH_ARG(prior_sd > 0.0f, "prior sd must be positive");
    if (!(walkers >= 512 && walkers % 512 == 0 && walkers <= 65536))
        return fail(BH_ERR_RANGE, "batched GLM samplers need a multiple of 512 walkers, got %u",
                    walkers);
    BH_HIP(hipSetDevice(ctx->device));
    auto b = std::make_unique<bh_glm_batch>();
    b->ctx = ctx;
    b->P = problems;
    b->W = walkers;
    b->K = walkers / 2;
    b->prior_prec = 1.0 / ((double)prior_sd * (double)prior_sd);
    b->dims_h.assign(dims, dims + problems);
    float dec[256];
    for (int c = 0; c < 256; ++c) dec[c] = fp8_e4m3_value((uint8_t)c);
    // padded data, scales, linear terms and units
    const uint8_t* xin = static_cast<const uint8_t*>(X_fp8);
    std::vector<unsigned long long> xoff(problems);
    unsigned long long total = 0, xin_off = 0, yoff = 0, soff = 0;
    for (uint32_t p = 0; p < problems; ++p) {
        BH_ARG(dims[p] >= 1 && dims[p] <= 64, "problem %u: dimension %u outside [1, 64]", p, dims[p]);
        BH_ARG(rows[p] >= 1 && rows[p] < (1ull << 31), "problem %u: bad row count", p);
        const unsigned D = dims[p] <= 32 ? 32 : 64;
        xoff[p] = total;
        total += (rows[p] * D + 255) / 256 * 256;
    }
    std::vector<uint8_t> xpad(total, 0);
    std::vector<float> sc((size_t)problems * 64, 0.0f), ts((size_t)problems * 64, 0.0f);
    std::vector<double> lin((size_t)problems * 64, 0.0), linh((size_t)problems * 64, 0.0);
    for (uint32_t p = 0; p < problems; ++p) {
        const unsigned d = dims[p], D = d <= 32 ? 32 : 64;
        const unsigned long long N = rows[p];
        for (unsigned j = 0; j < d; ++j) {
            const float s = col_scales[soff + j];
            BH_ARG(std::isfinite(s), "problem %u: column scale %u not finite", p, j);
            sc[(size_t)p * 64 + j] = s;
            ts[(size_t)p * 64 + j] = (float)((double)s * 1.4426950408889634);
        }
        std::vector<double> l(d, 0.0), cs(d, 0.0);
        for (unsigned long long i = 0; i < N; ++i) {
            const uint8_t* r = xin + xin_off + i * d;
            std::memcpy(&xpad[xoff[p] + i * D], r, d);
            const double yi = y[yoff + i];
            for (unsigned j = 0; j < d; ++j) {
                const float q = dec[r[j]];
                BH_ARG(std::isfinite(q), "problem %u: fp8 NaN in row %llu", p, i);
                const double x = (double)sc[(size_t)p * 64 + j] * (double)q;
                l[j] += yi * x;
                cs[j] += x;
            }
        }
        for (unsigned j = 0; j < d; ++j) {
            lin[(size_t)p * 64 + j] = l[j];
            linh[(size_t)p * 64 + j] = l[j] - 0.5 * cs[j];
        }
        xin_off += N * d;
        yoff += N;
        soff += d;
    }
    b->rows_h.assign(rows, rows + problems);
    const uint32_t P = problems, W = walkers, K = walkers / 2;
    BH_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BH_HIP(hipEventCreate(&b->ev0));
    BH_HIP(hipEventCreate(&b->ev1));
    auto up = [&](DevBuf& d, const void* h, size_t bytes) -> int {
        BH_TRY(d.alloc(std::max<size_t>(bytes, 16)));
        if (bytes) BH_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, b->stream));
        return BH_OK;
    };
    BH_TRY(up(b->dims, dims, (size_t)P * 4));
    BH_TRY(up(b->scales, sc.data(), sc.size() * 4));
    BH_TRY(up(b->tscale, ts.data(), ts.size() * 4));
    BH_TRY(up(b->X8, xpad.data(), xpad.size()));
    BH_TRY(up(b->xoff, xoff.data(), xoff.size() * 8));
    BH_TRY(up(b->lin, lin.data(), lin.size() * 8));
    BH_TRY(up(b->lin_half, linh.data(), linh.size() * 8));
    BH_TRY(b->xs.alloc((size_t)P * W * 64 * 4));
    BH_TRY(b->lp.alloc((size_t)P * W * 8));
    BH_TRY(b->Y.alloc((size_t)P * K * 64 * 4));
    BH_TRY(b->z.alloc((size_t)P * K * 4));
    BH_TRY(b->u2.alloc((size_t)P * K * 4));
    BH_TRY(b->ysc.alloc((size_t)P * W * 64 * 4));
    BH_TRY(b->ysplit.alloc((size_t)3 * P * W * 64 * 2));
    BH_TRY(gb_build_units(b.get()));
    BH_TRY(b->acc_cnt.alloc((size_t)P * 8));
    BH_TRY(b->msum.alloc((size_t)P * (W / 256) * 64 * 4));
    BH_TRY(b->msum2.alloc((size_t)P * (W / 256) * 64 * 4));
    BH_TRY(b->acc1.alloc((size_t)P * 64 * 8));
    BH_TRY(b->acc2.alloc((size_t)P * 64 * 8));
    BH_HIP(hipMemsetAsync(b->xs.p, 0, (size_t)P * W * 64 * 4, b->stream));
    BH_TRY(sync(b->stream));
    *out = b.release();
    return BH_OK;
}

int bh_glm_batch_destroy(bh_glm_batch* b) {
    if (!b) return BH_OK;
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipStreamDestroy(b->stream);
    }
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
    return BH_OK;
}

int bh_glm_batch_set_precision(bh_glm_batch* b, int precision) {
    BH_ARG(b, "NULL batch");
    BH_ARG(precision >= BH_GLM_FP64 && precision <= BH_GLM_FP16X2, "bad precision %d", precision);
    b->prec = precision;
    return BH_OK;
}

static int gb_check_tiles(const bh_glm_batch* b, uint32_t row_stride, uint32_t unit_rows,
                          uint32_t col_blocks) {
    BH_ARG(row_stride == 32 || row_stride == 64, "row stride %u is not 32 or 64", row_stride);
    BH_ARG(unit_rows >= 256 && unit_rows <= 65536 && unit_rows % 16 == 0,
           "unit rows %u outside [256, 65536] or not a multiple of 16", unit_rows);
    BH_ARG(col_blocks == 4 || (col_blocks == 8 && row_stride == 32),
           "column blocks %u: 4, or 8 for row stride 32", col_blocks);
    BH_ARG(b->K % (64 * col_blocks) == 0, "%u column blocks need a multiple of %u walkers",
           col_blocks, 128 * col_blocks);
    return BH_OK;
}

int bh_glm_batch_set_tiles(bh_glm_batch* b, uint32_t row_stride, uint32_t unit_rows,
                           uint32_t col_blocks) {
    BH_ARG(b, "NULL batch");
    BH_TRY(gb_check_tiles(b, row_stride, unit_rows, col_blocks));
    BH_HIP(hipSetDevice(b->ctx->device));
    const int cls = row_stride == 32 ? 0 : 1;
    b->unit_rows[cls] = unit_rows;
    b->col_blocks[cls] = col_blocks;
    BH_TRY(gb_build_units(b));
    // the stored log posteriors are sums over units: recomputed over the
    // new ones
    BH_TRY(gb_logpost_all(b));
    return sync(b->stream);
}

int bh_glm_batch_autotune(bh_glm_batch* b, uint32_t reps, uint32_t* unit_rows,
                          uint32_t* col_blocks, float* best_ms) {
    BH_ARG(b, "NULL batch");
    BH_ARG(reps >= 1 && reps <= 1000, "reps %u outside [1, 1000]", reps);
    BH_HIP(hipSetDevice(b->ctx->device));
    static const uint32_t kRows[] = {1024, 2048, 4096, 8192, 16384};
    const uint32_t keep_rows[2] = {b->unit_rows[0], b->unit_rows[1]};
    const uint32_t keep_cb[2] = {b->col_blocks[0], b->col_blocks[1]};
    uint32_t best_rows[2] = {keep_rows[0], keep_rows[1]}, best_cb[2] = {keep_cb[0], keep_cb[1]};
    float best[2] = {0.0f, 0.0f};
    // the shape of a half-step's contraction: K points (the current
    // positions' first K of every problem, split as the precision needs)
    const unsigned n = b->K;
    BH_HIP(bh::launch_gb_points(b->stream, b->xs.as<float>(), b->dims.as<unsigned>(),
                                b->tscale.as<float>(), b->P, n, b->ysc.as<float>(), b->ysplit.p,
                                glm_nsplit(b->prec), glm_f16(b->prec)));
    for (int cls = 0; cls < 2; ++cls) {
        for (uint32_t ur : kRows) {
            for (uint32_t cb : {4u, 8u}) {
                if (cb == 8 && (cls == 1 || !gb_mfma(b->prec) || b->K % 512 != 0)) continue;
                b->unit_rows[cls] = ur;
                b->col_blocks[cls] = cb;
                BH_TRY(gb_build_units(b));
                if ((cls == 0 ? b->n32 : b->n64) == 0) continue;
                BH_TRY(gb_contract_class(b, cls, b->xs.as<float>(), n));  // warm-up
                BH_HIP(hipEventRecord(b->ev0, b->stream));
                for (uint32_t r = 0; r < reps; ++r)
                    BH_TRY(gb_contract_class(b, cls, b->xs.as<float>(), n));
                BH_HIP(hipEventRecord(b->ev1, b->stream));
                BH_HIP(hipEventSynchronize(b->ev1));
                float ms = 0.0f;
                BH_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
                ms /= (float)reps;
                if (best[cls] == 0.0f || ms < best[cls]) {
                    best[cls] = ms;
                    best_rows[cls] = ur;
                    best_cb[cls] = cb;
                }
            }
        }
        b->unit_rows[cls] = best_rows[cls];
        b->col_blocks[cls] = best_cb[cls];
    }
    BH_TRY(gb_build_units(b));
    BH_TRY(gb_logpost_all(b));
    BH_TRY(sync(b->stream));
    for (int cls = 0; cls < 2; ++cls) {
        if (unit_rows) unit_rows[cls] = b->unit_rows[cls];
        if (col_blocks) col_blocks[cls] = b->col_blocks[cls];
        if (best_ms) best_ms[cls] = best[cls];
    }
    return BH_OK;
}

int bh_glm_batch_init_position(bh_glm_batch* b, uint32_t seed, const float* limits) {
    BH_ARG(b && limits, "NULL argument");
    BH_HIP(hipSetDevice(b->ctx->device));
    for (uint32_t p = 0; p < b->P; ++p)
        BH_ARG((uint64_t)b->dims_h[p] * b->W % 4 == 0, "problem %u: dim*walkers not a multiple of 4", p);
    DevIn dl;
    BH_TRY(dl.init(limits, (size_t)b->P * 64 * 2 * sizeof(float), b->stream));
    BH_HIP(hipMemsetAsync(b->xs.p, 0, (size_t)b->P * b->W * 64 * 4, b->stream));
    BH_HIP(bh::launch_gb_init(b->stream, seed, b->dims.as<unsigned>(), dl.as<float>(),
                              b->xs.as<float>(), b->P, b->W));
    BH_TRY(gb_logpost_all(b));
    b->bare_counter = 0;
    return sync(b->stream);
}

int bh_glm_batch_burn_in(bh_glm_batch* b, uint32_t seed0, uint64_t n, float a) {
    BH_ARG(b, "NULL batch");
    BH_HIP(hipSetDevice(b->ctx->device));
    return gb_steps(b, seed0, n, a, 0, nullptr, nullptr);
}

int bh_glm_batch_run(bh_glm_batch* b, uint32_t seed0, uint64_t n, float a,
                     double* acceptance_rate, double* mean, double* sd, float* series) {
    BH_ARG(b, "NULL batch");
    BH_ARG(n > 0 && n < (1u << 30), "bad step count");
    BH_HIP(hipSetDevice(b->ctx->device));
    const uint32_t P = b->P;
    BH_HIP(hipMemsetAsync(b->acc_cnt.p, 0, (size_t)P * 8, b->stream));
    BH_HIP(hipMemsetAsync(b->acc1.p, 0, (size_t)P * 64 * 8, b->stream));
    BH_HIP(hipMemsetAsync(b->acc2.p, 0, (size_t)P * 64 * 8, b->stream));
    DevOut ser;
    DevBuf soff;
    if (series) {
        std::vector<unsigned long long> off(P);
        unsigned long long o = 0;
        for (uint32_t p = 0; p < P; ++p) {
            off[p] = o;
            o += (unsigned long long)b->dims_h[p] * n;
        }
        BH_TRY(soff.alloc(off.size() * 8));
        BH_HIP(hipMemcpyAsync(soff.p, off.data(), off.size() * 8, hipMemcpyHostToDevice,
                              b->stream));
        BH_TRY(ser.init(series, (size_t)o * sizeof(float)));
    }
    BH_TRY(gb_steps(b, seed0, n, a, 1, series ? soff.as<unsigned long long>() : nullptr,
                    series ? ser.as<float>() : nullptr));
    if (series) BH_TRY(ser.finish(b->stream));
    std::vector<unsigned long long> acc(P);
    std::vector<double> s1((size_t)P * 64), s2((size_t)P * 64);
    BH_HIP(hipMemcpyAsync(acc.data(), b->acc_cnt.p, acc.size() * 8, hipMemcpyDeviceToHost,
                          b->stream));
    BH_HIP(hipMemcpyAsync(s1.data(), b->acc1.p, s1.size() * 8, hipMemcpyDeviceToHost, b->stream));
    BH_HIP(hipMemcpyAsync(s2.data(), b->acc2.p, s2.size() * 8, hipMemcpyDeviceToHost, b->stream));
    BH_TRY(sync(b->stream));
    const double cnt = (double)b->W * (double)n;
    for (uint32_t p = 0; p < P; ++p) {
        if (acceptance_rate) acceptance_rate[p] = (double)acc[p] / cnt;
        for (uint32_t i = 0; i < 64; ++i) {
            const size_t o = (size_t)p * 64 + i;
            const bool valid = i < b->dims_h[p];
            const double m = valid ? s1[o] / cnt : 0.0;
            const double v = valid ? std::max(0.0, s2[o] / cnt - m * m) : 0.0;
            if (mean) mean[o] = m;
            if (sd) sd[o] = std::sqrt(v);
        }
    }
    return BH_OK;
}

int bh_glm_batch_state_get(const bh_glm_batch* b, float* xs, double* lp) {
    BH_ARG(b, "NULL batch");
    BH_HIP(hipSetDevice(b->ctx->device));
    if (xs)
        BH_HIP(hipMemcpyAsync(xs, b->xs.p, (size_t)b->P * b->W * 64 * 4, hipMemcpyDefault,
                              b->stream));
    if (lp)
        BH_HIP(hipMemcpyAsync(lp, b->lp.p, (size_t)b->P * b->W * 8, hipMemcpyDefault, b->stream));
    return sync(b->stream);
}

int bh_glm_batch_log_density(bh_glm_batch* b, int precision, double* lp) {
    BH_ARG(b && lp, "NULL argument");
    BH_ARG(precision >= BH_GLM_FP64 && precision <= BH_GLM_FP16X2, "bad precision %d", precision);
    BH_HIP(hipSetDevice(b->ctx->device));
    DevBuf keep;
    BH_TRY(keep.alloc((size_t)b->P * b->W * 8));
    BH_HIP(hipMemcpyAsync(keep.p, b->lp.p, (size_t)b->P * b->W * 8, hipMemcpyDeviceToDevice,
                          b->stream));
    const int prec = b->prec;
    b->prec = precision;
    const int rc = gb_logpost_all(b);
    b->prec = prec;
    BH_TRY(rc);
    BH_HIP(hipMemcpyAsync(lp, b->lp.p, (size_t)b->P * b->W * 8, hipMemcpyDefault, b->stream));
    BH_HIP(hipMemcpyAsync(b->lp.p, keep.p, (size_t)b->P * b->W * 8, hipMemcpyDeviceToDevice,
                          b->stream));
    return sync(b->stream);
}

int bh_glm_batch_last_gpu_ms(const bh_glm_batch* b, double* ms) {
    BH_ARG(b && ms, "NULL argument");
    *ms = b->last_ms;
    return BH_OK;
}

int bh_stretch_info(const bh_stretch* s, uint32_t* walkers, uint64_t* iteration_counter) {
    BH_ARG(s, "NULL sampler");
    if (walkers) *walkers = s->W;
    if (iteration_counter) *iteration_counter = s->iter;
    return BH_OK;
}

int bh_stretch_init(bh_stretch* s, uint32_t seed) {
    BH_ARG(s, "NULL sampler");
    s->bare_seed = seed;
    s->bare_a = 2.0f;
    s->beta = 1.0f;
    s->bare_counter = 0;
    s->move_seed = seed;
    return BH_OK;
}

static int logfn_all(bh_stretch* s) {
    if (s->f->glm) {
        s->g0_age = -1;  // new positions: the next step computes the gradient afresh
        return glm_logpost_points(s, s->xs.as<float>(), s->W, s->lp.as<double>());
    }
    unsigned data_len = s->data_len, params_len = s->params_len, n = s->W;
    const float* params = s->params.as<float>();
    const float* xs = s->xs.as<float>();
    float* lf = s->logfn.as<float>();
    void* args[] = {&data_len, &params_len, &params, &xs, &lf, &n};
    return launch(s->f->logfn, (n + s->wgs - 1) / s->wgs, s->wgs, 0, s->stream, args);
}

int bh_stretch_init_position(bh_stretch* s, uint32_t seed, const float* limits) {
    BH_ARG(s && limits, "NULL argument");
    BH_HIP(hipSetDevice(s->ctx->device));
    DevIn dl;
    BH_TRY(dl.init(limits, (size_t)2 * s->dim * sizeof(float), s->stream));
    BH_HIP(bh::launch_init_walkers(s->stream, seed, dl.as<float>(), s->dim, s->W, s->xs.as<float>()));
    BH_TRY(logfn_all(s));
    s->iter = 0;
    return sync(s->stream);
}

int bh_stretch_init_position_from(bh_stretch* s, const bh_stretch* p) {
    BH_ARG(s && p, "NULL argument");
    BH_ARG(p->dim == s->dim && p->W == s->W, "position sampler has another shape");
    BH_HIP(hipSetDevice(s->ctx->device));
    BH_HIP(hipMemcpyAsync(s->xs.p, p->xs.p, (size_t)s->W * s->dim * sizeof(float),
                          hipMemcpyDeviceToDevice, s->stream));
    BH_TRY(logfn_all(s));
    s->iter = 0;
    return sync(s->stream);
}

int bh_stretch_move_bare(bh_stretch* s) {
    BH_ARG(s, "NULL sampler");
    BH_HIP(hipSetDevice(s->ctx->device));
    Timed t(s);
    BH_TRY(bare_pair(s, s->beta));
    BH_TRY(t.stop());
    s->bare_counter += 1;
    return sync(s->stream);
}

int bh_stretch_set_temperature(bh_stretch* s, double t) {
    BH_ARG(s, "NULL sampler");
    s->beta = (float)(1.0 / t);
    return BH_OK;
}

int bh_stretch_burn_in(bh_stretch* s, uint64_t n, float a) {
    BH_ARG(s, "NULL sampler");
    BH_HIP(hipSetDevice(s->ctx->device));
    s->bare_a = a;
    s->beta = 1.0f;
    Timed t(s);
    BH_TRY(bare_steps(s, n, nullptr));
    BH_TRY(t.stop());
    s->iter += n;
    return BH_OK;
}

int bh_stretch_anneal(bh_stretch* s, const double* temperatures, uint64_t n, float a) {
    BH_ARG(s && (temperatures || n == 0), "NULL argument");
    BH_HIP(hipSetDevice(s->ctx->device));
    s->bare_a = a;
    std::vector<float> betas(n);
    for (uint64_t i = 0; i < n; ++i) betas[i] = (float)(1.0 / temperatures[i]);
    Timed t(s);
    BH_TRY(bare_steps(s, n, betas.data()));
    BH_TRY(t.stop());
    s->iter += n;
    return BH_OK;
}

int bh_stretch_init_move(bh_stretch* s, float a, uint64_t max_steps) {
    BH_ARG(s, "NULL sampler");
    BH_HIP(hipSetDevice(s->ctx->device));
    BH_TRY(init_move(s, a, max_steps));
    return sync(s->stream);
}

int bh_stretch_move(bh_stretch* s) {
    BH_ARG(s, "NULL sampler");
    BH_HIP(hipSetDevice(s->ctx->device));
    if (!s->move_ready) return fail(BH_ERR_STATE, "move! before init-move!");
    if (s->move_counter + 1ull > s->means_cap)
        return fail(BH_ERR_STATE, "move! beyond the %llu steps reserved by init-move!",
                    (unsigned long long)s->means_cap);
    Timed t(s);
    BH_TRY(accu_pair(s));
    BH_TRY(t.stop());
    s->move_counter += 1;
    return sync(s->stream);
}

int bh_stretch_read_accounting(bh_stretch* s, uint32_t* accept, float* means_acc,
                               uint64_t steps) {
    BH_ARG(s, "NULL sampler");
    if (!s->move_ready) return fail(BH_ERR_STATE, "no init-move! yet");
    BH_ARG(steps <= s->means_cap, "more steps than reserved");
    if (accept)
        BH_HIP(hipMemcpyAsync(accept, s->accept.p, (size_t)s->groups * sizeof(unsigned),
                              hipMemcpyDefault, s->stream));
    if (means_acc)
        BH_HIP(hipMemcpyAsync(means_acc, s->means_acc.p,
                              (size_t)s->dim * s->groups * steps * sizeof(float), hipMemcpyDefault,
                              s->stream));
    return sync(s->stream);
}

static int accept_total(bh_stretch* s, unsigned long long* total) {
    DevBuf t;
    BH_TRY(t.alloc(sizeof(unsigned long long)));
    BH_HIP(bh::launch_sum_u32(s->stream, s->accept.as<unsigned>(), s->groups,
                              t.as<unsigned long long>()));
    BH_HIP(hipMemcpyAsync(total, t.p, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          s->stream));
    return sync(s->stream);
}

int bh_stretch_run(bh_stretch* s, uint64_t n, float a, double* acceptance_rate, float* means,
                   float* tau, float* mean, float* sigma, uint32_t* lag) {
    BH_ARG(s, "NULL sampler");
    BH_ARG(n > 0 && n < (1ull << 31), "steps %llu out of range", (unsigned long long)n);
    BH_HIP(hipSetDevice(s->ctx->device));
    BH_TRY(init_move(s, a, n));
    {
        Timed t(s);
        BH_TRY(accu_steps(s, n));
        BH_TRY(t.stop());
    }
    s->iter += n;
    // per-step ensemble means: sum_means_vertical + sum_reduction_horizontal
    // over the groups of each (step, dim) row, then scal! 0.5/(WGS*groups)
    // (amd_gcn.clj:437-439)
    const uint32_t rows = (uint32_t)(n * s->dim);
    const uint32_t local_n = std::min(s->groups, s->wgs);
    DevBuf dmeans, work;
    BH_TRY(dmeans.alloc((size_t)rows * sizeof(float)));
    BH_TRY(work.alloc((size_t)rows * ((s->groups + local_n - 1) / local_n + 1) * sizeof(float)));
    const float scale = (float)(0.5 / ((double)s->wgs * s->groups));
    BH_HIP(bh::launch_row_tree_sum(s->stream, s->means_acc.as<float>(), rows, s->groups, local_n,
                                   s->groups, work.as<float>(), scale, dmeans.as<float>(), 1));
    unsigned long long total = 0;
    BH_TRY(accept_total(s, &total));
    if (acceptance_rate) *acceptance_rate = (double)total / ((double)s->W * (double)n);
    if (means)
        BH_HIP(hipMemcpyAsync(means, dmeans.p, (size_t)rows * sizeof(float), hipMemcpyDefault,
                              s->stream));
    BH_TRY(sync(s->stream));
    return acor_device(s->ctx, s->stream, dmeans.as<float>(), s->dim, (uint32_t)n, tau, mean,
                       sigma, lag);
}

int bh_stretch_acc_rate(bh_stretch* s, float a, double* out) {
    BH_ARG(s && out, "NULL argument");
    BH_HIP(hipSetDevice(s->ctx->device));
    BH_TRY(init_move(s, a, 1));
    BH_TRY(accu_steps(s, 1));
    s->iter += 1;
    unsigned long long total = 0;
    BH_TRY(accept_total(s, &total));
    *out = (double)total / (double)s->W;
    return BH_OK;
}

int bh_stretch_sample(bh_stretch* s, float* out, uint64_t ncols) {
    BH_ARG(s && out, "NULL argument");
    BH_ARG(ncols > 0, "no columns requested");
    BH_HIP(hipSetDevice(s->ctx->device));
    s->beta = 1.0f;  // set-temperature! 1.0 (amd_gcn.clj:366)
    const uint64_t avail = (uint64_t)s->W * s->dim;
    const uint64_t req = ncols * s->dim;
    const uint64_t steps = (req + avail - 1) / avail;
    DevOut o;
    BH_TRY(o.init(out, req * sizeof(float)));
    Timed t(s);
    if (use_resident(s, 0) && steps < (1ull << 31)) {
        BH_TRY(run_resident(s, 0, s->bare_seed, s->bare_seed + 1u, 3333, 4444, s->bare_a, 1.0f,
                            nullptr, s->bare_counter, (uint32_t)steps, o.as<float>(), req));
        s->bare_counter += (uint32_t)steps;
    } else {
        uint64_t ofst = 0;
        for (uint64_t i = 0; i < steps; ++i) {
            BH_TRY(bare_pair(s, 1.0f));
            s->bare_counter += 1;
            const uint64_t chunk = std::min(avail, req - ofst);
            BH_HIP(hipMemcpyAsync(o.as<float>() + ofst, s->xs.p, chunk * sizeof(float),
                                  hipMemcpyDeviceToDevice, s->stream));
            ofst += chunk;
        }
    }
    BH_TRY(t.stop());
    s->iter += steps;
    BH_TRY(o.finish(s->stream));
    return sync(s->stream);
}

int bh_stretch_histogram(bh_stretch* s, uint32_t cycles, float* limits, float* pdf,
                         float* bin_ranks) {
    BH_ARG(s && limits && pdf && bin_ranks, "NULL argument");
    BH_ARG(cycles >= 1, "cycles must be >= 1");
    BH_HIP(hipSetDevice(s->ctx->device));
    bh_ctx* c = s->ctx;
    const uint32_t bins = s->wgs, m = s->dim;
    DevBuf counts;
    BH_TRY(counts.alloc((size_t)m * bins * sizeof(unsigned)));
    BH_HIP(hipMemsetAsync(counts.p, 0, (size_t)m * bins * sizeof(unsigned), s->stream));
    DevOut dl, dp, dr;
    BH_TRY(dl.init(limits, (size_t)2 * m * sizeof(float)));
    BH_TRY(dp.init(pdf, (size_t)bins * m * sizeof(float)));
    BH_TRY(dr.init(bin_ranks, (size_t)bins * m * sizeof(float)));
    // limits from the first cycle only (amd_gcn.clj:469-476)
    BH_TRY(histogram_device(c, s->stream, s->xs.as<float>(), m, s->W, m, dl.as<float>(),
                            counts.as<unsigned>(), true));
    s->beta = 1.0f;
    for (uint32_t i = 1; i < cycles; ++i) {
        BH_TRY(bare_pair(s, 1.0f));
        s->bare_counter += 1;
        BH_TRY(histogram_device(c, s->stream, s->xs.as<float>(), m, s->W, m, dl.as<float>(),
                                counts.as<unsigned>(), false));
    }
    s->iter += cycles - 1;
    const float alpha = (float)((double)bins / ((double)cycles * s->W));
    BH_HIP(bh::launch_histogram_finalize(s->stream, counts.as<unsigned>(), dl.as<float>(), alpha,
                                         m, bins, dp.as<float>(), dr.as<float>()));
    BH_TRY(dl.finish(s->stream));
    BH_TRY(dp.finish(s->stream));
    BH_TRY(dr.finish(s->stream));
    return sync(s->stream);
}

static int stretch_moments(bh_stretch* s, float* out, bool variance) {
    BH_ARG(s && out, "NULL argument");
    BH_HIP(hipSetDevice(s->ctx->device));
    DevBuf mv;
    BH_TRY(mv.alloc((size_t)2 * s->dim * sizeof(float)));
    void* scratch = nullptr;
    BH_TRY(ctx_scratch(s->ctx, bh::dataset_scratch_bytes(s->dim, s->W), &scratch));
    float* mean = mv.as<float>();
    float* var = mean + s->dim;
    BH_HIP(bh::launch_moments(s->stream, s->xs.as<float>(), s->dim, s->W, s->dim, mean,
                              variance ? var : nullptr, scratch));
    BH_HIP(hipMemcpyAsync(out, variance ? var : mean, (size_t)s->dim * sizeof(float),
                          hipMemcpyDefault, s->stream));
    return sync(s->stream);
}

int bh_stretch_mean(bh_stretch* s, float* out) { return stretch_moments(s, out, false); }

int bh_stretch_variance(bh_stretch* s, float* out) { return stretch_moments(s, out, true); }

int bh_stretch_state_get(const bh_stretch* s, float* xs, float* logfn, uint32_t counters[6]) {
    BH_ARG(s, "NULL sampler");
    BH_HIP(hipSetDevice(s->ctx->device));
    if (xs)
        BH_HIP(hipMemcpyAsync(xs, s->xs.p, (size_t)s->W * s->dim * sizeof(float), hipMemcpyDefault,
                              s->stream));
    if (logfn && s->f->glm) {
        std::vector<double> h(s->W);
        BH_HIP(hipMemcpyAsync(h.data(), s->lp.p, (size_t)s->W * sizeof(double),
                              hipMemcpyDeviceToHost, s->stream));
        BH_TRY(sync(s->stream));
        std::vector<float> hf(h.begin(), h.end());
        BH_HIP(hipMemcpyAsync(logfn, hf.data(), (size_t)s->W * sizeof(float), hipMemcpyDefault,
                              s->stream));
        BH_TRY(sync(s->stream));
    } else if (logfn) {
        BH_HIP(hipMemcpyAsync(logfn, s->logfn.p, (size_t)s->W * sizeof(float), hipMemcpyDefault,
                              s->stream));
    }
    if (counters) {
        counters[0] = s->bare_seed;
        counters[1] = s->move_seed;
        counters[2] = s->bare_counter;
        counters[3] = s->move_counter;
        counters[4] = (uint32_t)s->iter;
        counters[5] = (uint32_t)(s->iter >> 32);
    }
    return sync(s->stream);
}

int bh_stretch_state_set(bh_stretch* s, const float* xs, const float* logfn,
                         const uint32_t counters[6]) {
    BH_ARG(s, "NULL sampler");
    BH_HIP(hipSetDevice(s->ctx->device));
    if (xs)
        BH_HIP(hipMemcpyAsync(s->xs.p, xs, (size_t)s->W * s->dim * sizeof(float), hipMemcpyDefault,
                              s->stream));
    if (s->f->glm) {
        // log posteriors are kept in double: recompute them from positions
        if (xs) BH_TRY(logfn_all(s));
    } else if (logfn) {
        BH_HIP(hipMemcpyAsync(s->logfn.p, logfn, (size_t)s->W * sizeof(float), hipMemcpyDefault,
                              s->stream));
    } else if (xs) {
        BH_TRY(logfn_all(s));
    }
    if (counters) {
        s->bare_seed = counters[0];
        s->move_seed = counters[1];
        s->bare_counter = counters[2];
        s->move_counter = counters[3];
        s->iter = (uint64_t)counters[4] | ((uint64_t)counters[5] << 32);
    }
    return sync(s->stream);
}

}  // extern "C"
