// Model-independent kernels of the HIP engine, compiled ahead of time for
// gfx950: walker initialisation, the dataset engine (min/max, histogram with
// LDS-privatised bins, pdf + bin ranks, FP64 mean/variance), the run-sampler
// accounting reductions and the autocorrelation estimator.
//
// Matrices are the reference's DIM x n column-major blocks (element (d, j)
// at a[j*ld + d]).  The dataset kernels map a 256-thread block onto a tile of
// TM rows x TN columns (TM*TN = 256, TM the next power of two >= m, capped at
// 64 rows per pass), so every thread owns one row and consecutive threads
// read consecutive addresses of a column.
#include <hip/hip_runtime.h>

#include <mutex>

#include "bh_kernels.hpp"
#include "bh_philox.hpp"

// The working memory of bh_dataset_variance_noalloc (partial sums and row
// means), for callers that cannot allocate.  Each GPU holds its own copy at
// its own address; host code reaches it only through the address that
// noalloc_scratch_init looks up (the array's name is a host-side stand-in).
static __device__ __align__(256) unsigned char scratch[bh::kNoAllocScratchBytes];

namespace bh {

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices];
}  // namespace

hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_noalloc_ptr[device] = p;
    return hipSuccess;
}

hipError_t noalloc_scratch(int device, void** ptr) {
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    *ptr = g_noalloc_ptr[device];
    return *ptr ? hipSuccess : hipErrorNotInitialized;
}

// --------------------------------------------------------------------------
// init_walkers (amd-gcn-mcmc-stretch.cl:138-159)
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
init_walkers_kernel(unsigned seed, const float* __restrict__ limits, unsigned dim,
                    unsigned items, float* __restrict__ xs) {
    const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= items) return;
    const bh_f32x4 u = bh_u01x4(
        bh_philox4x32_10(bh_u32x4{g, BH_CTR1_INIT, BH_CTR2_INIT, BH_CTR3}, seed, BH_KEY1_INIT));
    const float uu[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (unsigned m = 0; m < 4; ++m) {
        const unsigned i = 4 * g + m;
        const unsigned r = i % dim;
        const float lo = limits[2 * r], hi = limits[2 * r + 1];
        xs[i] = __builtin_fmaf(uu[m], hi, (1.0f - uu[m]) * lo);
    }
}

// Uniform direct sampler (uniform-sampler.cl:19-34): x = fma(u, hi-lo, lo).
__global__ void __launch_bounds__(256)
uniform_sample_kernel(unsigned seed, float lower, float upper, unsigned quads,
                      float* __restrict__ x) {
    const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= quads) return;
    const bh_f32x4 u = bh_u01x4(
        bh_philox4x32_10(bh_u32x4{g, BH_CTR1_INIT, BH_CTR2_INIT, BH_CTR3}, seed, BH_KEY1_INIT));
    const float range = upper - lower;
    float4 v;
    v.x = __builtin_fmaf(u.x, range, lower);
    v.y = __builtin_fmaf(u.y, range, lower);
    v.z = __builtin_fmaf(u.z, range, lower);
    v.w = __builtin_fmaf(u.w, range, lower);
    reinterpret_cast<float4*>(x)[g] = v;
}

// Gaussian (Box-Muller), exponential and Erlang direct samplers
// (gaussian-sampler.cl:17-44, exponential-sampler.cl:16-29,
// erlang-sampler.cl:20-40); kinds as BH_DIRECT_* in bayadera_hip.h.
__global__ void __launch_bounds__(256)
direct_sample_kernel(int kind, unsigned seed, float p0, float p1, unsigned quads,
                     float* __restrict__ x) {
    const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= quads) return;
    float r[4];
    if (kind == 3) {
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (unsigned i = 0; (float)i < p1; ++i) {
            const bh_f32x4 u = bh_u01x4(
                bh_philox4x32_10(bh_u32x4{g, BH_CTR1_INIT, BH_CTR2_INIT, i}, seed, BH_KEY1_INIT));
            acc[0] += logf(u.x);
            acc[1] += logf(u.y);
            acc[2] += logf(u.z);
            acc[3] += logf(u.w);
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) r[m] = -(acc[m] / p0);
    } else {
        const bh_f32x4 u = bh_u01x4(
            bh_philox4x32_10(bh_u32x4{g, BH_CTR1_INIT, BH_CTR2_INIT, BH_CTR3}, seed, BH_KEY1_INIT));
        if (kind == 1) {
            const float two_pi = 6.2831855f;
            const float ra = sqrtf(-2.0f * logf(u.y)), rb = sqrtf(-2.0f * logf(u.w));
            r[0] = __builtin_fmaf(sinf(two_pi * u.x) * ra, p1, p0);
            r[1] = __builtin_fmaf(cosf(two_pi * u.x) * ra, p1, p0);
            r[2] = __builtin_fmaf(sinf(two_pi * u.z) * rb, p1, p0);
            r[3] = __builtin_fmaf(cosf(two_pi * u.z) * rb, p1, p0);
        } else if (kind == 2) {
            const float c = -1.0f / p0;
            r[0] = c * logf(1.0f - u.x);
            r[1] = c * logf(1.0f - u.y);
            r[2] = c * logf(1.0f - u.z);
            r[3] = c * logf(1.0f - u.w);
        } else {
            const float range = p1 - p0;
            r[0] = __builtin_fmaf(u.x, range, p0);
            r[1] = __builtin_fmaf(u.y, range, p0);
            r[2] = __builtin_fmaf(u.z, range, p0);
            r[3] = __builtin_fmaf(u.w, range, p0);
        }
    }
    reinterpret_cast<float4*>(x)[g] = make_float4(r[0], r[1], r[2], r[3]);
}

// --------------------------------------------------------------------------
// Dataset engine (amd-gcn-estimate.cl; host amd_gcn.clj:137-214)
// --------------------------------------------------------------------------
__device__ __forceinline__ float ordered_min(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float ordered_max(float a, float b) { return b > a ? b : a; }

// Per-block [min, max] of rows [r0, r0+tm) -> partial[(blk*m + r)*2 + {0,1}].
__global__ void __launch_bounds__(256)
minmax_partial_kernel(const float* __restrict__ a, unsigned m, unsigned long long n,
                      unsigned ld, unsigned r0, unsigned tm, float* __restrict__ partial) {
    const unsigned t = threadIdx.x;
    const unsigned tn = 256 / tm;
    const unsigned rl = t % tm, cl = t / tm;
    const unsigned r = r0 + rl;
    float lo = __builtin_inff(), hi = -__builtin_inff();
    if (r < m) {
        for (unsigned long long j = (unsigned long long)blockIdx.x * tn + cl; j < n;
             j += (unsigned long long)gridDim.x * tn) {
            const float x = a[j * ld + r];
            lo = ordered_min(lo, x);
            hi = ordered_max(hi, x);
        }
    }
    __shared__ float slo[256], shi[256];
    slo[t] = lo;
    shi[t] = hi;
    __syncthreads();
    for (unsigned s = tn / 2; s >= 1; s >>= 1) {
        if (cl < s) {
            slo[t] = ordered_min(slo[t], slo[t + s * tm]);
            shi[t] = ordered_max(shi[t], shi[t + s * tm]);
        }
        __syncthreads();
    }
    if (cl == 0 && r < m) {
        partial[((size_t)blockIdx.x * m + r) * 2] = slo[t];
        partial[((size_t)blockIdx.x * m + r) * 2 + 1] = shi[t];
    }
}

__global__ void __launch_bounds__(256)
minmax_final_kernel(const float* __restrict__ partial, unsigned blocks, unsigned m,
                    unsigned r0, float* __restrict__ limits) {
    const unsigned r = r0 + blockIdx.x;
    float lo = __builtin_inff(), hi = -__builtin_inff();
    for (unsigned b = threadIdx.x; b < blocks; b += blockDim.x) {
        lo = ordered_min(lo, partial[((size_t)b * m + r) * 2]);
        hi = ordered_max(hi, partial[((size_t)b * m + r) * 2 + 1]);
    }
    __shared__ float slo[256], shi[256];
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (unsigned s = blockDim.x / 2; s >= 1; s >>= 1) {
        if (threadIdx.x < s) {
            slo[threadIdx.x] = ordered_min(slo[threadIdx.x], slo[threadIdx.x + s]);
            shi[threadIdx.x] = ordered_max(shi[threadIdx.x], shi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        limits[2 * r] = slo[0];
        limits[2 * r + 1] = shi[0];
    }
}

// bin = (uint)((x-lo)/(hi-lo)*bins) clamped to bins-1 (amd-gcn-estimate.cl:17-18);
// v_cvt_u32_f32 saturates negative values and NaN to 0.
__device__ __forceinline__ unsigned hist_bin(float x, float lo, float hi, unsigned bins) {
    const float f = (x - lo) / (hi - lo) * (float)bins;
    unsigned b = f > 0.0f ? (f >= 4294967296.0f ? 0xffffffffu : (unsigned)f) : 0u;
    return b < bins ? b : bins - 1;
}

// LDS-privatised histogram of rows [r0, r0+tm): counts[r*bins + b] += n_rb.
// Each block keeps tm x bins counters in LDS (ds_add_u32), then adds its
// non-zero counters to global memory once.
__global__ void __launch_bounds__(256)
histogram_kernel(const float* __restrict__ a, unsigned m, unsigned long long n, unsigned ld,
                 unsigned r0, unsigned tm, const float* __restrict__ limits, unsigned bins,
                 unsigned* __restrict__ counts) {
    extern __shared__ unsigned hist[];  // tm * bins
    const unsigned t = threadIdx.x;
    const unsigned tn = 256 / tm;
    const unsigned rl = t % tm, cl = t / tm;
    const unsigned r = r0 + rl;
    for (unsigned i = t; i < tm * bins; i += 256) hist[i] = 0;
    __syncthreads();
    if (r < m) {
        const float lo = limits[2 * r], hi = limits[2 * r + 1];
        unsigned* h = hist + rl * bins;
        for (unsigned long long j = (unsigned long long)blockIdx.x * tn + cl; j < n;
             j += (unsigned long long)gridDim.x * tn) {
            atomicAdd(&h[hist_bin(a[j * ld + r], lo, hi, bins)], 1u);
        }
    }
    __syncthreads();
    for (unsigned i = t; i < tm * bins; i += 256) {
        const unsigned c = hist[i];
        const unsigned rr = r0 + i / bins;
        if (c && rr < m) atomicAdd(&counts[(size_t)rr * bins + (i % bins)], c);
    }
}

// uint_to_real + bitonic_local (amd-gcn-estimate.cl:27-34, 95-123) for one
// row per block of `bins` threads: pdf = alpha/(hi-lo) * count, then the
// reference's bitonic network on (index, pdf) pairs, which orders the bins
// by descending pdf; ties keep the network's positional order.
__global__ void __launch_bounds__(1024)
histogram_finalize_kernel(const unsigned* __restrict__ counts, const float* __restrict__ limits,
                          float alpha, unsigned bins, float* __restrict__ pdf,
                          float* __restrict__ ranks) {
    const unsigned r = blockIdx.x;
    const unsigned lid = threadIdx.x;
    __shared__ float ax[1024], ay[1024];
    const float p = alpha / (limits[2 * r + 1] - limits[2 * r]) * (float)counts[(size_t)r * bins + lid];
    pdf[(size_t)r * bins + lid] = p;
    float vx = (float)lid, vy = p;
    ax[lid] = vx;
    ay[lid] = vy;
    __syncthreads();
    for (unsigned length = 1; length < bins; length <<= 1) {
        const bool direction = (lid & (length << 1)) != 0;
        for (unsigned inc = length; inc > 0; inc >>= 1) {
            const unsigned j = lid ^ inc;
            const float ox = ax[j], oy = ay[j];
            const bool smaller = (vy < oy) || (oy == vy && j < lid);
            const bool swap = smaller ^ (j < lid) ^ direction;
            if (swap) {
                vx = ox;
                vy = oy;
            }
            __syncthreads();
            ax[lid] = vx;
            ay[lid] = vy;
            __syncthreads();
        }
    }
    ranks[(size_t)r * bins + lid] = vx;
}

// Row sums in double: pass 0 sums x, pass 1 sums (x - (float)mu)^2 with the
// difference taken in float as variance_reduce does (amd-gcn-estimate.cl:146-157).
__global__ void __launch_bounds__(256)
moments_partial_kernel(const float* __restrict__ a, unsigned m, unsigned long long n,
                       unsigned ld, unsigned r0, unsigned tm, const float* __restrict__ mu,
                       double* __restrict__ partial) {
    const unsigned t = threadIdx.x;
    const unsigned tn = 256 / tm;
    const unsigned rl = t % tm, cl = t / tm;
    const unsigned r = r0 + rl;
    double s = 0.0;
    if (r < m) {
        const float c = mu ? mu[r] : 0.0f;
        for (unsigned long long j = (unsigned long long)blockIdx.x * tn + cl; j < n;
             j += (unsigned long long)gridDim.x * tn) {
            const float x = a[j * ld + r];
            if (mu) {
                const double d = (double)(x - c);
                s = __builtin_fma(d, d, s);
            } else {
                s += (double)x;
            }
        }
    }
    __shared__ double ss[256];
    ss[t] = s;
    __syncthreads();
    for (unsigned k = tn / 2; k >= 1; k >>= 1) {
        if (cl < k) ss[t] += ss[t + k * tm];
        __syncthreads();
    }
    if (cl == 0 && r < m) partial[(size_t)blockIdx.x * m + r] = ss[t];
}

// out[r] = (float)(sum of partials / n); also keeps the float mean for pass 1.
__global__ void __launch_bounds__(256)
moments_final_kernel(const double* __restrict__ partial, unsigned blocks, unsigned m,
                     unsigned r0, double inv_n, float* __restrict__ out) {
    const unsigned r = r0 + blockIdx.x;
    double s = 0.0;
    for (unsigned b = threadIdx.x; b < blocks; b += blockDim.x) s += partial[(size_t)b * m + r];
    __shared__ double ss[256];
    ss[threadIdx.x] = s;
    __syncthreads();
    for (unsigned k = blockDim.x / 2; k >= 1; k >>= 1) {
        if (threadIdx.x < k) ss[threadIdx.x] += ss[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[r] = (float)(ss[0] * inv_n);
}

// --------------------------------------------------------------------------
// Work-group tree of the reference (amd-gcn-mcmc-stretch.cl:5-29) for a
// group of lsize <= blockDim items, including the odd-size step: while i > 0,
// item i/2-1 first adds item i/2+i/2... (2*(i/2)) when i is odd, then every
// item l < i/2 adds item l+i/2.  lacc has blockDim floats.  Returns the total
// in every thread.
// --------------------------------------------------------------------------
__device__ float group_tree(float v, unsigned lsize, float* lacc) {
    const unsigned l = threadIdx.x;
    __syncthreads();
    if (l < lsize) lacc[l] = v;
    __syncthreads();
    float pacc = v;
    unsigned i = lsize;
    while (i > 0) {
        const bool odd = (i & 1u) && (l == (i >> 1) - 1);
        i >>= 1;
        if (odd) pacc += lacc[l + i + 1];
        if (l < i) {
            pacc += lacc[l + i];
            lacc[l] = pacc;
        }
        __syncthreads();
    }
    const float r = lacc[0];
    __syncthreads();
    return r;
}

// clojurecl's enq-reduce! along the columns of one row (sum_means_vertical +
// sum_reduction_horizontal, amd_gcn.clj:437-438): groups of local_n
// consecutive values reduce with the tree, the partials reduce again until
// one is left.  One block per row; work holds a row-sized scratch area.
__global__ void __launch_bounds__(kMaxGroupThreads)
row_tree_sum_kernel(const float* __restrict__ a, unsigned count, unsigned local_n,
                    unsigned long long row_stride, float* __restrict__ work, float scale,
                    float* __restrict__ out, unsigned long long out_stride) {
    __shared__ float lacc[kMaxGroupThreads];
    const unsigned row = blockIdx.x;
    const float* src = a + row * row_stride;
    float* w = work + (size_t)row * ((count + local_n - 1) / local_n + 1);
    unsigned c = count;
    bool first = true;
    while (true) {
        const unsigned groups = (c + local_n - 1) / local_n;
        for (unsigned g = 0; g < groups; ++g) {
            const unsigned start = g * local_n;
            const unsigned len = (c - start < local_n) ? c - start : local_n;
            const float v = threadIdx.x < len ? (first ? src[start + threadIdx.x] : w[start + threadIdx.x]) : 0.0f;
            const float s = group_tree(v, len, lacc);
            // all reads of this pass read indices >= start > g unless g = 0;
            // group 0 reads before writing slot 0
            if (threadIdx.x == 0) w[g] = s;
            __syncthreads();
        }
        first = false;
        if (groups == 1) break;
        c = groups;
    }
    if (threadIdx.x == 0) out[row * out_stride] = w[0] * scale;
}

__global__ void __launch_bounds__(256)
sum_u32_kernel(const unsigned* __restrict__ v, unsigned n, unsigned long long* __restrict__ out) {
    unsigned long long s = 0;
    for (unsigned i = threadIdx.x; i < n; i += blockDim.x) s += v[i];
    __shared__ unsigned long long ss[256];
    ss[threadIdx.x] = s;
    __syncthreads();
    for (unsigned k = 128; k >= 1; k >>= 1) {
        if (threadIdx.x < k) ss[threadIdx.x] += ss[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = ss[0];
}

__global__ void __launch_bounds__(256)
sum_f64_kernel(const double* __restrict__ v, unsigned n, double* __restrict__ out) {
    double s = 0.0;
    for (unsigned i = threadIdx.x; i < n; i += blockDim.x) s += v[i];
    __shared__ double ss[256];
    ss[threadIdx.x] = s;
    __syncthreads();
    for (unsigned k = 128; k >= 1; k >>= 1) {
        if (threadIdx.x < k) ss[threadIdx.x] += ss[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = ss[0];
}

// --------------------------------------------------------------------------
// Autocorrelation (GCNAcorEngine.acor, amd_gcn.clj:216-267; amd-gcn-acor.cl)
// One block of wgs threads per dimension.  The reference's device-side
// enqueue recursion (acor.cl:143-168) becomes a loop; each of its
// work-groups becomes a pass of this block over a slice of the series with
// the same loads, lag sums and trees, so the float results are the same.
// series: n x dim (series[t*dim + d]), a device copy the kernel centres and
// pair-sums in place.  scratch: >= 2*(ceil(n/wgs)+1) floats per dimension.
// --------------------------------------------------------------------------
// One column: element t at col[t*cs]; its results go to tau[0], mean[0],
// sigma[0]; c0acc: scratch_per_dim floats of scratch.
__device__ __forceinline__ void acor_column(float* __restrict__ col, unsigned cs, unsigned n,
                                            unsigned lag, float inv_n, unsigned local_n_mean,
                                            float* __restrict__ c0acc, unsigned scratch_per_dim,
                                            float* __restrict__ tau, float* __restrict__ mean,
                                            float* __restrict__ sigma) {
    const unsigned wgs = blockDim.x;
    const unsigned l = threadIdx.x;
    const unsigned win_mult = 5, min_lag = 10;
    __shared__ float lacc[kMaxGroupThreads];
    __shared__ float lm[2 * kMaxGroupThreads];
    float* dacc = c0acc + scratch_per_dim / 2;

    // mean: enq-reduce! with local (min(16,dim), wgs/that) along n, scal! 1/n
    {
        unsigned c = n;
        bool first = true;
        while (true) {
            const unsigned groups = (c + local_n_mean - 1) / local_n_mean;
            for (unsigned g = 0; g < groups; ++g) {
                const unsigned start = g * local_n_mean;
                const unsigned len = (c - start < local_n_mean) ? c - start : local_n_mean;
                float v = 0.0f;
                if (l < len) v = first ? col[(size_t)(start + l) * cs] : c0acc[start + l];
                const float s = group_tree(v, len, lacc);
                if (l == 0) c0acc[g] = s;
                __syncthreads();
            }
            first = false;
            if (groups == 1) break;
            c = groups;
        }
    }
    const float mu = c0acc[0] * inv_n;
    __syncthreads();
    for (unsigned t = l; t < n; t += wgs) col[(size_t)t * cs] -= mu;
    __syncthreads();

    // one acor_1d level (acor.cl:26-86) over gsize points at stride
    auto level = [&](unsigned stride, unsigned lg, unsigned gsize, unsigned lw) -> unsigned {
        const unsigned groups = (gsize + lw - 1) / lw;
        for (unsigned g = 0; g < groups; ++g) {
            const unsigned start = g * lw;
            const unsigned lsize = (gsize - start < lw) ? gsize - start : lw;
            const bool has_next = g + 1 < groups;
            __syncthreads();
            lm[l] = 0.0f;
            lm[l + wgs] = 0.0f;
            __syncthreads();
            if (l < lsize) {
                const unsigned gid = start + l;
                const bool compute = gid + lg < gsize;
                const bool load_lag = (l < lg) && has_next && (gid + lsize < gsize);
                lm[l] = compute ? col[(size_t)gid * stride] : 0.0f;
                lm[l + lsize] = load_lag ? col[(size_t)(gid + lsize) * stride] : 0.0f;
            }
            __syncthreads();
            float xv = 0.0f, xacc = 0.0f;
            if (l < lsize) {
                xv = lm[l];
                for (unsigned s = 0; s < lg; ++s) {
                    const unsigned idx = l + s + 1;
                    xacc += idx < 2 * lsize ? lm[idx] : 0.0f;
                }
                xacc = xv * __builtin_fmaf(2.0f, xacc, xv);
            }
            const float c0 = group_tree(xv * xv, lsize, lacc);
            const float dd = group_tree(xacc, lsize, lacc);
            if (l == 0) {
                c0acc[g] = c0;
                dacc[g] = dd;
            }
        }
        __syncthreads();
        return groups;
    };

    // sum_reduction_acor's reduction (acor.cl:88-130)
    auto reduce = [&](unsigned groups, float& c0, float& dd) {
        const unsigned lsize = groups < wgs ? groups : wgs;
        float a = 0.0f, b = 0.0f;
        if (l < lsize) {
            for (unsigned t = l; t < groups; t += lsize) {
                a += c0acc[t];
                b += dacc[t];
            }
        }
        c0 = group_tree(a, lsize, lacc);
        dd = group_tree(b, lsize, lacc);
    };

    unsigned groups = level(cs, lag, n, wgs);
    float c0, dd;
    reduce(groups, c0, dd);
    unsigned stride = 1, cur_lag = lag, prev_n = n;
    float orig_c0 = 0.0f, prev_c0 = 0.0f;
    while (true) {
        const float start_c0 = (stride == 1) ? c0 : orig_c0;
        const float tv = dd / ((stride == 1) ? c0 : prev_c0);
        const unsigned lag2 =
            (cur_lag * win_mult < prev_n) ? cur_lag : ((prev_n / win_mult) > 10u ? prev_n / win_mult : 10u);
        if (min_lag < lag2 && (float)lag2 < tv * (float)win_mult) {
            const unsigned n2 = prev_n / 2;
            const unsigned w2 = wgs < n2 ? wgs : n2;
            const size_t step = (size_t)stride * cs;
            __syncthreads();
            for (unsigned g = l; g < n2; g += wgs) col[2 * g * step] += col[(2 * (size_t)g + 1) * step];
            __syncthreads();
            groups = level((unsigned)(2 * step), lag2, n2, w2);
            float c0n, dn;
            reduce(groups, c0n, dn);
            orig_c0 = start_c0;
            prev_c0 = c0;
            c0 = c0n;
            dd = dn;
            stride *= 2;
            cur_lag = lag2;
            prev_n = n2;
        } else {
            if (l == 0) {
                const float scale = (float)stride * (float)(prev_n - lag2);
                tau[0] = dd * (float)(n - cur_lag) / (scale * start_c0);
                sigma[0] = sqrtf(dd / (scale * (float)n));
                mean[0] = mu;
            }
            break;
        }
    }
}

__global__ void __launch_bounds__(kMaxGroupThreads)
acor_kernel(float* __restrict__ series, unsigned dim, unsigned n, unsigned lag, float inv_n,
            unsigned local_n_mean, float* __restrict__ scratch, unsigned scratch_per_dim,
            float* __restrict__ tau, float* __restrict__ mean, float* __restrict__ sigma) {
    const unsigned d = blockIdx.x;
    acor_column(series + d, dim, n, lag, inv_n, local_n_mean,
                scratch + (size_t)d * scratch_per_dim, scratch_per_dim, tau + d, mean + d,
                sigma + d);
}

// Many series of n steps in one launch, one block per column: block b takes
// the column starting at series + col_off[b] with stride col_stride[b] (its
// series' dimension, which also sets the mean's reduction shape as in
// acor_kernel), results at index b.
__global__ void __launch_bounds__(kMaxGroupThreads)
acor_batch_kernel(float* __restrict__ series, const unsigned long long* __restrict__ col_off,
                  const unsigned* __restrict__ col_stride, unsigned n, unsigned lag, float inv_n,
                  float* __restrict__ scratch, unsigned scratch_per_dim, float* __restrict__ tau,
                  float* __restrict__ mean, float* __restrict__ sigma) {
    const unsigned b = blockIdx.x;
    const unsigned cs = col_stride[b];
    const unsigned wgs = blockDim.x;
    unsigned wgsm = cs < 16 ? cs : 16;
    if (wgs < wgsm) wgsm = wgs;
    acor_column(series + col_off[b], cs, n, lag, inv_n, wgs / wgsm,
                scratch + (size_t)b * scratch_per_dim, scratch_per_dim, tau + b, mean + b,
                sigma + b);
}

// --------------------------------------------------------------------------
// Streaming copy (the measured HBM denominator): each thread copies four
// 16-byte pieces, 4 KB apart within its block's 16 KB chunk, the four loads
// in flight before the stores; NT: non-temporal loads and stores.
// --------------------------------------------------------------------------
typedef float bh_copy_f4 __attribute__((ext_vector_type(4)));

template <bool NT>
__global__ void __launch_bounds__(256)
copy16_kernel(const bh_copy_f4* __restrict__ src, bh_copy_f4* __restrict__ dst,
              unsigned long long n16) {
    const unsigned long long base = (unsigned long long)blockIdx.x * 1024u + threadIdx.x;
    bh_copy_f4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const unsigned long long i = base + 256u * u;
        if (i < n16) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const unsigned long long i = base + 256u * u;
        if (i < n16) {
            if (NT) __builtin_nontemporal_store(v[u], dst + i);
            else dst[i] = v[u];
        }
    }
}

// --------------------------------------------------------------------------
// Host-side launchers (all on the caller's stream)
// --------------------------------------------------------------------------
static unsigned tile_rows(unsigned m) {
    unsigned tm = 1;
    while (tm < m && tm < 64) tm <<= 1;
    return tm;
}

static unsigned blocks_for(unsigned long long n, unsigned tn, unsigned cap) {
    unsigned long long b = (n + tn - 1) / tn;
    if (b > cap) b = cap;
    return b ? (unsigned)b : 1u;
}

hipError_t launch_copy16(hipStream_t s, const void* src, void* dst, unsigned long long n16,
                         bool nt) {
    const unsigned long long blocks = (n16 + 1023) / 1024;
    if (blocks == 0) return hipSuccess;
    if (nt)
        copy16_kernel<true><<<(unsigned)blocks, 256, 0, s>>>(
            static_cast<const bh_copy_f4*>(src), static_cast<bh_copy_f4*>(dst), n16);
    else
        copy16_kernel<false><<<(unsigned)blocks, 256, 0, s>>>(
            static_cast<const bh_copy_f4*>(src), static_cast<bh_copy_f4*>(dst), n16);
    return hipGetLastError();
}

hipError_t launch_init_walkers(hipStream_t s, unsigned seed, const float* limits, unsigned dim,
                               unsigned walkers, float* xs) {
    const unsigned items = dim * walkers / 4;
    init_walkers_kernel<<<(items + 255) / 256, 256, 0, s>>>(seed, limits, dim, items, xs);
    return hipGetLastError();
}

hipError_t launch_uniform_sample(hipStream_t s, unsigned seed, float lo, float hi, unsigned quads,
                                 float* x) {
    uniform_sample_kernel<<<(quads + 255) / 256, 256, 0, s>>>(seed, lo, hi, quads, x);
    return hipGetLastError();
}

hipError_t launch_direct_sample(hipStream_t s, int kind, unsigned seed, float p0, float p1,
                                unsigned quads, float* x) {
    direct_sample_kernel<<<(quads + 255) / 256, 256, 0, s>>>(kind, seed, p0, p1, quads, x);
    return hipGetLastError();
}

size_t dataset_scratch_bytes(unsigned m, unsigned long long n) {
    (void)n;
    return (size_t)kMaxDatasetBlocks * m * 2 * sizeof(double) + 64;
}

hipError_t launch_min_max(hipStream_t s, const float* a, unsigned m, unsigned long long n,
                          unsigned ld, float* limits, void* scratch) {
    float* partial = (float*)scratch;
    for (unsigned r0 = 0; r0 < m; r0 += 64) {
        const unsigned rows = m - r0 < 64 ? m - r0 : 64;
        const unsigned tm = tile_rows(rows);
        const unsigned blocks = blocks_for(n, 256 / tm, kMaxDatasetBlocks);
        minmax_partial_kernel<<<blocks, 256, 0, s>>>(a, m, n, ld, r0, tm, partial);
        minmax_final_kernel<<<rows, 256, 0, s>>>(partial, blocks, m, r0, limits);
    }
    return hipGetLastError();
}

hipError_t launch_histogram(hipStream_t s, const float* a, unsigned m, unsigned long long n,
                            unsigned ld, const float* limits, unsigned bins, unsigned* counts) {
    // rows per pass so that rows x bins counters take at most 32 KB of LDS
    unsigned per_pass = 8192 / bins;
    per_pass = per_pass < 1 ? 1 : (per_pass > 64 ? 64 : per_pass);
    for (unsigned r0 = 0; r0 < m; r0 += per_pass) {
        const unsigned rows = m - r0 < per_pass ? m - r0 : per_pass;
        const unsigned tm = tile_rows(rows);
        const unsigned blocks = blocks_for(n, 256 / tm, kMaxDatasetBlocks);
        const size_t lds = (size_t)tm * bins * sizeof(unsigned);
        histogram_kernel<<<blocks, 256, lds, s>>>(a, m, n, ld, r0, tm, limits, bins, counts);
    }
    return hipGetLastError();
}

hipError_t launch_histogram_finalize(hipStream_t s, const unsigned* counts, const float* limits,
                                     float alpha, unsigned m, unsigned bins, float* pdf,
                                     float* ranks) {
    histogram_finalize_kernel<<<m, bins, 0, s>>>(counts, limits, alpha, bins, pdf, ranks);
    return hipGetLastError();
}

hipError_t launch_moments(hipStream_t s, const float* a, unsigned m, unsigned long long n,
                          unsigned ld, float* mean, float* var, void* scratch) {
    double* partial = (double*)scratch;
    const double inv_n = 1.0 / (double)n;
    for (int pass = 0; pass < (var ? 2 : 1); ++pass) {
        for (unsigned r0 = 0; r0 < m; r0 += 64) {
            const unsigned rows = m - r0 < 64 ? m - r0 : 64;
            const unsigned tm = tile_rows(rows);
            const unsigned blocks = blocks_for(n, 256 / tm, kMaxDatasetBlocks);
            moments_partial_kernel<<<blocks, 256, 0, s>>>(a, m, n, ld, r0, tm,
                                                          pass ? mean : nullptr, partial);
            moments_final_kernel<<<rows, 256, 0, s>>>(partial, blocks, m, r0, inv_n,
                                                      pass ? var : mean);
        }
    }
    return hipGetLastError();
}

hipError_t launch_row_tree_sum(hipStream_t s, const float* a, unsigned rows, unsigned count,
                               unsigned local_n, unsigned long long row_stride, float* work,
                               float scale, float* out, unsigned long long out_stride) {
    // the tree of a group of local_n values needs local_n threads (local_n <= WGS)
    const unsigned threads = local_n > 256 ? local_n : 256;
    row_tree_sum_kernel<<<rows, threads, 0, s>>>(a, count, local_n, row_stride, work, scale, out,
                                                 out_stride);
    return hipGetLastError();
}

hipError_t launch_sum_u32(hipStream_t s, const unsigned* v, unsigned n, unsigned long long* out) {
    sum_u32_kernel<<<1, 256, 0, s>>>(v, n, out);
    return hipGetLastError();
}

hipError_t launch_sum_f64(hipStream_t s, const double* v, unsigned n, double* out) {
    sum_f64_kernel<<<1, 256, 0, s>>>(v, n, out);
    return hipGetLastError();
}

hipError_t launch_acor_batch(hipStream_t s, float* series, const unsigned long long* col_off,
                             const unsigned* col_stride, unsigned cols, unsigned n, unsigned lag,
                             unsigned wgs, float* scratch, unsigned scratch_per_dim, float* tau,
                             float* mean, float* sigma) {
    const float inv_n = (float)(1.0 / (double)n);
    acor_batch_kernel<<<cols, wgs, 0, s>>>(series, col_off, col_stride, n, lag, inv_n, scratch,
                                           scratch_per_dim, tau, mean, sigma);
    return hipGetLastError();
}

hipError_t launch_acor(hipStream_t s, float* series, unsigned dim, unsigned n, unsigned lag,
                       unsigned wgs, float* scratch, unsigned scratch_per_dim, float* tau,
                       float* mean, float* sigma) {
    unsigned wgsm = dim < 16 ? dim : 16;
    if (wgs < wgsm) wgsm = wgs;
    const unsigned local_n_mean = wgs / wgsm;
    const float inv_n = (float)(1.0 / (double)n);
    acor_kernel<<<dim, wgs, 0, s>>>(series, dim, n, lag, inv_n, local_n_mean, scratch,
                                    scratch_per_dim, tau, mean, sigma);
    return hipGetLastError();
}

}  // namespace bh
