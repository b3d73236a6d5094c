// Logistic-regression log-likelihood of a batch of n parameter vectors,
//   ll(b) = y.(X b) - sum_i softplus(x_i . b),
// the dense part of which is the contraction U = X.B (N x n) over the d
// features.  The fp64 kernel produces, for each row range r and point k,
//   partial[r][k] = sum_{i in r} softplus(x_i . b_k)        (natural units).
// The fp32 and MFMA kernels contract against log2(e)*b, so u = log2(e) x.b,
// and use softplus2(u) = max(u,0) + log2(1 + 2^-|u|) with max(u,0) =
// (u + |u|)/2; the u/2 part sums to (log2 e / 2) (sum_i x_i).b, which is
// linear in b and added per point, so these kernels produce
//   partial[r][k] = sum_{i in r} |u_i|/2 + log2(1 + 2^-|u_i|)   (log2 units).
// The linear terms y.(X b) = (X'y).b and -(1/2)(X'1).b and the prior are
// added per point by the walker-side kernels.
//
// X is bf16, row-major N x D (the data as stored).  U is never materialised:
// each tile is reduced through the softplus epilogue as it comes out of the
// matrix cores.
#include <hip/hip_runtime.h>

#include "bh_glm.hpp"
#include "bh_glm_core.hpp"

namespace bh {

// Block index -> (walker group, row range), XCD-aware: blocks are dealt to
// the 8 XCDs round-robin, so the consecutive blocks one XCD receives take all
// walker groups of one row range in turn and read the same rows of X through
// that XCD's L2.
__device__ __forceinline__ void glm_tile(unsigned groups, unsigned& group, unsigned& range) {
    const unsigned b = blockIdx.x;
    const unsigned idx = b >> 3;
    group = idx % groups;
    range = (idx / groups) * 8 + (b & 7);
}

// A fragments from 16-bit X (bf16 or fp16, row-major N x D): lane row lr,
// features kc*32 + lg*8 .. +7.
template <int D, class T, class V>
struct Rows16 {
    using vec = V;
    const T* X;
    unsigned lr, lg;
    __device__ void full(unsigned rb, V (&a)[D / 32]) const {
        const T* p = X + (size_t)(rb + lr) * D + lg * 8;
#pragma unroll
        for (int kc = 0; kc < D / 32; ++kc) a[kc] = *reinterpret_cast<const V*>(p + kc * 32);
    }
    __device__ void masked(unsigned rb, unsigned r1, V (&a)[D / 32]) const {
        const unsigned row = rb + lr;
#pragma unroll
        for (int kc = 0; kc < D / 32; ++kc) {
            if (row < r1) {
                a[kc] = *reinterpret_cast<const V*>(X + (size_t)row * D + kc * 32 + lg * 8);
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) a[kc][j] = (T)0.0f;
            }
        }
    }
};

// A fragments from the fp16 copy of X, stored in fragment order: for each
// 16-row block b and 32-feature chunk kc, the 64 lanes' eight values (lane
// l: row 16 b + l % 16, features 32 kc + 8 (l / 16) .. +7) one after the
// other, rows past N zero.  One fragment load of a wave reads 1 KB of
// consecutive bytes (the row-major layout reads half of each of 16 lines).
template <int D, class T, class V>
struct Frag16 {
    using vec = V;
    const T* X;
    unsigned lr, lg;
    __device__ const T* at(unsigned rb, int kc) const {
        return X + ((size_t)(rb / 16) * (D / 32) + kc) * 512 + (lg * 16 + lr) * 8;
    }
    __device__ void full(unsigned rb, V (&a)[D / 32]) const {
#pragma unroll
        for (int kc = 0; kc < D / 32; ++kc) a[kc] = *reinterpret_cast<const V*>(at(rb, kc));
    }
    __device__ void masked(unsigned rb, unsigned r1, V (&a)[D / 32]) const {
        const bool live = rb + lr < r1;
#pragma unroll
        for (int kc = 0; kc < D / 32; ++kc) {
            a[kc] = *reinterpret_cast<const V*>(at(rb, kc));
            if (!live) {
#pragma unroll
                for (int j = 0; j < 8; ++j) a[kc][j] = (T)0.0f;
            }
        }
    }
};

// MFMA kernel: v_mfma_f32_16x16x32_bf16 (T = __bf16) or _f16 (T =
// _Float16).  A block of 4 waves owns 256 points, 64 per wave as four
// 16-point column blocks whose B fragments (all D features, NS split terms)
// stay in registers for the whole row range.  Lane layout (16x16x32, checked
// on the device): A lane l holds row l%16, features (l/16)*8..+7 of a
// 32-feature chunk; B lane l holds point l%16, the same features; the
// accumulator lane l holds rows (l/16)*4..+3 of point l%16.  The row loop
// (glm_mfma_rows) is software-pipelined: the matrix-core products of row
// block i+1 are issued with the softplus epilogue of block i placed between
// them.  Split terms are accumulated smallest first.  Rows are masked only in
// the last block.  The fp16 kernels read X in fragment order (Frag16), the
// bf16 ones the caller's row-major bf16 X (Rows16).
template <int D, int NS, class T, class V>
__global__ void __launch_bounds__(256)
glm_logistic_mfma_kernel(const T* __restrict__ X, unsigned N, const T* __restrict__ ys, unsigned n,
                         unsigned groups, unsigned nranges, unsigned rows_per_range,
                         double* __restrict__ partial) {
    constexpr int KC = D / 32;
    const unsigned lane = threadIdx.x & 63u;
    const unsigned wave = threadIdx.x >> 6;
    unsigned group, range;
    glm_tile(groups, group, range);
    if (range >= nranges) return;
    const unsigned w0 = group * 256u + wave * 64u;
    const unsigned lr = lane & 15u, lg = lane >> 4;

    V bfrag[4][KC][NS];
#pragma unroll
    for (int wb = 0; wb < 4; ++wb)
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
#pragma unroll
            for (int s = 0; s < NS; ++s)
                bfrag[wb][kc][s] = *reinterpret_cast<const V*>(
                    ys + ((size_t)s * n + w0 + wb * 16 + lr) * D + kc * 32 + lg * 8);

    const unsigned r0 = range * rows_per_range;
    const unsigned r1 = (r0 >= N) ? r0 : ((N - r0 < rows_per_range) ? N : r0 + rows_per_range);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (r0 < r1) {
        if constexpr (__is_same(T, _Float16))
            glm_mfma_rows<D, NS, 4>(Frag16<D, T, V>{X, lr, lg}, bfrag, r0, r1, lg, acc);
        else
            glm_mfma_rows<D, NS, 4>(Rows16<D, T, V>{X, lr, lg}, bfrag, r0, r1, lg, acc);
    }
#pragma unroll
    for (int wb = 0; wb < 4; ++wb) {
        double v = acc[wb];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (lg == 0) partial[(size_t)range * n + w0 + wb * 16 + lr] = v;
    }
}

// fp16 copy of bf16 X (N x D row-major) for the fp16 matrix-core kernels,
// in Frag16's fragment order, eight values (one 16-byte piece of a row) per
// thread; counts the values outside the fp16 range (|x| > 65504).  The
// padding rows of the last block are left to the caller (zeroed).
template <int D>
__global__ void __launch_bounds__(256)
glm_x_to_f16_kernel(const __bf16* __restrict__ X, unsigned long long n8,
                    _Float16* __restrict__ out, unsigned* __restrict__ overflow) {
    const unsigned long long i = (unsigned long long)blockIdx.x * 256u + threadIdx.x;
    if (i >= n8) return;
    const unsigned long long r = i / (D / 8);
    const unsigned c = (unsigned)(i % (D / 8));
    const unsigned long long o = ((r / 16) * (D / 32) + c / 4) * 64 + (c % 4) * 16 + r % 16;
    const bf16x8 v = reinterpret_cast<const bf16x8*>(X)[i];
    f16x8 h;
    unsigned bad = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float f = (float)v[j];
        bad += __builtin_fabsf(f) > 65504.0f ? 1u : 0u;
        h[j] = (_Float16)f;
    }
    reinterpret_cast<f16x8*>(out)[o] = h;
    if (bad) atomicAdd(overflow, bad);
}

// Plain fp32 kernel (no matrix cores): one point per thread, its D
// coefficients (times log2 e) in registers; tiles of 32 rows of X are
// converted to fp32 in LDS and read back as broadcasts.
template <int D>
__global__ void __launch_bounds__(256)
glm_logistic_fp32_kernel(const __bf16* __restrict__ X, unsigned long long N,
                         const float* __restrict__ P, unsigned n, unsigned groups,
                         unsigned nranges, unsigned long long rows_per_range, float yscale,
                         double* __restrict__ partial) {
    constexpr int TR = 32;
    __shared__ float4 xs[TR * D / 4];
    unsigned group, range;
    glm_tile(groups, group, range);
    if (range >= nranges) return;
    const unsigned k = group * 256u + threadIdx.x;
    float y[D];
#pragma unroll
    for (int j = 0; j < D; ++j) y[j] = P[(size_t)k * D + j] * yscale;
    const unsigned long long r0 = (unsigned long long)range * rows_per_range;
    unsigned long long r1 = r0 + rows_per_range;
    if (r1 > N) r1 = N;
    double acc = 0.0;
    for (unsigned long long rb = r0; rb < r1; rb += TR) {
        __syncthreads();
        for (unsigned e = threadIdx.x; e < TR * D; e += 256) {
            const unsigned long long row = rb + e / D;
            reinterpret_cast<float*>(xs)[e] = row < r1 ? (float)X[row * D + e % D] : 0.0f;
        }
        __syncthreads();
        // rows past r1 are zero in the tile: each adds exactly log2(2) = 1,
        // taken out below
        const int pad = (int)((r1 - rb) < TR ? TR - (r1 - rb) : 0);
        for (int r = 0; r < TR; r += 4) {
            float u[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float t = 0.0f;
#pragma unroll
                for (int j4 = 0; j4 < D / 4; ++j4) {
                    const float4 x = xs[(r + i) * (D / 4) + j4];
                    t = __builtin_fmaf(x.x, y[4 * j4 + 0], t);
                    t = __builtin_fmaf(x.y, y[4 * j4 + 1], t);
                    t = __builtin_fmaf(x.z, y[4 * j4 + 2], t);
                    t = __builtin_fmaf(x.w, y[4 * j4 + 3], t);
                }
                u[i] = t;
            }
            acc += (double)softplus2_rest4(u[0], u[1], u[2], u[3]);
        }
        acc -= (double)pad;
    }
    partial[(size_t)range * n + k] = acc;
}

// fp64 reference kernel: the products of bf16 data and fp32 coefficients
// are exact in double, the sum and softplus(eta) = max(eta,0) +
// log1p(exp(-|eta|)) run in double.  Natural-log units.
template <int D>
__global__ void __launch_bounds__(256)
glm_logistic_fp64_kernel(const __bf16* __restrict__ X, unsigned long long N,
                         const float* __restrict__ P, unsigned n, unsigned groups,
                         unsigned nranges, unsigned long long rows_per_range,
                         double* __restrict__ partial) {
    constexpr int TR = 16;
    __shared__ double xs[TR * D];
    unsigned group, range;
    glm_tile(groups, group, range);
    if (range >= nranges) return;
    const unsigned k = group * 256u + threadIdx.x;
    double y[D];
#pragma unroll
    for (int j = 0; j < D; ++j) y[j] = (double)P[(size_t)k * D + j];
    const unsigned long long r0 = (unsigned long long)range * rows_per_range;
    unsigned long long r1 = r0 + rows_per_range;
    if (r1 > N) r1 = N;
    double acc = 0.0;
    for (unsigned long long rb = r0; rb < r1; rb += TR) {
        __syncthreads();
        for (unsigned e = threadIdx.x; e < TR * D; e += 256) {
            const unsigned long long row = rb + e / D;
            xs[e] = row < r1 ? (double)(float)X[row * D + e % D] : 0.0;
        }
        __syncthreads();
        const int rows = (int)((r1 - rb) < TR ? (r1 - rb) : TR);
        for (int r = 0; r < rows; ++r) {
            double eta = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) eta = __builtin_fma(xs[r * D + j], y[j], eta);
            acc += __builtin_fmax(eta, 0.0) + log1p(exp(-__builtin_fabs(eta)));
        }
    }
    partial[(size_t)range * n + k] = acc;
}

// Linear regression (Gaussian likelihood with unknown sigma): for each row
// range r and point k (coefficients b_k = the first d of its DIM = d + 1
// coordinates, row stride pstride),
//   partial[r][k] = sum_{i in r} (y_i - x_i . b_k)^2,
// X fp32 N x D (columns past d zero), y fp32.  The plain fp32 kernel: PPT
// points per thread (their coefficients in registers), tiles of 32 rows of
// X and y in LDS read as broadcasts, each read feeding every point's fma;
// fp32 residuals, squares summed four rows at a time into a double.
template <int D, int PPT>
__global__ void __launch_bounds__(256)
glm_gauss_fp32_kernel(const float* __restrict__ X, const float* __restrict__ yv,
                      unsigned long long N, const float* __restrict__ P, unsigned pstride,
                      unsigned d, unsigned n, unsigned groups, unsigned nranges,
                      unsigned long long rows_per_range, double* __restrict__ partial) {
    constexpr int TR = 32;
    __shared__ float4 xs[TR * D / 4];
    __shared__ float ys[TR];
    unsigned group, range;
    glm_tile(groups, group, range);
    if (range >= nranges) return;
    // PPT points per thread, 256 apart: each broadcast read of a row feeds
    // every point's fma
    const unsigned k = group * (256u * PPT) + threadIdx.x;
    float b[PPT][D];
#pragma unroll
    for (int p = 0; p < PPT; ++p)
#pragma unroll
        for (int j = 0; j < D; ++j)
            b[p][j] = (unsigned)j < d ? P[(size_t)(k + 256u * p) * pstride + j] : 0.0f;
    const unsigned long long r0 = (unsigned long long)range * rows_per_range;
    unsigned long long r1 = r0 + rows_per_range;
    if (r1 > N) r1 = N;
    double acc[PPT];
#pragma unroll
    for (int p = 0; p < PPT; ++p) acc[p] = 0.0;
    for (unsigned long long rb = r0; rb < r1; rb += TR) {
        __syncthreads();
        for (unsigned e = threadIdx.x; e < TR * D / 4; e += 256) {
            const unsigned long long row = rb + (4 * e) / D;
            xs[e] = row < r1 ? *reinterpret_cast<const float4*>(X + rb * D + 4 * e)
                             : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        if (threadIdx.x < TR) ys[threadIdx.x] = rb + threadIdx.x < r1 ? yv[rb + threadIdx.x] : 0.0f;
        __syncthreads();
        // rows past r1 are zero with y = 0: residual 0
        for (int r = 0; r < TR; r += 4) {
            float s4[PPT];
#pragma unroll
            for (int p = 0; p < PPT; ++p) s4[p] = 0.0f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float t[PPT];
#pragma unroll
                for (int p = 0; p < PPT; ++p) t[p] = 0.0f;
#pragma unroll
                for (int j4 = 0; j4 < D / 4; ++j4) {
                    const float4 x = xs[(r + i) * (D / 4) + j4];
#pragma unroll
                    for (int p = 0; p < PPT; ++p) {
                        t[p] = __builtin_fmaf(x.x, b[p][4 * j4 + 0], t[p]);
                        t[p] = __builtin_fmaf(x.y, b[p][4 * j4 + 1], t[p]);
                        t[p] = __builtin_fmaf(x.z, b[p][4 * j4 + 2], t[p]);
                        t[p] = __builtin_fmaf(x.w, b[p][4 * j4 + 3], t[p]);
                    }
                }
#pragma unroll
                for (int p = 0; p < PPT; ++p) {
                    const float res = ys[r + i] - t[p];
                    s4[p] = __builtin_fmaf(res, res, s4[p]);
                }
            }
#pragma unroll
            for (int p = 0; p < PPT; ++p) acc[p] += (double)s4[p];
        }
    }
#pragma unroll
    for (int p = 0; p < PPT; ++p) partial[(size_t)range * n + k + 256u * p] = acc[p];
}

// fp64 reference of the same sums: exact products of fp32 data and
// coefficients, double residuals and squares.
template <int D>
__global__ void __launch_bounds__(256)
glm_gauss_fp64_kernel(const float* __restrict__ X, const float* __restrict__ yv,
                      unsigned long long N, const float* __restrict__ P, unsigned pstride,
                      unsigned d, unsigned n, unsigned groups, unsigned nranges,
                      unsigned long long rows_per_range, double* __restrict__ partial) {
    constexpr int TR = 16;
    __shared__ double xs[TR * D];
    __shared__ double ys[TR];
    unsigned group, range;
    glm_tile(groups, group, range);
    if (range >= nranges) return;
    const unsigned k = group * 256u + threadIdx.x;
    double b[D];
#pragma unroll
    for (int j = 0; j < D; ++j) b[j] = (unsigned)j < d ? (double)P[(size_t)k * pstride + j] : 0.0;
    const unsigned long long r0 = (unsigned long long)range * rows_per_range;
    unsigned long long r1 = r0 + rows_per_range;
    if (r1 > N) r1 = N;
    double acc = 0.0;
    for (unsigned long long rb = r0; rb < r1; rb += TR) {
        __syncthreads();
        for (unsigned e = threadIdx.x; e < TR * D; e += 256) {
            const unsigned long long row = rb + e / D;
            xs[e] = row < r1 ? (double)X[rb * D + e] : 0.0;
        }
        if (threadIdx.x < TR) ys[threadIdx.x] = rb + threadIdx.x < r1 ? (double)yv[rb + threadIdx.x] : 0.0;
        __syncthreads();
        const int rows = (int)((r1 - rb) < TR ? (r1 - rb) : TR);
        for (int r = 0; r < rows; ++r) {
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) t = __builtin_fma(xs[r * D + j], b[j], t);
            const double res = ys[r] - t;
            acc = __builtin_fma(res, res, acc);
        }
    }
    partial[(size_t)range * n + k] = acc;
}

// The softplus sums of the guard in double: softplus(eta) = max(eta, 0) +
// log1p(exp(-|eta|)), summed over a run of up to 16 rows as the sum of the
// max(eta, 0) plus the logarithm of the product of the factors
// 1 + exp(-|eta|) (each in (1, 2], the product in [1, 2^16], one rounding
// per factor), to an absolute error of a few 1e-16 per row (the guard sums
// over up to 2^32 rows and needs the total to ~1e-9, far above this; the
// library's exp and log1p aim at last-bit accuracy and cost about three
// times as many instructions).
//   t = exp(-a): a = n ln2 + r with n = rint(a / ln2), ln2 split in a
//     32-bit head (n ln2_hi exact) and a tail, |r| <= ln2 / 2; exp(-r) by
//     its Taylor series to degree 13 (remainder < 5e-18), scaled by 2^-n.
//   log P, P in [1, 2^16]: P = 2^k m with m in [sqrt(1/2), sqrt(2)) (frexp,
//     then one doubling), f = m - 1 exact, log m = 2 atanh(s), s = f / (2 +
//     f), |s| <= 0.172, by its series to s^23 (remainder < 1e-18), plus
//     k ln2.
// Its constants, read from memory by the kernel that uses it (as immediates
// the compiler keeps them all in scalar registers, which the list kernel's
// coefficient loads need): ln2 head (32 significant bits) and tail, 1/ln2,
// the Taylor coefficients 1/13! .. 1/0!, then 2/23 .. 2/1.
__constant__ double kGlmSoftplus64[3 + 14 + 12] = {
    0.6931471803691238, 1.9082149292705877e-10, 1.4426950408889634,
    1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
    1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0, 1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0,
    1.0 / 6.0, 0.5, 1.0, 1.0,
    2.0 / 23.0, 2.0 / 21.0, 2.0 / 19.0, 2.0 / 17.0, 2.0 / 15.0, 2.0 / 13.0,
    2.0 / 11.0, 2.0 / 9.0, 2.0 / 7.0, 2.0 / 5.0, 2.0 / 3.0, 2.0};

// exp(-a) for a >= 0 (0 past the double range; a NaN gives 0, the caller
// carries it through the max(eta, 0) sum)
__device__ __forceinline__ double glm_exp_neg64(double a, const double* __restrict__ k) {
    const double n = __builtin_rint(a * k[2]);
    double r = __builtin_fma(-n, k[0], a);
    r = __builtin_fma(-n, k[1], r);
    // exp(-r) = sum_j (-r)^j / j!, Horner from degree 13
    const double x = -r;
    double q = k[3];
#pragma unroll
    for (int j = 4; j < 17; ++j) q = __builtin_fma(q, x, k[j]);
    const bool small = a < 745.0;  // else exp(-a) underflows to 0
    return small ? __builtin_ldexp(q, -(int)(small ? n : 0.0)) : 0.0;
}

// log P for P in [1, 2^16] (log 1 = 0 exactly)
__device__ __forceinline__ double glm_log64(double P, const double* __restrict__ k) {
    int e;
    double m = __builtin_frexp(P, &e);  // P = m 2^e, m in [1/2, 1)
    if (m < 0.70710678118654752) {
        m *= 2.0;
        e -= 1;
    }
    const double f = m - 1.0;
    const double sv = f / (2.0 + f);
    const double s2 = sv * sv;
    // 2 atanh(s) = s (2 + 2 s^2/3 + ... + 2 s^22/23)
    double g = k[17];
#pragma unroll
    for (int j = 18; j < 29; ++j) g = __builtin_fma(g, s2, k[j]);
    const double de = (double)e;
    return __builtin_fma(de, k[0], __builtin_fma(de, k[1], sv * g));
}

// fp64 log-likelihood sums of the points listed by the guarded accept test.
// pts holds the 2c listed points in double (c = min(*count, lmax), read on
// the device): row 2i the proposal of listed walker i, row 2i + 1 its
// position.  A work-group takes one row range and the points eight at a
// time (its rows come from L2 after the first eight); each thread a row at
// a time, the row's pieces loaded together, converted to double eight
// features at a time, against eight points whose coefficients are uniform
// across the wave (scalar loads): fma chains over the features in order as
// the fp64 kernel, then the run softplus above.  partial64[r * 2 lmax + j]
// gets the work-group's sum for point row j < 2c.
template <int D>
__global__ void __launch_bounds__(256)
glm_fp64_list_kernel(const __bf16* __restrict__ X, unsigned long long N,
                     const double* __restrict__ pts, const unsigned* __restrict__ count,
                     unsigned lmax, unsigned long long rows_per_range,
                     double* __restrict__ partial64) {
    constexpr int PC = 8;
    __shared__ double red[PC][4];
    const unsigned n = *count;
    const unsigned c = n < lmax ? n : lmax;
    const unsigned m = 2 * c;
    if (m == 0) return;
    const unsigned t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const unsigned long long r0 = (unsigned long long)blockIdx.x * rows_per_range;
    unsigned long long r1 = r0 + rows_per_range;
    if (r1 > N) r1 = N;
    const double* k = kGlmSoftplus64;
    for (unsigned p0 = 0; p0 < m; p0 += PC) {
        const double* bp = pts + (size_t)p0 * D;
        double lsum[PC], prod[PC];
#pragma unroll
        for (int q = 0; q < PC; ++q) {
            lsum[q] = 0.0;
            prod[q] = 1.0;
        }
        unsigned run = 0;
        for (unsigned long long row = r0 + t; row < r1; row += 256) {
            const bf16x8* xr = reinterpret_cast<const bf16x8*>(X + row * D);
            bf16x8 xv[D / 8];
#pragma unroll
            for (int v = 0; v < D / 8; ++v) xv[v] = xr[v];
            // the row's loads are issued before any of its arithmetic, and
            // the coefficients are read again for every row (the barrier
            // also keeps all 8 x D of them from being hoisted into scalar
            // registers, which would spill)
            asm volatile("" ::: "memory");
            double eta[PC];
#pragma unroll
            for (int q = 0; q < PC; ++q) eta[q] = 0.0;
#pragma unroll
            for (int v = 0; v < D / 8; ++v) {
                double x[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] = (double)(float)xv[v][e];
#pragma unroll
                for (int q = 0; q < PC; ++q)
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        eta[q] = __builtin_fma(x[e], bp[q * D + v * 8 + e], eta[q]);
            }
            // the softplus constants are read after the coefficients
            asm volatile("" ::: "memory");
#pragma unroll
            for (int q = 0; q < PC; ++q) {
                const double et = eta[q];
                prod[q] = __builtin_fma(glm_exp_neg64(__builtin_fabs(et), k), prod[q], prod[q]);
                lsum[q] += et != et ? et : __builtin_fmax(et, 0.0);
            }
            if (++run == 16) {
#pragma unroll
                for (int q = 0; q < PC; ++q) {
                    lsum[q] += glm_log64(prod[q], k);
                    prod[q] = 1.0;
                }
                run = 0;
            }
        }
#pragma unroll
        for (int q = 0; q < PC; ++q) {
            double v = lsum[q] + glm_log64(prod[q], k);
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) red[q][wave] = v;
        }
        __syncthreads();
        if (t < PC && p0 + t < m)
            partial64[(size_t)blockIdx.x * 2 * lmax + p0 + t] =
                (red[t][0] + red[t][1]) + (red[t][2] + red[t][3]);
        __syncthreads();
    }
}

hipError_t launch_glm_fp64_list(hipStream_t s, const void* X, unsigned long long N, unsigned d,
                                const double* pts, const unsigned* count, unsigned lmax,
                                unsigned nranges, double* partial64) {
    if (lmax % 8 || nranges == 0) return hipErrorInvalidValue;
    const unsigned long long rpr = (N + nranges - 1) / nranges;
    const __bf16* x = static_cast<const __bf16*>(X);
    if (d == 64)
        glm_fp64_list_kernel<64><<<nranges, 256, 0, s>>>(x, N, pts, count, lmax, rpr, partial64);
    else if (d == 32)
        glm_fp64_list_kernel<32><<<nranges, 256, 0, s>>>(x, N, pts, count, lmax, rpr, partial64);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// Gradient of the logistic softplus sum at a reference point, for the
// first-order correction of the split precisions:
// g0 = sum_r sigma(x_r . b) x_r (natural units), b the mean of the first
// nref (1 .. 256) points of pts (row stride D).  The split precisions
// evaluate the likelihood at coefficients y' = y + dy that differ from the
// proposal y by the split's rounding (|dy| ~ 2^-22 |y| for fp16x2), and
// g(y) . dy is nearly all of their log posterior's error against fp64;
// with g0 from a point of the ensemble in place of g(y) the correction
// leaves (g(y) - g0) . dy.  Only ~1e-3 relative accuracy is needed of g0:
// rows in fp32 (u, sigma, the thread's sums), blocks and the total in
// double.  Three launches: the reference point (one work-group), the rows
// (one per thread, the next row's loads in flight; partial[block * D + j]),
// the sum over blocks.
template <int D>
__global__ void __launch_bounds__(1024)
glm_grad_ref_kernel(const float* __restrict__ pts, unsigned nref, float* __restrict__ bref) {
    constexpr unsigned G = 1024 / D, PER = 256 / G;
    __shared__ double red[1024];  // G x D
    const unsigned t = threadIdx.x, j = t % D, g = t / D;
    float v[PER];
#pragma unroll
    for (unsigned i = 0; i < PER; ++i) {
        const unsigned p = g + i * G;
        v[i] = p < nref ? pts[(size_t)p * D + j] : 0.0f;
    }
    double m = 0.0;
#pragma unroll
    for (unsigned i = 0; i < PER; ++i) m += (double)v[i];
    red[g * D + j] = m;
    __syncthreads();
    if (t < D) {
        double sm = 0.0;
        for (unsigned q = 0; q < G; ++q) sm += red[q * D + t];
        bref[t] = (float)(sm / (double)nref);
    }
}

template <int D>
__global__ void __launch_bounds__(256)
glm_grad_partial_kernel(const __bf16* __restrict__ X, unsigned long long N,
                        const float* __restrict__ bref_g, unsigned long long rows_per_block,
                        double* __restrict__ partial) {
    __shared__ double red[4 * D];
    __shared__ float bref[D];
    const unsigned t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    if (t < D) bref[t] = bref_g[t];
    __syncthreads();
    const unsigned long long r0 = (unsigned long long)blockIdx.x * rows_per_block;
    unsigned long long r1 = r0 + rows_per_block;
    if (r1 > N) r1 = N;
    float acc[D];
#pragma unroll
    for (int j = 0; j < D; ++j) acc[j] = 0.0f;
    // one row per thread, the next row's pieces in flight meanwhile
    bf16x8 xn[D / 8];
    if (r0 + t < r1) {
#pragma unroll
        for (int v = 0; v < D / 8; ++v) xn[v] = reinterpret_cast<const bf16x8*>(X + (r0 + t) * D)[v];
    }
    for (unsigned long long row = r0 + t; row < r1; row += 256) {
        float x[D];
#pragma unroll
        for (int v = 0; v < D / 8; ++v)
#pragma unroll
            for (int e = 0; e < 8; ++e) x[v * 8 + e] = (float)xn[v][e];
        if (row + 256 < r1) {
#pragma unroll
            for (int v = 0; v < D / 8; ++v)
                xn[v] = reinterpret_cast<const bf16x8*>(X + (row + 256) * D)[v];
        }
        float u = 0.0f;
#pragma unroll
        for (int j = 0; j < D; ++j) u = __builtin_fmaf(x[j], bref[j], u);
        const float sg = 1.0f / (1.0f + __expf(-u));
#pragma unroll
        for (int j = 0; j < D; ++j) acc[j] = __builtin_fmaf(sg, x[j], acc[j]);
    }
    double mine = 0.0;  // lane j < D of each wave keeps the wave's sum for feature j
#pragma unroll
    for (int j = 0; j < D; ++j) {
        double v = (double)acc[j];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == (unsigned)j) mine = v;
    }
    if (lane < D) red[wave * D + lane] = mine;
    __syncthreads();
    if (t < D)
        partial[(size_t)blockIdx.x * D + t] =
            (red[t] + red[D + t]) + (red[2 * D + t] + red[3 * D + t]);
}

// g0[j] = sum over the blocks' partials (minus off[j], when given): thread
// (g, j) sums blocks g, g + G, ... (all its loads in flight), then one
// thread per feature the G sums in order.
template <int D>
__global__ void __launch_bounds__(1024)
glm_grad_final_kernel(const double* __restrict__ partial, const double* __restrict__ off,
                      double* __restrict__ g0) {
    constexpr unsigned G = 1024 / D, PER = kGlmGradBlocks / G;
    __shared__ double red[1024];
    const unsigned t = threadIdx.x, j = t % D, g = t / D;
    double v[PER];
#pragma unroll
    for (unsigned i = 0; i < PER; ++i) v[i] = partial[(size_t)(g + i * G) * D + j];
    double s = 0.0;
#pragma unroll
    for (unsigned i = 0; i < PER; ++i) s += v[i];
    red[t] = s;
    __syncthreads();
    if (t < D) {
        double sm = 0.0;
        for (unsigned q = 0; q < G; ++q) sm += red[q * D + t];
        g0[t] = off ? sm - off[t] : sm;
    }
}

hipError_t launch_glm_grad(hipStream_t s, const void* X, unsigned long long N, unsigned d,
                           const float* pts, unsigned nref, const double* off, double* work,
                           double* g0) {
    if (nref == 0 || nref > 256) return hipErrorInvalidValue;
    static_assert(kGlmGradBlocks % 32 == 0, "the final sum's grouping");
    const unsigned long long rpb = (N + kGlmGradBlocks - 1) / kGlmGradBlocks;
    const __bf16* x = static_cast<const __bf16*>(X);
    // work: kGlmGradBlocks x d partials, then the reference point (d floats)
    float* bref = reinterpret_cast<float*>(work + (size_t)kGlmGradBlocks * d);
    if (d == 64) {
        glm_grad_ref_kernel<64><<<1, 1024, 0, s>>>(pts, nref, bref);
        glm_grad_partial_kernel<64><<<kGlmGradBlocks, 256, 0, s>>>(x, N, bref, rpb, work);
        glm_grad_final_kernel<64><<<1, 1024, 0, s>>>(work, off, g0);
    } else if (d == 32) {
        glm_grad_ref_kernel<32><<<1, 1024, 0, s>>>(pts, nref, bref);
        glm_grad_partial_kernel<32><<<kGlmGradBlocks, 256, 0, s>>>(x, N, bref, rpb, work);
        glm_grad_final_kernel<32><<<1, 1024, 0, s>>>(work, off, g0);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Hessian of the logistic softplus sum at the reference point b:
// H0 = sum_r sigma'(x_r . b) x_r x_r^T (natural units), for the second-order
// correction of the one-term fp16 precision.  A work-group stages RB rows
// in LDS as floats with their weights sigma'(u_r), and each of its threads
// accumulates a TI x TI tile of H in fp32 over the work-group's rows (the
// correction needs H0 to ~1e-3 relative); the tiles are written in double
// and summed over work-groups in two passes (32 at a time, then the 32-sums
// in order).
template <int D>
__global__ void __launch_bounds__(256)
glm_hess_partial_kernel(const __bf16* __restrict__ X, unsigned long long N,
                        const float* __restrict__ bref_g, unsigned long long rows_per_block,
                        double* __restrict__ partial) {
    constexpr int TI = D / 16;     // 4 (D = 64) or 2 (D = 32)
    constexpr int PPR = D / 8;     // 16-byte pieces per row
    constexpr int RB = 256 / PPR;  // rows per stage: 32 or 64
    __shared__ float xs[RB][D + 4];
    __shared__ float ws[RB];
    __shared__ float bref[D];
    const unsigned t = threadIdx.x, ti = t / 16, tj = t % 16;
    if (t < D) bref[t] = bref_g[t];
    const unsigned long long r0 = (unsigned long long)blockIdx.x * rows_per_block;
    unsigned long long r1 = r0 + rows_per_block;
    if (r1 > N) r1 = N;
    float acc[TI][TI];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TI; ++j) acc[i][j] = 0.0f;
    for (unsigned long long base = r0; base < r1; base += RB) {
        __syncthreads();  // the previous stage is consumed (and bref visible)
        {
            const unsigned rr = t / PPR, pc = t % PPR;
            const unsigned long long row = base + rr;
            float v[8];
            if (row < r1) {
                const bf16x8 xv = reinterpret_cast<const bf16x8*>(X + row * D)[pc];
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (float)xv[e];
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = 0.0f;
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) xs[rr][pc * 8 + e] = v[e];
        }
        __syncthreads();
        if (t < RB) {
            float u = 0.0f;
#pragma unroll
            for (int j = 0; j < D; ++j) u = __builtin_fmaf(xs[t][j], bref[j], u);
            const float sg = 1.0f / (1.0f + __expf(-u));
            ws[t] = base + t < r1 ? sg * (1.0f - sg) : 0.0f;
        }
        __syncthreads();
#pragma unroll 4
        for (int r = 0; r < RB; ++r) {
            const float w = ws[r];
            float a[TI], bb[TI];
#pragma unroll
            for (int i = 0; i < TI; ++i) {
                a[i] = xs[r][ti * TI + i];
                bb[i] = w * xs[r][tj * TI + i];
            }
#pragma unroll
            for (int i = 0; i < TI; ++i)
#pragma unroll
                for (int j = 0; j < TI; ++j) acc[i][j] = __builtin_fmaf(a[i], bb[j], acc[i][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TI; ++j)
            partial[(size_t)blockIdx.x * D * D + (ti * TI + i) * D + tj * TI + j] =
                (double)acc[i][j];
}

// out[c * DD + e] = sum of in[(c * 32 + q) * DD + e], q < 32 (all loads in
// flight); with final set, H0[e] = (float) of the sum of the nc 32-sums.
__global__ void __launch_bounds__(256)
glm_hess_sum_kernel(const double* __restrict__ in, unsigned dd, double* __restrict__ out) {
    const unsigned e = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (e >= dd) return;
    double v[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) v[q] = in[((size_t)c * 32 + q) * dd + e];
    double sm = 0.0;
#pragma unroll
    for (int q = 0; q < 32; ++q) sm += v[q];
    out[(size_t)c * dd + e] = sm;
}

__global__ void __launch_bounds__(256)
glm_hess_final_kernel(const double* __restrict__ in, unsigned nc, unsigned dd,
                      float* __restrict__ H0) {
    const unsigned e = blockIdx.x * 256 + threadIdx.x;
    if (e >= dd) return;
    double sm = 0.0;
    for (unsigned c = 0; c < nc; ++c) sm += in[(size_t)c * dd + e];
    H0[e] = (float)sm;
}

hipError_t launch_glm_hess(hipStream_t s, const void* X, unsigned long long N, unsigned d,
                           const float* bref, double* work, float* H0) {
    static_assert(kGlmHessBlocks % 32 == 0, "the sums over work-groups go 32 at a time");
    const unsigned long long rpb = (N + kGlmHessBlocks - 1) / kGlmHessBlocks;
    const __bf16* x = static_cast<const __bf16*>(X);
    const unsigned dd = d * d, nc = kGlmHessBlocks / 32;
    double* part2 = work + (size_t)kGlmHessBlocks * dd;
    if (d == 64)
        glm_hess_partial_kernel<64><<<kGlmHessBlocks, 256, 0, s>>>(x, N, bref, rpb, work);
    else if (d == 32)
        glm_hess_partial_kernel<32><<<kGlmHessBlocks, 256, 0, s>>>(x, N, bref, rpb, work);
    else
        return hipErrorInvalidValue;
    glm_hess_sum_kernel<<<dim3((dd + 255) / 256, nc), 256, 0, s>>>(work, dd, part2);
    glm_hess_final_kernel<<<(dd + 255) / 256, 256, 0, s>>>(part2, nc, dd, H0);
    return hipGetLastError();
}

// lin[j] = sum_i y_i x_ij (or sum_i x_ij when yv is null) in double: one
// block per chunk of rows, one thread per feature; chunk partials summed in
// chunk order by lin_final.
__global__ void __launch_bounds__(256)
glm_lin_partial_kernel(const __bf16* __restrict__ X, const float* __restrict__ yv,
                       unsigned long long N, unsigned d, unsigned long long rows_per_chunk,
                       double* __restrict__ part) {
    const unsigned j = threadIdx.x;
    if (j >= d) return;
    const unsigned long long r0 = (unsigned long long)blockIdx.x * rows_per_chunk;
    unsigned long long r1 = r0 + rows_per_chunk;
    if (r1 > N) r1 = N;
    double s = 0.0;
    if (yv) {
        for (unsigned long long i = r0; i < r1; ++i)
            s = __builtin_fma((double)yv[i], (double)(float)X[i * d + j], s);
    } else {
        for (unsigned long long i = r0; i < r1; ++i) s += (double)(float)X[i * d + j];
    }
    part[(size_t)blockIdx.x * d + j] = s;
}

__global__ void __launch_bounds__(256)
glm_lin_final_kernel(const double* __restrict__ part, unsigned chunks, unsigned d,
                     double* __restrict__ lin) {
    const unsigned j = threadIdx.x;
    if (j >= d) return;
    double s = 0.0;
    for (unsigned c = 0; c < chunks; ++c) s += part[(size_t)c * d + j];
    lin[j] = s;
}

// --------------------------------------------------------------------------
template <int D>
static hipError_t launch_d(int precision, hipStream_t s, const void* Xv, unsigned long long N,
                           const float* P, const void* ysv, unsigned n, unsigned nranges,
                           unsigned long long rpr, float yscale, double* partial) {
    const __bf16* X = static_cast<const __bf16*>(Xv);
    const __bf16* ys = static_cast<const __bf16*>(ysv);
    const _Float16* X16 = static_cast<const _Float16*>(Xv);
    const _Float16* ys16 = static_cast<const _Float16*>(ysv);
    switch (precision) {
    case kGlmBf16: {
        const unsigned groups = n / 256;
        glm_logistic_mfma_kernel<D, 1, __bf16, bf16x8><<<groups * nranges, 256, 0, s>>>(
            X, (unsigned)N, ys, n, groups, nranges, (unsigned)rpr, partial);
        break;
    }
    case kGlmBf16x3: {
        const unsigned groups = n / 256;
        glm_logistic_mfma_kernel<D, 3, __bf16, bf16x8><<<groups * nranges, 256, 0, s>>>(
            X, (unsigned)N, ys, n, groups, nranges, (unsigned)rpr, partial);
        break;
    }
    case kGlmFp16x2: {
        const unsigned groups = n / 256;
        glm_logistic_mfma_kernel<D, 2, _Float16, f16x8><<<groups * nranges, 256, 0, s>>>(
            X16, (unsigned)N, ys16, n, groups, nranges, (unsigned)rpr, partial);
        break;
    }
    case kGlmFp16: {
        const unsigned groups = n / 256;
        glm_logistic_mfma_kernel<D, 1, _Float16, f16x8><<<groups * nranges, 256, 0, s>>>(
            X16, (unsigned)N, ys16, n, groups, nranges, (unsigned)rpr, partial);
        break;
    }
    case kGlmFp32: {
        const unsigned groups = n / 256;
        glm_logistic_fp32_kernel<D><<<groups * nranges, 256, 0, s>>>(X, N, P, n, groups, nranges,
                                                                     rpr, yscale, partial);
        break;
    }
    case kGlmFp64: {
        const unsigned groups = n / 256;
        glm_logistic_fp64_kernel<D><<<groups * nranges, 256, 0, s>>>(X, N, P, n, groups, nranges,
                                                                     rpr, partial);
        break;
    }
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool glm_supported_dim(unsigned d) { return d == 32 || d == 64; }

unsigned glm_point_quantum(int precision) {
    (void)precision;
    return 256u;
}

unsigned glm_blocks_per_cu(int precision, unsigned d) {
    const void* k = nullptr;
    switch (precision) {
    case kGlmBf16:
        k = d == 64 ? (const void*)glm_logistic_mfma_kernel<64, 1, __bf16, bf16x8>
                    : (const void*)glm_logistic_mfma_kernel<32, 1, __bf16, bf16x8>;
        break;
    case kGlmBf16x3:
        k = d == 64 ? (const void*)glm_logistic_mfma_kernel<64, 3, __bf16, bf16x8>
                    : (const void*)glm_logistic_mfma_kernel<32, 3, __bf16, bf16x8>;
        break;
    case kGlmFp16x2:
        k = d == 64 ? (const void*)glm_logistic_mfma_kernel<64, 2, _Float16, f16x8>
                    : (const void*)glm_logistic_mfma_kernel<32, 2, _Float16, f16x8>;
        break;
    case kGlmFp16:
        k = d == 64 ? (const void*)glm_logistic_mfma_kernel<64, 1, _Float16, f16x8>
                    : (const void*)glm_logistic_mfma_kernel<32, 1, _Float16, f16x8>;
        break;
    case kGlmFp32:
        k = d == 64 ? (const void*)glm_logistic_fp32_kernel<64>
                    : (const void*)glm_logistic_fp32_kernel<32>;
        break;
    case kGlmFp64:
        k = d == 64 ? (const void*)glm_logistic_fp64_kernel<64>
                    : (const void*)glm_logistic_fp64_kernel<32>;
        break;
    default:
        return 1;
    }
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 256, 0) != hipSuccess ||
        blocks < 1)
        return 1;
    return (unsigned)blocks;
}

hipError_t launch_glm_logistic(int precision, hipStream_t s, const void* X, unsigned long long N,
                               unsigned d, const float* P, const void* ys, unsigned n,
                               unsigned nranges, unsigned long long rows_per_range, float yscale,
                               double* partial) {
    if (d == 64) return launch_d<64>(precision, s, X, N, P, ys, n, nranges, rows_per_range, yscale, partial);
    if (d == 32) return launch_d<32>(precision, s, X, N, P, ys, n, nranges, rows_per_range, yscale, partial);
    return hipErrorInvalidValue;
}

// Issue-rate kernels: eight independent chains per thread of v_exp_f32
// (the softplus epilogue's transcendental) or of v_fma_f32.  Each thread
// writes one float.
__global__ void __launch_bounds__(256) glm_rate_exp_kernel(float* __restrict__ out, int iters) {
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = 0.001f * (float)(threadIdx.x + i);
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = __builtin_amdgcn_exp2f(-__builtin_fabsf(x[i]));
    out[blockIdx.x * 256 + threadIdx.x] = ((x[0] + x[1]) + (x[2] + x[3])) + ((x[4] + x[5]) + (x[6] + x[7]));
}

__global__ void __launch_bounds__(256) glm_rate_fma_kernel(float* __restrict__ out, int iters) {
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = 0.001f * (float)(threadIdx.x + i);
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = __builtin_fmaf(x[i], 0.999f, 1e-4f);
    out[blockIdx.x * 256 + threadIdx.x] = ((x[0] + x[1]) + (x[2] + x[3])) + ((x[4] + x[5]) + (x[6] + x[7]));
}

// Eight independent accumulators per wave of v_mfma_f32_16x16x32_f16 (the
// fp16 likelihood kernel's product).  Each thread writes one float.
__global__ void __launch_bounds__(256) glm_rate_mfma_kernel(float* __restrict__ out, int iters) {
    f16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = (_Float16)(0.001f * (float)(threadIdx.x + i));
        b[i] = (_Float16)(0.002f * (float)i);
    }
    f32x4 c[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) c[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int j = 0; j < 8; ++j) c[j] = mfma_16x16x32(a, b, c[j]);
    float sum = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) sum += c[j][0] + c[j][3];
    out[blockIdx.x * 256 + threadIdx.x] = sum;
}

// Random 16-byte gathers: thread k reads record (k * 2654435761) mod n16
// of src (n16 a power of two; 16-byte records) and writes their sum; NT:
// the loads are non-temporal.
template <bool NT>
__global__ void __launch_bounds__(256) glm_rate_gather_kernel(const f32x4* __restrict__ src,
                                                              float* __restrict__ out,
                                                              unsigned n16) {
    const unsigned k = blockIdx.x * 256 + threadIdx.x;
    const f32x4* p = src + ((k * 2654435761u) & (n16 - 1u));
    const f32x4 v = NT ? __builtin_nontemporal_load(p) : *p;
    out[k] = (v.x + v.y) + (v.z + v.w);
}

hipError_t launch_glm_gather(hipStream_t s, const void* src, float* out, unsigned n16, bool nt) {
    if (n16 == 0 || (n16 & (n16 - 1)) || n16 % 256) return hipErrorInvalidValue;
    const f32x4* q = static_cast<const f32x4*>(src);
    if (nt)
        glm_rate_gather_kernel<true><<<n16 / 256, 256, 0, s>>>(q, out, n16);
    else
        glm_rate_gather_kernel<false><<<n16 / 256, 256, 0, s>>>(q, out, n16);
    return hipGetLastError();
}

hipError_t launch_glm_rate(hipStream_t s, int which, float* out, unsigned blocks, int iters) {
    if (which == 0)
        glm_rate_exp_kernel<<<blocks, 256, 0, s>>>(out, iters);
    else if (which == 1)
        glm_rate_fma_kernel<<<blocks, 256, 0, s>>>(out, iters);
    else
        glm_rate_mfma_kernel<<<blocks, 256, 0, s>>>(out, iters);
    return hipGetLastError();
}

size_t glm_x_f16_bytes(unsigned long long N, unsigned d) {
    return (size_t)((N + 15) / 16) * 16 * d * sizeof(_Float16);
}

hipError_t launch_glm_x_to_f16(hipStream_t s, const void* X, unsigned long long N, unsigned d,
                               void* out, unsigned* overflow) {
    if (d != 32 && d != 64) return hipErrorInvalidValue;
    const unsigned long long n8 = N * d / 8;
    if (n8 == 0) return hipSuccess;
    const hipError_t e = hipMemsetAsync(out, 0, glm_x_f16_bytes(N, d), s);
    if (e != hipSuccess) return e;
    const unsigned blocks = (unsigned)((n8 + 255) / 256);
    if (d == 64)
        glm_x_to_f16_kernel<64><<<blocks, 256, 0, s>>>(static_cast<const __bf16*>(X), n8,
                                                       static_cast<_Float16*>(out), overflow);
    else
        glm_x_to_f16_kernel<32><<<blocks, 256, 0, s>>>(static_cast<const __bf16*>(X), n8,
                                                       static_cast<_Float16*>(out), overflow);
    return hipGetLastError();
}

// Points per thread of the fp32 linear-regression kernel: 2 for up to 32
// features (4 points at 16 features measured no faster, 2 at 64 spill) when
// n divides into whole work-groups of 512 points.
unsigned glm_gauss_points_per_thread(int precision, unsigned D, unsigned n) {
    return precision == kGlmFp32 && D <= 32 && n % 512 == 0 ? 2u : 1u;
}

template <int D>
hipError_t launch_gauss_d(int precision, hipStream_t s, const float* X, const float* yv,
                          unsigned long long N, const float* P, unsigned pstride, unsigned d,
                          unsigned n, unsigned nranges, unsigned long long rows_per_range,
                          double* partial) {
    const unsigned groups = n / 256;
    const unsigned blocks = groups * nranges;
    if (precision == kGlmFp32) {
        if (glm_gauss_points_per_thread(precision, D, n) == 2) {
            if constexpr (D <= 32)
                glm_gauss_fp32_kernel<D, 2><<<blocks / 2, 256, 0, s>>>(
                    X, yv, N, P, pstride, d, n, groups / 2, nranges, rows_per_range, partial);
        } else {
            glm_gauss_fp32_kernel<D, 1><<<blocks, 256, 0, s>>>(X, yv, N, P, pstride, d, n,
                                                               groups, nranges, rows_per_range,
                                                               partial);
        }
    } else if (precision == kGlmFp64) {
        glm_gauss_fp64_kernel<D><<<blocks, 256, 0, s>>>(X, yv, N, P, pstride, d, n, groups,
                                                        nranges, rows_per_range, partial);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_glm_gauss(int precision, hipStream_t s, const float* X, const float* yv,
                            unsigned long long N, unsigned D, const float* P, unsigned pstride,
                            unsigned d, unsigned n, unsigned nranges,
                            unsigned long long rows_per_range, double* partial) {
    if (D == 16)
        return launch_gauss_d<16>(precision, s, X, yv, N, P, pstride, d, n, nranges,
                                  rows_per_range, partial);
    if (D == 32)
        return launch_gauss_d<32>(precision, s, X, yv, N, P, pstride, d, n, nranges,
                                  rows_per_range, partial);
    if (D == 64)
        return launch_gauss_d<64>(precision, s, X, yv, N, P, pstride, d, n, nranges,
                                  rows_per_range, partial);
    return hipErrorInvalidValue;
}

hipError_t launch_glm_lin(hipStream_t s, const void* X, const float* yv, unsigned long long N,
                          unsigned d, double* part, unsigned chunks, double* lin) {
    const unsigned long long rpc = (N + chunks - 1) / chunks;
    glm_lin_partial_kernel<<<chunks, 256, 0, s>>>(static_cast<const __bf16*>(X), yv, N, d, rpc,
                                                  part);
    glm_lin_final_kernel<<<1, 256, 0, s>>>(part, chunks, d, lin);
    return hipGetLastError();
}

}  // namespace bh
