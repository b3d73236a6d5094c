// Batched GLM posteriors: many independent logistic regressions with uneven
// row counts N_p and dimensions d_p <= 64 advance in lockstep, one set of
// launches per half-step for all of them.  X_p is stored as fp8 e4m3 (the
// OCP format gfx950's conversions read) with a fp32 scale per column, row
// stride D_p = 32 (d_p <= 32) or 64, padding columns zero; walkers are
// padded to 64 floats.  The contraction runs over "units" (a problem's row
// range of at most kGbUnitRows rows) x 256-point groups: the MFMA kernel
// decodes fp8 rows to bf16 or fp16 in registers (exact: every e4m3 value is
// a bf16 and an fp16 value) and multiplies them with the split proposals
// (three bf16 or two fp16 terms), whose column scales are folded in
// (x_ij b_j = q_ij (s_j b_j)).  The walker-side kernels
// follow the single-posterior GLM sampler: the same proposal and accept
// arithmetic, with each problem's own dimension and seeds.  Partners are
// whole walkers of the complementary half.
#include <hip/hip_runtime.h>

#include "bh_glm.hpp"
#include "bh_glm_batch.hpp"
#include "bh_glm_core.hpp"
#include "bh_philox.hpp"

namespace bh {

constexpr int kGbD = 64;  // walker stride (floats)

// Four fp8 bytes -> four floats (v_cvt_pk_f32_fp8, low then high word).
__device__ __forceinline__ void fp8x4_to_f32(unsigned v, float (&f)[4]) {
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    const f32x2 lo = __builtin_amdgcn_cvt_pk_f32_fp8(v, false);
    const f32x2 hi = __builtin_amdgcn_cvt_pk_f32_fp8(v, true);
    f[0] = lo.x;
    f[1] = lo.y;
    f[2] = hi.x;
    f[3] = hi.y;
}

// Eight fp8 bytes -> eight bf16: the floats of fp8 values have zero low 16
// bits, so truncation is exact.
__device__ __forceinline__ bf16x8 fp8x8_to_bf16x8(uint2 v) {
    float f[8];
    fp8x4_to_f32(v.x, *reinterpret_cast<float(*)[4]>(&f[0]));
    fp8x4_to_f32(v.y, *reinterpret_cast<float(*)[4]>(&f[4]));
    bf16x8 r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = __builtin_bit_cast(__bf16, (unsigned short)(__builtin_bit_cast(unsigned, f[i]) >> 16));
    return r;
}

// The fp16 matrix-core path reads fp8 values times 2^-8 (and the proposals
// times 2^8): every e4m3 value times 2^-8 is still an fp16 value (its lowest
// bit, 2^-17, is above fp16's 2^-24), and the proposals' column-scaled
// coefficients, scaled up, keep their low split term out of fp16's
// subnormal range.
constexpr float kGbF16Down = 0x1p-8f;
constexpr float kGbF16Up = 0x1p8f;

// Eight fp8 bytes -> eight fp16 values times 2^-8 (v_cvt_scalef32_pk_f16_fp8,
// two bytes per instruction: word 0 then word 1 of each dword).
__device__ __forceinline__ f16x8 fp8x8_to_f16x8(uint2 v) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    u32x4 w;
    w[0] = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_scalef32_pk_f16_fp8(v.x, kGbF16Down, false));
    w[1] = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_scalef32_pk_f16_fp8(v.x, kGbF16Down, true));
    w[2] = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_scalef32_pk_f16_fp8(v.y, kGbF16Down, false));
    w[3] = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_scalef32_pk_f16_fp8(v.y, kGbF16Down, true));
    return __builtin_bit_cast(f16x8, w);
}

__device__ __forceinline__ void fp8x8_decode(uint2 v, bf16x8& a) { a = fp8x8_to_bf16x8(v); }
__device__ __forceinline__ void fp8x8_decode(uint2 v, f16x8& a) { a = fp8x8_to_f16x8(v); }

// A fragments from fp8 X (row stride D bytes), decoded to V (bf16x8 or
// f16x8).
template <int D, class V>
struct Fp8Rows {
    using vec = V;
    const unsigned char* X;
    unsigned lr, lg;
    __device__ void full(unsigned rb, V (&a)[D / 32]) const {
        const unsigned char* p = X + (size_t)(rb + lr) * D + lg * 8;
#pragma unroll
        for (int kc = 0; kc < D / 32; ++kc)
            fp8x8_decode(*reinterpret_cast<const uint2*>(p + kc * 32), a[kc]);
    }
    __device__ void masked(unsigned rb, unsigned r1, V (&a)[D / 32]) const {
        const unsigned row = rb + lr;
#pragma unroll
        for (int kc = 0; kc < D / 32; ++kc) {
            const uint2 v = row < r1 ? *reinterpret_cast<const uint2*>(X + (size_t)row * D +
                                                                       kc * 32 + lg * 8)
                                     : uint2{0u, 0u};
            fp8x8_decode(v, a[kc]);
        }
    }
};

// The device's decoding of fp8 bytes, four per thread: f16 = 0 the fp32
// path's (v_cvt_pk_f32_fp8), f16 = 1 the fp16 matrix-core path's (scaled
// fp16 values, returned times 2^8 as floats).
__global__ void __launch_bounds__(256)
gb_fp8_decode_kernel(const unsigned char* __restrict__ in, float* __restrict__ out,
                     unsigned long long n, int f16) {
    const unsigned long long i = (unsigned long long)blockIdx.x * 256 + threadIdx.x;
    if (4 * i + 3 < n) {
        const unsigned v = *reinterpret_cast<const unsigned*>(in + 4 * i);
        float f[4];
        if (f16) {
            const uint2 vv{v, 0u};
            const f16x8 h = fp8x8_to_f16x8(vv);
#pragma unroll
            for (int m = 0; m < 4; ++m) f[m] = (float)h[m] * kGbF16Up;
        } else {
            fp8x4_to_f32(v, f);
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) out[4 * i + m] = f[m];
    }
}

// Initial positions: problem p runs init_walkers' stream (key {seed + p,
// 0xdecafaaa}, amd-gcn-mcmc-stretch.cl:138-159) over its W x d_p values,
// limits row i % d_p, written to the padded layout.
__global__ void __launch_bounds__(256)
gb_init_kernel(unsigned seed, const unsigned* __restrict__ dims,
               const float* __restrict__ limits, float* __restrict__ xs, unsigned W) {
    const unsigned p = blockIdx.y;
    const unsigned d = dims[p];
    float* gx = xs + (size_t)p * W * kGbD;
    const float* lim = limits + (size_t)p * kGbD * 2;
    const unsigned items = d * W / 4;
    for (unsigned g = blockIdx.x * 256 + threadIdx.x; g < items; g += gridDim.x * 256) {
        const bh_f32x4 u = bh_u01x4(bh_philox4x32_10(
            bh_u32x4{g, BH_CTR1_INIT, BH_CTR2_INIT, BH_CTR3}, seed + p, BH_KEY1_INIT));
        const float uu[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (unsigned m = 0; m < 4; ++m) {
            const unsigned i = 4 * g + m;
            const unsigned w = i / d, r = i % d;
            gx[(size_t)w * kGbD + r] = __builtin_fmaf(uu[m], lim[2 * r + 1], (1.0f - uu[m]) * lim[2 * r]);
        }
    }
}

// Writes r * t (t = log2(e) * column scale) as fp32 and as nsplit 16-bit
// terms (bf16; or fp16 of r * t * 2^8 when f16 is set; hi, mid, lo), for
// point k of problem p: the first D of its 64 slots (D = 32 or 64, the
// problem's row stride, all that the contraction kernels read; slots from
// d on hold zeros), in 16-byte stores.  r is overwritten (the scaled
// values, then the split residuals).
template <int D>
__device__ __forceinline__ void gb_store_point(float (&r)[D], const float* __restrict__ t,
                                               unsigned d, unsigned P, unsigned p, unsigned n,
                                               unsigned k, float* __restrict__ ysc,
                                               void* __restrict__ ysplit, int nsplit, int f16) {
#pragma unroll
    for (int i = 0; i < D; ++i) r[i] = (unsigned)i < d ? r[i] * t[i] : 0.0f;
    float4* o = reinterpret_cast<float4*>(ysc + ((size_t)p * n + k) * kGbD);
#pragma unroll
    for (int c = 0; c < D / 4; ++c) o[c] = make_float4(r[4 * c], r[4 * c + 1], r[4 * c + 2], r[4 * c + 3]);
    if (f16) {
#pragma unroll
        for (int i = 0; i < D; ++i) r[i] *= kGbF16Up;
    }
    for (int sp = 0; sp < nsplit; ++sp) {
        const size_t at = (((size_t)sp * P + p) * n + k) * kGbD;
#pragma unroll
        for (int c = 0; c < D / 8; ++c) {
            if (f16) {
                f16x8 h;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    h[e] = (_Float16)r[8 * c + e];
                    r[8 * c + e] -= (float)h[e];
                }
                reinterpret_cast<f16x8*>(static_cast<_Float16*>(ysplit) + at)[c] = h;
            } else {
                bf16x8 h;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    h[e] = (__bf16)r[8 * c + e];
                    r[8 * c + e] -= (float)h[e];
                }
                reinterpret_cast<bf16x8*>(static_cast<__bf16*>(ysplit) + at)[c] = h;
            }
        }
    }
}

template <int D>
__device__ __forceinline__ void gb_points_one(const float* __restrict__ x, const float* t,
                                              unsigned d, unsigned P, unsigned p, unsigned n,
                                              unsigned k, float* ysc, void* ysplit, int nsplit,
                                              int f16) {
    float y[D];
#pragma unroll
    for (int c = 0; c < D / 4; ++c) {
        const float4 v = reinterpret_cast<const float4*>(x)[c];
        y[4 * c] = v.x;
        y[4 * c + 1] = v.y;
        y[4 * c + 2] = v.z;
        y[4 * c + 3] = v.w;
    }
    gb_store_point<D>(y, t, d, P, p, n, k, ysc, ysplit, nsplit, f16);
}

// Scaled points (and their splits) of all walkers, for the initial log
// posteriors.
__global__ void __launch_bounds__(256)
gb_points_kernel(const float* __restrict__ xs, const unsigned* __restrict__ dims,
                 const float* __restrict__ tscale, unsigned P, unsigned n,
                 float* __restrict__ ysc, void* __restrict__ ysplit, int nsplit, int f16) {
    const unsigned p = blockIdx.y;
    const unsigned k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const unsigned d = dims[p];
    const float* x = xs + ((size_t)p * n + k) * kGbD;
    const float* t = tscale + (size_t)p * kGbD;
    if (d <= 32)
        gb_points_one<32>(x, t, d, P, p, n, k, ysc, ysplit, nsplit, f16);
    else
        gb_points_one<64>(x, t, d, P, p, n, k, ysc, ysplit, nsplit, f16);
}

// The proposals of a block's 256 walkers from their partners (indices jl,
// stretch factors zl), one 16-byte piece per thread: D/4 consecutive threads
// take one walker's record, so the block reads and writes its records, and
// each partner's, as consecutive pieces: Y = fma(z, X - S, S) (zero beyond
// d), then the scaled point and its split as gb_store_point makes them.
template <int D>
__device__ __forceinline__ void gb_propose_pieces(const float* __restrict__ X,
                                                  const float* __restrict__ S,
                                                  const unsigned* jl, const float* zl,
                                                  unsigned d, unsigned p, unsigned K, unsigned kb,
                                                  float* __restrict__ Y,
                                                  const float* __restrict__ t, unsigned P,
                                                  float* __restrict__ ysc,
                                                  void* __restrict__ ysplit, int nsplit,
                                                  int f16) {
    typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    constexpr unsigned T = D / 4, WP = 256 / T;
    const unsigned c = threadIdx.x % T;
    const float4 tv = reinterpret_cast<const float4*>(t)[c];
    const float ta[4] = {tv.x, tv.y, tv.z, tv.w};
    for (unsigned it = 0; it < T; ++it) {
        const unsigned w = it * WP + threadIdx.x / T;
        const unsigned k = kb + w;
        const float4 xv = reinterpret_cast<const float4*>(X + (size_t)k * kGbD)[c];
        const float4 sv = reinterpret_cast<const float4*>(S + (size_t)jl[w] * kGbD)[c];
        const float z = zl[w];
        const float xa[4] = {xv.x, xv.y, xv.z, xv.w};
        const float sa[4] = {sv.x, sv.y, sv.z, sv.w};
        float y[4], r[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const unsigned i = 4 * c + e;
            y[e] = i < d ? __builtin_fmaf(z, xa[e] - sa[e], sa[e]) : 0.0f;
            r[e] = i < d ? y[e] * ta[e] : 0.0f;
        }
        const size_t row = ((size_t)p * K + k) * kGbD;
        reinterpret_cast<float4*>(Y + row)[c] = make_float4(y[0], y[1], y[2], y[3]);
        reinterpret_cast<float4*>(ysc + row)[c] = make_float4(r[0], r[1], r[2], r[3]);
        if (f16) {
#pragma unroll
            for (int e = 0; e < 4; ++e) r[e] *= kGbF16Up;
        }
        for (int sp = 0; sp < nsplit; ++sp) {
            const size_t at = (((size_t)sp * P + p) * K + k) * kGbD;
            if (f16) {
                f16x4 h;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    h[e] = (_Float16)r[e];
                    r[e] -= (float)h[e];
                }
                reinterpret_cast<f16x4*>(static_cast<_Float16*>(ysplit) + at)[c] = h;
            } else {
                bf16x4 h;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    h[e] = (__bf16)r[e];
                    r[e] -= (float)h[e];
                }
                reinterpret_cast<bf16x4*>(static_cast<__bf16*>(ysplit) + at)[c] = h;
            }
        }
    }
}

// Proposals of one half-step for all problems (the move of
// amd-gcn-mcmc-stretch.cl:31-65 with whole-walker partners): problem p uses
// seed seed0 + 2p (+1 for the second half) and counter word `oe`.  Each
// thread draws one walker's uniforms; the records are then built piece by
// piece (gb_propose_pieces).
__global__ void __launch_bounds__(256)
gb_propose_kernel(unsigned seed0, int half, unsigned oe, unsigned step, float a,
                  const unsigned* __restrict__ dims, const float* __restrict__ tscale,
                  const float* __restrict__ xs, unsigned P, unsigned K, float* __restrict__ Y,
                  float* __restrict__ zbuf, float* __restrict__ u2buf, float* __restrict__ ysc,
                  void* __restrict__ ysplit, int nsplit, int f16) {
    __shared__ unsigned jl[256];
    __shared__ float zl[256];
    const unsigned p = blockIdx.y;
    const unsigned kb = blockIdx.x * 256;
    const unsigned k = kb + threadIdx.x;
    const unsigned d = dims[p];
    const float* X = xs + ((size_t)p * 2 * K + (half ? K : 0)) * kGbD;
    const float* S = xs + ((size_t)p * 2 * K + (half ? 0 : K)) * kGbD;
    const bh_f32x4 u = bh_u01x4(bh_philox4x32_10(bh_u32x4{k, step, oe, BH_CTR3},
                                                 seed0 + 2u * p + (half ? 1u : 0u),
                                                 BH_KEY1_STRETCH));
    const float inv_a = 1.0f / a;
    const float ca = a - 2.0f + inv_a;
    const float cb = 2.0f * (1.0f - inv_a);
    const float z = __builtin_fmaf(ca * u.y, u.y, cb * u.y) + inv_a;
    unsigned j = (unsigned)(u.x * (float)K);
    j = j < K ? j : K - 1;
    jl[threadIdx.x] = j;
    zl[threadIdx.x] = z;
    zbuf[(size_t)p * K + k] = z;
    u2buf[(size_t)p * K + k] = u.z;
    __syncthreads();
    const float* t = tscale + (size_t)p * kGbD;
    if (d <= 32)
        gb_propose_pieces<32>(X, S, jl, zl, d, p, K, kb, Y, t, P, ysc, ysplit, nsplit, f16);
    else
        gb_propose_pieces<64>(X, S, jl, zl, d, p, K, kb, Y, t, P, ysc, ysplit, nsplit, f16);
}

// Block -> (unit, 256-point group), XCD-aware as glm_tile: the consecutive
// blocks one XCD receives take all point groups of one unit.
__device__ __forceinline__ bool gb_tile(unsigned nunits, unsigned groups, unsigned& u,
                                        unsigned& g) {
    const unsigned b = blockIdx.x;
    const unsigned idx = b >> 3;
    g = idx % groups;
    u = (idx / groups) * 8 + (b & 7);
    return u < nunits;
}

// MFMA contraction of the units of one row stride D (units: {problem, r0,
// r1, global unit}); partial[unit][point] as the single-posterior kernel
// (log2 units, |u|/2 + log2(1 + 2^-|u|) summed over the unit's rows).  Each
// wave holds the B fragments of CB 16-point column blocks (a block of four
// waves: 64 CB points); CB and the unit rows are the tile shape the batch's
// autotuner chooses per row stride.
template <int D, int NS, int CB, class T, class V>
__global__ void __launch_bounds__(256)
gb_mfma_kernel(const unsigned char* __restrict__ X8, const unsigned long long* __restrict__ xoff,
               const uint4* __restrict__ units, unsigned nunits, unsigned P,
               const T* __restrict__ ysplit, unsigned n, unsigned groups,
               double* __restrict__ partial) {
    constexpr int KC = D / 32;
    unsigned ui, g;
    if (!gb_tile(nunits, groups, ui, g)) return;
    const uint4 unit = units[ui];
    const unsigned p = unit.x;
    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const unsigned lr = lane & 15u, lg = lane >> 4;
    const unsigned w0 = g * (64u * CB) + wave * (16u * CB);
    V bfrag[CB][KC][NS];
#pragma unroll
    for (int wb = 0; wb < CB; ++wb)
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
#pragma unroll
            for (int s = 0; s < NS; ++s)
                bfrag[wb][kc][s] = *reinterpret_cast<const V*>(
                    ysplit + (((size_t)s * P + p) * n + w0 + wb * 16 + lr) * kGbD + kc * 32 +
                    lg * 8);
    double acc[CB];
#pragma unroll
    for (int wb = 0; wb < CB; ++wb) acc[wb] = 0.0;
    if (unit.y < unit.z)
        glm_mfma_rows<D, NS, CB>(Fp8Rows<D, V>{X8 + xoff[p], lr, lg}, bfrag, unit.y, unit.z, lg,
                                 acc);
#pragma unroll
    for (int wb = 0; wb < CB; ++wb) {
        double v = acc[wb];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (lg == 0) partial[(size_t)unit.w * n + w0 + wb * 16 + lr] = v;
    }
}

// Plain fp32 contraction: one point per thread, tiles of 32 rows decoded to
// fp32 in LDS (the scaled, log2(e)-multiplied coefficients in registers).
template <int D>
__global__ void __launch_bounds__(256)
gb_fp32_kernel(const unsigned char* __restrict__ X8, const unsigned long long* __restrict__ xoff,
               const uint4* __restrict__ units, unsigned nunits, unsigned P,
               const float* __restrict__ ysc, unsigned n, unsigned groups,
               double* __restrict__ partial) {
    constexpr int TR = 32;
    __shared__ float4 xs[TR * D / 4];
    unsigned ui, g;
    if (!gb_tile(nunits, groups, ui, g)) return;
    const uint4 unit = units[ui];
    const unsigned p = unit.x;
    const unsigned k = g * 256u + threadIdx.x;
    float y[D];
#pragma unroll
    for (int j = 0; j < D; ++j) y[j] = ysc[((size_t)p * n + k) * kGbD + j];
    const unsigned char* X = X8 + xoff[p];
    double acc = 0.0;
    for (unsigned rb = unit.y; rb < unit.z; rb += TR) {
        __syncthreads();
        for (unsigned e = threadIdx.x; e < TR * D / 4; e += 256) {
            const unsigned row = rb + (4 * e) / D;
            float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (row < unit.z) fp8x4_to_f32(*reinterpret_cast<const unsigned*>(X + (size_t)rb * D + 4 * e), f);
            xs[e] = make_float4(f[0], f[1], f[2], f[3]);
        }
        __syncthreads();
        // rows past the unit are zero: each adds exactly log2(2) = 1
        const int pad = (int)(unit.z - rb < TR ? TR - (unit.z - rb) : 0);
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
    partial[(size_t)unit.w * n + k] = acc;
}

// fp64 reference contraction: x_ij = s_j q_ij and eta = sum_j x_ij y_j in
// double, softplus(eta) = max(eta, 0) + log1p(exp(-|eta|)) (natural units).
template <int D>
__global__ void __launch_bounds__(256)
gb_fp64_kernel(const unsigned char* __restrict__ X8, const unsigned long long* __restrict__ xoff,
               const uint4* __restrict__ units, unsigned nunits, unsigned P,
               const float* __restrict__ scales, const float* __restrict__ Ypts, unsigned n,
               unsigned groups, double* __restrict__ partial) {
    constexpr int TR = 16;
    __shared__ double xs[TR * D];
    unsigned ui, g;
    if (!gb_tile(nunits, groups, ui, g)) return;
    const uint4 unit = units[ui];
    const unsigned p = unit.x;
    const unsigned k = g * 256u + threadIdx.x;
    double y[D];
#pragma unroll
    for (int j = 0; j < D; ++j) y[j] = (double)Ypts[((size_t)p * n + k) * kGbD + j];
    const unsigned char* X = X8 + xoff[p];
    const float* sc = scales + (size_t)p * kGbD;
    double acc = 0.0;
    for (unsigned rb = unit.y; rb < unit.z; rb += TR) {
        __syncthreads();
        for (unsigned e = threadIdx.x; e < TR * D / 4; e += 256) {
            const unsigned row = rb + (4 * e) / D;
            float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (row < unit.z) fp8x4_to_f32(*reinterpret_cast<const unsigned*>(X + (size_t)rb * D + 4 * e), f);
            const unsigned c = (4 * e) % D;
#pragma unroll
            for (int m = 0; m < 4; ++m) xs[4 * e + m] = (double)sc[c + m] * (double)f[m];
        }
        __syncthreads();
        const int rows = (int)(unit.z - rb < TR ? unit.z - rb : TR);
        for (int r = 0; r < rows; ++r) {
            double eta = 0.0;
#pragma unroll
            for (int j = 0; j < D; ++j) eta = __builtin_fma(xs[r * D + j], y[j], eta);
            acc += __builtin_fmax(eta, 0.0) + log1p(exp(-__builtin_fabs(eta)));
        }
    }
    partial[(size_t)unit.w * n + k] = acc;
}

// lp of point k of problem p from its units' partial sums (in unit order):
// lin.y - scale * s - |y|^2 / (2 s0^2), in double.
__device__ __forceinline__ double gb_logpost(const float* __restrict__ y, unsigned d,
                                             const double* __restrict__ partial, unsigned ub,
                                             unsigned ue, unsigned n, unsigned k,
                                             const double* __restrict__ lin, double scale,
                                             double prior_prec) {
    // the units' partial sums in unit order, loaded eight at a time so the
    // loads are in flight together
    double s = 0.0;
    unsigned u = ub;
    for (; u + 8 <= ue; u += 8) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = partial[(size_t)(u + e) * n + k];
#pragma unroll
        for (int e = 0; e < 8; ++e) s += v[e];
    }
    for (; u < ue; ++u) s += partial[(size_t)u * n + k];
    double l = 0.0, q = 0.0;
    for (unsigned i = 0; i < d; ++i) {
        l = __builtin_fma(lin[i], (double)y[i], l);
        q = __builtin_fma((double)y[i], (double)y[i], q);
    }
    return l - scale * s - 0.5 * prior_prec * q;
}

__global__ void __launch_bounds__(256)
gb_logpost_kernel(const float* __restrict__ xs, const unsigned* __restrict__ dims,
                  const unsigned* __restrict__ ubeg, const double* __restrict__ partial,
                  unsigned n, const double* __restrict__ lin, double scale, double prior_prec,
                  double* __restrict__ lp) {
    const unsigned p = blockIdx.y;
    const unsigned k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    lp[(size_t)p * n + k] = gb_logpost(xs + ((size_t)p * n + k) * kGbD, dims[p], partial, ubeg[p],
                                       ubeg[p + 1], n, k, lin + (size_t)p * kGbD, scale,
                                       prior_prec);
}

// The accept test and accounting of walker k of problem p (dimension
// d <= D), in 16-byte accesses over the first D slots of its proposal and
// position (the slots from d on are zeros in both, and so are the linear
// coefficients there).
// Sum over the wave for lane 0 in the butterfly's pairing (lane t adds
// lane t + i, i = 32, 16, 8, 4, 2, 1): v_permlane32_swap / v_permlane16_swap
// bring lanes t + 32 / t + 16 to the low lanes (the builtins' second result),
// DPP row_shl:i lane t + i within each 16-lane row.
__device__ __forceinline__ float gb_wave_sum_lane0(float x) {
    const unsigned u = __float_as_uint(x);
    x = x + __uint_as_float(__builtin_amdgcn_permlane32_swap(u, u, false, false)[1]);
    const unsigned u16 = __float_as_uint(x);
    x = x + __uint_as_float(__builtin_amdgcn_permlane16_swap(u16, u16, false, false)[1]);
    x = x + __uint_as_float(__builtin_amdgcn_update_dpp(0u, __float_as_uint(x), 0x108, 0xf, 0xf, false));
    x = x + __uint_as_float(__builtin_amdgcn_update_dpp(0u, __float_as_uint(x), 0x104, 0xf, 0xf, false));
    x = x + __uint_as_float(__builtin_amdgcn_update_dpp(0u, __float_as_uint(x), 0x102, 0xf, 0xf, false));
    x = x + __uint_as_float(__builtin_amdgcn_update_dpp(0u, __float_as_uint(x), 0x101, 0xf, 0xf, false));
    return x;
}

template <int D>
__device__ __forceinline__ void gb_accept_one(int half, unsigned p, unsigned k, unsigned d,
                                              const unsigned* __restrict__ ubeg,
                                              const float* __restrict__ Y,
                                              const float* __restrict__ zbuf,
                                              const float* __restrict__ u2buf,
                                              const double* __restrict__ partial,
                                              const double* __restrict__ lin, double scale,
                                              double prior_prec, float* __restrict__ xs,
                                              double* __restrict__ lp, float beta, unsigned K,
                                              int accu, unsigned long long* __restrict__ acc_cnt,
                                              float* __restrict__ msum,
                                              float* __restrict__ msum2) {
    __shared__ float red[2][4][kGbD];
    __shared__ unsigned cnt[4];
    __shared__ double lin_s[D];
    if (threadIdx.x < D) lin_s[threadIdx.x] = lin[(size_t)p * kGbD + threadIdx.x];
    __syncthreads();
    const float4* y4 = reinterpret_cast<const float4*>(Y + ((size_t)p * K + k) * kGbD);
    // lp(Y): the units' partial sums in unit order, eight loads in flight
    double s = 0.0;
    unsigned u = ubeg[p];
    const unsigned ue = ubeg[p + 1];
    for (; u + 8 <= ue; u += 8) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = partial[(size_t)(u + e) * K + k];
#pragma unroll
        for (int e = 0; e < 8; ++e) s += v[e];
    }
    for (; u < ue; ++u) s += partial[(size_t)u * K + k];
    double l = 0.0, q = 0.0;
#pragma unroll 4
    for (int c = 0; c < D / 4; ++c) {
        const float4 v = y4[c];
        const float a[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            l = __builtin_fma(lin_s[4 * c + e], (double)a[e], l);
            q = __builtin_fma((double)a[e], (double)a[e], q);
        }
    }
    const double lpy = l - scale * s - 0.5 * prior_prec * q;
    const size_t w = (size_t)p * 2 * K + (half ? K : 0) + k;
    float4* xk4 = reinterpret_cast<float4*>(xs + w * kGbD);
    const float z = zbuf[(size_t)p * K + k];
    float zp = 1.0f;
    for (unsigned i = 0; i + 1 < d; ++i) zp *= z;
    const float dl = (float)(lpy - lp[w]);
    const float qa = __builtin_isfinite(lpy) ? zp * expf(beta * dl) : 0.0f;
    const bool accepted = u2buf[(size_t)p * K + k] <= qa;
    if (accepted) {
#pragma unroll 4
        for (int c = 0; c < D / 4; ++c) xk4[c] = y4[c];
        lp[w] = lpy;
    }
    if (!accu) return;
    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const unsigned long long ball = __ballot(accepted);
    if (lane == 0) cnt[wave] = (unsigned)__popcll(ball);
    // the walker's coordinates after the move, four at a time, reduced over
    // the wave coordinate by coordinate
    const float4* src = accepted ? y4 : xk4;
    for (unsigned c = 0; 4 * c < d; ++c) {
        const float4 v = src[c];
        const float a[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float s1 = gb_wave_sum_lane0(a[e]);
            const float s2 = gb_wave_sum_lane0(a[e] * a[e]);
            if (lane == 0 && 4 * c + e < d) {
                red[0][wave][4 * c + e] = s1;
                red[1][wave][4 * c + e] = s2;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
        atomicAdd(&acc_cnt[p], (unsigned long long)((cnt[0] + cnt[1]) + (cnt[2] + cnt[3])));
    const unsigned nslots = 2 * gridDim.x;
    const unsigned slot = (half ? gridDim.x : 0u) + blockIdx.x;
    if (threadIdx.x < d) {
        const unsigned i = threadIdx.x;
        const size_t o = ((size_t)p * nslots + slot) * kGbD + i;
        msum[o] = (red[0][0][i] + red[0][1][i]) + (red[0][2][i] + red[0][3][i]);
        msum2[o] = (red[1][0][i] + red[1][1][i]) + (red[1][2][i] + red[1][3][i]);
    }
}

// Accept test of one half-step for all problems (amd-gcn-mcmc-stretch.cl:
// 66-78 with z^(d_p - 1) by repeated multiplication), and with accu != 0
// the accounting: accepted walkers per problem (acc_cnt), and per 256-walker
// block the sums of x and x^2 of every coordinate over the block's walkers
// after the move (msum/msum2[p][slot][64], slot = half * K/256 + block of the
// W/256 slots; a fixed shuffle-and-LDS tree).
__global__ void __launch_bounds__(256)
gb_accept_kernel(int half, const unsigned* __restrict__ dims, const unsigned* __restrict__ ubeg,
                 const float* __restrict__ Y, const float* __restrict__ zbuf,
                 const float* __restrict__ u2buf, const double* __restrict__ partial,
                 const double* __restrict__ lin, double scale, double prior_prec,
                 float* __restrict__ xs, double* __restrict__ lp, float beta, unsigned K, int accu,
                 unsigned long long* __restrict__ acc_cnt, float* __restrict__ msum,
                 float* __restrict__ msum2) {
    const unsigned p = blockIdx.y;
    const unsigned k = blockIdx.x * 256 + threadIdx.x;
    const unsigned d = dims[p];
    if (d <= 32)
        gb_accept_one<32>(half, p, k, d, ubeg, Y, zbuf, u2buf, partial, lin, scale, prior_prec,
                          xs, lp, beta, K, accu, acc_cnt, msum, msum2);
    else
        gb_accept_one<64>(half, p, k, d, ubeg, Y, zbuf, u2buf, partial, lin, scale, prior_prec,
                          xs, lp, beta, K, accu, acc_cnt, msum, msum2);
}

// End of an accounted step: per problem and coordinate, the step's
// ensemble sums (its W/256 slots, in slot order) into the double
// accumulators, and the step's ensemble mean into the recorded series
// (problem p's n x d_p block at series + soff[p]) when there is one.
__global__ void __launch_bounds__(64)
gb_step_kernel(const unsigned* __restrict__ dims, const float* __restrict__ msum,
               const float* __restrict__ msum2, unsigned W, double* __restrict__ acc1,
               double* __restrict__ acc2, const unsigned long long* __restrict__ soff,
               float* __restrict__ series, unsigned st) {
    const unsigned p = blockIdx.x;
    const unsigned i = threadIdx.x;
    if (i >= dims[p]) return;
    const unsigned nslots = W / 256;
    const float* m1 = msum + (size_t)p * nslots * kGbD + i;
    const float* m2 = msum2 + (size_t)p * nslots * kGbD + i;
    float s1 = 0.0f, s2 = 0.0f;
    for (unsigned sl = 0; sl < nslots; ++sl) {
        s1 += m1[(size_t)sl * kGbD];
        s2 += m2[(size_t)sl * kGbD];
    }
    acc1[(size_t)p * kGbD + i] += (double)s1;
    acc2[(size_t)p * kGbD + i] += (double)s2;
    if (series) series[soff[p] + (size_t)st * dims[p] + i] = s1 / (float)W;
}

// --------------------------------------------------------------------------
hipError_t launch_gb_fp8_decode(hipStream_t s, const void* in, float* out, unsigned long long n,
                                int f16) {
    const unsigned long long quads = n / 4;
    gb_fp8_decode_kernel<<<(unsigned)((quads + 255) / 256), 256, 0, s>>>(
        static_cast<const unsigned char*>(in), out, n, f16);
    return hipGetLastError();
}

hipError_t launch_gb_init(hipStream_t s, unsigned seed, const unsigned* dims, const float* limits,
                          float* xs, unsigned P, unsigned W) {
    const unsigned gx = (kGbD * W / 4 + 255) / 256;
    gb_init_kernel<<<dim3(gx, P), 256, 0, s>>>(seed, dims, limits, xs, W);
    return hipGetLastError();
}

hipError_t launch_gb_points(hipStream_t s, const float* xs, const unsigned* dims,
                            const float* tscale, unsigned P, unsigned n, float* ysc, void* ysplit,
                            int nsplit, int f16) {
    gb_points_kernel<<<dim3((n + 255) / 256, P), 256, 0, s>>>(xs, dims, tscale, P, n, ysc, ysplit,
                                                             nsplit, f16);
    return hipGetLastError();
}

hipError_t launch_gb_propose(hipStream_t s, unsigned seed0, int half, unsigned oe, unsigned step,
                             float a, const unsigned* dims, const float* tscale, const float* xs,
                             unsigned P, unsigned K, float* Y, float* z, float* u2, float* ysc,
                             void* ysplit, int nsplit, int f16) {
    gb_propose_kernel<<<dim3(K / 256, P), 256, 0, s>>>(seed0, half, oe, step, a, dims, tscale,
                                                       xs, P, K, Y, z, u2, ysc, ysplit, nsplit,
                                                       f16);
    return hipGetLastError();
}

hipError_t launch_gb_contract(int precision, int D, int CB, hipStream_t s, const void* X8,
                              const unsigned long long* xoff, const void* units, unsigned nunits,
                              unsigned P, const float* scales, const float* Ypts,
                              const float* ysc, const void* ysplit, unsigned n,
                              double* partial) {
    if (nunits == 0) return hipSuccess;
    // CB = 8 (512 points per block) only for the matrix-core kernels of
    // stride 32; the fp32 and fp64 kernels take 256 points per block
    const bool mfma = precision == kGlmBf16x3 || precision == kGlmFp16x2 || precision == kGlmBf16;
    if (CB != 4 && !(CB == 8 && D == 32 && mfma)) return hipErrorInvalidValue;
    const unsigned pts = mfma ? 64u * CB : 256u;
    if (n % pts) return hipErrorInvalidValue;
    const unsigned groups = n / pts;
    const unsigned blocks = (nunits + 7) / 8 * 8 * groups;
    const unsigned char* x = static_cast<const unsigned char*>(X8);
    const uint4* un = static_cast<const uint4*>(units);
    const __bf16* ys = static_cast<const __bf16*>(ysplit);
    const _Float16* ys16 = static_cast<const _Float16*>(ysplit);
#define BH_GB_MFMA(DD, CBV)                                                                     \
    switch (precision) {                                                                        \
    case kGlmBf16x3:                                                                            \
        gb_mfma_kernel<DD, 3, CBV, __bf16, bf16x8><<<blocks, 256, 0, s>>>(                      \
            x, xoff, un, nunits, P, ys, n, groups, partial);                                    \
        break;                                                                                  \
    case kGlmFp16x2:                                                                            \
        gb_mfma_kernel<DD, 2, CBV, _Float16, f16x8><<<blocks, 256, 0, s>>>(                     \
            x, xoff, un, nunits, P, ys16, n, groups, partial);                                  \
        break;                                                                                  \
    default:                                                                                    \
        gb_mfma_kernel<DD, 1, CBV, __bf16, bf16x8><<<blocks, 256, 0, s>>>(                      \
            x, xoff, un, nunits, P, ys, n, groups, partial);                                    \
        break;                                                                                  \
    }
    if (mfma && CB == 8) {
        BH_GB_MFMA(32, 8)
        return hipGetLastError();
    }
#define BH_GB_CASES(DD)                                                                         \
    switch (precision) {                                                                        \
    case kGlmBf16x3:                                                                            \
    case kGlmFp16x2:                                                                            \
    case kGlmBf16:                                                                              \
        BH_GB_MFMA(DD, 4)                                                                       \
        break;                                                                                  \
    case kGlmFp32:                                                                              \
        gb_fp32_kernel<DD><<<blocks, 256, 0, s>>>(x, xoff, un, nunits, P, ysc, n, groups,       \
                                                  partial);                                     \
        break;                                                                                  \
    case kGlmFp64:                                                                              \
        gb_fp64_kernel<DD><<<blocks, 256, 0, s>>>(x, xoff, un, nunits, P, scales, Ypts, n,      \
                                                  groups, partial);                             \
        break;                                                                                  \
    default:                                                                                    \
        return hipErrorInvalidValue;                                                            \
    }
    if (D == 64) {
        BH_GB_CASES(64)
    } else if (D == 32) {
        BH_GB_CASES(32)
    } else {
        return hipErrorInvalidValue;
    }
#undef BH_GB_CASES
#undef BH_GB_MFMA
    return hipGetLastError();
}

hipError_t launch_gb_logpost(hipStream_t s, const float* xs, const unsigned* dims,
                             const unsigned* ubeg, const double* partial, unsigned P, unsigned n,
                             const double* lin, double scale, double prior_prec, double* lp) {
    gb_logpost_kernel<<<dim3((n + 255) / 256, P), 256, 0, s>>>(xs, dims, ubeg, partial, n, lin,
                                                               scale, prior_prec, lp);
    return hipGetLastError();
}

hipError_t launch_gb_accept(hipStream_t s, int half, const unsigned* dims, const unsigned* ubeg,
                            const float* Y, const float* z, const float* u2,
                            const double* partial, const double* lin, double scale,
                            double prior_prec, float* xs, double* lp, float beta, unsigned P,
                            unsigned K, int accu, unsigned long long* acc_cnt, float* msum,
                            float* msum2) {
    gb_accept_kernel<<<dim3(K / 256, P), 256, 0, s>>>(half, dims, ubeg, Y, z, u2, partial, lin,
                                                      scale, prior_prec, xs, lp, beta, K, accu,
                                                      acc_cnt, msum, msum2);
    return hipGetLastError();
}

hipError_t launch_gb_step(hipStream_t s, const unsigned* dims, const float* msum,
                          const float* msum2, unsigned P, unsigned W, double* acc1, double* acc2,
                          const unsigned long long* soff, float* series, unsigned st) {
    gb_step_kernel<<<P, 64, 0, s>>>(dims, msum, msum2, W, acc1, acc2, soff, series, st);
    return hipGetLastError();
}

}  // namespace bh
