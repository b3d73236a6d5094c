// Model-dependent kernels of the HIP engine, compiled at run time by hipRTC
// together with the user's model source (the reference compiles its OpenCL
// engine the same way: amd_gcn.clj:573-631, with -DLOGFN/-DLOGPDF/-DLOGLIK,
// -DDIM and -DWGS).  The host defines:
//   WGS     reference work-group size: accounting groups and block size
//   DIM     model dimension (stretch kernels only)
//   LOGFN   mcmc log-density  REAL f(uint data_len, uint params_len, const REAL* params, uint dim, const REAL* x)
//   LOGPDF  log-density       (same signature)
//   LOGLIK  log-likelihood    REAL f(uint data_len, const REAL* data, uint dim, const REAL* x)
//
// Walker state is AoS, x[walker*DIM + d], the two half-ensembles contiguous,
// exactly the layout the reference hands across its boundary (DIM x W
// column-major).  The stretch move reads the partner walker's DIM floats as
// one contiguous run, so a random partner costs one cache line rather than
// DIM of them.

#include "bh_philox.hpp"

#define BH_PARTNER_FLAT 0
#define BH_PARTNER_WALKER 1

// The last six levels of the work-group tree inside one wave, for lane 0:
// level i adds lane t + i to lane t (i = 32, 16, 8, 4, 2, 1), the pairing
// of the reference's tree; v_permlane32_swap / v_permlane16_swap bring
// lanes t + 32 / t + 16 to lanes t < 32 / t < 16 (the builtins' second
// result), DPP row_shl:i lanes t + i within each 16-lane row.
__device__ __forceinline__ float bh_wave_tree_sum(float x) {
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

// Sum v over each aligned slice of WGS threads with the reference's
// work-group tree (amd-gcn-mcmc-stretch.cl:5-29 for power-of-two WGS: level
// i adds item l+i into item l, i = WGS/2 ... 1).  Float addition is not
// associative, so the tree shape is kept exactly: the result is
// bit-identical to the reference's.  Values are staged in
// buf[d*blockDim + t]; after one barrier lane l < 64 of the slice reads
// items l + 64j (j < WGS/64) and applies the levels i >= 64 to them in the
// tree's pairing (item l + 64j gets item l + 64(j + i/64)), then the slice's
// first wave applies levels 32 ... 1 with cross-lane moves.  The total of
// dim d is returned in out[d] by the slice's lane 0.  With count non-null
// the slice's set flags are counted too (ballot counts staged with the
// values; order-independent integers), returned in *count by lane 0.
// tail_sync: a barrier at the end, before the caller writes buf or
// wave_counts again (not needed when a barrier or the kernel's end comes
// first).
template <int ND>
__device__ __forceinline__ void bh_slice_tree(const float (&v)[ND], float* buf, float (&out)[ND],
                                              bool flag, unsigned* wave_counts,
                                              unsigned* count, bool tail_sync) {
    static_assert(WGS >= 64 && (WGS & (WGS - 1)) == 0, "WGS: a power of two >= 64");
    constexpr unsigned R = WGS / 64;
    const unsigned t = threadIdx.x;
    const unsigned nb = blockDim.x;
    const unsigned l = t & (WGS - 1);
    if (count) {
        const unsigned long long m = __ballot(flag);
        if ((t & 63u) == 0) wave_counts[t >> 6] = (unsigned)__popcll(m);
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) buf[d * nb + t] = v[d];
    __syncthreads();
    if (l < 64) {
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            float x[R];
#pragma unroll
            for (unsigned j = 0; j < R; ++j) x[j] = buf[d * nb + t + 64 * j];
#pragma unroll
            for (unsigned h = R / 2; h >= 1; h >>= 1)
#pragma unroll
                for (unsigned j = 0; j < h; ++j) x[j] = x[j] + x[j + h];
            out[d] = bh_wave_tree_sum(x[0]);
        }
        if (count && l == 0) {
            unsigned total = 0;
            const unsigned w0 = t >> 6;
            for (unsigned w = 0; w < R; ++w) total += wave_counts[w0 + w];
            *count = total;
        }
    }
    if (tail_sync) __syncthreads();
}

template <int ND>
__device__ __forceinline__ void bh_slice_tree_sums(const float (&v)[ND], float* buf,
                                                   float (&out)[ND], bool tail_sync = true) {
    bh_slice_tree<ND>(v, buf, out, false, nullptr, nullptr, tail_sync);
}

template <int ND>
__device__ __forceinline__ void bh_slice_count_tree_sums(bool flag, const float (&v)[ND],
                                                         float* buf, unsigned* wave_counts,
                                                         float (&out)[ND], unsigned& count,
                                                         bool tail_sync = true) {
    count = 0;
    bh_slice_tree<ND>(v, buf, out, flag, wave_counts, &count, tail_sync);
}

#if defined(LOGFN) || defined(BH_GLM)

// A walker's DIM coordinates.  Walkers start at multiples of DIM floats from
// 256-byte aligned bases (and half-ensembles at K*DIM floats, K a multiple of
// 256), so for DIM % 4 == 0 they are 16-byte aligned: one 128-bit access per
// four coordinates.
__device__ __forceinline__ void bh_load_walker(const float* p, float (&v)[DIM]) {
    if constexpr (DIM % 4 == 0) {
        const float4* q = reinterpret_cast<const float4*>(__builtin_assume_aligned(p, 16));
#pragma unroll
        for (int i = 0; i < DIM / 4; ++i) {
            const float4 t = q[i];
            v[4 * i] = t.x;
            v[4 * i + 1] = t.y;
            v[4 * i + 2] = t.z;
            v[4 * i + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < DIM; ++i) v[i] = p[i];
    }
}

__device__ __forceinline__ void bh_store_walker(float* p, const float (&v)[DIM]) {
    if constexpr (DIM % 4 == 0) {
        float4* q = reinterpret_cast<float4*>(__builtin_assume_aligned(p, 16));
#pragma unroll
        for (int i = 0; i < DIM / 4; ++i)
            q[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < DIM; ++i) p[i] = v[i];
    }
}

// The draws of the stretch move (amd-gcn-mcmc-stretch.cl:46-57):
//   counter {k, step, odd_or_even, 0xbeeff00d}, key {seed, 0xdecafbad}
//   z  = fma((a-2+1/a)*u1, u1, (2(1-1/a))*u1) + 1/a                     (:53-54)
// Partner: FLAT reads the complementary half at element offset
// j0 = (uint)((u0*K)*DIM) (:57), wrapping modulo K*DIM where the reference
// reads out of bounds; WALKER reads walker min((uint)(u0*K), K-1), i.e.
// element offset j*DIM.  Sets z and the partner's first element offset
// (< K*DIM); returns u2, the uniform of the accept test.
template <int PARTNER>
__device__ __forceinline__ float bh_stretch_draw(unsigned k, unsigned K, unsigned seed,
                                                 unsigned odd_or_even, unsigned step, float a,
                                                 float& z, unsigned& start) {
    const bh_f32x4 u = bh_u01x4(
        bh_philox4x32_10(bh_u32x4{k, step, odd_or_even, BH_CTR3}, seed, BH_KEY1_STRETCH));
    const float inv_a = 1.0f / a;
    const float ca = a - 2.0f + inv_a;
    const float cb = 2.0f * (1.0f - inv_a);
    z = __builtin_fmaf(ca * u.y, u.y, cb * u.y) + inv_a;
    if constexpr (PARTNER == BH_PARTNER_FLAT) {
        const unsigned total = K * DIM;
        unsigned e = (unsigned)((u.x * (float)K) * (float)DIM);
        if (e >= total) e -= total;
        start = e;
    } else {
        unsigned j = (unsigned)(u.x * (float)K);
        j = j < K ? j : K - 1;
        start = j * DIM;
    }
    return u.z;
}

// The proposal half of the stretch move (amd-gcn-mcmc-stretch.cl:46-65):
// the draws above, then Y = fma(z, Xk - Xj, Xj) (:64) with the partner's
// coordinates read from S (FLAT: element i at (start + i) mod K*DIM).
// Returns u2, the uniform of the accept test.
template <int PARTNER>
__device__ __forceinline__ float bh_propose(unsigned k, unsigned K, unsigned seed,
                                           unsigned odd_or_even, unsigned step, float a,
                                           const float* S, const float (&xk)[DIM],
                                           float (&y)[DIM], float& z) {
    unsigned e;
    const float u2 = bh_stretch_draw<PARTNER>(k, K, seed, odd_or_even, step, a, z, e);
    if constexpr (PARTNER == BH_PARTNER_FLAT) {
        const unsigned total = K * DIM;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            const float xj = S[e];
            y[i] = __builtin_fmaf(z, xk[i] - xj, xj);
            e = (e + 1 == total) ? 0u : e + 1;
        }
    } else {
        float sj[DIM];
        bh_load_walker(S + e, sj);
#pragma unroll
        for (int i = 0; i < DIM; ++i) y[i] = __builtin_fmaf(z, xk[i] - sj[i], sj[i]);
    }
    return u2;
}

// z^(DIM-1) by repeated multiplication
__device__ __forceinline__ float bh_zpow(float z) {
    float zp = 1.0f;
#pragma unroll
    for (int i = 0; i < DIM - 1; ++i) zp *= z;
    return zp;
}

// Accounting of stretch_move_accu (amd-gcn-mcmc-stretch.cl:98-114) for the
// slice of WGS threads holding group g: accept[g] += accepted count;
// means[groups*DIM*step + i*groups + g] += tree sum of coordinate i.
// Dimensions are reduced BH_ACCU_CHUNK at a time so the LDS staging area
// stays at BH_ACCU_CHUNK floats per thread whatever DIM is.
#define BH_ACCU_CHUNK (DIM < 8 ? DIM : 8)

__device__ __forceinline__ void bh_accu(const float (&xk)[DIM], bool accepted, unsigned g,
                                        unsigned groups, unsigned mstep, unsigned* accept,
                                        float* means, float* red, unsigned* wave_counts) {
    const bool leader = (threadIdx.x & (WGS - 1)) == 0;
    float* m = means + (size_t)groups * DIM * mstep + g;
#pragma unroll
    for (int d0 = 0; d0 < DIM; d0 += BH_ACCU_CHUNK) {
        float v[BH_ACCU_CHUNK], sums[BH_ACCU_CHUNK];
#pragma unroll
        for (int c = 0; c < BH_ACCU_CHUNK; ++c) v[c] = (d0 + c < DIM) ? xk[d0 + c] : 0.0f;
        // every caller ends its kernel here: no barrier after the last chunk
        const bool more = d0 + BH_ACCU_CHUNK < DIM;
        if (d0 == 0) {
            unsigned cnt;
            bh_slice_count_tree_sums<BH_ACCU_CHUNK>(accepted, v, red, wave_counts, sums, cnt,
                                                    more);
            if (leader) accept[g] += cnt;
        } else {
            bh_slice_tree_sums<BH_ACCU_CHUNK>(v, red, sums, more);
        }
        if (leader) {
#pragma unroll
            for (int c = 0; c < BH_ACCU_CHUNK; ++c)
                if (d0 + c < DIM) m[(size_t)(d0 + c) * groups] += sums[c];
        }
    }
}

// The accounting's LDS staging, one area for every instantiation of a
// half-step in a kernel.
struct BhAccuLds {
    float red[BH_ACCU_CHUNK * WGS];
    unsigned wave_counts[WGS / 64];
};

__device__ __forceinline__ BhAccuLds& bh_accu_lds() {
    __shared__ BhAccuLds st;
    return st;
}

#endif  // LOGFN || BH_GLM

#ifdef LOGFN

// One walker's stretch move (amd-gcn-mcmc-stretch.cl:31-81).
//   counter {k, step, odd_or_even, 0xbeeff00d}, key {seed, 0xdecafbad}
//   z  = fma((a-2+1/a)*u1, u1, (2(1-1/a))*u1) + 1/a                     (:53-54)
//   Y  = fma(z, Xk - Xj, Xj)                                           (:64)
//   q  = isfinite(ly) ? z^(DIM-1) exp(beta (ly - lX)) : 0; accept iff u2 <= q (:68-71)
// Partner: FLAT reads the complementary half at element offset
// j0 = (uint)((u0*K)*DIM) (:57), wrapping modulo K*DIM where the reference
// reads out of bounds; WALKER reads walker min((uint)(u0*K), K-1).
// xk holds the walker's coordinates on entry and its new ones on exit.
template <int PARTNER>
__device__ __forceinline__ bool bh_stretch_walker(unsigned k, unsigned K, unsigned seed,
                                                  unsigned odd_or_even, unsigned step, float a,
                                                  float beta, const float* S,
                                                  float (&xk)[DIM], float& lx,
                                                  unsigned data_len, unsigned params_len,
                                                  const float* params) {
    float y[DIM];
    float z;
    const float u2 = bh_propose<PARTNER>(k, K, seed, odd_or_even, step, a, S, xk, y, z);
    const float ly = LOGFN(data_len, params_len, params, DIM, y);
    const float q = __builtin_isfinite(ly) ? bh_zpow(z) * expf(beta * (ly - lx)) : 0.0f;
    const bool accepted = u2 <= q;
    if (accepted) {
#pragma unroll
        for (int i = 0; i < DIM; ++i) xk[i] = y[i];
        lx = ly;
    }
    return accepted;
}

// logfn kernel (amd-gcn-mcmc-stretch.cl:161-167): res[w] = LOGFN(x + DIM*w).
extern "C" __global__ void __launch_bounds__(WGS)
bh_logfn(unsigned data_len, unsigned params_len, const float* __restrict__ params,
         const float* __restrict__ x, float* __restrict__ res, unsigned n) {
    const unsigned w = blockIdx.x * WGS + threadIdx.x;
    if (w < n) res[w] = LOGFN(data_len, params_len, params, DIM, x + (size_t)w * DIM);
}

// The same accounting inside the resident kernel, without global
// read-modify-writes in the step loop: a slice's accept count goes to LDS
// (acc_lds[g], added to accept[g] once at the end of the launch), the odd
// half's per-group sums wait in LDS (odd_lds[g][d]) and the even half stores
// (0 + odd) + even to the step's means slot -- the value the two
// read-modify-writes of the per-half-step kernels leave in a zeroed slot.
__device__ __forceinline__ void bh_accu_resident(const float (&xk)[DIM], bool accepted, unsigned g,
                                                 unsigned groups, unsigned mstep, int half,
                                                 unsigned* acc_lds, float* odd_lds, float* means,
                                                 float* red, unsigned* wave_counts,
                                                 bool last_pass) {
    const bool leader = (threadIdx.x & (WGS - 1)) == 0;
    float* m = means + (size_t)groups * DIM * mstep + g;
#pragma unroll
    for (int d0 = 0; d0 < DIM; d0 += BH_ACCU_CHUNK) {
        float v[BH_ACCU_CHUNK], sums[BH_ACCU_CHUNK];
#pragma unroll
        for (int c = 0; c < BH_ACCU_CHUNK; ++c) v[c] = (d0 + c < DIM) ? xk[d0 + c] : 0.0f;
        // the half-step's own barrier follows its last pass over walkers
        const bool sync = d0 + BH_ACCU_CHUNK < DIM || !last_pass;
        if (d0 == 0) {
            unsigned cnt;
            bh_slice_count_tree_sums<BH_ACCU_CHUNK>(accepted, v, red, wave_counts, sums, cnt,
                                                    sync);
            if (leader) acc_lds[g] += cnt;
        } else {
            bh_slice_tree_sums<BH_ACCU_CHUNK>(v, red, sums, sync);
        }
        if (leader) {
#pragma unroll
            for (int c = 0; c < BH_ACCU_CHUNK; ++c) {
                if (d0 + c < DIM) {
                    float* o = odd_lds + (size_t)g * DIM + d0 + c;
                    if (half == 0) *o = sums[c];
                    else m[(size_t)(d0 + c) * groups] = (0.0f + *o) + sums[c];
                }
            }
        }
    }
}

// One half-step over K walkers, one thread per walker, WGS walkers per block
// (stretch_move_bare / stretch_move_accu, amd-gcn-mcmc-stretch.cl:83-134).
// The walker's own record and log density are read and written once per
// half-step: streamed with non-temporal loads and stores, so that the
// caches keep the complementary half, which is read at random.
__device__ __forceinline__ void bh_load_walker_nt(const float* p, float (&v)[DIM]) {
    if constexpr (DIM % 4 == 0) {
        typedef float bh_f4 __attribute__((ext_vector_type(4)));
        const bh_f4* q = reinterpret_cast<const bh_f4*>(__builtin_assume_aligned(p, 16));
#pragma unroll
        for (int i = 0; i < DIM / 4; ++i) {
            const bh_f4 t = __builtin_nontemporal_load(q + i);
            v[4 * i] = t.x;
            v[4 * i + 1] = t.y;
            v[4 * i + 2] = t.z;
            v[4 * i + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < DIM; ++i) v[i] = __builtin_nontemporal_load(p + i);
    }
}

__device__ __forceinline__ void bh_store_walker_nt(float* p, const float (&v)[DIM]) {
    if constexpr (DIM % 4 == 0) {
        typedef float bh_f4 __attribute__((ext_vector_type(4)));
        bh_f4* q = reinterpret_cast<bh_f4*>(__builtin_assume_aligned(p, 16));
#pragma unroll
        for (int i = 0; i < DIM / 4; ++i)
            __builtin_nontemporal_store(bh_f4{v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]},
                                        q + i);
    } else {
#pragma unroll
        for (int i = 0; i < DIM; ++i) __builtin_nontemporal_store(v[i], p + i);
    }
}

// Staged half-step for DIM >= 8: a thread's own record and its partner's
// span DIM/4 16-byte pieces (DIM floats when DIM % 4 != 0), and reading them
// thread by thread makes every load instruction of a wave touch up to 64
// separate lines (4 to 16 times the requests of a contiguous read).  Here
// the block's WGS records, contiguous in X, are read and written as
// consecutive pieces by consecutive threads, and each partner's pieces by
// consecutive threads too; an LDS tile transposes between that order and
// one record per thread.  Rows of the tile are an odd number of 16-byte
// units (odd number of floats for DIM % 4 != 0) apart, so threads reading
// consecutive rows hit distinct banks.  Arithmetic and results are those of
// the thread-per-record path.
constexpr int bh_stage_stride() {
    return DIM % 4 == 0 ? ((DIM / 4) % 2 ? DIM : DIM + 4) : (DIM % 2 ? DIM : DIM + 1);
}
constexpr int kBhStageStride = bh_stage_stride();
constexpr bool kBhStage = DIM >= 8 && WGS * kBhStageStride * 4 <= 49152;

// One LDS staging area for every instantiation of the staged half-step in a
// kernel (a __shared__ array declared in each instantiation would be
// allocated once per instantiation, halving the blocks that fit on a CU).
struct BhStageLds {
    float tile[kBhStage ? WGS * kBhStageStride : 1];
    unsigned pstart[WGS];
    unsigned long long acc_mask[WGS / 64];
};

__device__ __forceinline__ BhStageLds& bh_stage_lds() {
    __shared__ __attribute__((aligned(16))) BhStageLds st;
    return st;
}

template <bool ACCU, int PARTNER>
__device__ __forceinline__ void bh_stretch_half_staged(unsigned seed, unsigned odd_or_even,
                                                       unsigned data_len, unsigned params_len,
                                                       const float* __restrict__ params,
                                                       const float* __restrict__ S,
                                                       float* __restrict__ X,
                                                       float* __restrict__ logfn_X, float a,
                                                       float beta, unsigned step, unsigned K,
                                                       unsigned* __restrict__ accept,
                                                       float* __restrict__ means) {
    typedef float bh_f4 __attribute__((ext_vector_type(4)));
    constexpr int ST = kBhStageStride;
    constexpr bool V4 = DIM % 4 == 0;                     // records as 16-byte pieces
    constexpr bool SV4 = V4 && PARTNER == BH_PARTNER_WALKER;  // partners too
    constexpr int NP = V4 ? DIM / 4 : DIM;                 // pieces per record
    BhStageLds& lds = bh_stage_lds();
    float* tile = lds.tile;
    unsigned* pstart = lds.pstart;
    unsigned long long* acc_mask = lds.acc_mask;
    const unsigned t = threadIdx.x;
    const unsigned k = blockIdx.x * WGS + t;
    float* xb = X + (size_t)blockIdx.x * WGS * DIM;
    // own records: piece g = i*WGS + t is piece g % NP of record g / NP
    float xs[NP * (V4 ? 4 : 1)];
    if constexpr (V4) {
        const bh_f4* q = reinterpret_cast<const bh_f4*>(__builtin_assume_aligned(xb, 16));
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const bh_f4 v = __builtin_nontemporal_load(q + i * WGS + t);
            xs[4 * i] = v.x;
            xs[4 * i + 1] = v.y;
            xs[4 * i + 2] = v.z;
            xs[4 * i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < NP; ++i) xs[i] = __builtin_nontemporal_load(xb + i * WGS + t);
    }
    float lx = __builtin_nontemporal_load(logfn_X + k);
    float z;
    unsigned e;
    const float u2 = bh_stretch_draw<PARTNER>(k, K, seed, odd_or_even, step, a, z, e);
    pstart[t] = e;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const unsigned g = i * WGS + t, w = g / NP, c = g % NP;
        if constexpr (V4) {
            *reinterpret_cast<bh_f4*>(&tile[w * ST + 4 * c]) =
                bh_f4{xs[4 * i], xs[4 * i + 1], xs[4 * i + 2], xs[4 * i + 3]};
        } else {
            tile[w * ST + c] = xs[i];
        }
    }
    __syncthreads();
    // partner pieces, consecutive threads on consecutive pieces of a partner
    constexpr int NPS = SV4 ? DIM / 4 : DIM;
    float ss[DIM];
    if constexpr (SV4) {
        const bh_f4* q = reinterpret_cast<const bh_f4*>(__builtin_assume_aligned(S, 16));
#pragma unroll
        for (int i = 0; i < NPS; ++i) {
            const unsigned g = i * WGS + t, w = g / NPS, c = g % NPS;
            const bh_f4 v = q[pstart[w] / 4 + c];
            ss[4 * i] = v.x;
            ss[4 * i + 1] = v.y;
            ss[4 * i + 2] = v.z;
            ss[4 * i + 3] = v.w;
        }
    } else {
        const unsigned total = K * DIM;
#pragma unroll
        for (int i = 0; i < NPS; ++i) {
            const unsigned g = i * WGS + t, w = g / NPS, c = g % NPS;
            unsigned idx = pstart[w] + c;
            if (PARTNER == BH_PARTNER_FLAT && idx >= total) idx -= total;
            ss[i] = S[idx];
        }
    }
    float xk[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) xk[i] = tile[t * ST + i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NPS; ++i) {
        const unsigned g = i * WGS + t, w = g / NPS, c = g % NPS;
        if constexpr (SV4) {
            *reinterpret_cast<bh_f4*>(&tile[w * ST + 4 * c]) =
                bh_f4{ss[4 * i], ss[4 * i + 1], ss[4 * i + 2], ss[4 * i + 3]};
        } else {
            tile[w * ST + c] = ss[i];
        }
    }
    __syncthreads();
    float y[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        const float xj = tile[t * ST + i];
        y[i] = __builtin_fmaf(z, xk[i] - xj, xj);
    }
    const float ly = LOGFN(data_len, params_len, params, DIM, y);
    const float qv = __builtin_isfinite(ly) ? bh_zpow(z) * expf(beta * (ly - lx)) : 0.0f;
    const bool accepted = u2 <= qv;
    if (accepted) {
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            xk[i] = y[i];
            tile[t * ST + i] = y[i];
        }
        __builtin_nontemporal_store(ly, logfn_X + k);
    }
    const unsigned long long m = __ballot(accepted);
    if ((t & 63u) == 0) acc_mask[t >> 6] = m;
    __syncthreads();
    // accepted records back to X as consecutive pieces
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const unsigned g = i * WGS + t, w = g / NP, c = g % NP;
        if ((acc_mask[w >> 6] >> (w & 63u)) & 1ull) {
            if constexpr (V4) {
                bh_f4* q = reinterpret_cast<bh_f4*>(__builtin_assume_aligned(xb, 16));
                __builtin_nontemporal_store(*reinterpret_cast<const bh_f4*>(&tile[w * ST + 4 * c]),
                                            q + g);
            } else {
                __builtin_nontemporal_store(tile[w * ST + c], xb + g);
            }
        }
    }
    if (ACCU) {
        BhAccuLds& al = bh_accu_lds();
        bh_accu(xk, accepted, blockIdx.x, gridDim.x, step, accept, means, al.red,
                al.wave_counts);
    }
}

template <bool ACCU, int PARTNER>
__device__ __forceinline__ void bh_stretch_half(unsigned seed, unsigned odd_or_even,
                                                unsigned data_len, unsigned params_len,
                                                const float* __restrict__ params,
                                                const float* __restrict__ S, float* __restrict__ X,
                                                float* __restrict__ logfn_X, float a, float beta,
                                                unsigned step, unsigned K,
                                                unsigned* __restrict__ accept,
                                                float* __restrict__ means) {
    if constexpr (kBhStage) {
        bh_stretch_half_staged<ACCU, PARTNER>(seed, odd_or_even, data_len, params_len, params, S,
                                              X, logfn_X, a, beta, step, K, accept, means);
        return;
    } else {
        const unsigned k = blockIdx.x * WGS + threadIdx.x;
        float xk[DIM];
        float* xp = X + (size_t)k * DIM;
        bh_load_walker_nt(xp, xk);
        float lx = __builtin_nontemporal_load(logfn_X + k);
        const bool accepted = bh_stretch_walker<PARTNER>(k, K, seed, odd_or_even, step, a, beta, S,
                                                         xk, lx, data_len, params_len, params);
        if (accepted) {
            bh_store_walker_nt(xp, xk);
            __builtin_nontemporal_store(lx, logfn_X + k);
        }
        if (ACCU) {
            BhAccuLds& al = bh_accu_lds();
            bh_accu(xk, accepted, blockIdx.x, gridDim.x, step, accept, means, al.red,
                    al.wave_counts);
        }
    }
}

extern "C" __global__ void __launch_bounds__(WGS)
bh_stretch_move_bare(unsigned seed, unsigned odd_or_even, unsigned data_len, unsigned params_len,
                     const float* __restrict__ params, const float* __restrict__ S,
                     float* __restrict__ X, float* __restrict__ logfn_X, float a, float beta,
                     unsigned step, unsigned K, int partner) {
    if (partner == BH_PARTNER_FLAT)
        bh_stretch_half<false, BH_PARTNER_FLAT>(seed, odd_or_even, data_len, params_len, params, S,
                                                X, logfn_X, a, beta, step, K, nullptr, nullptr);
    else
        bh_stretch_half<false, BH_PARTNER_WALKER>(seed, odd_or_even, data_len, params_len, params,
                                                  S, X, logfn_X, a, beta, step, K, nullptr,
                                                  nullptr);
}

extern "C" __global__ void __launch_bounds__(WGS)
bh_stretch_move_accu(unsigned seed, unsigned odd_or_even, unsigned data_len,
                     unsigned params_len, const float* __restrict__ params,
                     const float* __restrict__ S, float* __restrict__ X,
                     float* __restrict__ logfn_X, unsigned* __restrict__ accept,
                     float* __restrict__ means, float a, unsigned step, unsigned K,
                     int partner) {
    if (partner == BH_PARTNER_FLAT)
        bh_stretch_half<true, BH_PARTNER_FLAT>(seed, odd_or_even, data_len, params_len, params, S,
                                               X, logfn_X, a, 1.0f, step, K, accept, means);
    else
        bh_stretch_half<true, BH_PARTNER_WALKER>(seed, odd_or_even, data_len, params_len, params,
                                                 S, X, logfn_X, a, 1.0f, step, K, accept, means);
}

// Resident multi-step kernel: one work-group holds the whole ensemble
// (positions and logfn values) in LDS and runs nsteps full steps -- both
// half-steps each -- without returning to the host.  Every walker's
// arithmetic and every accounting tree is the same as in the per-half-step
// kernels, so the results are bit-identical; what goes away is two kernel
// launches per step, which bound small ensembles (W = 1024: 16 KB of state).
//   mode 0: bare steps (burn-in!, anneal! with betas[s], sample! with out)
//   mode 1: accounting steps (move!): accept/means as stretch_move_accu
// blockDim.x is a multiple of WGS dividing K.
template <int PARTNER>
__device__ __forceinline__ void bh_resident_impl(
    unsigned seed_odd, unsigned seed_even, unsigned c_odd, unsigned c_even, unsigned data_len,
    unsigned params_len, const float* __restrict__ params, float* __restrict__ xs,
    float* __restrict__ logfn, unsigned K, float a, float beta, const float* __restrict__ betas,
    unsigned step0, unsigned nsteps, int mode, unsigned* __restrict__ accept,
    float* __restrict__ means, float* __restrict__ sample_out, unsigned long long sample_floats) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const unsigned W = 2 * K;
    const unsigned nb = blockDim.x;
    const unsigned t = threadIdx.x;
    float* sx = lds;                       // W*DIM
    float* sl = sx + (size_t)W * DIM;      // W
    float* red = sl + W;                   // BH_ACCU_CHUNK*nb (mode 1)
    const unsigned groups = K / WGS;
    float* odd_lds = red + (mode == 1 ? BH_ACCU_CHUNK * nb : 0);       // groups*DIM (mode 1)
    unsigned* acc_lds = (unsigned*)(odd_lds + (mode == 1 ? groups * DIM : 0));  // groups
    unsigned* wave_counts = acc_lds + (mode == 1 ? groups : 0);
    for (unsigned i = t; i < W * DIM; i += nb) sx[i] = xs[i];
    for (unsigned i = t; i < W; i += nb) sl[i] = logfn[i];
    if (mode == 1)
        for (unsigned i = t; i < groups; i += nb) acc_lds[i] = 0u;
    __syncthreads();
    unsigned long long written = 0;
    for (unsigned s = 0; s < nsteps; ++s) {
        const unsigned step = step0 + s;
        const float b = betas ? betas[s] : beta;
        for (int h = 0; h < 2; ++h) {
            float* X = sx + (h ? (size_t)K * DIM : 0);
            const float* S = sx + (h ? 0 : (size_t)K * DIM);
            float* L = sl + (h ? K : 0);
            const unsigned seed = h ? seed_even : seed_odd;
            const unsigned oe = h ? c_even : c_odd;
            for (unsigned k0 = 0; k0 < K; k0 += nb) {
                const unsigned k = k0 + t;
                float xk[DIM];
#pragma unroll
                for (int i = 0; i < DIM; ++i) xk[i] = X[(size_t)k * DIM + i];
                float lx = L[k];
                const bool acc = bh_stretch_walker<PARTNER>(k, K, seed, oe, step, a,
                                                            mode == 1 ? 1.0f : b, S, xk, lx,
                                                            data_len, params_len, params);
                if (acc) {
#pragma unroll
                    for (int i = 0; i < DIM; ++i) X[(size_t)k * DIM + i] = xk[i];
                    L[k] = lx;
                }
                if (mode == 1)
                    bh_accu_resident(xk, acc, k / WGS, groups, step, h, acc_lds, odd_lds, means,
                                     red, wave_counts, k0 + nb >= K);
            }
            __syncthreads();
        }
        if (sample_out && written < sample_floats) {
            const unsigned long long left = sample_floats - written;
            const unsigned chunk = left < (unsigned long long)W * DIM ? (unsigned)left : W * DIM;
            float* dst = sample_out + written;
            if (((unsigned long long)dst & 15ull) == 0 && chunk % 4 == 0) {
                // 16-byte pieces (the LDS ensemble starts 16-byte aligned)
                const float4* src4 = reinterpret_cast<const float4*>(sx);
                float4* dst4 = reinterpret_cast<float4*>(dst);
                for (unsigned i = t; i < chunk / 4; i += nb) dst4[i] = src4[i];
            } else {
                for (unsigned i = t; i < chunk; i += nb) dst[i] = sx[i];
            }
            written += chunk;
            // the next step's first half writes accepted proposals into sx:
            // every wave finishes copying this step's ensemble first (the
            // condition is uniform over the work-group)
            __syncthreads();
        }
    }
    for (unsigned i = t; i < W * DIM; i += nb) xs[i] = sx[i];
    for (unsigned i = t; i < W; i += nb) logfn[i] = sl[i];
    if (mode == 1) {
        __syncthreads();
        for (unsigned i = t; i < groups; i += nb) accept[i] += acc_lds[i];
    }
}

extern "C" __global__ void __launch_bounds__(1024)
bh_stretch_resident(unsigned seed_odd, unsigned seed_even, unsigned c_odd, unsigned c_even,
                    unsigned data_len, unsigned params_len, const float* __restrict__ params,
                    float* __restrict__ xs, float* __restrict__ logfn, unsigned K, float a,
                    float beta, const float* __restrict__ betas, unsigned step0,
                    unsigned nsteps, int partner, int mode, unsigned* __restrict__ accept,
                    float* __restrict__ means, float* __restrict__ sample_out,
                    unsigned long long sample_floats) {
    if (partner == BH_PARTNER_FLAT)
        bh_resident_impl<BH_PARTNER_FLAT>(seed_odd, seed_even, c_odd, c_even, data_len,
                                          params_len, params, xs, logfn, K, a, beta, betas, step0,
                                          nsteps, mode, accept, means, sample_out, sample_floats);
    else
        bh_resident_impl<BH_PARTNER_WALKER>(seed_odd, seed_even, c_odd, c_even, data_len,
                                            params_len, params, xs, logfn, K, a, beta, betas,
                                            step0, nsteps, mode, accept, means, sample_out,
                                            sample_floats);
}

#endif  // LOGFN

#ifdef BH_BATCH
// Many independent posteriors of one model source in one launch: one
// work-group per problem, its W walkers advanced for all steps of the call
// inside the work-group.  A problem b with dimension dims[b] <= DIM and
// params params + poff[b] runs exactly the single-sampler algorithm
// (move-bare!, constants 3333/4444) with bare seed seed0 + 2b: the same
// streams, moves and accept tests as a sampler created for that problem
// alone.  Positions are packed with the problem's own dimension; a proposal
// is built in the thread's LDS slot, where the model reads it.
// State placement (gstate): 0 copies the problem's positions and log
// densities into LDS for the call (partner reads from LDS; the LDS a
// problem needs bounds how many share a CU), 1 works on them in place in
// the problem's global slot (L2-resident; many more problems per CU).
// Mode 1 also accumulates per problem, in double, the sums over steps of the
// ensemble sums of x and x^2; each step's ensemble sum of a coordinate is a
// float sum in a fixed order (threads own (coordinate, walker residue)
// pairs, P = nb / dim residues: a serial sum over walkers p, p + P, ..., then
// the P partial sums in order).  With series non-null it also writes each
// step's ensemble mean (that sum over W), problem b's n x dims[b] block
// starting at series + soff[b].  Work-group i of a launch runs problem
// order[base + i] (problems are launched in classes of similar dimension,
// each with the LDS it needs).
// LDS: 2*DIM doubles + 2*nb floats + nb*dim floats (+ W*dim + W floats for
// gstate 0).
template <bool GSTATE>
__device__ __forceinline__ void bh_batch_impl(
    unsigned seed0, const unsigned* __restrict__ dims, const unsigned long long* __restrict__ poff,
    const unsigned* __restrict__ dlen, const unsigned* __restrict__ plen,
    const float* __restrict__ params, float* __restrict__ xs, float* __restrict__ logfn,
    unsigned W, float a, unsigned step0, unsigned nsteps, int partner, int mode,
    unsigned long long* __restrict__ accept, double* __restrict__ sums,
    const unsigned long long* __restrict__ soff, float* __restrict__ series,
    const unsigned* __restrict__ order, unsigned base) {
    extern __shared__ double lds_d[];
    const unsigned b = order[base + blockIdx.x];
    const unsigned dim = dims[b];
    const unsigned K = W / 2;
    const unsigned t = threadIdx.x, nb = blockDim.x;
    const float* prm = params + poff[b];
    const unsigned data_len = dlen[b], params_len = plen[b];
    float* gx = xs + (size_t)b * W * DIM;  // problem slot: W x DIM floats, packed by dim
    float* gl = logfn + (size_t)b * W;
    double* acc1 = lds_d;                  // DIM
    double* acc2 = acc1 + DIM;             // DIM
    float* red = reinterpret_cast<float*>(acc2 + DIM);  // 2 * nb
    float* sx;                             // W * dim
    float* sl;                             // W
    float* yb;                             // nb * dim
    if constexpr (GSTATE) {
        sx = gx;
        sl = gl;
        yb = red + 2 * nb;
    } else {
        sx = red + 2 * nb;
        sl = sx + (size_t)W * dim;
        yb = sl + W;
        for (unsigned i = t; i < W * dim; i += nb) sx[i] = gx[i];
        for (unsigned i = t; i < W; i += nb) sl[i] = gl[i];
    }
    for (unsigned i = t; i < 2 * DIM; i += nb) acc1[i] = 0.0;
    __syncthreads();
    const unsigned seed = seed0 + 2u * b;
    unsigned long long nacc = 0;
    const float inv_a = 1.0f / a;
    const float ca = a - 2.0f + inv_a;
    const float cb = 2.0f * (1.0f - inv_a);
    float* y = yb + (size_t)t * dim;
    const unsigned P = dim <= nb ? nb / dim : 1u;
    for (unsigned st = 0; st < nsteps; ++st) {
        const unsigned step = step0 + st;
        for (int h = 0; h < 2; ++h) {
            float* X = sx + (h ? (size_t)K * dim : 0);
            const float* S = sx + (h ? 0 : (size_t)K * dim);
            float* L = sl + (h ? K : 0);
            for (unsigned k = t; k < K; k += nb) {
                const bh_f32x4 u = bh_u01x4(bh_philox4x32_10(
                    bh_u32x4{k, step, h ? 4444u : 3333u, BH_CTR3}, seed + (unsigned)h,
                    BH_KEY1_STRETCH));
                const float z = __builtin_fmaf(ca * u.y, u.y, cb * u.y) + inv_a;
                const float* xk = X + (size_t)k * dim;
                // partner element i at S[(e + i) mod K*dim] (FLAT) or
                // S[j*dim + i] (WALKER); eight coordinates' loads in flight
                const unsigned total = K * dim;
                unsigned e;
                if (partner == BH_PARTNER_FLAT) {
                    e = (unsigned)((u.x * (float)K) * (float)dim);
                    if (e >= total) e -= total;
                } else {
                    unsigned j = (unsigned)(u.x * (float)K);
                    j = j < K ? j : K - 1;
                    e = j * dim;  // e + i < total for i < dim
                }
                unsigned i = 0;
                for (; i + 8 <= dim; i += 8) {
                    float xa[8], sa[8];
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        unsigned ec = e + c;
                        if (ec >= total) ec -= total;
                        xa[c] = xk[i + c];
                        sa[c] = S[ec];
                    }
#pragma unroll
                    for (int c = 0; c < 8; ++c) y[i + c] = __builtin_fmaf(z, xa[c] - sa[c], sa[c]);
                    e += 8;
                    if (e >= total) e -= total;
                }
                for (; i < dim; ++i) {
                    const float xj = S[e];
                    y[i] = __builtin_fmaf(z, xk[i] - xj, xj);
                    e = (e + 1 == total) ? 0u : e + 1;
                }
                const float ly = LOGFN(data_len, params_len, prm, dim, y);
                float zp = 1.0f;
                for (unsigned i = 1; i < dim; ++i) zp *= z;
                const float q = __builtin_isfinite(ly) ? zp * expf(ly - L[k]) : 0.0f;
                if (u.z <= q) {
                    for (unsigned i = 0; i < dim; ++i) X[(size_t)k * dim + i] = y[i];
                    L[k] = ly;
                    nacc += 1;
                }
            }
            __syncthreads();
        }
        if (mode == 1) {
            if (dim <= nb) {
                if (t < P * dim) {
                    const unsigned d = t % dim, p = t / dim;
                    float s1 = 0.0f, s2 = 0.0f;
                    unsigned w = p;
                    // walkers p, p + P, ... in order, eight loads in flight
                    for (; w + 7 * P < W; w += 8 * P) {
                        float xv[8];
#pragma unroll
                        for (int c = 0; c < 8; ++c) xv[c] = sx[(size_t)(w + c * P) * dim + d];
#pragma unroll
                        for (int c = 0; c < 8; ++c) {
                            s1 += xv[c];
                            s2 = __builtin_fmaf(xv[c], xv[c], s2);
                        }
                    }
                    for (; w < W; w += P) {
                        const float x = sx[(size_t)w * dim + d];
                        s1 += x;
                        s2 = __builtin_fmaf(x, x, s2);
                    }
                    red[t] = s1;
                    red[nb + t] = s2;
                }
                __syncthreads();
                if (t < dim) {
                    float s1 = 0.0f, s2 = 0.0f;
                    for (unsigned p = 0; p < P; ++p) {
                        s1 += red[p * dim + t];
                        s2 += red[nb + p * dim + t];
                    }
                    acc1[t] += (double)s1;
                    acc2[t] += (double)s2;
                    if (series) series[soff[b] + (size_t)st * dim + t] = s1 / (float)W;
                }
                // red is written again only after the next step's two
                // half-step barriers
            } else {
                for (unsigned d = t; d < dim; d += nb) {
                    float s1 = 0.0f, s2 = 0.0f;
                    for (unsigned w = 0; w < W; ++w) {
                        const float x = sx[(size_t)w * dim + d];
                        s1 += x;
                        s2 = __builtin_fmaf(x, x, s2);
                    }
                    acc1[d] += (double)s1;
                    acc2[d] += (double)s2;
                    if (series) series[soff[b] + (size_t)st * dim + d] = s1 / (float)W;
                }
                // the next step's first half writes accepted proposals into
                // sx, which these sums read
                __syncthreads();
            }
        }
    }
    if constexpr (!GSTATE) {
        for (unsigned i = t; i < W * dim; i += nb) gx[i] = sx[i];
        for (unsigned i = t; i < W; i += nb) gl[i] = sl[i];
    }
    // accept count: block reduce in the (now free) staging area
    __syncthreads();
    unsigned long long* ra = reinterpret_cast<unsigned long long*>(red);
    ra[t] = nacc;
    __syncthreads();
    for (unsigned r = nb / 2; r >= 1; r >>= 1) {
        if (t < r) ra[t] += ra[t + r];
        __syncthreads();
    }
    if (t == 0) accept[b] += ra[0];
    if (mode == 1) {
        for (unsigned i = t; i < dim; i += nb) {
            sums[((size_t)b * DIM + i) * 2] += acc1[i];
            sums[((size_t)b * DIM + i) * 2 + 1] += acc2[i];
        }
    }
}

// State in LDS (bh_batch_run) or in place in global memory (bh_batch_run_g):
// two kernels, so that each has the registers of one placement.
#define BH_BATCH_RUN_KERNEL(NAME, GSTATE)                                                        \
    extern "C" __global__ void __launch_bounds__(1024) NAME(                                     \
        unsigned seed0, const unsigned* __restrict__ dims,                                        \
        const unsigned long long* __restrict__ poff, const unsigned* __restrict__ dlen,           \
        const unsigned* __restrict__ plen, const float* __restrict__ params,                      \
        float* __restrict__ xs, float* __restrict__ logfn, unsigned W, float a, unsigned step0,   \
        unsigned nsteps, int partner, int mode, unsigned long long* __restrict__ accept,          \
        double* __restrict__ sums, const unsigned long long* __restrict__ soff,                   \
        float* __restrict__ series, const unsigned* __restrict__ order, unsigned base) {          \
        bh_batch_impl<GSTATE>(seed0, dims, poff, dlen, plen, params, xs, logfn, W, a, step0,      \
                              nsteps, partner, mode, accept, sums, soff, series, order, base);    \
    }
BH_BATCH_RUN_KERNEL(bh_batch_run, false)
BH_BATCH_RUN_KERNEL(bh_batch_run_g, true)
#undef BH_BATCH_RUN_KERNEL

// Initial positions and log-densities of every problem: init_walkers'
// stream (key {seed0 + b, 0xdecafaaa}) over the problem's W*dim values,
// limits (DIM x 2 per problem) taken row i % dim.
extern "C" __global__ void __launch_bounds__(256)
bh_batch_init(unsigned seed0, const unsigned* __restrict__ dims,
              const unsigned long long* __restrict__ poff, const unsigned* __restrict__ dlen,
              const unsigned* __restrict__ plen, const float* __restrict__ params,
              const float* __restrict__ limits, float* __restrict__ xs, float* __restrict__ logfn,
              unsigned W) {
    const unsigned b = blockIdx.y;
    const unsigned dim = dims[b];
    float* gx = xs + (size_t)b * W * DIM;
    const float* lim = limits + (size_t)b * DIM * 2;
    const unsigned items = dim * W / 4;
    for (unsigned g = blockIdx.x * 256 + threadIdx.x; g < items; g += gridDim.x * 256) {
        const bh_f32x4 u = bh_u01x4(bh_philox4x32_10(
            bh_u32x4{g, BH_CTR1_INIT, BH_CTR2_INIT, BH_CTR3}, seed0 + b, BH_KEY1_INIT));
        const float uu[4] = {u.x, u.y, u.z, u.w};
        for (unsigned m = 0; m < 4; ++m) {
            const unsigned i = 4 * g + m;
            const unsigned r = i % dim;
            gx[i] = __builtin_fmaf(uu[m], lim[2 * r + 1], (1.0f - uu[m]) * lim[2 * r]);
        }
    }
}

extern "C" __global__ void __launch_bounds__(256)
bh_batch_logfn(const unsigned* __restrict__ dims, const unsigned long long* __restrict__ poff,
               const unsigned* __restrict__ dlen, const unsigned* __restrict__ plen,
               const float* __restrict__ params, const float* __restrict__ xs,
               float* __restrict__ logfn, unsigned W) {
    const unsigned b = blockIdx.y;
    const unsigned w = blockIdx.x * 256 + threadIdx.x;
    if (w >= W) return;
    const unsigned dim = dims[b];
    const float* x = xs + (size_t)b * W * DIM + (size_t)w * dim;
    logfn[(size_t)b * W + w] = LOGFN(dlen[b], plen[b], params + poff[b], dim, x);
}
#endif  // BH_BATCH

#ifdef BH_GLM
#ifndef BH_GLM_KIND
#define BH_GLM_KIND 1
#endif
// Generalised linear model posteriors evaluated for a whole batch of walkers
// at once (the likelihood is a dense contraction X.Y over data rows, run by
// the ahead-of-time kernels of bh_glm.hip).  These kernels are the
// walker-side halves: the proposal, and the accept test with its accounting,
// with the same arithmetic as bh_stretch_walker.

// Writes yscale*Y as nsplit 16-bit terms (bf16, or fp16 when f16 is set),
// hi + mid + lo, each the nearest value of what the previous terms leave
// (exact residuals in fp32).  With g0 (the logistic softplus sum's gradient
// at a reference point b of the ensemble, natural units) it returns the
// correction of the log posterior for the split's rounding: with y' =
// lnscale * terms (lnscale = 1 / yscale in the epilogue's units) and
// dl = y - y', the first-order -g0 . dl, and given H0 (the sum's Hessian at
// b) also -dl . H0 (y' - b + dl / 2), so that the split's log posterior plus
// it is the log posterior at y up to terms of third order in (y' - b, dl)
// (0 when not finite).
__device__ __forceinline__ double bh_glm_split(const float (&y)[DIM], float yscale, int nsplit,
                                               int f16, void* out, unsigned k, unsigned n,
                                               const double* __restrict__ g0, double lnscale,
                                               const float* __restrict__ H0,
                                               const float* __restrict__ bref) {
    double c = 0.0;
    double dl[DIM];  // y - y', y' what the terms represent (natural units)
    float w[DIM];    // y' - b + dl / 2
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        float r = y[i] * yscale;
        double hs = 0.0;  // what the terms represent, in units of 1 / yscale
        const size_t o = (size_t)k * DIM + i;
        for (int sp = 0; sp < nsplit; ++sp) {
            if (f16) {
                const _Float16 h = (_Float16)r;
                static_cast<_Float16*>(out)[(size_t)sp * n * DIM + o] = h;
                r = r - (float)h;
                hs += (double)(float)h;
            } else {
                const __bf16 h = (__bf16)r;
                static_cast<__bf16*>(out)[(size_t)sp * n * DIM + o] = h;
                r = r - (float)h;
                hs += (double)(float)h;
            }
        }
        const double yp = hs * lnscale;
        dl[i] = (double)y[i] - yp;
        if (g0) c = __builtin_fma(g0[i], -dl[i], c);
        if (H0) w[i] = (float)(yp - (double)bref[i] + 0.5 * dl[i]);
    }
    if (g0 && H0) {
        for (int i = 0; i < DIM; ++i) {
            float v = 0.0f;
            for (int j = 0; j < DIM; ++j) v = __builtin_fmaf(H0[i * DIM + j], w[j], v);
            c = __builtin_fma(-dl[i], (double)v, c);
        }
    }
    return __builtin_isfinite(c) ? c : 0.0;
}

// Proposals of one half-step: Y (fp32), z and u2 per walker, and the 16-bit
// split of log2(e)*Y for the MFMA kernels with, given g0 (and H0, bref), each
// proposal's correction for it in corr (as bh_glm_split); clears the guard
// list's count.  Given H0 and gscale, gscale[k] = max(1, q / (2 DIM)), q the
// larger H0-norm squared of the proposal's and the position's distances
// from b, by which the screen widens the guard for walker k.  BH_GLM_TPW
// threads per walker, thread `sub` taking coordinates sub, sub + BH_GLM_TPW,
// ... (so a walker's threads read and write consecutive coordinates); each
// draws the walker's uniforms (the same Philox call), and the corrections'
// parts are summed over the walker's threads.
#define BH_GLM_TPW 16
extern "C" __global__ void __launch_bounds__(WGS)
bh_glm_propose(unsigned seed, unsigned odd_or_even, const float* __restrict__ S,
               const float* __restrict__ X, float a, unsigned step, unsigned K, int partner,
               float* __restrict__ Y, float* __restrict__ zbuf, float* __restrict__ u2buf,
               void* __restrict__ ysplit, float yscale, int nsplit, int f16,
               unsigned* __restrict__ guard_count, const double* __restrict__ g0, double lnscale,
               double* __restrict__ corr, const float* __restrict__ H0,
               const float* __restrict__ bref, float* __restrict__ gscale) {
    constexpr int CPT = (DIM + BH_GLM_TPW - 1) / BH_GLM_TPW;  // coordinates per thread
    // H0 rows padded to a multiple of 4 floats plus 4 (16-byte reads; the
    // 16 rows a wave reads at once start in different banks)
    constexpr int HP = (DIM + 3) / 4 * 4 + 4;
    __shared__ __attribute__((aligned(16))) float wsh[WGS / BH_GLM_TPW][HP];
    __shared__ __attribute__((aligned(16))) float wxs[WGS / BH_GLM_TPW][HP];
    __shared__ __attribute__((aligned(16))) float hsh[DIM * HP];
    const unsigned gt = blockIdx.x * WGS + threadIdx.x;
    const unsigned k = gt / BH_GLM_TPW, sub = gt % BH_GLM_TPW;
    if (guard_count && gt == 0) *guard_count = 0u;
    float z;
    unsigned e;
    const float u2 = partner == BH_PARTNER_FLAT
                         ? bh_stretch_draw<BH_PARTNER_FLAT>(k, K, seed, odd_or_even, step, a, z, e)
                         : bh_stretch_draw<BH_PARTNER_WALKER>(k, K, seed, odd_or_even, step, a, z,
                                                              e);
    const unsigned total = K * DIM;
    // the loads of all the thread's coordinates (partner's and own position,
    // g0 and the reference point) issued before any of them is used; where
    // g0 or the reference point is absent the load reads the ensemble S
    // instead (at least 4 K DIM >= 8 DIM bytes; the value is never used), so
    // that no load sits in a branch of its own
    const double* g0p = g0 ? g0 : reinterpret_cast<const double*>(S);
    const float* brp = H0 ? bref : S;
    float xjv[CPT], xiv[CPT], brv[CPT];
    double g0v[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const unsigned i = sub + q * BH_GLM_TPW;
        xjv[q] = 0.0f;
        xiv[q] = 0.0f;
        brv[q] = 0.0f;
        g0v[q] = 0.0;
        if (i < DIM) {
            unsigned ei = e + i;
            if (partner == BH_PARTNER_FLAT && ei >= total) ei -= total;
            xjv[q] = S[ei];
            xiv[q] = X[(size_t)k * DIM + i];
            g0v[q] = g0p[i];
            brv[q] = brp[i];
        }
    }
    asm volatile("" ::: "memory");
    double c = 0.0;
    double dl[CPT];
    float wv[CPT], wx[CPT];
    int m = 0;
    for (unsigned i = sub; i < DIM; i += BH_GLM_TPW, ++m) {
        const float xj = xjv[m];
        const float xi = xiv[m];
        const float yi = __builtin_fmaf(z, xi - xj, xj);
        Y[(size_t)k * DIM + i] = yi;
        if (nsplit > 0) {
            float r = yi * yscale;
            double hs = 0.0;
            const size_t o = (size_t)k * DIM + i;
            for (int sp = 0; sp < nsplit; ++sp) {
                if (f16) {
                    const _Float16 h = (_Float16)r;
                    static_cast<_Float16*>(ysplit)[(size_t)sp * K * DIM + o] = h;
                    r = r - (float)h;
                    hs += (double)(float)h;
                } else {
                    const __bf16 h = (__bf16)r;
                    static_cast<__bf16*>(ysplit)[(size_t)sp * K * DIM + o] = h;
                    r = r - (float)h;
                    hs += (double)(float)h;
                }
            }
            const double yp = hs * lnscale;
            dl[m] = (double)yi - yp;
            if (g0) c = __builtin_fma(g0v[m], -dl[m], c);
            if (H0) {
                wv[m] = (float)(yp - (double)brv[m] + 0.5 * dl[m]);
                wx[m] = xi - brv[m];
            }
        }
    }
    float qy = 0.0f, qx = 0.0f;
    if (nsplit > 0 && g0 && H0) {
        // the second-order term: the walker's y' - b + dl / 2 through LDS,
        // each thread the rows of H0 of its coordinates (H0 staged in LDS
        // once per work-group); with it the H0-norms of the proposal's and
        // the position's distances from b
        const unsigned wl = threadIdx.x / BH_GLM_TPW;
        // rows padded with zeros past DIM; eight loads in flight per thread
        constexpr unsigned NH = (DIM * HP + WGS - 1) / WGS;
        for (unsigned t0 = 0; t0 < NH; t0 += 8) {
            float hv[8];
#pragma unroll
            for (unsigned t = 0; t < 8; ++t) {
                const unsigned q = threadIdx.x + (t0 + t) * WGS, r = q / HP, j = q % HP;
                hv[t] = q < DIM * HP && j < DIM ? H0[r * DIM + j] : 0.0f;
            }
            asm volatile("" ::: "memory");
#pragma unroll
            for (unsigned t = 0; t < 8; ++t) {
                const unsigned q = threadIdx.x + (t0 + t) * WGS;
                if (q < DIM * HP) hsh[q] = hv[t];
            }
        }
        for (unsigned j = DIM + sub; j < HP; j += BH_GLM_TPW) {
            wsh[wl][j] = 0.0f;
            wxs[wl][j] = 0.0f;
        }
        m = 0;
        for (unsigned i = sub; i < DIM; i += BH_GLM_TPW, ++m) {
            wsh[wl][i] = wv[m];
            wxs[wl][i] = wx[m];
        }
        __syncthreads();
        m = 0;
        for (unsigned i = sub; i < DIM; i += BH_GLM_TPW, ++m) {
            float v = 0.0f, vx = 0.0f;
            const float4* hr = reinterpret_cast<const float4*>(hsh + i * HP);
            const float4* ws = reinterpret_cast<const float4*>(wsh[wl]);
            const float4* wxr = reinterpret_cast<const float4*>(wxs[wl]);
#pragma unroll 4
            for (int j4 = 0; j4 < (DIM + 3) / 4; ++j4) {
                const float4 h = hr[j4], wy = ws[j4], wxx = wxr[j4];
                v = __builtin_fmaf(h.x, wy.x, v);
                v = __builtin_fmaf(h.y, wy.y, v);
                v = __builtin_fmaf(h.z, wy.z, v);
                v = __builtin_fmaf(h.w, wy.w, v);
                vx = __builtin_fmaf(h.x, wxx.x, vx);
                vx = __builtin_fmaf(h.y, wxx.y, vx);
                vx = __builtin_fmaf(h.z, wxx.z, vx);
                vx = __builtin_fmaf(h.w, wxx.w, vx);
            }
            c = __builtin_fma(-dl[m], (double)v, c);
            qy = __builtin_fmaf(wv[m], v, qy);
            qx = __builtin_fmaf(wx[m], vx, qx);
        }
    }
#pragma unroll
    for (int off = BH_GLM_TPW / 2; off >= 1; off >>= 1) {
        c += __shfl_xor(c, off, 64);
        qy += __shfl_xor(qy, off, 64);
        qx += __shfl_xor(qx, off, 64);
    }
    if (sub == 0) {
        zbuf[k] = z;
        u2buf[k] = u2;
        if (nsplit > 0 && corr) corr[k] = __builtin_isfinite(c) ? c : 0.0;
        // the guard's scale for this walker: the correction's remainder grows
        // with the square of the distance from b, which near equilibrium is
        // about DIM in the H0-norm
        if (gscale) {
            const float q = __builtin_fmaxf(qy, qx) / (2.0f * (float)DIM);
            gscale[k] = q > 1.0f ? q : 1.0f;
        }
    }
}

// 16-bit split of n given points (for the initial log-densities), with their
// corrections given g0.
extern "C" __global__ void __launch_bounds__(WGS)
bh_glm_split_points(const float* __restrict__ P, unsigned n, void* __restrict__ ysplit,
                    float yscale, int nsplit, int f16, const double* __restrict__ g0,
                    double lnscale, double* __restrict__ corr, const float* __restrict__ H0,
                    const float* __restrict__ bref) {
    const unsigned k = blockIdx.x * WGS + threadIdx.x;
    if (k >= n) return;
    float y[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) y[i] = P[(size_t)k * DIM + i];
    const double c = bh_glm_split(y, yscale, nsplit, f16, ysplit, k, n, g0, lnscale, H0, bref);
    if (corr) corr[k] = c;
}

// Log posterior of point y from the sum s over data rows of the likelihood
// kernel's per-row terms, in double.
// BH_GLM_KIND 1, logistic: lin.y - scale * s - |y|^2 / (2 s0^2), with s the
//   softplus sum and lin the linear coefficients of the precision.
// BH_GLM_KIND 2, linear regression, y = [beta (DIM-1), log sigma], s the sum
//   of squared residuals, lin = {N, N log(2 pi) / 2, 1 / s_beta^2,
//   1 / s_logsigma^2}: -N log sigma - N log(2 pi)/2 - s / (2 sigma^2)
//   - |beta|^2 / (2 s_beta^2) - (log sigma)^2 / (2 s_logsigma^2).
__device__ __forceinline__ double bh_glm_logpost_sum(const float (&y)[DIM], double s,
                                                     const double* __restrict__ lin, double scale,
                                                     double prior_prec) {
#if BH_GLM_KIND == 2
    (void)scale;
    (void)prior_prec;
    const double ls = (double)y[DIM - 1];
    double q = 0.0;
#pragma unroll
    for (int i = 0; i < DIM - 1; ++i) q = __builtin_fma((double)y[i], (double)y[i], q);
    return -lin[0] * ls - lin[1] - 0.5 * exp(-2.0 * ls) * s - 0.5 * (lin[2] * q + lin[3] * ls * ls);
#else
    double l = 0.0, q = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        l = __builtin_fma(lin[i], (double)y[i], l);
        q = __builtin_fma((double)y[i], (double)y[i], q);
    }
    return l - scale * s - 0.5 * prior_prec * q;
#endif
}

// The same from the per-range row sums of the likelihood kernel (the sum
// over ranges runs in range order).
__device__ __forceinline__ double bh_glm_logpost(const float (&y)[DIM], unsigned k, unsigned n,
                                                 const double* __restrict__ partial,
                                                 unsigned nranges, const double* __restrict__ lin,
                                                 double scale, double prior_prec,
                                                 const double* __restrict__ corr) {
    // in range order, eight loads in flight at a time
    double s = 0.0;
    unsigned r = 0;
    for (; r + 8 <= nranges; r += 8) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = partial[(size_t)(r + e) * n + k];
#pragma unroll
        for (int e = 0; e < 8; ++e) s += v[e];
    }
    for (; r < nranges; ++r) s += partial[(size_t)r * n + k];
    const double l = bh_glm_logpost_sum(y, s, lin, scale, prior_prec);
    return corr ? l + corr[k] : l;
}

// Sum over r < nranges of partial[r * stride + col] by one wave: lane-strided
// partial sums (eight loads in flight), then a butterfly (a fixed order).
__device__ __forceinline__ double bh_wave_range_sum(const double* __restrict__ partial,
                                                    unsigned nranges, unsigned stride,
                                                    unsigned col) {
    const unsigned lane = threadIdx.x & 63u;
    double s = 0.0;
    unsigned r = lane;
    for (; r + 7 * 64 < nranges; r += 8 * 64) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = partial[(size_t)(r + e * 64) * stride + col];
#pragma unroll
        for (int e = 0; e < 8; ++e) s += v[e];
    }
    for (; r < nranges; r += 64) s += partial[(size_t)r * stride + col];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// Initial log posteriors of n points.
extern "C" __global__ void __launch_bounds__(WGS)
bh_glm_logpost(const float* __restrict__ P, unsigned n, const double* __restrict__ partial,
               unsigned nranges, const double* __restrict__ lin, double scale, double prior_prec,
               double* __restrict__ lp, const double* __restrict__ corr) {
    const unsigned k = blockIdx.x * WGS + threadIdx.x;
    if (k >= n) return;
    float y[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) y[i] = P[(size_t)k * DIM + i];
    lp[k] = bh_glm_logpost(y, k, n, partial, nranges, lin, scale, prior_prec, corr);
}

// The accept decision of stretch_move (amd-gcn-mcmc-stretch.cl:66-71): q =
// z^(DIM-1) exp(beta (lp(Y) - lp(X))), u2 <= q, the difference taken in
// double and rounded to float as the kernels of every precision do.
__device__ __forceinline__ bool bh_glm_decide(double lpy, double lpx, float z, float u2,
                                              float beta) {
    const float d = (float)(lpy - lpx);
    const float q = __builtin_isfinite(lpy) ? bh_zpow(z) * expf(beta * d) : 0.0f;
    return u2 <= q;
}

// Guarded accept test of the reduced-precision likelihoods.  The decision is
// a monotone step in d = lp(Y) - lp(X); a walker whose log margin
//   m = log u2 - (DIM-1) log z - beta d
// is within `guard` of zero (times gscale[k] when given) could be decided
// differently by an error of that size in d, so it is listed (up to lmax
// walkers, the rest counting as overflow; the engine passes lmax = K, so
// none overflow) for an fp64 evaluation of lp(Y) and lp(X) and decided on
// those.
// BH_GLM_STPW threads per walker.
// flag[k]: 0 decided here, 3 listed (bh_glm_resolve turns it into 1 = reject
// or 2 = accept).
#define BH_GLM_STPW 8
extern "C" __global__ void __launch_bounds__(WGS)
bh_glm_screen(const float* __restrict__ Y, const float* __restrict__ zbuf,
              const float* __restrict__ u2buf, const double* __restrict__ partial,
              unsigned nranges, const double* __restrict__ lin, double scale, double prior_prec,
              const double* __restrict__ lp, float beta, unsigned K, double guard,
              unsigned* __restrict__ flag, unsigned* __restrict__ list,
              unsigned* __restrict__ count, unsigned lmax, const float* __restrict__ Xpos,
              double* __restrict__ pts64, double* __restrict__ lpy_out,
              const double* __restrict__ corr, const float* __restrict__ gscale) {
    const unsigned gt = blockIdx.x * WGS + threadIdx.x;
    const unsigned k = gt / BH_GLM_STPW, sub = gt % BH_GLM_STPW;
    const unsigned base = (threadIdx.x & 63u) & ~(unsigned)(BH_GLM_STPW - 1);
#if BH_GLM_KIND == 2
    // (the guard is for logistic GLMs; this form is not launched)
    double lpy = 0.0;
    if (sub == 0) {
        float y[DIM];
#pragma unroll
        for (int i = 0; i < DIM; ++i) y[i] = Y[(size_t)k * DIM + i];
        lpy = bh_glm_logpost(y, k, K, partial, nranges, lin, scale, prior_prec, corr);
    }
#else
    // the walker's BH_GLM_STPW threads (consecutive lanes) share its sums:
    // row ranges sub, sub + STPW, ... (in that order, eight loads in
    // flight) and coordinates likewise.  The thread's coordinates and the
    // walker's scalars are loaded first, all together (where corr or gscale
    // is absent the load reads lp or zbuf instead; the value is unused).
    constexpr int SC = (DIM + BH_GLM_STPW - 1) / BH_GLM_STPW;
    float ycv[SC];
    double lcv[SC];
#pragma unroll
    for (int c = 0; c < SC; ++c) {
        const unsigned i = sub + c * BH_GLM_STPW;
        ycv[c] = i < DIM ? Y[(size_t)k * DIM + i] : 0.0f;
        lcv[c] = i < DIM ? lin[i] : 0.0;
    }
    const float u2k = u2buf[k], zk = zbuf[k];
    const double lpk = lp[k];
    const double ck = (corr ? corr : lp)[k];
    const float gsk = (gscale ? gscale : zbuf)[k];
    double sm = 0.0, l = 0.0, q = 0.0;
    unsigned r = sub;
    for (; r + 7 * BH_GLM_STPW < nranges; r += 8 * BH_GLM_STPW) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = partial[(size_t)(r + e * BH_GLM_STPW) * K + k];
#pragma unroll
        for (int e = 0; e < 8; ++e) sm += v[e];
    }
    for (; r < nranges; r += BH_GLM_STPW) sm += partial[(size_t)r * K + k];
#pragma unroll
    for (int c = 0; c < SC; ++c) {
        if (sub + c * BH_GLM_STPW < DIM) {
            const double yi = (double)ycv[c];
            l = __builtin_fma(lcv[c], yi, l);
            q = __builtin_fma(yi, yi, q);
        }
    }
#pragma unroll
    for (int off = BH_GLM_STPW / 2; off >= 1; off >>= 1) {
        sm += __shfl_xor(sm, off, 64);
        l += __shfl_xor(l, off, 64);
        q += __shfl_xor(q, off, 64);
    }
    double lpy = l - scale * sm - 0.5 * prior_prec * q;
    if (corr) lpy += ck;
#endif
    // the walker's first thread decides or lists it; all its threads stage
    // the listed points
#if BH_GLM_KIND == 2
    const float u2k = u2buf[k], zk = zbuf[k];
    const double lpk = lp[k];
    const float gsk = gscale ? gscale[k] : 1.0f;
#endif
    unsigned slot = 0xffffffffu;
    if (sub == 0) {
        lpy_out[k] = lpy;
        unsigned f = 0;
        if (__builtin_isfinite(lpy)) {
            const double m = log((double)u2k) - (double)(DIM - 1) * log((double)zk) -
                             (double)beta * (lpy - lpk);
            const double gk = gscale ? guard * (double)gsk : guard;
            if (__builtin_fabs(m) <= gk) {
                const unsigned i = atomicAdd(count, 1u);
                if (i < lmax) {
                    list[i] = k;
                    f = 3;
                    slot = i;
                }
            }
        }
        // walkers not listed are decided here, on this precision's values
        // (1 reject, 2 accept; bh_glm_resolve decides the listed ones)
        if (f == 0) f = bh_glm_decide(lpy, lpk, zk, u2k, beta) ? 2u : 1u;
        flag[k] = f;
    }
    slot = __shfl(slot, base, 64);
    if (slot != 0xffffffffu) {
        // the two points in double for the fp64 likelihood kernel: row 2i
        // the proposal, row 2i + 1 the current position
        for (unsigned j = sub; j < DIM; j += BH_GLM_STPW) {
            pts64[(size_t)2 * slot * DIM + j] = (double)Y[(size_t)k * DIM + j];
            pts64[((size_t)2 * slot + 1) * DIM + j] = (double)Xpos[(size_t)k * DIM + j];
        }
    }
}

// Decides the listed walkers on their fp64 log posteriors, one wave per
// walker: partial64 holds, per row range r (row stride `stride`), column
// 2i the softplus sum of Y[list[i]] and column 2i + 1 that of X[list[i]]
// (natural units).
// totals[0] += listed walkers, totals[1] += walkers past lmax.
extern "C" __global__ void __launch_bounds__(WGS)
bh_glm_resolve(const float* __restrict__ Y, const float* __restrict__ X,
               const float* __restrict__ zbuf, const float* __restrict__ u2buf,
               const double* __restrict__ partial64, unsigned nranges64, unsigned stride,
               const double* __restrict__ lin64, double prior_prec, float beta,
               const unsigned* __restrict__ list, const unsigned* __restrict__ count,
               unsigned lmax, unsigned* __restrict__ flag,
               unsigned long long* __restrict__ totals) {
    const unsigned n = *count;
    const unsigned c = n < lmax ? n : lmax;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&totals[0], (unsigned long long)c);
        atomicAdd(&totals[1], (unsigned long long)(n - c));
    }
    const unsigned waves = WGS / 64;
    for (unsigned i = blockIdx.x * waves + threadIdx.x / 64; i < c; i += gridDim.x * waves) {
#if BH_GLM_KIND != 2
        // both columns of a range in one 16-byte load (the row stride is
        // even), eight loads in flight per lane; the coordinates' terms
        // across the lanes; then one butterfly for all six sums
        const unsigned lane = threadIdx.x & 63u;
        double sy = 0.0, sx = 0.0;
        unsigned r = lane;
        for (; r + 7 * 64 < nranges64; r += 8 * 64) {
            double2 v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e)
                v[e] = *reinterpret_cast<const double2*>(partial64 + (size_t)(r + e * 64) * stride +
                                                         2 * i);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                sy += v[e].x;
                sx += v[e].y;
            }
        }
        for (; r < nranges64; r += 64) {
            const double2 v =
                *reinterpret_cast<const double2*>(partial64 + (size_t)r * stride + 2 * i);
            sy += v.x;
            sx += v.y;
        }
        const unsigned k = list[i];
        double ly = 0.0, lx = 0.0, qy = 0.0, qx = 0.0;
        for (unsigned j = lane; j < DIM; j += 64) {
            const double yj = (double)Y[(size_t)k * DIM + j];
            const double xj = (double)X[(size_t)k * DIM + j];
            ly = __builtin_fma(lin64[j], yj, ly);
            lx = __builtin_fma(lin64[j], xj, lx);
            qy = __builtin_fma(yj, yj, qy);
            qx = __builtin_fma(xj, xj, qx);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            sy += __shfl_xor(sy, off, 64);
            sx += __shfl_xor(sx, off, 64);
            ly += __shfl_xor(ly, off, 64);
            lx += __shfl_xor(lx, off, 64);
            qy += __shfl_xor(qy, off, 64);
            qx += __shfl_xor(qx, off, 64);
        }
        if (lane == 0) {
            // the logistic log posterior (bh_glm_logpost_sum, scale 1)
            const double lpy = ly - sy - 0.5 * prior_prec * qy;
            const double lpx = lx - sx - 0.5 * prior_prec * qx;
            flag[k] = bh_glm_decide(lpy, lpx, zbuf[k], u2buf[k], beta) ? 2u : 1u;
        }
#else
        const double sy = bh_wave_range_sum(partial64, nranges64, stride, 2 * i);
        const double sx = bh_wave_range_sum(partial64, nranges64, stride, 2 * i + 1);
        if ((threadIdx.x & 63u) == 0) {
            const unsigned k = list[i];
            float y[DIM], xk[DIM];
#pragma unroll
            for (int j = 0; j < DIM; ++j) {
                y[j] = Y[(size_t)k * DIM + j];
                xk[j] = X[(size_t)k * DIM + j];
            }
            const double ly = bh_glm_logpost_sum(y, sy, lin64, 1.0, prior_prec);
            const double lx = bh_glm_logpost_sum(xk, sx, lin64, 1.0, prior_prec);
            flag[k] = bh_glm_decide(ly, lx, zbuf[k], u2buf[k], beta) ? 2u : 1u;
        }
#endif
    }
}

// Accept test of one half-step (amd-gcn-mcmc-stretch.cl:66-78) with accu !=
// 0 also the accounting of stretch_move_accu.  With flag non-null, walkers
// resolved in fp64 take that decision.
extern "C" __global__ void __launch_bounds__(WGS)
bh_glm_accept(const float* __restrict__ Y, const float* __restrict__ zbuf,
              const float* __restrict__ u2buf, const double* __restrict__ partial,
              unsigned nranges, const double* __restrict__ lin, double scale, double prior_prec,
              float* __restrict__ X, double* __restrict__ lp, float beta, unsigned K, int accu,
              unsigned* __restrict__ accept, float* __restrict__ means, unsigned mstep,
              const unsigned* __restrict__ flag, const double* __restrict__ corr) {
    const unsigned k = blockIdx.x * WGS + threadIdx.x;
    float y[DIM], xk[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) y[i] = Y[(size_t)k * DIM + i];
    // the stored lp stays in this sampler's precision in every case, so that
    // its common offset from the fp64 values cancels in later differences
    const double lpy = bh_glm_logpost(y, k, K, partial, nranges, lin, scale, prior_prec, corr);
    const unsigned f = flag ? flag[k] : 0u;
    const bool accepted = (f == 1u || f == 2u)
                              ? f == 2u
                              : bh_glm_decide(lpy, lp[k], zbuf[k], u2buf[k], beta);
    if (accepted) {
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            X[(size_t)k * DIM + i] = y[i];
            xk[i] = y[i];
        }
        lp[k] = lpy;
    } else if (accu) {
#pragma unroll
        for (int i = 0; i < DIM; ++i) xk[i] = X[(size_t)k * DIM + i];
    }
    if (accu) {
        __shared__ float red[BH_ACCU_CHUNK * WGS];
        __shared__ unsigned wave_counts[WGS / 64];
        bh_accu(xk, accepted, blockIdx.x, gridDim.x, mstep, accept, means, red, wave_counts);
    }
}
// The guarded path's accept step, with every decision already in flag (1
// reject, 2 accept) and the proposals' log posteriors in lpy: work-group
// (g, c) handles group g's WGS walkers and coordinates
// [c*BH_ACCU_CHUNK, (c+1)*BH_ACCU_CHUNK): it copies accepted proposals of
// those coordinates into X and adds their group sums to the means (the
// same tree per coordinate as bh_accu); chunk 0 also stores the accepted
// log posteriors and the group's accept count.
extern "C" __global__ void __launch_bounds__(WGS)
bh_glm_accept_split(const float* __restrict__ Y, const double* __restrict__ lpy,
                    float* __restrict__ X, double* __restrict__ lp, int accu,
                    unsigned* __restrict__ accept, float* __restrict__ means, unsigned mstep,
                    const unsigned* __restrict__ flag) {
    const unsigned g = blockIdx.x, c = blockIdx.y;
    const unsigned k = g * WGS + threadIdx.x;
    const bool accepted = flag[k] == 2u;
    const unsigned d0 = c * BH_ACCU_CHUNK;
    float v[BH_ACCU_CHUNK];
#pragma unroll
    for (int i = 0; i < BH_ACCU_CHUNK; ++i) {
        const unsigned d = d0 + i;
        v[i] = 0.0f;
        if (d < DIM) {
            if (accepted) {
                v[i] = Y[(size_t)k * DIM + d];
                X[(size_t)k * DIM + d] = v[i];
            } else if (accu) {
                v[i] = X[(size_t)k * DIM + d];
            }
        }
    }
    if (c == 0 && accepted) lp[k] = lpy[k];
    if (accu) {
        __shared__ float red[BH_ACCU_CHUNK * WGS];
        __shared__ unsigned wave_counts[WGS / 64];
        const unsigned groups = gridDim.x;
        const bool leader = threadIdx.x == 0;
        float sums[BH_ACCU_CHUNK];
        unsigned cnt;
        bh_slice_count_tree_sums<BH_ACCU_CHUNK>(accepted, v, red, wave_counts, sums, cnt, false);
        if (c == 0 && leader) accept[g] += cnt;
        if (leader) {
            float* m = means + (size_t)groups * DIM * mstep + g;
#pragma unroll
            for (int i = 0; i < BH_ACCU_CHUNK; ++i)
                if (d0 + i < DIM) m[(size_t)(d0 + i) * groups] += sums[i];
        }
    }
}
#endif  // BH_GLM

#ifdef LOGPDF
// logpdf / pdf kernels (amd-gcn-distribution.cl:1-14).
extern "C" __global__ void __launch_bounds__(WGS)
bh_logpdf(unsigned data_len, unsigned params_len, const float* __restrict__ params, unsigned dim,
          const float* __restrict__ x, float* __restrict__ res, unsigned long long n) {
    const unsigned long long j = (unsigned long long)blockIdx.x * WGS + threadIdx.x;
    if (j < n) res[j] = LOGPDF(data_len, params_len, params, dim, x + j * dim);
}

extern "C" __global__ void __launch_bounds__(WGS)
bh_pdf(unsigned data_len, unsigned params_len, const float* __restrict__ params, unsigned dim,
       const float* __restrict__ x, float* __restrict__ res, unsigned long long n) {
    const unsigned long long j = (unsigned long long)blockIdx.x * WGS + threadIdx.x;
    if (j < n) res[j] = native_exp(LOGPDF(data_len, params_len, params, dim, x + j * dim));
}
#endif  // LOGPDF

#ifdef LOGLIK
// loglik / lik / evidence kernels (amd-gcn-likelihood.cl:1-29); the
// evidence accumulates in double (ACCUMULATOR=double, amd_gcn.clj:580).
extern "C" __global__ void __launch_bounds__(WGS)
bh_loglik(unsigned data_len, const float* __restrict__ data, unsigned dim,
          const float* __restrict__ x, float* __restrict__ res, unsigned long long n) {
    const unsigned long long j = (unsigned long long)blockIdx.x * WGS + threadIdx.x;
    if (j < n) res[j] = LOGLIK(data_len, data, dim, x + j * dim);
}

extern "C" __global__ void __launch_bounds__(WGS)
bh_lik(unsigned data_len, const float* __restrict__ data, unsigned dim,
       const float* __restrict__ x, float* __restrict__ res, unsigned long long n) {
    const unsigned long long j = (unsigned long long)blockIdx.x * WGS + threadIdx.x;
    if (j < n) res[j] = native_exp(LOGLIK(data_len, data, dim, x + j * dim));
}

extern "C" __global__ void __launch_bounds__(WGS)
bh_evidence_partial(unsigned data_len, const float* __restrict__ data, unsigned dim,
                    const float* __restrict__ x, unsigned long long n,
                    double* __restrict__ partial) {
    const unsigned long long j = (unsigned long long)blockIdx.x * WGS + threadIdx.x;
    double v = 0.0;
    if (j < n) v = (double)native_exp(LOGLIK(data_len, data, dim, x + j * dim));
#pragma unroll
    for (int i = 32; i >= 1; i >>= 1) v += __shfl_down(v, i, 64);
    __shared__ double wsum[WGS / 64];
    if ((threadIdx.x & 63u) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < WGS / 64; ++w) s += wsum[w];
        partial[blockIdx.x] = s;
    }
}
#endif  // LOGLIK
