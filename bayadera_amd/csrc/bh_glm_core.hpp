// Device building blocks of the GLM likelihood kernels, shared by the
// single-posterior kernels (bh_glm.hip) and the batched ones
// (bh_glm_batch.hip): the softplus epilogue, the 16x16x32 matrix-core
// products of one 16-row block (bf16 or fp16 operands, fp32 accumulators),
// and the software-pipelined loop over the row blocks of a range with the
// epilogue spread between the products.
#pragma once
#include <hip/hip_runtime.h>

namespace bh {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x32_{bf16,f16}: the same lane layout for both operand types
__device__ __forceinline__ f32x4 mfma_16x16x32(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_16x16x32(const f16x8& a, const f16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// sum over four values of |u|/2 + log2(1 + 2^-|u|): four v_exp_f32 and one
// v_log_f32 of the product of the four factors 1 + 2^-|u|, which lies in
// [1, 16] (each factor is in (1, 2]); every transcendental runs where it is
// accurate to about an ulp.
__device__ __forceinline__ float softplus2_rest4(float u0, float u1, float u2, float u3) {
    const float p01 = (1.0f + __builtin_amdgcn_exp2f(-__builtin_fabsf(u0))) *
                      (1.0f + __builtin_amdgcn_exp2f(-__builtin_fabsf(u1)));
    const float p23 = (1.0f + __builtin_amdgcn_exp2f(-__builtin_fabsf(u2))) *
                      (1.0f + __builtin_amdgcn_exp2f(-__builtin_fabsf(u3)));
    const float a = (__builtin_fabsf(u0) + __builtin_fabsf(u1)) +
                    (__builtin_fabsf(u2) + __builtin_fabsf(u3));
    return __builtin_fmaf(0.5f, a, __builtin_amdgcn_logf(p01 * p23));
}

template <int D, int NS, int CB, class V>
__device__ __forceinline__ void glm_mfma_block(const V (&a)[D / 32], const V (&bfrag)[CB][D / 32][NS],
                                               f32x4 (&c)[CB]) {
#pragma unroll
    for (int wb = 0; wb < CB; ++wb) {
        c[wb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = NS - 1; s >= 0; --s)
#pragma unroll
            for (int kc = 0; kc < D / 32; ++kc)
                c[wb] = mfma_16x16x32(a[kc], bfrag[wb][kc][s], c[wb]);
    }
}

template <int CB>
__device__ __forceinline__ void glm_epilogue(const f32x4 (&c)[CB], double (&acc)[CB]) {
#pragma unroll
    for (int wb = 0; wb < CB; ++wb)
        acc[wb] += (double)softplus2_rest4(c[wb][0], c[wb][1], c[wb][2], c[wb][3]);
}

// rows at or past r1 contribute nothing: u = +inf gives the factor 1 + 0,
// and its |u|/2 is taken out by zeroing it in the sum
template <int CB>
__device__ __forceinline__ void glm_epilogue_masked(const f32x4 (&c)[CB], double (&acc)[CB],
                                                    unsigned orow, unsigned r1) {
#pragma unroll
    for (int wb = 0; wb < CB; ++wb) {
        float p = 1.0f, a = 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (orow + i < r1) {
                p *= 1.0f + __builtin_amdgcn_exp2f(-__builtin_fabsf(c[wb][i]));
                a += __builtin_fabsf(c[wb][i]);
            }
        }
        acc[wb] += (double)__builtin_fmaf(0.5f, a, __builtin_amdgcn_logf(p));
    }
}

// The epilogue of a run of two or four consecutive 16-row blocks, eight or
// sixteen values per lane and column block, with one v_log_f32 for the run:
// the product of its factors 1 + 2^-|u| (each in (1, 2], so at most 2^16)
// is built as a chain p <- fma(e, p, p) = p (1 + e), one rounding per
// factor.  It is cut into pieces of a few VALU instructions so that they can
// be placed between matrix-core products: five per column block for a block
// that starts the run (H = 0) or continues it (H = 2) (exponentials, product
// and |u| sum carried in st), seven for the block that ends it (H = 1: the
// same, then the logarithm and the double accumulation).
template <int CB>
struct GlmEpiState {
    float p[CB];  // product of the factors so far, per column block
    float a[CB];  // sum of the |u| so far
};

template <int H, class St>
__device__ __forceinline__ void glm_epilogue_piece(int ph, int wb, const f32x4& c, float (&e)[4],
                                                   St& st, double& acc) {
    switch (ph) {
    case 0:
        e[0] = __builtin_amdgcn_exp2f(-__builtin_fabsf(c[0]));
        e[1] = __builtin_amdgcn_exp2f(-__builtin_fabsf(c[1]));
        break;
    case 1:
        e[2] = __builtin_amdgcn_exp2f(-__builtin_fabsf(c[2]));
        e[3] = __builtin_amdgcn_exp2f(-__builtin_fabsf(c[3]));
        break;
    case 2:
        if (H == 0) {
            st.p[wb] = 1.0f + e[0];
        } else {
            st.p[wb] = __builtin_fmaf(e[0], st.p[wb], st.p[wb]);
        }
        st.p[wb] = __builtin_fmaf(e[1], st.p[wb], st.p[wb]);
        break;
    case 3:
        st.p[wb] = __builtin_fmaf(e[2], st.p[wb], st.p[wb]);
        st.p[wb] = __builtin_fmaf(e[3], st.p[wb], st.p[wb]);
        break;
    case 4:
        e[1] = (__builtin_fabsf(c[0]) + __builtin_fabsf(c[1])) +
               (__builtin_fabsf(c[2]) + __builtin_fabsf(c[3]));
        if (H == 0) st.a[wb] = e[1];
        else st.a[wb] = st.a[wb] + e[1];
        break;
    case 5:  // H = 1 only
        e[3] = __builtin_fmaf(0.5f, st.a[wb], __builtin_amdgcn_logf(st.p[wb]));
        break;
    default:
        acc += (double)e[3];
        break;
    }
}

// The products of one 16-row block into cn (zeroed by the first product of
// each chain) with part H of the run epilogue of the previous block's c
// spread between them: products of the CB column blocks alternate, so
// consecutive products are independent, and scheduling barriers keep each
// epilogue piece between two products, where it runs on the VALU while the
// matrix core works.  The per-column-block product order is that of
// glm_mfma_block (split terms smallest first, chunks in order).
template <int D, int NS, int H, int CB, class V>
__device__ __forceinline__ void glm_mfma_overlap(const V (&a)[D / 32],
                                                 const V (&bfrag)[CB][D / 32][NS],
                                                 f32x4 (&cn)[CB], const f32x4 (&c)[CB],
                                                 double (&acc)[CB], GlmEpiState<CB>& st) {
    constexpr int KC = D / 32;
    constexpr int NM = CB * KC * NS;
    constexpr int PW = H == 1 ? 7 : 5;  // pieces per column block
    constexpr int NP = CB * PW;
    float e[CB][4];
    int m = 0;
#pragma unroll
    for (int s = NS - 1; s >= 0; --s)
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
#pragma unroll
            for (int wb = 0; wb < CB; ++wb) {
                const f32x4 acc0 = (s == NS - 1 && kc == 0) ? f32x4{0.0f, 0.0f, 0.0f, 0.0f}
                                                            : cn[wb];
                cn[wb] = mfma_16x16x32(a[kc], bfrag[wb][kc][s], acc0);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int pc = m * NP / NM; pc < (m + 1) * NP / NM; ++pc)
                    glm_epilogue_piece<H>(pc % PW, pc / PW, c[pc / PW], e[pc / PW], st,
                                          acc[pc / PW]);
                __builtin_amdgcn_sched_barrier(0);
                ++m;
            }
}

// Half 1 of the pair epilogue without products (the last pair of a range).
template <int CB>
__device__ __forceinline__ void glm_epilogue_second(const f32x4 (&c)[CB], double (&acc)[CB],
                                                    GlmEpiState<CB>& st) {
    float e[CB][4];
#pragma unroll
    for (int wb = 0; wb < CB; ++wb)
#pragma unroll
        for (int ph = 0; ph < 7; ++ph) glm_epilogue_piece<1>(ph, wb, c[wb], e[wb], st, acc[wb]);
}

// The row blocks [r0, r1) of one range for the 16 CB points whose B
// fragments the wave holds: full 16-row blocks through the software pipeline, then one
// masked block.  `ld` loads the A fragments of the block starting at a row
// (ld.full(rb, a): all 16 rows valid; ld.masked(rb, r1, a): rows >= r1 read
// as zeros; Loader::vec is the operand type).  acc[wb] accumulates this
// lane's four rows of column block wb.  CB = 4 (64 points per wave) or 8.
template <int D, int NS, int CB, class Loader>
__device__ __forceinline__ void glm_mfma_rows(const Loader& ld,
                                              const typename Loader::vec (&bfrag)[CB][D / 32][NS],
                                              unsigned r0,
                                              unsigned r1, unsigned lg, double (&acc)[CB]) {
    constexpr int KC = D / 32;
    const unsigned rfull = r0 + (r1 - r0) / 16 * 16;
    f32x4 c0[CB], c1[CB];
    if (rfull > r0) {
        // software pipeline over 16-row blocks, the accumulator sets c0/c1
        // alternating without copies: the products of block i+1 run with the
        // epilogue of block i.  Runs of four blocks share a logarithm, then
        // runs of two for what is left.  Invariant at each loop head: c0
        // holds the products of block rb - 16 (epilogue pending), a1 the
        // rows of block rb.  Row blocks past the last are clamped to it
        // (loaded, never used).
        const unsigned rlast = rfull - 16;
        auto clamp = [&](unsigned rb) { return rb <= rlast ? rb : rlast; };
        typename Loader::vec a0[KC], a1[KC];
        ld.full(r0, a0);
        ld.full(clamp(r0 + 16), a1);
        glm_mfma_block<D, NS, CB>(a0, bfrag, c0);
        GlmEpiState<CB> st;
        unsigned rb = r0 + 16;
        // runs of four blocks rb - 16 .. rb + 32 (products of rb .. rb + 48)
        for (; rb + 48 < rfull; rb += 64) {
            ld.full(clamp(rb + 16), a0);
            glm_mfma_overlap<D, NS, 0, CB>(a1, bfrag, c1, c0, acc, st);
            ld.full(clamp(rb + 32), a1);
            glm_mfma_overlap<D, NS, 2, CB>(a0, bfrag, c0, c1, acc, st);
            ld.full(clamp(rb + 48), a0);
            glm_mfma_overlap<D, NS, 2, CB>(a1, bfrag, c1, c0, acc, st);
            ld.full(clamp(rb + 64), a1);
            glm_mfma_overlap<D, NS, 1, CB>(a0, bfrag, c0, c1, acc, st);
        }
        // pairs of blocks (rb - 16, rb): the first's epilogue runs with the
        // products of rb, the second's with those of rb + 16
        for (; rb + 16 < rfull; rb += 32) {
            ld.full(clamp(rb + 16), a0);
            glm_mfma_overlap<D, NS, 0, CB>(a1, bfrag, c1, c0, acc, st);
            ld.full(clamp(rb + 32), a1);
            glm_mfma_overlap<D, NS, 1, CB>(a0, bfrag, c0, c1, acc, st);
        }
        if (rb < rfull) {
            glm_mfma_overlap<D, NS, 0, CB>(a1, bfrag, c1, c0, acc, st);
            glm_epilogue_second(c1, acc, st);
        } else {
            glm_epilogue(c0, acc);
        }
    }
    if (rfull < r1) {
        typename Loader::vec a[KC];
        ld.masked(rfull, r1, a);
        f32x4 c[CB];
        glm_mfma_block<D, NS, CB>(a, bfrag, c);
        glm_epilogue_masked(c, acc, rfull + lg * 4, r1);
    }
}

}  // namespace bh
