// Launchers of the batched GLM kernels (bh_glm_batch.hip).  Layouts: walkers
// xs[P][W][64] (padded floats, halves contiguous), per-half proposals
// Y[P][K][64], z/u2[P][K], scaled points ysc[P][n][64] and their 16-bit
// splits [NS][P][n][64] (bf16, or fp16 with f16 set); units {problem, r0, r1, global unit id} (uint4) per row
// stride; partial[unit][n] (double).
#pragma once
#include <hip/hip_runtime.h>

namespace bh {

// rows of one unit of the batched contraction (a multiple of 16) until the
// batch's tile shapes are set or autotuned
constexpr unsigned kGbUnitRows = 2048;

// f16 = 0: the fp32 path's decoding; 1: the fp16 matrix-core path's.
hipError_t launch_gb_fp8_decode(hipStream_t s, const void* in, float* out, unsigned long long n,
                                int f16);
hipError_t launch_gb_init(hipStream_t s, unsigned seed, const unsigned* dims, const float* limits,
                          float* xs, unsigned P, unsigned W);
hipError_t launch_gb_points(hipStream_t s, const float* xs, const unsigned* dims,
                            const float* tscale, unsigned P, unsigned n, float* ysc, void* ysplit,
                            int nsplit, int f16);
hipError_t launch_gb_propose(hipStream_t s, unsigned seed0, int half, unsigned oe, unsigned step,
                             float a, const unsigned* dims, const float* tscale, const float* xs,
                             unsigned P, unsigned K, float* Y, float* z, float* u2, float* ysc,
                             void* ysplit, int nsplit, int f16);
// precision: kGlmFp64 / kGlmFp32 / kGlmBf16 / kGlmBf16x3 / kGlmFp16x2; D: 32
// or 64; CB: 16-point column blocks per wave of the matrix-core kernels (4,
// or 8 for D = 32; n a multiple of 64 CB; the fp32 and fp64 kernels take
// CB = 4 and n a multiple of 256).
// Ypts: the unscaled points (fp64 kernel), ysc / ysplit the scaled ones.
hipError_t launch_gb_contract(int precision, int D, int CB, hipStream_t s, const void* X8,
                              const unsigned long long* xoff, const void* units, unsigned nunits,
                              unsigned P, const float* scales, const float* Ypts,
                              const float* ysc, const void* ysplit, unsigned n, double* partial);
hipError_t launch_gb_logpost(hipStream_t s, const float* xs, const unsigned* dims,
                             const unsigned* ubeg, const double* partial, unsigned P, unsigned n,
                             const double* lin, double scale, double prior_prec, double* lp);
hipError_t launch_gb_accept(hipStream_t s, int half, const unsigned* dims, const unsigned* ubeg,
                            const float* Y, const float* z, const float* u2,
                            const double* partial, const double* lin, double scale,
                            double prior_prec, float* xs, double* lp, float beta, unsigned P,
                            unsigned K, int accu, unsigned long long* acc_cnt, float* msum,
                            float* msum2);
hipError_t launch_gb_step(hipStream_t s, const unsigned* dims, const float* msum,
                          const float* msum2, unsigned P, unsigned W, double* acc1, double* acc2,
                          const unsigned long long* soff, float* series, unsigned st);

}  // namespace bh
