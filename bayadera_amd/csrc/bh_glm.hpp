// Launchers of the GLM likelihood kernels (bh_glm.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace bh {

// precision of the contraction X.B (values of BH_GLM_* in bayadera_hip.h)
constexpr int kGlmFp64 = 0;
constexpr int kGlmFp32 = 1;
constexpr int kGlmBf16 = 2;
constexpr int kGlmBf16x3 = 3;
constexpr int kGlmFp16x2 = 4;
constexpr int kGlmFp16 = 5;

bool glm_supported_dim(unsigned d);
// the number of points n of a launch must be a multiple of this
unsigned glm_point_quantum(int precision);
// points per thread of the linear-regression likelihood kernel for n points
// of D padded features (1 or 2); its work-groups take 256 times as many
unsigned glm_gauss_points_per_thread(int precision, unsigned D, unsigned n);
// work-groups of the logistic likelihood kernel of a precision that one
// compute unit holds at once (the device's occupancy for its registers)
unsigned glm_blocks_per_cu(int precision, unsigned d);
// partial[r][k] for r < nranges (a multiple of 8), k < n; X bf16 N x d
// (fp16 for kGlmFp16x2), P fp32 n x d (fp32/fp64 kernels), ys the 16-bit
// split of yscale*P (NS x n x d, MFMA kernels; fp16 terms for kGlmFp16x2).
hipError_t launch_glm_logistic(int precision, hipStream_t s, const void* X, unsigned long long N,
                               unsigned d, const float* P, const void* ys, unsigned n,
                               unsigned nranges, unsigned long long rows_per_range, float yscale,
                               double* partial);
// Issue-rate kernels (which = 0: v_exp_f32, 1: v_fma_f32, 8 x iters
// operations per thread; 2: v_mfma_f32_16x16x32_f16, 8 x iters per wave),
// out: blocks x 256 floats.
hipError_t launch_glm_rate(hipStream_t s, int which, float* out, unsigned blocks, int iters);
// n16 random 16-byte gathers from src (n16 x 16 bytes; n16 a power of two and
// a multiple of 256), one per thread, sums to out (n16 floats); nt:
// non-temporal loads.
hipError_t launch_glm_gather(hipStream_t s, const void* src, float* out, unsigned n16,
                             bool nt = false);
// fp16 copy of the N x d bf16 X (d = 32 or 64) in the fragment order the fp16
// matrix-core kernels read (glm_x_f16_bytes bytes: whole 16-row blocks, the
// padding rows zero); *overflow is incremented by the number of values
// outside the fp16 range.
size_t glm_x_f16_bytes(unsigned long long N, unsigned d);
hipError_t launch_glm_x_to_f16(hipStream_t s, const void* X, unsigned long long N, unsigned d,
                               void* out, unsigned* overflow);
// fp64 softplus sums of the points listed by the guarded accept test: pts
// (double, 2 lmax x d) holds in row 2i the proposal of listed walker i and
// in row 2i + 1 its position (c = min(*count, lmax) listed, read on the
// device); partial64[r * 2 lmax + row] for r < nranges, rows < 2c, one
// work-group per row range.  Does nothing when the list is empty.  lmax a
// multiple of 8.
// Gradient of the logistic softplus sum at the mean of the first nref
// (1 .. 256) points of pts (device, row stride d = 32 or 64), natural units,
// minus off[d] when given, into g0[d] (double); work holds
// glm_grad_work_doubles(d) doubles.  The split precisions' first-order
// correction: their log posterior at y plus g0 . (y' - y), y' the
// coefficients the split represents (off: the linear half (1/2) X'1, which
// the kernels add at the exact y).
constexpr unsigned kGlmGradBlocks = 512;
inline size_t glm_grad_work_doubles(unsigned d) { return (size_t)kGlmGradBlocks * d + d; }
hipError_t launch_glm_grad(hipStream_t s, const void* X, unsigned long long N, unsigned d,
                           const float* pts, unsigned nref, const double* off, double* work,
                           double* g0);
// The reference point launch_glm_grad used (d floats inside its work).
inline const float* glm_grad_bref(const double* work, unsigned d) {
    return reinterpret_cast<const float*>(work + (size_t)kGlmGradBlocks * d);
}
// Hessian of the logistic softplus sum at bref (device, d floats), natural
// units, into H0 (d x d floats); work holds glm_hess_work_doubles(d)
// doubles.  The one-term fp16 precision's second-order correction.
constexpr unsigned kGlmHessBlocks = 1024;
inline size_t glm_hess_work_doubles(unsigned d) {
    return (size_t)(kGlmHessBlocks + kGlmHessBlocks / 32) * d * d;
}
hipError_t launch_glm_hess(hipStream_t s, const void* X, unsigned long long N, unsigned d,
                           const float* bref, double* work, float* H0);
hipError_t launch_glm_fp64_list(hipStream_t s, const void* X, unsigned long long N, unsigned d,
                                const double* pts, const unsigned* count, unsigned lmax,
                                unsigned nranges, double* partial64);
// Linear regression sums of squared residuals (fp32 or fp64 kernels): X fp32
// N x D (D = 16, 32 or 64, columns past d zero), points P with row stride
// pstride whose first d coordinates are the coefficients.
hipError_t launch_glm_gauss(int precision, hipStream_t s, const float* X, const float* yv,
                            unsigned long long N, unsigned D, const float* P, unsigned pstride,
                            unsigned d, unsigned n, unsigned nranges,
                            unsigned long long rows_per_range, double* partial);
hipError_t launch_glm_lin(hipStream_t s, const void* X, const float* yv, unsigned long long N,
                          unsigned d, double* part, unsigned chunks, double* lin);

}  // namespace bh
