// Launchers of the ahead-of-time kernels (bh_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace bh {

constexpr unsigned kMaxDatasetBlocks = 2048;  // 8 blocks per CU over 256 CUs
// largest work-group the context accepts (WGS); the reduction trees of acor and
// the per-step means are sized for it
constexpr unsigned kMaxGroupThreads = 1024;

// bh_dataset_variance_noalloc: the rows one call may have, and the device
// array (one per GPU, held until the process ends) that takes the place of
// the buffers bh_dataset_variance allocates: kMaxDatasetBlocks x rows double
// partial sums, then the row means.
constexpr unsigned kNoAllocMaxRows = 256;
constexpr size_t kNoAllocPartialBytes = (size_t)kMaxDatasetBlocks * kNoAllocMaxRows * sizeof(double);
constexpr size_t kNoAllocScratchBytes = kNoAllocPartialBytes + kNoAllocMaxRows * sizeof(float);
// Looks up the array's address on `device`, which must be the current
// device; once per device, later calls return at once.
hipError_t noalloc_scratch_init(int device);
// The address noalloc_scratch_init found for `device`.
hipError_t noalloc_scratch(int device, void** ptr);

// dst[i] = src[i] for n16 16-byte pieces (nt: non-temporal loads and stores)
hipError_t launch_copy16(hipStream_t s, const void* src, void* dst, unsigned long long n16,
                         bool nt);
hipError_t launch_init_walkers(hipStream_t s, unsigned seed, const float* limits, unsigned dim,
                               unsigned walkers, float* xs);
hipError_t launch_uniform_sample(hipStream_t s, unsigned seed, float lo, float hi, unsigned quads,
                                 float* x);
hipError_t launch_direct_sample(hipStream_t s, int kind, unsigned seed, float p0, float p1,
                                unsigned quads, float* x);
size_t dataset_scratch_bytes(unsigned m, unsigned long long n);
hipError_t launch_min_max(hipStream_t s, const float* a, unsigned m, unsigned long long n,
                          unsigned ld, float* limits, void* scratch);
hipError_t launch_histogram(hipStream_t s, const float* a, unsigned m, unsigned long long n,
                            unsigned ld, const float* limits, unsigned bins, unsigned* counts);
hipError_t launch_histogram_finalize(hipStream_t s, const unsigned* counts, const float* limits,
                                     float alpha, unsigned m, unsigned bins, float* pdf,
                                     float* ranks);
hipError_t launch_moments(hipStream_t s, const float* a, unsigned m, unsigned long long n,
                          unsigned ld, float* mean, float* var, void* scratch);
hipError_t launch_row_tree_sum(hipStream_t s, const float* a, unsigned rows, unsigned count,
                               unsigned local_n, unsigned long long row_stride, float* work,
                               float scale, float* out, unsigned long long out_stride);
hipError_t launch_sum_u32(hipStream_t s, const unsigned* v, unsigned n, unsigned long long* out);
hipError_t launch_sum_f64(hipStream_t s, const double* v, unsigned n, double* out);
hipError_t launch_acor_batch(hipStream_t s, float* series, const unsigned long long* col_off,
                             const unsigned* col_stride, unsigned cols, unsigned n, unsigned lag,
                             unsigned wgs, float* scratch, unsigned scratch_per_dim, float* tau,
                             float* mean, float* sigma);
hipError_t launch_acor(hipStream_t s, float* series, unsigned dim, unsigned n, unsigned lag,
                       unsigned wgs, float* scratch, unsigned scratch_per_dim, float* tau,
                       float* mean, float* sigma);

}  // namespace bh
