// Model-source prelude for hipRTC.
//
// User models keep the reference's plugin dialect: C99 functions over REAL
// with the signatures of distributions/distribution.cl:1-4 (logpdf) and
// models.clj:46-51 (loglik), calling the OpenCL built-ins native_exp,
// native_log, pown, lgamma and the NAN macro.  The host compiles them
// unchanged between `#pragma clang force_cuda_host_device begin/end`, so they
// become device functions; this prelude supplies the built-ins.
#pragma once

typedef unsigned int uint;
typedef unsigned long ulong;

#ifndef REAL
#define REAL float
#endif
#ifndef REAL2
#define REAL2 float2
#endif

#ifndef NAN
#define NAN __builtin_nanf("")
#endif
#ifndef INFINITY
#define INFINITY __builtin_inff()
#endif

// native_* are the hardware approximations in OpenCL C; the gfx950
// equivalents are v_exp_f32 / v_log_f32: __expf is exp2(x log2 e) on
// v_exp_f32, and native_log is log2 on v_log_f32 times ln 2 (HIP's __logf
// is the full-precision logf sequence).
#define native_exp(x) __expf(x)
#define native_log(x) (__builtin_amdgcn_logf(x) * 0.693147182f)
#define native_sqrt(x) __builtin_sqrtf(x)
#define native_recip(x) (1.0f / (x))
#define native_divide(x, y) ((x) / (y))

#pragma clang force_cuda_host_device begin
// pown(x, n): x to an integer power by repeated squaring.
inline float pown(float x, int n) {
    float r = 1.0f;
    float b = x;
    unsigned int e = n < 0 ? (unsigned int)(-n) : (unsigned int)n;
    while (e) {
        if (e & 1u) r *= b;
        b *= b;
        e >>= 1;
    }
    return n < 0 ? 1.0f / r : r;
}
#pragma clang force_cuda_host_device end
