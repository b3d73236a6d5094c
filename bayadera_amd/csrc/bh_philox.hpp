// Philox4x32-10 counter-based RNG and the (0,1) float mapping, callable from
// host and device code.  Compiled into the AOT kernels and handed to hipRTC as
// a named header for the per-model kernels.
//
// Algorithm: Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As Easy as
// 1, 2, 3", SC'11 -- philox4x32 with 10 rounds, multipliers 0xD2511F53 and
// 0xCD9E8D57, Weyl key increments 0x9E3779B9 and 0xBB67AE85.  The reference
// calls it through Random123's philox.h (amd-gcn-mcmc-stretch.cl:48-50,
// uniform-sampler.cl:24-31).  On gfx950 each round is two v_mul_hi_u32 /
// v_mul_lo_u32 pairs and four XORs, all in VGPRs.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#define BH_HD __host__ __device__ __forceinline__
#else
#define BH_HD inline
#endif

struct bh_u32x4 {
    unsigned int x, y, z, w;
};

struct bh_f32x4 {
    float x, y, z, w;
};

BH_HD bh_u32x4 bh_philox4x32_10(bh_u32x4 c, unsigned int k0, unsigned int k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const unsigned long long p0 = (unsigned long long)0xD2511F53u * c.x;
        const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c.z;
        const unsigned int hi0 = (unsigned int)(p0 >> 32), lo0 = (unsigned int)p0;
        const unsigned int hi1 = (unsigned int)(p1 >> 32), lo1 = (unsigned int)p1;
        c = bh_u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// (0.5 + (i >> 9)) * 2^-23: exact in fp32, strictly inside (0, 1)
// (uniform-sampler.cl:8-10).
BH_HD float bh_u01(unsigned int i) {
    return (0.5f + (float)(i >> 9)) * 1.1920928955078125E-7f;
}

BH_HD bh_f32x4 bh_u01x4(bh_u32x4 r) {
    return bh_f32x4{bh_u01(r.x), bh_u01(r.y), bh_u01(r.z), bh_u01(r.w)};
}

// Key/counter layouts of the reference's streams.
#define BH_KEY1_STRETCH 0xdecafbadu  // stretch move   (amd-gcn-mcmc-stretch.cl:48)
#define BH_KEY1_INIT 0xdecafaaau     // init / uniform (amd-gcn-mcmc-stretch.cl:150)
#define BH_CTR3 0xbeeff00du
#define BH_CTR1_INIT 0xf00dcafeu
#define BH_CTR2_INIT 0xdeadbeefu
