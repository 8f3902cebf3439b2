// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): ten rounds of the published round function on a
// 4 x 32-bit counter with a 2 x 32-bit key, and the Box-Muller transform of
// fillWithNormalDistribution (programs/bec_wigner_3d.cuda/wigner.cu:15-34,
// deviation 0.5) evaluated in FP64 with u = (x + 1) / 2^32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bec {

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
    }
}

// Four normal deviates (deviation 0.5) for counter {cell, traj, stream, 0}.
__device__ __forceinline__ void noise4(uint32_t cell, uint32_t traj, uint32_t stream, uint32_t k0, uint32_t k1,
                                       double eta[4])
{
    uint32_t c[4] = {cell, traj, stream, 0u};
    philox4x32_10(c, k0, k1);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const double u1 = ((double)c[2 * p] + 1.0) * (1.0 / 4294967296.0);
        const double u2 = ((double)c[2 * p + 1] + 1.0) * (1.0 / 4294967296.0);
        const double rad = 0.5 * sqrt(-2.0 * log(u1));
        double sn, cs;
        sincospi(2.0 * u2, &sn, &cs);
        eta[2 * p] = rad * cs;
        eta[2 * p + 1] = rad * sn;
    }
}

} // namespace bec

namespace bec {

// Four standard normal deviates for the loss noise of step `step`: Philox
// counter {cell, traj, step, 1 + group}, Box-Muller on 24-bit uniforms
// u = ((x >> 8) + 1) 2^-24, which FP32 represents exactly, so the FP32 and
// FP64 paths start from the same uniforms.
template <typename T>
__device__ __forceinline__ void loss_noise4(uint32_t cell, uint32_t traj, uint32_t step, uint32_t group, uint32_t k0,
                                            uint32_t k1, T z[4])
{
    uint32_t c[4] = {cell, traj, step, 1u + group};
    philox4x32_10(c, k0, k1);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const T u1 = (T)((c[2 * p] >> 8) + 1u) * (T)(1.0 / 16777216.0);
        const T u2 = (T)((c[2 * p + 1] >> 8) + 1u) * (T)(1.0 / 16777216.0);
        T rad, sn, cs;
        if constexpr (sizeof(T) == 4) {
            rad = sqrtf(-2.0f * __logf(u1));
            sincospif(2.0f * u2, &sn, &cs);
        } else {
            rad = sqrt(-2.0 * log(u1));
            sincospi(2.0 * u2, &sn, &cs);
        }
        z[2 * p] = rad * cs;
        z[2 * p + 1] = rad * sn;
    }
}

} // namespace bec
