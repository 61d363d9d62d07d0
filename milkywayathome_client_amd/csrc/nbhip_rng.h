/*
 * Counter-based random numbers for the initial-condition sampler:
 * Philox4x32-10 (J. Salmon, M. Moraes, R. Dror, D. Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64-bit
 * multiplications with the multipliers 0xD2511F53 and 0xCD9E8D57 and a Weyl
 * key schedule (0x9E3779B9, 0xBB67AE85).  A stream is a key plus a counter,
 * so every body owns an independent sequence that any thread can compute.
 */
#ifndef NBHIP_RNG_H
#define NBHIP_RNG_H

#include <hip/hip_runtime.h>

__host__ __device__ inline void nbh_philox4x32_10(const unsigned* ctr, const unsigned* key,
                                                   unsigned* out)
{
    unsigned c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    unsigned k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const unsigned long long p0 = (unsigned long long) 0xD2511F53u * c0;
        const unsigned long long p1 = (unsigned long long) 0xCD9E8D57u * c2;
        const unsigned n0 = (unsigned) (p1 >> 32) ^ c1 ^ k0;
        const unsigned n2 = (unsigned) (p0 >> 32) ^ c3 ^ k1;
        c1 = (unsigned) p1;
        c3 = (unsigned) p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* A double in (0, 1) from 53 bits of two words: (k + 1/2) 2^-53. */
__host__ __device__ inline double nbh_u01(unsigned hi, unsigned lo)
{
    const unsigned long long k = (((unsigned long long) hi << 32) | lo) >> 11;
    return ((double) k + 0.5) * 1.1102230246251565404e-16;
}

#endif
