/*
 * rng.h -- the linear congruential generator behind Chameleon's matrix
 * generators (coreblas/include/coreblas/random.h:32-73): state
 * x <- a x + c with a = 6364136223846793005, c = 1, jump-ahead by binary
 * powering of the affine map, value 0.5 - x * 2^-64.
 */
#ifndef CHAMELEON_AMD_RNG_H
#define CHAMELEON_AMD_RNG_H

#include <hip/hip_runtime.h>

#define RND_A 6364136223846793005ULL
#define RND_C 1ULL

/* state after n steps from x */
__device__ __forceinline__ unsigned long long rnd_jump(unsigned long long n, unsigned long long x)
{
    unsigned long long a = RND_A, c = RND_C;
    while (n) {
        if (n & 1ULL) x = a * x + c;
        c = c * (a + 1ULL);
        a = a * a;
        n >>= 1;
    }
    return x;
}

__device__ __forceinline__ double rnd_val(unsigned long long x)
{
    return 0.5 - (double)x * 5.4210108624275222e-20;
}

#endif
