/*
 * mfma.h -- CDNA4 (gfx950) matrix-core helpers shared by the tile kernels.
 *
 * v_mfma_f64_16x16x4_f64 (one wavefront of 64 lanes):
 *   A operand  lane l supplies A[row = l % 16][k = l / 16]   (16 x 4)
 *   B operand  lane l supplies B[k = l / 16][col = l % 16]   (4 x 16)
 *   result     slot s of lane l holds D[row = l/16 + 4 s][col = l % 16]
 * v_mfma_f32_16x16x4_f32 has the same operand layout; its result slot s of
 * lane l holds D[row = 4 (l/16) + s][col = l % 16].
 * (Both layouts were checked on the device with a known-answer product.)
 */
#ifndef CHAMELEON_AMD_MFMA_H
#define CHAMELEON_AMD_MFMA_H

#include <hip/hip_runtime.h>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float  v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));   /* native vector: stays in registers */

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4f mfma_f32(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

/* C -= v for an element that this launch updates exactly once: either a
 * load, subtract and store, or (atomic) one FP64 atomic add of -v done in
 * the L2 without a return value -- the same rounding, bit for bit, and the
 * wave does not wait for C's old value, so the tile's read-modify-write
 * streams while the next blocks compute. */
__device__ __forceinline__ void c_sub(double *p, double v, int atomic)
{
    if (atomic) atomicAdd(p, -v);
    else *p -= v;
}

/* result row of slot s for lane l */
__device__ __forceinline__ int mfma64_row(int l, int s) { return (l >> 4) + 4 * s; }
__device__ __forceinline__ int mfma32_row(int l, int s) { return 4 * (l >> 4) + s; }

/* Value of v in lane l (uniform l), as a scalar broadcast: the readlane
 * builtin moves 32 bits, so the double travels as its two halves. */
__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ v4d zero4d() { v4d z = {0.0, 0.0, 0.0, 0.0}; return z; }
__device__ __forceinline__ v4f zero4f() { v4f z = {0.f, 0.f, 0.f, 0.f}; return z; }

#endif
