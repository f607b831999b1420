/*
 * mixed.hip -- kernels of the mixed-precision Cholesky solve with FP64
 * iterative refinement (no counterpart in Chameleon: its dsposv is only a
 * commented-out prototype, include/chameleon/chameleon_zc.h:34,42; the
 * per-tile demotion machinery it has is compute/pzhered.c:176-288).
 *
 * The working matrix stays in FP64 storage; after the FP64 panel of step k
 * is factored, its tiles are demoted into a low-precision panel buffer, and
 * the trailing update multiplies demoted operands on the low-precision MFMA
 * units with FP32 accumulation, subtracting the result from the FP64 tiles.
 *   LP_F32    FP32 operands, v_mfma_f32_16x16x4_f32
 *   LP_BF16   BF16 operands, v_mfma_f32_16x16x32_bf16
 *   LP_BF16X3 each operand split x = hi + lo (two BF16), three products
 *             hi*hi + hi*lo + lo*hi, relative error ~2^-16 per product
 *   LP_FP8    OCP E4M3 operands with one power-of-two scale per tile,
 *             v_mfma_scale_f32_16x16x128_f8f6f4 (the scale applied by the
 *             MFMA)
 * Also: the FP64 symmetric residual R = B - A X and column max-norms used by
 * the refinement loop.
 */
#include "../internal.h"
#include "mfma.h"

#include <algorithm>

namespace chx {
namespace k {

typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned short v8u16 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int xcd_remap_m(int id, int total)
{
    const int nx = 8;
    int per = total / nx, rem = total % nx;
    int x = id % nx, local = id / nx;
    return (x < rem) ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
}

/* round-to-nearest-even FP32 -> BF16 bits (finite inputs) */
__device__ __forceinline__ unsigned short bf16_bits(float f)
{
    unsigned int u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_float(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }

/* ------------------------------------------------------------------ */
/* Demote tiles (m, k), m0 <= m < m1, of the FP64 mirror into the panel
 * buffer.  F32: column-major floats, tile m at P + m*ts.  BF16 / BF16X3:
 * row-major BF16 (element (i, p) at p + i*ldt), hi plane at P + m*ts and,
 * for X3, lo plane at P + (NT + m)*ts (in BF16 elements). */
__global__ __launch_bounds__(256) void panel_demote_kernel(const double *A, int NT, int ldt, int k, int m0,
                                                            void *P, int mode)
{
    const size_t ts = (size_t)ldt * ldt;
    const int m = m0 + blockIdx.z;
    const double *T = A + ((size_t)k * NT + m) * ts;
    const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    if (mode == 0) {
        float *D = (float *)P + (size_t)m * ts;
        for (int jj = ty; jj < 32; jj += 8) {
            if (r0 + tx >= ldt || c0 + jj >= ldt) continue;
            const size_t e = (r0 + tx) + (size_t)(c0 + jj) * ldt;
            D[e] = (float)T[e];
        }
        return;
    }
    unsigned short *H = (unsigned short *)P + (size_t)m * ts;
    unsigned short *Lo = (unsigned short *)P + ((size_t)NT + m) * ts;
    for (int jj = ty; jj < 32; jj += 8) {
        const int i = r0 + jj, p = c0 + tx;
        if (i >= ldt || p >= ldt) continue;
        /* split from the FP64 value for the lo part: hi = bf16(x), lo = bf16(x - hi) */
        const double x = T[i + (size_t)p * ldt];
        const unsigned short hb = bf16_bits((float)x);
        H[p + (size_t)i * ldt] = hb;
        if (mode == 2) Lo[p + (size_t)i * ldt] = bf16_bits((float)(x - (double)bf16_to_float(hb)));
    }
}

void panel_demote(const double *A, int NT, int ldt, int k, int m0, int m1, void *P, int mode, hipStream_t s)
{
    if (m1 <= m0) return;
    dim3 grid((ldt + 31) / 32, (ldt + 31) / 32, m1 - m0);
    hipLaunchKernelGGL(panel_demote_kernel, grid, dim3(256), 0, s, A, NT, ldt, k, m0, P, mode);
}

/* Tile-pair decoding shared by the update kernels (same enumeration as the
 * FP64 trailing_update_kernel: ncol columns from n0, or the triangle
 * n0..NT-1 when ncol is 0). */
__device__ __forceinline__ void decode_pair(int pair, int n0, int ncol, int NT, int &m, int &n)
{
    if (ncol > 0) {
        int q = pair, len = NT - n0;
        n = n0;
        while (q >= len && n + 1 < n0 + ncol) {
            q -= len;
            ++n;
            --len;
        }
        m = n + q;
        return;
    }
    int i = (int)((sqrt(8.0 * (double)pair + 1.0) - 1.0) * 0.5);
    while ((long long)(i + 1) * (i + 2) / 2 <= pair) ++i;
    while ((long long)i * (i + 1) / 2 > pair) --i;
    m = n0 + i;
    n = n0 + (pair - i * (i + 1) / 2);
}

/* ------------------------------------------------------------------ */
/* FP32-operand trailing update: A(m,n) -= P(m) P(n)^T, 128 x 128 block, 8
 * wavefronts of 64 x 32, K in 32-deep LDS slices (double-buffered). */
template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, 2) void update_f32_kernel(
    double *A, const float *P, const float *P2, int lmt, int ldt, int n0, int ncol, int total_blocks,
    const int *info, const unsigned char *pmap, int atomic_c)
{
    constexpr int NW = (BM / WM) * (BN / WN), NTH = NW * 64, BK = 32;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LA = (BK * BM) / (4 * NTH), LB = (BK * BN) / (4 * NTH);
    static_assert(LA >= 1 && LB >= 1, "slice too small");
    __shared__ float As[2][BK * BM];
    __shared__ float Bs[2][BK * BN];
    if (*info != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bid = xcd_remap_m(blockIdx.x, total_blocks);
    const int tbm = ldt / BM, tbn = ldt / BN, bpt = tbm * tbn;
    int m, n;
    decode_pair(bid / bpt, n0, ncol, lmt, m, n);   /* the tile grid is lmt x lmt, lmt = NT */
    if (pmap && pmap[m + (size_t)n * lmt] != 1) return;   /* per-tile precision map: FP32 tiles */
    const int b = bid % bpt, row0 = (b % tbm) * BM, col0 = (b / tbm) * BN;
    const bool diag = (m == n);
    if (diag && col0 >= row0 + BM) return;
    const size_t ts = (size_t)ldt * ldt;
    const float *Am = P + (size_t)m * ts + row0;
    const float *An = P + (size_t)n * ts + col0;
    double *C = A + ((size_t)n * lmt + m) * ts;
    const int wm = (wave % (BM / WM)) * WM, wn = (wave / (BM / WM)) * WN;
    v4f acc[FM][FN];
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = zero4f();
    v4f ra[LA], rb[LB];
    /* slices of the first panel, then (P2 given) of the second */
    const int SK = ldt / BK;
    const ptrdiff_t d2 = P2 ? P2 - P : 0;
    auto gload = [&](int st) {
        const int pi = st / SK, kb = (st - pi * SK) * BK;
        const float *am = Am + pi * d2, *an = An + pi * d2;
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int e = (it * NTH + tid) * 4;
            ra[it] = *reinterpret_cast<const v4f *>(am + (e % BM) + (size_t)(kb + e / BM) * ldt);
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int e = (it * NTH + tid) * 4;
            rb[it] = *reinterpret_cast<const v4f *>(an + (e % BN) + (size_t)(kb + e / BN) * ldt);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int it = 0; it < LA; ++it) *reinterpret_cast<v4f *>(&As[buf][(it * NTH + tid) * 4]) = ra[it];
#pragma unroll
        for (int it = 0; it < LB; ++it) *reinterpret_cast<v4f *>(&Bs[buf][(it * NTH + tid) * 4]) = rb[it];
    };
    const int S = (P2 ? 2 : 1) * SK;
    gload(0);
    lstore(0);
    if (S > 1) gload(1);
    __syncthreads();
    for (int st = 0; st < S; ++st) {
        const int cur = st & 1;
        if (st + 1 < S) {
            lstore(cur ^ 1);
            if (st + 2 < S) gload(st + 2);
        }
        const float *as = As[cur], *bs = Bs[cur];
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            float a[FM], bv[FN];
#pragma unroll
            for (int fm = 0; fm < FM; ++fm) a[fm] = as[(kk + (lane >> 4)) * BM + wm + fm * 16 + (lane & 15)];
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) bv[fn] = bs[(kk + (lane >> 4)) * BN + wn + fn * 16 + (lane & 15)];
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = mfma_f32(bv[fn], a[fm], acc[fm][fn]);
        }
        __syncthreads();
    }
    /* FP32 16x16x4 result slot s of lane l: row 4 (l/16) + s of D = C^T */
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int gi = row0 + wm + fm * 16 + (lane & 15);
                const int gj = col0 + wn + fn * 16 + 4 * (lane >> 4) + s;
                if (!diag || gi >= gj) c_sub(&C[gi + (size_t)gj * ldt], (double)acc[fm][fn][s], atomic_c);
            }
}

/* ------------------------------------------------------------------ */
/* BF16-operand trailing update (SPLIT = 3: hi*hi + hi*lo + lo*hi).  Panel
 * planes are row-major (k contiguous).  Each lane of a 16x16x32 MFMA
 * supplies 8 consecutive k of one row (the same k assignment on both
 * operands, so the dot product is exact in any order).
 *
 * LDS rows are 128 bytes (64 BF16) with the 16-byte chunk q of row r
 * stored at chunk q ^ ((r >> 1) & 7): a slice store (two rows per 16 lanes)
 * and an operand read (16 rows, one chunk each per 16 lanes) both touch 16
 * different 4-bank groups.  The next slice is loaded into registers while
 * the current one is multiplied.  Both measured together: 246 -> 274
 * TFLOP/s FP64-equivalent on one paired launch (neither alone helps). */
template <int SPLIT>
__global__ __launch_bounds__(512, 4) void update_bf16_kernel(double *A, const unsigned short *P,
                                                             const unsigned short *P2, int NT, int lmt, int ldt,
                                                             int n0, int ncol, int total_blocks, const int *info,
                                                             const unsigned char *pmap, int atomic_c)
{
    constexpr int BM = 128, BN = 128, WM = 64, WN = 32, BK = 64, LDK = BK;
    constexpr int FM = WM / 16, FN = WN / 16, NTH = 512;
    constexpr int NP = (SPLIT == 3) ? 2 : 1;          /* planes per operand */
    constexpr int LA = (BM * BK) / (8 * NTH), LB = (BN * BK) / (8 * NTH);
    __shared__ unsigned short As[NP][BM * LDK];
    __shared__ unsigned short Bs[NP][BN * LDK];
    if (*info != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int bid = xcd_remap_m(blockIdx.x, total_blocks);
    const int tbm = ldt / BM, tbn = ldt / BN, bpt = tbm * tbn;
    int m, n;
    decode_pair(bid / bpt, n0, ncol, NT, m, n);
    if (pmap && pmap[m + (size_t)n * NT] != 2) return;   /* per-tile precision map: BF16 tiles */
    const int b = bid % bpt, row0 = (b % tbm) * BM, col0 = (b / tbm) * BN;
    const bool diag = (m == n);
    if (diag && col0 >= row0 + BM) return;
    const size_t ts = (size_t)ldt * ldt;
    double *C = A + ((size_t)n * lmt + m) * ts;
    const int wm = (wave % (BM / WM)) * WM, wn = (wave / (BM / WM)) * WN;
    /* element offset of 16-byte chunk q8 (k / 8) of LDS row r */
    auto soff = [](int r, int q8) { return r * LDK + 8 * (q8 ^ ((r >> 1) & 7)); };
    v4f acc[FM][FN];
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = zero4f();
    const int K = (P2 ? 2 : 1) * ldt;      /* the second panel (P2) follows the first */
    v8u16 ra[NP][LA], rb[NP][LB];
    /* slice kb2: BM (BN) rows x BK k of each plane, 8 BF16 per load */
    auto gload = [&](int kb2) {
        const unsigned short *Pp = kb2 < ldt ? P : P2;
        const int kb = kb2 < ldt ? kb2 : kb2 - ldt;
#pragma unroll
        for (int pl = 0; pl < NP; ++pl) {
            const unsigned short *Pm = Pp + ((size_t)pl * NT + m) * ts;
            const unsigned short *Pn = Pp + ((size_t)pl * NT + n) * ts;
#pragma unroll
            for (int it = 0; it < LA; ++it) {
                const int e = (it * NTH + tid) * 8, r = e / BK, q = e % BK;
                ra[pl][it] = *reinterpret_cast<const v8u16 *>(Pm + (size_t)(row0 + r) * ldt + kb + q);
            }
#pragma unroll
            for (int it = 0; it < LB; ++it) {
                const int e = (it * NTH + tid) * 8, r = e / BK, q = e % BK;
                rb[pl][it] = *reinterpret_cast<const v8u16 *>(Pn + (size_t)(col0 + r) * ldt + kb + q);
            }
        }
    };
    gload(0);
    for (int kb2 = 0; kb2 < K; kb2 += BK) {
#pragma unroll
        for (int pl = 0; pl < NP; ++pl) {
#pragma unroll
            for (int it = 0; it < LA; ++it) {
                const int e = (it * NTH + tid) * 8;
                *reinterpret_cast<v8u16 *>(&As[pl][soff(e / BK, (e % BK) / 8)]) = ra[pl][it];
            }
#pragma unroll
            for (int it = 0; it < LB; ++it) {
                const int e = (it * NTH + tid) * 8;
                *reinterpret_cast<v8u16 *>(&Bs[pl][soff(e / BK, (e % BK) / 8)]) = rb[pl][it];
            }
        }
        __syncthreads();
        if (kb2 + BK < K) gload(kb2 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 32) {
            v8bf ah[FM], bh[FN], al[FM], bl[FN];
#pragma unroll
            for (int fm = 0; fm < FM; ++fm) {
                ah[fm] = *reinterpret_cast<const v8bf *>(&As[0][soff(wm + fm * 16 + li, kk / 8 + lk)]);
                if (SPLIT == 3) al[fm] = *reinterpret_cast<const v8bf *>(&As[NP - 1][soff(wm + fm * 16 + li, kk / 8 + lk)]);
            }
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) {
                bh[fn] = *reinterpret_cast<const v8bf *>(&Bs[0][soff(wn + fn * 16 + li, kk / 8 + lk)]);
                if (SPLIT == 3) bl[fn] = *reinterpret_cast<const v8bf *>(&Bs[NP - 1][soff(wn + fn * 16 + li, kk / 8 + lk)]);
            }
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < FN; ++fn)
                    acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[fn], ah[fm], acc[fm][fn], 0, 0, 0);
            if (SPLIT == 3) {
#pragma unroll
                for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                    for (int fn = 0; fn < FN; ++fn)
                        acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[fn], ah[fm], acc[fm][fn], 0, 0, 0);
#pragma unroll
                for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                    for (int fn = 0; fn < FN; ++fn)
                        acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[fn], al[fm], acc[fm][fn], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int gi = row0 + wm + fm * 16 + li;
                const int gj = col0 + wn + fn * 16 + 4 * lk + s;
                if (!diag || gi >= gj) c_sub(&C[gi + (size_t)gj * ldt], (double)acc[fm][fn][s], atomic_c);
            }
}

/* ------------------------------------------------------------------ */
/* FP8 (OCP E4M3: 4 exponent bits, bias 7, 3 mantissa bits, largest finite
 * 448, subnormals from 2^-9) operands with one power-of-two scale per
 * demoted tile.  Tile m of panel k is stored as bytes e4m3(x 2^e_m),
 * row-major (k contiguous), e_m = 7 - floor(log2 max|tile|) so the scaled
 * maximum lies in [128, 256); the update kernel hands 2^-e_m to the MFMA as
 * its per-operand E8M0 scale (v_mfma_scale_f32_16x16x128_f8f6f4), so the
 * products come out unscaled in FP32.  Panel buffer layout: NT tiles of
 * ldt^2 bytes, then NT tile maxima (float bits), then NT exponents e_m.  */

/* round-to-nearest-even into E4M3; |x| >= 464 saturates to 448 */
__device__ __forceinline__ unsigned char e4m3_rne(double x)
{
    const unsigned char sign = (x < 0.0) ? 0x80 : 0x00;
    const double a = fabs(x);
    if (!(a < 464.0)) return sign | 0x7E;
    if (a < 0x1p-6) return sign | (unsigned char)rint(a * 0x1p9);   /* 8 encodes 2^-6 */
    int E = ilogb(a);
    double r = rint(ldexp(a, 3 - E));
    if (r == 16.0) {
        r = 8.0;
        ++E;
    }
    return sign | (unsigned char)(((E + 7) << 3) | ((int)r - 8));
}

__global__ __launch_bounds__(256) void tile_absmax_kernel(const double *A, int NT, int ldt, int k, int m0,
                                                           unsigned *maxbits)
{
    const size_t ts = (size_t)ldt * ldt;
    const int m = m0 + blockIdx.z;
    const double *T = A + ((size_t)k * NT + m) * ts;
    const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    float mx = 0.f;
    for (int jj = ty; jj < 32; jj += 8)
        if (r0 + tx < ldt && c0 + jj < ldt) mx = nan_maxf(mx, (float)fabs(T[(r0 + tx) + (size_t)(c0 + jj) * ldt]));
    /* non-negative floats order like their bits; a NaN compares above inf */
    unsigned u = __float_as_uint(mx);
    if (mx != mx) u = 0x7fc00000u;
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned o = __shfl_xor(u, off);
        u = o > u ? o : u;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(&maxbits[m], u);
}

__global__ __launch_bounds__(256) void demote_fp8_kernel(const double *A, int NT, int ldt, int k, int m0,
                                                          unsigned char *P, const unsigned *maxbits, int *ex,
                                                          int *info)
{
    const size_t ts = (size_t)ldt * ldt;
    const int m = m0 + blockIdx.z;
    const double *T = A + ((size_t)k * NT + m) * ts;
    const float mx = __uint_as_float(maxbits[m]);
    int e = 0;
    if (mx > 0.f && mx <= 3.0e38f) e = 7 - ilogbf(mx);
    e = max(-126, min(126, e));
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        ex[m] = e;
        if (!(mx <= 3.0e38f)) atomicCAS(info, 0, LP_RANGE);    /* inf / NaN: no FP8 scale fits */
    }
    unsigned char *D = P + (size_t)m * ts;
    const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int jj = ty; jj < 32; jj += 8) {
        const int i = r0 + jj, p = c0 + tx;
        if (i >= ldt || p >= ldt) continue;
        D[p + (size_t)i * ldt] = e4m3_rne(ldexp(T[i + (size_t)p * ldt], e));
    }
}

void panel_demote_fp8(const double *A, int NT, int ldt, int k, int m0, int m1, void *P, int *info, hipStream_t s)
{
    if (m1 <= m0) return;
    const size_t ts = (size_t)ldt * ldt;
    unsigned *maxbits = (unsigned *)((unsigned char *)P + (size_t)NT * ts);
    int *ex = (int *)(maxbits + NT);
    zero_ints((int *)maxbits + m0, m1 - m0, s);
    dim3 grid((ldt + 31) / 32, (ldt + 31) / 32, m1 - m0);
    hipLaunchKernelGGL(tile_absmax_kernel, grid, dim3(256), 0, s, A, NT, ldt, k, m0, maxbits);
    hipLaunchKernelGGL(demote_fp8_kernel, grid, dim3(256), 0, s, A, NT, ldt, k, m0, (unsigned char *)P, maxbits, ex,
                       info);
}

typedef int v8i32 __attribute__((ext_vector_type(8)));
typedef int v4i32 __attribute__((ext_vector_type(4)));

/* FP8-operand trailing update: 128 x 128 blocks, 8 wavefronts of 64 x 32,
 * k in 128-byte slices (one 16x16x128 MFMA deep), LDS rows of 128 bytes
 * with the 16-byte chunk q of row r at q ^ ((r >> 1) & 7) (the slice
 * stores and the operand reads of 16 rows x 32 bytes are free of bank
 * conflicts), next slice prefetched into registers.  Lane (li, lk)
 * supplies bytes 32 lk .. 32 lk + 31 of its row for both operands (the
 * same k assignment on both, so the sum over k is complete). */
__global__ __launch_bounds__(512, 2) void update_fp8_kernel(double *A, const unsigned char *P,
                                                             const unsigned char *P2, int NT, int lmt, int ldt,
                                                             int n0, int ncol, int total_blocks, const int *info,
                                                             int atomic_c)
{
    constexpr int BM = 128, BN = 128, WM = 64, WN = 32, BK = 128;
    constexpr int FM = WM / 16, FN = WN / 16, NTH = 512;
    constexpr int LA = (BM * BK) / (16 * NTH), LB = (BN * BK) / (16 * NTH);
    __shared__ unsigned char As[BM * BK];
    __shared__ unsigned char Bs[BN * BK];
    if (*info != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int bid = xcd_remap_m(blockIdx.x, total_blocks);
    const int tbm = ldt / BM, tbn = ldt / BN, bpt = tbm * tbn;
    int m, n;
    decode_pair(bid / bpt, n0, ncol, NT, m, n);
    const int b = bid % bpt, row0 = (b % tbm) * BM, col0 = (b / tbm) * BN;
    const bool diag = (m == n);
    if (diag && col0 >= row0 + BM) return;
    const size_t ts = (size_t)ldt * ldt;
    double *C = A + ((size_t)n * lmt + m) * ts;
    const int wm = (wave % (BM / WM)) * WM, wn = (wave / (BM / WM)) * WN;
    auto soff = [](int r, int q) { return r * BK + 16 * (q ^ ((r >> 1) & 7)); };
    /* E8M0 descale bytes of each panel: 2^-e as 127 - e */
    const int *ex1 = (const int *)(P + (size_t)NT * ts + (size_t)NT * 4);
    const int sa1 = 127 - ex1[m], sb1 = 127 - ex1[n];
    int sa2 = sa1, sb2 = sb1;
    if (P2) {
        const int *ex2 = (const int *)(P2 + (size_t)NT * ts + (size_t)NT * 4);
        sa2 = 127 - ex2[m];
        sb2 = 127 - ex2[n];
    }
    v4f acc[FM][FN];
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = zero4f();
    const int K = (P2 ? 2 : 1) * ldt;
    v4i32 ra[LA], rb[LB];
    auto gload = [&](int kb2) {
        const unsigned char *Pp = kb2 < ldt ? P : P2;
        const int kb = kb2 < ldt ? kb2 : kb2 - ldt;
        const unsigned char *Pm = Pp + (size_t)m * ts, *Pn = Pp + (size_t)n * ts;
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int e = (it * NTH + tid) * 16, r = e / BK, q = e % BK;
            ra[it] = *reinterpret_cast<const v4i32 *>(Pm + (size_t)(row0 + r) * ldt + kb + q);
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int e = (it * NTH + tid) * 16, r = e / BK, q = e % BK;
            rb[it] = *reinterpret_cast<const v4i32 *>(Pn + (size_t)(col0 + r) * ldt + kb + q);
        }
    };
    gload(0);
    for (int kb2 = 0; kb2 < K; kb2 += BK) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int e = (it * NTH + tid) * 16;
            *reinterpret_cast<v4i32 *>(&As[soff(e / BK, (e % BK) / 16)]) = ra[it];
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int e = (it * NTH + tid) * 16;
            *reinterpret_cast<v4i32 *>(&Bs[soff(e / BK, (e % BK) / 16)]) = rb[it];
        }
        __syncthreads();
        if (kb2 + BK < K) gload(kb2 + BK);
        const int sa = kb2 < ldt ? sa1 : sa2, sb = kb2 < ldt ? sb1 : sb2;
        v8i32 af[FM], bf[FN];
#pragma unroll
        for (int fm = 0; fm < FM; ++fm) {
            const int r = wm + fm * 16 + li;
            const v4i32 lo = *reinterpret_cast<const v4i32 *>(&As[soff(r, 2 * lk)]);
            const v4i32 hi = *reinterpret_cast<const v4i32 *>(&As[soff(r, 2 * lk + 1)]);
            af[fm] = v8i32{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        }
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) {
            const int r = wn + fn * 16 + li;
            const v4i32 lo = *reinterpret_cast<const v4i32 *>(&Bs[soff(r, 2 * lk)]);
            const v4i32 hi = *reinterpret_cast<const v4i32 *>(&Bs[soff(r, 2 * lk + 1)]);
            bf[fn] = v8i32{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        }
        /* first operand: the B rows (results land as consecutive rows of C) */
#pragma unroll
        for (int fm = 0; fm < FM; ++fm)
#pragma unroll
            for (int fn = 0; fn < FN; ++fn)
                acc[fm][fn] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(bf[fn], af[fm], acc[fm][fn], 0, 0, 0, sb,
                                                                              0, sa);
        __syncthreads();
    }
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int gi = row0 + wm + fm * 16 + li;
                const int gj = col0 + wn + fn * 16 + 4 * lk + s;
                if (!diag || gi >= gj) c_sub(&C[gi + (size_t)gj * ldt], (double)acc[fm][fn][s], atomic_c);
            }
}

/* ------------------------------------------------------------------ */
/* Symmetric diagonal scaling of the working copy: s_i = 1 / sqrt(a_ii)
 * (1 where a_ii <= 0: the factorisation reports it), then
 * A(i, j) <- s_i A(i, j) s_j on the lower tiles, and the scaling of
 * transposed right-hand sides (element (c, r) at c + r nrp) by s_r. */
__global__ void diag_rsqrt_kernel(const double *A, int NT, int ldt, double *sc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NT * ldt) return;
    const int t = i / ldt, r = i % ldt;
    const double d = A[((size_t)t * NT + t) * ldt * ldt + r + (size_t)r * ldt];
    sc[i] = d > 0.0 ? 1.0 / sqrt(d) : 1.0;
}

__global__ void scale_lower_kernel(double *A, int NT, int ldt, const double *sc)
{
    const size_t ts = (size_t)ldt * ldt;
    const int pair = blockIdx.y;
    int i = (int)((sqrt(8.0 * (double)pair + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= pair) ++i;
    while (i * (i + 1) / 2 > pair) --i;
    const int m = i, n = pair - i * (i + 1) / 2;
    double *T = A + ((size_t)n * NT + m) * ts;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < ts; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % ldt), c = (int)(e / ldt);
        T[e] *= sc[m * ldt + r] * sc[n * ldt + c];
    }
}

__global__ void scale_bt_kernel(double *BT, int nrp, int rows, const double *sc)
{
    const size_t total = (size_t)nrp * rows;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
        BT[e] *= sc[e / nrp];
}

void diag_rsqrt(const double *A, int NT, int ldt, double *sc, hipStream_t s)
{
    hipLaunchKernelGGL(diag_rsqrt_kernel, dim3((NT * ldt + 255) / 256), dim3(256), 0, s, A, NT, ldt, sc);
}

void diag_scaling(double *A, int NT, int ldt, double *sc, hipStream_t s)
{
    diag_rsqrt(A, NT, ldt, sc, s);
    hipLaunchKernelGGL(scale_lower_kernel, dim3(std::min(64, (ldt * ldt + 255) / 256), NT * (NT + 1) / 2), dim3(256),
                       0, s, A, NT, ldt, sc);
}

void scale_bt(double *BT, int nrp, int rows, const double *sc, hipStream_t s)
{
    const size_t total = (size_t)nrp * rows;
    if (!total) return;
    hipLaunchKernelGGL(scale_bt_kernel, dim3((unsigned)std::min<size_t>((total + 255) / 256, 8192)), dim3(256), 0, s,
                       BT, nrp, rows, sc);
}

/* ------------------------------------------------------------------ */
/* Squared Frobenius norms of the lower tiles (m >= n) of the symmetric
 * matrix whose lower triangle the mirror holds: off-diagonal tiles whole,
 * diagonal tiles as the symmetric tile (strictly lower entries twice, the
 * upper part of the stored tile not read).  One block per tile, ss[m + n NT]. */
__global__ __launch_bounds__(256) void tile_sumsq_lower_kernel(const double *A, int NT, int ldt, double *ss)
{
    __shared__ double red[256];
    const int pair = blockIdx.x;
    int i = (int)((sqrt(8.0 * (double)pair + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= pair) ++i;
    while (i * (i + 1) / 2 > pair) --i;
    const int m = i, n = pair - i * (i + 1) / 2;
    const size_t ts = (size_t)ldt * ldt;
    const double *T = A + ((size_t)n * NT + m) * ts;
    double acc = 0.0;
    for (size_t e = threadIdx.x; e < ts; e += 256) {
        const int r = (int)(e % ldt), c = (int)(e / ldt);
        const double v = T[e];
        if (m != n) acc += v * v;
        else if (r > c) acc += 2.0 * v * v;
        else if (r == c) acc += v * v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) ss[m + (size_t)n * NT] = red[0];
}

void tile_sumsq_lower(const double *A, int NT, int ldt, double *ss, hipStream_t s)
{
    if (NT <= 0) return;
    hipLaunchKernelGGL(tile_sumsq_lower_kernel, dim3(NT * (NT + 1) / 2), dim3(256), 0, s, A, NT, ldt, ss);
}

void trailing_update_lp(double *A, const void *P, const void *P2, int mode, int lmt, int ldt, int n0, int n1, int NT,
                        const int *info, hipStream_t s, const unsigned char *pmap)
{
    if (n0 >= NT || n1 <= n0) return;
    int ncol = 0;
    const int npairs = update_pairs(n0, n1, NT, &ncol);
    const int bpt = (ldt / 128) * (ldt / 128);
    const long long total = (long long)npairs * bpt;
    if (total <= 0) return;
    const int ac = atomic_tile_updates();
    if (mode == 0)
        hipLaunchKernelGGL((update_f32_kernel<128, 128, 64, 32>), dim3((unsigned)total), dim3(512), 0, s, A,
                           (const float *)P, (const float *)P2, lmt, ldt, n0, ncol, (int)total, info, pmap, ac);
    else if (mode == 3)
        hipLaunchKernelGGL(update_fp8_kernel, dim3((unsigned)total), dim3(512), 0, s, A, (const unsigned char *)P,
                           (const unsigned char *)P2, NT, lmt, ldt, n0, ncol, (int)total, info, ac);
    else {
        auto *kern = mode == 1 ? update_bf16_kernel<1> : update_bf16_kernel<3>;
        hipLaunchKernelGGL(kern, dim3((unsigned)total), dim3(512), 0, s, A, (const unsigned short *)P,
                           (const unsigned short *)P2, NT, lmt, ldt, n0, ncol, (int)total, info, pmap, ac);
    }
}

/* ------------------------------------------------------------------ */
/* FP64 residual R = B - A X for the symmetric A whose lower triangle is the
 * tile mirror A (NT x NT tiles of ldt x ldt), right-hand sides transposed
 * (element (j, p) at j + p*nrp).
 *
 * The row range of A is cut in RES_KS pieces: wavefront (rb, jc, ks)
 * forms the (16 RG) x 64 block (rows 16 RG rb.., columns 64 jc..) of
 * A(:, piece ks) X(piece ks, :) into its own partial buffer, and a second
 * kernel adds the pieces in a fixed order (deterministic).  With RG = 2 a
 * wave's two 16-row groups share every X operand it loads (0.75 instead of
 * 1.25 vector loads per MFMA; ldt a multiple of 32 keeps both groups in
 * one tile row).  Blocks below the diagonal tile are read as stored,
 * blocks above as the transpose of the stored lower tile (32-byte row
 * vectors), the diagonal tile by triangle.  The k loop runs in 16-deep
 * chunks with the next chunk's loads in flight; the last chunk of a tile
 * is peeled off the loop, and with FULL (every column chunk 64 wide) the
 * column count is a constant, so the loop body has no branches: with
 * run-time tests around each load and product the compiler split it into
 * dozens of basic blocks. */
#define RES_KS 8
template <int RG, bool FULL>
__global__ __launch_bounds__(256, 2) void residual_kernel(const double *A, int NT, int ldt, const double *XT, int nrp,
                                                        double *part)
{
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int rows = NT * ldt;
    const int jchunks = (nrp + 63) / 64;
    const int nrb = rows / (16 * RG);
    /* the four waves of a workgroup take four consecutive row blocks of
     * one k range, so they read the same X chunks at the same time */
    const int groups = (nrb + 3) / 4;
    const int ks = blockIdx.x % RES_KS, rest = blockIdx.x / RES_KS;
    const int rb = (rest % groups) * 4 + (threadIdx.x >> 6), jc = rest / groups;
    if (jc >= jchunks || rb >= nrb) return;     /* no barriers below: waves are independent */
    const int r0 = rb * 16 * RG;
    const int m = r0 / ldt, i0 = r0 % ldt;
    const int j0 = jc * 64;
    const int nf = FULL ? 4 : min(4, (nrp - j0) / 16);
    const size_t ts = (size_t)ldt * ldt;
    /* this wave's k range: [kb, ke) in units of 16 */
    const int nk16 = rows / 16;
    const int kb = (int)((long long)nk16 * ks / RES_KS) * 16, ke = (int)((long long)nk16 * (ks + 1) / RES_KS) * 16;
    v4d acc[RG][4];
#pragma unroll
    for (int g = 0; g < RG; ++g)
#pragma unroll
        for (int f = 0; f < 4; ++f) acc[g][f] = zero4d();
    /* x operands of chunk p0: X(p0 + 4 lk + q, col(f, li)), where product f
     * covers columns j0 + 16 f + li, or with FULL columns j0 + 4 li + f (a
     * lane's four columns adjacent: one 32-byte load per k) */
    auto load_x = [&](int p0, double (&x)[4][4]) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if constexpr (FULL) {
                const v4d v = *(const v4d *)(XT + (size_t)(p0 + 4 * lk + q) * nrp + j0 + 4 * li);
#pragma unroll
                for (int f = 0; f < 4; ++f) x[q][f] = v[f];
            } else {
                const double *xr = XT + (size_t)(p0 + 4 * lk + q) * nrp + j0 + li;
#pragma unroll
                for (int f = 0; f < 4; ++f) x[q][f] = (f < nf) ? xr[16 * f] : 0.0;
            }
        }
    };
    auto mma = [&](const double (&a)[RG][4], const double (&x)[4][4]) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int g = 0; g < RG; ++g)
#pragma unroll
                for (int f = 0; f < 4; ++f)
                    if (f < nf) acc[g][f] = mfma_f64(a[g][q], x[q][f], acc[g][f]);
    };
    /* the k range, tile by tile: three loop shapes, each software-pipelined */
    for (int nt = kb / ldt; nt * ldt < ke; ++nt) {
        const int pb = max(kb, nt * ldt) - nt * ldt, pe = min(ke, (nt + 1) * ldt) - nt * ldt;
        if (nt < m) {
            /* stored tile (m, nt): element (i0 + 16 g + li, p) at T[16 g + p*ldt] */
            const double *T = A + ((size_t)nt * NT + m) * ts + i0 + li;
            double a[RG][4], x[4][4], an[RG][4], xn[4][4];
#pragma unroll
            for (int g = 0; g < RG; ++g)
#pragma unroll
                for (int q = 0; q < 4; ++q) a[g][q] = T[16 * g + (size_t)(pb + 4 * lk + q) * ldt];
            load_x(nt * ldt + pb, x);
            for (int p = pb; p + 16 < pe; p += 16) {
#pragma unroll
                for (int g = 0; g < RG; ++g)
#pragma unroll
                    for (int q = 0; q < 4; ++q) an[g][q] = T[16 * g + (size_t)(p + 16 + 4 * lk + q) * ldt];
                load_x(nt * ldt + p + 16, xn);
                mma(a, x);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
#pragma unroll
                    for (int g = 0; g < RG; ++g) a[g][q] = an[g][q];
#pragma unroll
                    for (int f = 0; f < 4; ++f) x[q][f] = xn[q][f];
                }
            }
            mma(a, x);
        } else if (nt > m) {
            /* transpose of stored tile (nt, m): element (i0 + 16 g + li, p) at
             * T[p + 16 g ldt] */
            const double *T = A + ((size_t)m * NT + nt) * ts + (size_t)(i0 + li) * ldt;
            double a[RG][4], x[4][4], an[RG][4], xn[4][4];
#pragma unroll
            for (int g = 0; g < RG; ++g) {
                const v4d v = *(const v4d *)(T + (size_t)16 * g * ldt + pb + 4 * lk);
#pragma unroll
                for (int q = 0; q < 4; ++q) a[g][q] = v[q];
            }
            load_x(nt * ldt + pb, x);
            for (int p = pb; p + 16 < pe; p += 16) {
#pragma unroll
                for (int g = 0; g < RG; ++g) {
                    const v4d v = *(const v4d *)(T + (size_t)16 * g * ldt + p + 16 + 4 * lk);
#pragma unroll
                    for (int q = 0; q < 4; ++q) an[g][q] = v[q];
                }
                load_x(nt * ldt + p + 16, xn);
                mma(a, x);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
#pragma unroll
                    for (int g = 0; g < RG; ++g) a[g][q] = an[g][q];
#pragma unroll
                    for (int f = 0; f < 4; ++f) x[q][f] = xn[q][f];
                }
            }
            mma(a, x);
        } else {
            /* diagonal tile, lower triangle stored */
            const double *T = A + ((size_t)m * NT + m) * ts;
            for (int p = pb; p < pe; p += 16) {
                double a[RG][4], x[4][4];
#pragma unroll
                for (int g = 0; g < RG; ++g) {
                    const int ii = i0 + 16 * g + li;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int pp = p + 4 * lk + q;
                        a[g][q] = T[(ii >= pp) ? ii + (size_t)pp * ldt : pp + (size_t)ii * ldt];
                    }
                }
                load_x(nt * ldt + p, x);
                mma(a, x);
            }
        }
    }
    /* slot s of acc[g][f]: row r0 + 16 g + lk + 4s, column col(f, li) */
    double *P = part + (size_t)ks * rows * nrp;
#pragma unroll
    for (int g = 0; g < RG; ++g) {
        if constexpr (FULL) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
                *(v4d *)&P[(j0 + 4 * li) + (size_t)(r0 + 16 * g + lk + 4 * s) * nrp] =
                    v4d{acc[g][0][s], acc[g][1][s], acc[g][2][s], acc[g][3][s]};
        } else {
#pragma unroll
            for (int f = 0; f < 4; ++f)
                if (f < nf) {
#pragma unroll
                    for (int s = 0; s < 4; ++s)
                        P[(j0 + 16 * f + li) + (size_t)(r0 + 16 * g + lk + 4 * s) * nrp] = acc[g][f][s];
                }
        }
    }
}

__global__ void residual_sum_kernel(const double *BT, const double *part, double *RT, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < RES_KS; ++k) s += part[(size_t)k * n + e];
        RT[e] = BT[e] - s;
    }
}

size_t residual_scratch(int NT, int ldt, int nrp) { return (size_t)RES_KS * NT * ldt * nrp; }

void residual(const double *A, int NT, int ldt, const double *XT, const double *BT, double *RT, int nrp, double *part,
              hipStream_t s)
{
    /* two 16-row groups per wave when both lie in one tile row */
    const int rg = (ldt % 32 == 0) ? 2 : 1;
    const bool full = nrp % 64 == 0;
    const long long blocks = (long long)((NT * ldt / (16 * rg) + 3) / 4) * ((nrp + 63) / 64) * RES_KS;
    auto kern = rg == 2 ? (full ? residual_kernel<2, true> : residual_kernel<2, false>)
                        : (full ? residual_kernel<1, true> : residual_kernel<1, false>);
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), 0, s, A, NT, ldt, XT, nrp, part);
    const size_t n = (size_t)NT * ldt * nrp;
    hipLaunchKernelGGL(residual_sum_kernel, dim3(2048), dim3(256), 0, s, BT, part, RT, n);
}

/* the column maxima max_p |V(j, p)| (V transposed: element (j, p) at
 * j + p*nrp) of X into out[j] and of R into out[nrp + j], in one launch;
 * out cleared by the caller.  Lane c of wave g: column 64 blockIdx.y + c,
 * rows g, g + 4, ... of this block's share: a wave reads 64 adjacent
 * columns of a row (coalesced; one column per workgroup had each lane
 * touch its own 512-byte line), the four waves meet in LDS, one atomic max
 * per column and array per workgroup (on the bit pattern: order-free and
 * NaN-propagating for these non-negative values) */
__global__ void colmax2_kernel(const double *X, const double *R, int nrp, int ncols, int nrows, double *out)
{
    __shared__ double sx[4][64], sr[4][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int j = blockIdx.y * 64 + c;
    double mx = 0.0, mr = 0.0;
    if (j < ncols) {
        for (int p = blockIdx.x * 4 + g; p < nrows; p += gridDim.x * 4) {
            mx = nan_max(mx, fabs(X[j + (size_t)p * nrp]));
            mr = nan_max(mr, fabs(R[j + (size_t)p * nrp]));
        }
    }
    sx[g][c] = mx;
    sr[g][c] = mr;
    __syncthreads();
    if (g == 0 && j < ncols) {
        mx = nan_max(nan_max(sx[0][c], sx[1][c]), nan_max(sx[2][c], sx[3][c]));
        mr = nan_max(nan_max(sr[0][c], sr[1][c]), nan_max(sr[2][c], sr[3][c]));
        atomicMax(reinterpret_cast<unsigned long long *>(&out[j]), (unsigned long long)__double_as_longlong(mx));
        atomicMax(reinterpret_cast<unsigned long long *>(&out[nrp + j]), (unsigned long long)__double_as_longlong(mr));
    }
}

void colmax2(const double *X, const double *R, int nrp, int ncols, int nrows, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(colmax2_kernel, dim3(128, (ncols + 63) / 64), dim3(256), 0, s, X, R, nrp, ncols, nrows, out);
}

__global__ void axpy_kernel(double *Y, const double *X, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        Y[e] += X[e];
}

void axpy(double *Y, const double *X, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(axpy_kernel, dim3(2048), dim3(256), 0, s, Y, X, n);
}

/* ------------------------------------------------------------------ */
/* Working copy of the lower tiles of a symmetric tile matrix, with the
 * row sums of its inf-norm formed in the same pass (one read of A instead
 * of a copy and a norm pass).  S and W hold NT x NT tiles of ldt x ldt,
 * tile (m, n) at (n NT + m) ldt^2; only tiles m >= n are read and written.
 * The valid part of tile (m, n) is min(mb, N - m mb) x min(mb, N - n mb);
 * of a diagonal tile only the lower triangle counts.  For tile t = n NT + m:
 *   part[2 t ldt + i]       = sum_j |S(i, j)|                  (rows)
 *   part[(2 t + 1) ldt + j] = sum_i |S(i, j)|, diagonal excluded (columns)
 * so logical row r = m mb + i of the symmetric matrix is the sum of the
 * tile rows i of tiles (m, 0..m) and the tile columns i of (m..NT-1, m).
 * One workgroup per tile: wave w takes rows w*128 + 512 c, two rows per
 * lane (16-byte loads), eight columns in flight; each column's sum goes
 * through a wave reduction into an LDS row of the wave's own (no atomics:
 * the result does not depend on scheduling).  ldt a multiple of 128. */
__global__ __launch_bounds__(256) void copy_lower_norm_kernel(const double *S, double *W, int NT, int ldt, int mb,
                                                              int N, double *part, const double *sc)
{
    extern __shared__ double cs[];      /* [4][ldt] column sums, one row per wave */
    const int m = blockIdx.x % NT, n = blockIdx.x / NT;
    if (m < n) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t ts = (size_t)ldt * ldt, t = (size_t)n * NT + m;
    const double *src = S + t * ts;
    double *dst = W + t * ts;
    const int vm = min(mb, N - m * mb), vn = min(mb, N - n * mb);
    const bool diag = (m == n);
    double *csw = cs + (size_t)wave * ldt;
    for (int j = lane; j < ldt; j += 64) csw[j] = 0.0;
    __syncthreads();            /* lane 0 adds below into entries other lanes cleared */
    for (int r0 = wave * 128; r0 < ldt; r0 += 512) {
        const int i = r0 + 2 * lane;
        double rs0 = 0.0, rs1 = 0.0;
        /* with sc: the copy is S A S (element (i, j) times s_i s_j, the
         * product of the two scales first), the sums stay those of A */
        const double sr0 = sc ? sc[m * ldt + i] : 1.0, sr1 = sc ? sc[m * ldt + i + 1] : 1.0;
        for (int j0 = 0; j0 < ldt; j0 += 8) {
            double2 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const double2 *>(src + (size_t)(j0 + u) * ldt + i);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int j = j0 + u;
                if (sc) {
                    const double sj = sc[n * ldt + j];
                    *reinterpret_cast<double2 *>(dst + (size_t)j * ldt + i) = double2{v[u].x * (sr0 * sj), v[u].y * (sr1 * sj)};
                } else {
                    *reinterpret_cast<double2 *>(dst + (size_t)j * ldt + i) = v[u];
                }
                const bool cj = j < vn;
                const double a0 = (cj && i < vm && (!diag || i >= j)) ? fabs(v[u].x) : 0.0;
                const double a1 = (cj && i + 1 < vm && (!diag || i + 1 >= j)) ? fabs(v[u].y) : 0.0;
                rs0 += a0;
                rs1 += a1;
                double c = (diag && i == j ? 0.0 : a0) + (diag && i + 1 == j ? 0.0 : a1);
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
                if (lane == 0) csw[j] += c;
            }
        }
        part[2 * t * ldt + i] = rs0;
        part[2 * t * ldt + i + 1] = rs1;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < ldt; j += 256)
        part[(2 * t + 1) * ldt + j] = ((cs[j] + cs[ldt + j]) + cs[2 * ldt + j]) + cs[3 * ldt + j];
}

/* rowsum[r], r < N: the symmetric matrix's absolute row sums from the
 * partials of copy_lower_norm_kernel, added in a fixed order */
__global__ void norm_rows_kernel(const double *part, int NT, int ldt, int mb, int N, double *rowsum)
{
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < N; r += gridDim.x * blockDim.x) {
        const int m = r / mb, i = r % mb;
        double s = 0.0;
        for (int n = 0; n <= m; ++n) s += part[2 * ((size_t)n * NT + m) * ldt + i];
        for (int q = m; q < NT; ++q) s += part[(2 * ((size_t)m * NT + q) + 1) * ldt + i];
        rowsum[r] = s;
    }
}

size_t copy_lower_norm_scratch(int NT, int ldt) { return (size_t)2 * NT * NT * ldt; }

int copy_lower_norm(const double *S, double *W, int NT, int ldt, int mb, int N, double *part, double *rowsum,
                    hipStream_t s, const double *sc)
{
    if (NT <= 0 || ldt % 128 != 0 || ldt > COPY_NORM_MAX_LDT || mb <= 0 || mb > ldt || N <= (NT - 1) * mb ||
        N > NT * mb)
        return -1;
    hipLaunchKernelGGL(copy_lower_norm_kernel, dim3((unsigned)(NT * NT)), dim3(256), (size_t)4 * ldt * sizeof(double),
                       s, S, W, NT, ldt, mb, N, part, sc);
    hipLaunchKernelGGL(norm_rows_kernel, dim3((unsigned)std::min(2048, (N + 255) / 256)), dim3(256), 0, s, part, NT,
                       ldt, mb, N, rowsum);
    return 0;
}

} /* namespace k */
} /* namespace chx */
