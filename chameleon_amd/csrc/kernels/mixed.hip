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
 * Also: the FP64 symmetric residual R = B - A X and column max-norms used by
 * the refinement loop.
 */
#include "../internal.h"
#include "mfma.h"

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
 * FP64 trailing_update_kernel: one column n0, or the triangle n0..NT-1). */
__device__ __forceinline__ void decode_pair(int pair, int n0, int colmode, int &m, int &n)
{
    if (colmode) {
        n = n0;
        m = n0 + pair;
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
    double *A, const float *P, int lmt, int ldt, int n0, int colmode, int total_blocks, const int *info)
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
    decode_pair(bid / bpt, n0, colmode, m, n);
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
    auto gload = [&](int kb) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int e = (it * NTH + tid) * 4;
            ra[it] = *reinterpret_cast<const v4f *>(Am + (e % BM) + (size_t)(kb + e / BM) * ldt);
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int e = (it * NTH + tid) * 4;
            rb[it] = *reinterpret_cast<const v4f *>(An + (e % BN) + (size_t)(kb + e / BN) * ldt);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int it = 0; it < LA; ++it) *reinterpret_cast<v4f *>(&As[buf][(it * NTH + tid) * 4]) = ra[it];
#pragma unroll
        for (int it = 0; it < LB; ++it) *reinterpret_cast<v4f *>(&Bs[buf][(it * NTH + tid) * 4]) = rb[it];
    };
    const int S = ldt / BK;
    gload(0);
    lstore(0);
    if (S > 1) gload(BK);
    __syncthreads();
    for (int st = 0; st < S; ++st) {
        const int cur = st & 1;
        if (st + 1 < S) {
            lstore(cur ^ 1);
            if (st + 2 < S) gload((st + 2) * BK);
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
                if (!diag || gi >= gj) C[gi + (size_t)gj * ldt] -= (double)acc[fm][fn][s];
            }
}

/* ------------------------------------------------------------------ */
/* BF16-operand trailing update (SPLIT = 3: hi*hi + hi*lo + lo*hi).  Panel
 * planes are row-major (k contiguous), LDS rows padded by 8 BF16. Each
 * lane of a 16x16x32 MFMA supplies 8 consecutive k of one row (the same k
 * assignment on both operands, so the dot product is exact in any order). */
template <int SPLIT>
__global__ __launch_bounds__(512, 2) void update_bf16_kernel(double *A, const unsigned short *P, int NT, int lmt,
                                                             int ldt, int n0, int colmode, int total_blocks,
                                                             const int *info)
{
    constexpr int BM = 128, BN = 128, WM = 64, WN = 32, BK = 64, LDK = BK + 8;
    constexpr int FM = WM / 16, FN = WN / 16, NTH = 512;
    constexpr int NP = (SPLIT == 3) ? 2 : 1;          /* planes per operand */
    __shared__ unsigned short As[NP][BM * LDK];
    __shared__ unsigned short Bs[NP][BN * LDK];
    if (*info != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int bid = xcd_remap_m(blockIdx.x, total_blocks);
    const int tbm = ldt / BM, tbn = ldt / BN, bpt = tbm * tbn;
    int m, n;
    decode_pair(bid / bpt, n0, colmode, m, n);
    const int b = bid % bpt, row0 = (b % tbm) * BM, col0 = (b / tbm) * BN;
    const bool diag = (m == n);
    if (diag && col0 >= row0 + BM) return;
    const size_t ts = (size_t)ldt * ldt;
    double *C = A + ((size_t)n * lmt + m) * ts;
    const int wm = (wave % (BM / WM)) * WM, wn = (wave / (BM / WM)) * WN;
    v4f acc[FM][FN];
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = zero4f();
    /* slice loader: BM rows x BK k of each plane = BM*BK/8 v8 loads = 2 per thread per operand plane */
    for (int kb = 0; kb < ldt; kb += BK) {
#pragma unroll
        for (int pl = 0; pl < NP; ++pl) {
            const unsigned short *Pm = P + ((size_t)pl * NT + m) * ts;
            const unsigned short *Pn = P + ((size_t)pl * NT + n) * ts;
#pragma unroll
            for (int it = 0; it < (BM * BK) / (8 * NTH); ++it) {
                const int e = (it * NTH + tid) * 8;
                const int r = e / BK, q = e % BK;
                *reinterpret_cast<v8u16 *>(&As[pl][r * LDK + q]) =
                    *reinterpret_cast<const v8u16 *>(Pm + (size_t)(row0 + r) * ldt + kb + q);
                *reinterpret_cast<v8u16 *>(&Bs[pl][r * LDK + q]) =
                    *reinterpret_cast<const v8u16 *>(Pn + (size_t)(col0 + r) * ldt + kb + q);
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 32) {
            v8bf ah[FM], bh[FN], al[FM], bl[FN];
#pragma unroll
            for (int fm = 0; fm < FM; ++fm) {
                ah[fm] = *reinterpret_cast<const v8bf *>(&As[0][(wm + fm * 16 + li) * LDK + kk + 8 * lk]);
                if (SPLIT == 3) al[fm] = *reinterpret_cast<const v8bf *>(&As[NP - 1][(wm + fm * 16 + li) * LDK + kk + 8 * lk]);
            }
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) {
                bh[fn] = *reinterpret_cast<const v8bf *>(&Bs[0][(wn + fn * 16 + li) * LDK + kk + 8 * lk]);
                if (SPLIT == 3) bl[fn] = *reinterpret_cast<const v8bf *>(&Bs[NP - 1][(wn + fn * 16 + li) * LDK + kk + 8 * lk]);
            }
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < FN; ++fn) {
                    acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[fn], ah[fm], acc[fm][fn], 0, 0, 0);
                    if (SPLIT == 3) {
                        acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[fn], ah[fm], acc[fm][fn], 0, 0, 0);
                        acc[fm][fn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[fn], al[fm], acc[fm][fn], 0, 0, 0);
                    }
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
                if (!diag || gi >= gj) C[gi + (size_t)gj * ldt] -= (double)acc[fm][fn][s];
            }
}

void trailing_update_lp(double *A, const void *P, int mode, int lmt, int ldt, int n0, int n1, int NT, const int *info,
                        hipStream_t s)
{
    if (n0 >= NT || n1 <= n0) return;
    const bool colmode = (n1 == n0 + 1);
    const int S = NT - n0;
    const int npairs = colmode ? S : S * (S + 1) / 2;
    const int bpt = (ldt / 128) * (ldt / 128);
    const long long total = (long long)npairs * bpt;
    if (total <= 0) return;
    if (mode == 0)
        hipLaunchKernelGGL((update_f32_kernel<128, 128, 64, 32>), dim3((unsigned)total), dim3(512), 0, s, A,
                           (const float *)P, lmt, ldt, n0, (int)colmode, (int)total, info);
    else if (mode == 1)
        hipLaunchKernelGGL((update_bf16_kernel<1>), dim3((unsigned)total), dim3(512), 0, s, A, (const unsigned short *)P,
                           NT, lmt, ldt, n0, (int)colmode, (int)total, info);
    else
        hipLaunchKernelGGL((update_bf16_kernel<3>), dim3((unsigned)total), dim3(512), 0, s, A, (const unsigned short *)P,
                           NT, lmt, ldt, n0, (int)colmode, (int)total, info);
}

/* ------------------------------------------------------------------ */
/* FP64 residual R = B - A X for the symmetric A whose lower triangle is the
 * tile mirror A (NT x NT tiles of ldt x ldt), right-hand sides transposed
 * (element (j, p) at j + p*nrp).  One wavefront per 16 rows x 64 columns;
 * blocks below the diagonal tile are read as stored, blocks above as the
 * transpose of the stored lower tile, the diagonal tile by triangle. */
__global__ __launch_bounds__(256) void residual_kernel(const double *A, int NT, int ldt, const double *XT,
                                                        const double *BT, double *RT, int nrp)
{
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int gw = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int rows = NT * ldt;
    const int jchunks = (nrp + 63) / 64;
    const int rb = gw / jchunks, jc = gw % jchunks;
    if (rb * 16 >= rows) return;
    const int r0 = rb * 16;
    const int m = r0 / ldt, i0 = r0 % ldt;
    const int j0 = jc * 64;
    const int nf = min(4, (nrp - j0) / 16);
    const size_t ts = (size_t)ldt * ldt;
    v4d acc[4] = {zero4d(), zero4d(), zero4d(), zero4d()};
    for (int nt = 0; nt < NT; ++nt) {
        const double *T;
        long long si, sp;   /* opA(i, p) at T[i*si + p*sp] */
        if (nt < m) { T = A + ((size_t)nt * NT + m) * ts + i0; si = 1; sp = ldt; }
        else if (nt > m) { T = A + ((size_t)m * NT + nt) * ts + (size_t)i0 * ldt; si = ldt; sp = 1; }
        else { T = A + ((size_t)m * NT + m) * ts; si = 0; sp = 0; }
        for (int p = 0; p < ldt; p += 4) {
            double a;
            if (nt != m) {
                a = T[(long long)li * si + (long long)(p + lk) * sp];
            } else {
                const int ii = i0 + li, pp = p + lk;
                a = (ii >= pp) ? T[ii + (size_t)pp * ldt] : T[pp + (size_t)ii * ldt];
            }
            const double *xr = XT + (size_t)(nt * ldt + p + lk) * nrp + j0 + li;
#pragma unroll
            for (int f = 0; f < 4; ++f)
                if (f < nf) acc[f] = mfma_f64(a, xr[16 * f], acc[f]);
        }
    }
    /* slot s of acc[f]: row i0 + lk + 4s, column j0 + 16 f + li */
#pragma unroll
    for (int f = 0; f < 4; ++f)
        if (f < nf) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const size_t e = (j0 + 16 * f + li) + (size_t)(r0 + lk + 4 * s) * nrp;
                RT[e] = BT[e] - acc[f][s];
            }
        }
}

void residual(const double *A, int NT, int ldt, const double *XT, const double *BT, double *RT, int nrp,
              hipStream_t s)
{
    const int waves = (NT * ldt / 16) * ((nrp + 63) / 64);
    hipLaunchKernelGGL(residual_kernel, dim3((waves + 3) / 4), dim3(256), 0, s, A, NT, ldt, XT, BT, RT, nrp);
}

/* per-column max |V(j, p)| over p < nrows (V transposed: j + p*nrp); out
 * zeroed by the caller */
__global__ void colmax_kernel(const double *V, int nrp, int nrows, double *out)
{
    const int j = blockIdx.y;
    double mx = 0.0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nrows; p += gridDim.x * blockDim.x)
        mx = fmax(mx, fabs(V[j + (size_t)p * nrp]));
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off));
    if ((threadIdx.x & 63) == 0)
        atomicMax(reinterpret_cast<unsigned long long *>(&out[j]), (unsigned long long)__double_as_longlong(mx));
}

void colmax(const double *V, int nrp, int ncols, int nrows, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(colmax_kernel, dim3(64, ncols), dim3(256), 0, s, V, nrp, nrows, out);
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

} /* namespace k */
} /* namespace chx */
