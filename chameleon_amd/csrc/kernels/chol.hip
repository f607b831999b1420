/*
 * chol.hip -- FP64 tile kernels of the tiled Cholesky factor-and-solve path
 * for gfx950 (MI355X).  They replace the four tile operations Chameleon
 * submits per step of its right-looking algorithm (compute/pzpotrf.c:84-135:
 * POTRF on A(k,k), TRSM on A(m,k), SYRK on A(n,n), GEMM on A(m,n)) and the
 * tile operations of its triangular solves (compute/pztrsm.c:134-201).
 *
 * All tiles are device mirror tiles: ldt x ldt doubles, column-major,
 * contiguous, ldt a multiple of 16 (see internal.h).
 *
 * Every kernel reads the device info flag first and returns when an earlier
 * diagonal block was found not positive definite, so a failed
 * factorisation stops doing work without a host round trip.
 */
#include "../internal.h"
#include "mfma.h"

namespace chx {
namespace k {

/* ------------------------------------------------------------------ */
/* XCD-aware block order.  Workgroups are dealt to the 8 XCDs of the chip
 * round-robin by id; this remap gives each XCD a contiguous range of logical
 * block ids so that blocks working on the same tiles share an L2.          */
__device__ __forceinline__ int xcd_remap(int id, int total)
{
    const int nx = 8;
    int per = total / nx, rem = total % nx;
    int x = id % nx, local = id / nx;
    return (x < rem) ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
}

/* Global -> register load of one K slice (BK rows of W contiguous
 * elements, leading dimension ld) as v2d per thread, and the matching
 * register -> LDS store into a [BK][W] array. */
template <int W, int L, int NTH>
__device__ __forceinline__ void gload_slice(v2d (&r)[L], const double *src, int kb, int ld, int tid)
{
#pragma unroll
    for (int it = 0; it < L; ++it) {
        const int e = (it * NTH + tid) * 2;
        r[it] = *reinterpret_cast<const v2d *>(src + (e % W) + (size_t)(kb + e / W) * ld);
    }
}

template <int W, int L, int NTH>
__device__ __forceinline__ void lstore_slice(double *lds, const v2d (&r)[L], int tid)
{
#pragma unroll
    for (int it = 0; it < L; ++it) {
        const int e = (it * NTH + tid) * 2;
        *reinterpret_cast<v2d *>(lds + e) = r[it];
    }
}

/* ================================================================== */
/* Trailing update of step k: A(m,n) -= A(m,k) A(n,k)^T for the tile pairs
 * (with kp = 2, the two panels k and k+1 at once: A(m,n) -= A(m,k) A(n,k)^T
 * + A(m,k+1) A(n,k+1)^T, one read and write of A(m,n) for both steps)
 * n <= m of either ncol columns n0 .. n0+ncol-1 (ncol > 0, the lookahead
 * columns) or the whole trailing triangle n0 <= n <= m < NT (ncol = 0).
 * Diagonal tiles (m == n) only write their lower triangle: that is the SYRK
 * of pzpotrf.c:112-117, the off-diagonal pairs are its GEMMs
 * (pzpotrf.c:123-129), all in one launch.
 *
 * Block tile BM x BN, K staged through LDS in BK-deep slices with the next
 * slice prefetched into registers while the current one is multiplied.
 * Each wavefront owns a WM x WN sub-tile of (WM/16)x(WN/16) MFMA blocks;
 * its operand fragments are read from LDS one 4-deep step ahead of the
 * MFMAs that use them (two register sets), and the first step of the next
 * slice right after the slice barrier, so LDS latency stays off the MFMA
 * pipe.  Operands are swapped (D = B A^T) so each lane's results are
 * consecutive rows of one column: stores of 16 lanes form one 128-byte
 * segment.                                                                */
template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, 2) void trailing_update_kernel(
    double *A, int lmt, int ldt, int k, int kp, int n0, int NT, int ncol, int total_blocks, const int *info,
    const unsigned char *pmap, int atomic_c)
{
    constexpr int NW = (BM / WM) * (BN / WN);
    constexpr int NTH = NW * 64;
    constexpr int BK = 16;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LA = (BK * BM) / (2 * NTH);   /* v2d loads per thread for A */
    constexpr int LB = (BK * BN) / (2 * NTH);
    constexpr int KS = BK / 4;
    static_assert(LA >= 1 && LB >= 1, "tile too small for the thread count");
    /* two LDS stages: stage s+1 is stored while stage s is multiplied,
     * one barrier per K slice */
    __shared__ double As[2][BK * BM];
    __shared__ double Bs[2][BK * BN];

    if (*info != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bid = xcd_remap(blockIdx.x, total_blocks);
    const int tbm = ldt / BM, tbn = ldt / BN;
    const int bpt = tbm * tbn;
    const int pair = bid / bpt, b = bid % bpt;
    int m, n;
    if (ncol > 0) {
        /* column n holds the NT - n tiles n .. NT-1 */
        int q = pair, len = NT - n0;
        n = n0;
        while (q >= len && n + 1 < n0 + ncol) {
            q -= len;
            ++n;
            --len;
        }
        m = n + q;
    } else {
        /* pair -> (i, j), i >= j, row-major over the lower triangle */
        int i = (int)((sqrt(8.0 * (double)pair + 1.0) - 1.0) * 0.5);
        while ((long long)(i + 1) * (i + 2) / 2 <= pair) ++i;
        while ((long long)i * (i + 1) / 2 > pair) --i;
        const int j = pair - i * (i + 1) / 2;
        m = n0 + i;
        n = n0 + j;
    }
    /* per-tile precision map: this launch updates the FP64 tiles only */
    if (pmap && pmap[m + (size_t)n * NT] != 0) return;
    const int bi = b % tbm, bj = b / tbm;
    const int row0 = bi * BM, col0 = bj * BN;
    const bool diag = (m == n);
    if (diag && col0 >= row0 + BM) return;

    const size_t ts = (size_t)ldt * ldt;
    const double *Am = A + ((size_t)k * lmt + m) * ts + row0;
    const double *An = A + ((size_t)k * lmt + n) * ts + col0;
    double *C = A + ((size_t)n * lmt + m) * ts;

    const int wm = (wave % (BM / WM)) * WM, wn = (wave / (BM / WM)) * WN;
    v4d acc[FM][FN];
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = zero4d();

    v2d ra[LA], rb[LB];
    /* kp panels (tile columns k .. k+kp-1) of ldt / BK slices each */
    const int SK = ldt / BK, S = kp * SK;
    const size_t pstride = (size_t)lmt * ts;
    auto gload = [&](int st) {
        const int pi = st / SK, kb = (st - pi * SK) * BK;
        gload_slice<BM, LA, NTH>(ra, Am + pi * pstride, kb, ldt, tid);
        gload_slice<BN, LB, NTH>(rb, An + pi * pstride, kb, ldt, tid);
    };
    /* operand fragments of one 4-deep step, two sets */
    double fa[2][FM], fb[2][FN];
    auto fload = [&](int buf, int kk, int set) {
        const double *as = As[buf], *bs = Bs[buf];
#pragma unroll
        for (int fm = 0; fm < FM; ++fm) fa[set][fm] = as[(kk + (lane >> 4)) * BM + wm + fm * 16 + (lane & 15)];
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) fb[set][fn] = bs[(kk + (lane >> 4)) * BN + wn + fn * 16 + (lane & 15)];
    };
    gload(0);
    lstore_slice<BM, LA, NTH>(As[0], ra, tid);
    lstore_slice<BN, LB, NTH>(Bs[0], rb, tid);
    if (S > 1) gload(1);
    __syncthreads();
    fload(0, 0, 0);
    for (int st = 0; st < S; ++st) {
        const int cur = st & 1;
        if (st + 1 < S) {
            lstore_slice<BM, LA, NTH>(As[cur ^ 1], ra, tid);
            lstore_slice<BN, LB, NTH>(Bs[cur ^ 1], rb, tid);
            if (st + 2 < S) gload(st + 2);
        }
#pragma unroll
        for (int q = 0; q < KS; ++q) {
            const int set = q & 1;
            if (q + 1 < KS) {
                fload(cur, 4 * (q + 1), set ^ 1);
            } else {
                /* slice st+1 is in LDS and every wave is done reading
                 * slice st (its buffer is refilled next iteration) */
                __syncthreads();
                if (st + 1 < S) fload(cur ^ 1, 0, set ^ 1);
            }
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = mfma_f64(fb[set][fn], fa[set][fm], acc[fm][fn]);
        }
    }
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int gi = row0 + wm + fm * 16 + (lane & 15);
                const int gj = col0 + wn + fn * 16 + (lane >> 4) + 4 * s;
                if (!diag || gi >= gj) c_sub(&C[gi + (size_t)gj * ldt], acc[fm][fn][s], atomic_c);
            }
}

template <int BM, int BN, int WM, int WN>
static void launch_update(double *A, int lmt, int ldt, int k, int kp, int n0, int NT, int ncol, int npairs,
                          const int *info, hipStream_t s, const unsigned char *pmap)
{
    const int bpt = (ldt / BM) * (ldt / BN);
    const long long total = (long long)npairs * bpt;
    if (total <= 0) return;
    constexpr int NTH = (BM / WM) * (BN / WN) * 64;
    hipLaunchKernelGGL((trailing_update_kernel<BM, BN, WM, WN>), dim3((unsigned)total), dim3(NTH), 0, s, A, lmt,
                       ldt, k, kp, n0, NT, ncol, (int)total, info, pmap, atomic_tile_updates());
}

/* the tile updates subtract with FP64 atomic adds in the L2 (c_sub);
 * CHAMAMD_ATOMIC_UPDATE=0: load, subtract and store */
int atomic_tile_updates()
{
    static const int v = [] {
        const char *e = getenv("CHAMAMD_ATOMIC_UPDATE");
        return (e && *e == '0') ? 0 : 1;
    }();
    return v;
}

int update_pairs(int n0, int n1, int NT, int *ncol)
{
    n1 = n1 < NT ? n1 : NT;
    const int S = NT - n0;
    /* columns n0 .. n1-1 only, unless that is the whole triangle */
    const int nc = (n1 < NT || n1 == n0 + 1) ? n1 - n0 : 0;
    *ncol = nc;
    if (nc == 0) return S * (S + 1) / 2;
    return nc * S - nc * (nc - 1) / 2;
}

void trailing_update(double *A, int lmt, int ldt, int k, int n0, int n1, int NT, const int *info, hipStream_t s,
                     int kp, const unsigned char *pmap)
{
    if (n0 >= NT || n1 <= n0) return;
    int ncol = 0;
    const int npairs = update_pairs(n0, n1, NT, &ncol);
    /* A column launch too small to give every compute unit two 128 x 128
     * blocks is latency-bound (each block walks the whole K alone): it
     * takes 64 x 64 blocks, four times as many with a quarter of the work
     * each (the same k order per element, so the same sums). */
    static const int ncu = [] {
        int dev = 0, n = 0;
        CHX_HIP(hipGetDevice(&dev));
        CHX_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        return n;
    }();
    static const bool use_small = [] {
        const char *e = getenv("CHAMAMD_SMALL_COLS");
        return !(e && atoi(e) == 0);
    }();
    const bool small_cols = use_small && ncol > 0 && (long long)npairs * (ldt / 128) * (ldt / 128) < 2LL * ncu;
    if (ldt % 128 == 0 && !small_cols)
        launch_update<128, 128, 32, 64>(A, lmt, ldt, k, kp, n0, NT, ncol, npairs, info, s, pmap);
    else if (ldt % 64 == 0)
        launch_update<64, 64, 32, 32>(A, lmt, ldt, k, kp, n0, NT, ncol, npairs, info, s, pmap);
    else if (ldt % 32 == 0)
        launch_update<32, 32, 16, 16>(A, lmt, ldt, k, kp, n0, NT, ncol, npairs, info, s, pmap);
    else
        launch_update<16, 16, 16, 16>(A, lmt, ldt, k, kp, n0, NT, ncol, npairs, info, s, pmap);
}

/* ================================================================== */
/* Strided batched GEMM for the solves:
 *   C_j = beta C_j + alpha op(A_j) B_j,   j < batch
 * op(A) is M x K (transA = 0: A(i,p) at A[i + p lda]; 1: at A[p + i lda]),
 * B is K x N with B(p,q) at B[p + q ldb].  M, N, K multiples of the block
 * sizes.  An extra block row (blockIdx.y == batch) copies B_0 into
 * copy_dst (same shape and leading dimension as B) when copy_dst != null,
 * which is how the right-looking solve writes the solved block back while
 * the blocks below are updated with it.                                  */
template <int BM, int BN, int WM, int WN, int TRANSA>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void gemm_batched_kernel(
    int M, int N, int K, double alpha, const double *A0, int lda, long long sA, const double *B0, int ldb,
    long long sB, double beta, double *C0, int ldc, long long sC, int batch, double *copy_dst, const int *info)
{
    constexpr int NW = (BM / WM) * (BN / WN);
    constexpr int NTH = NW * 64;
    constexpr int BK = 16;
    constexpr int FM = WM / 16, FN = WN / 16;
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    if (*info != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = blockIdx.y;
    const int tbm = M / BM;
    const int bi = blockIdx.x % tbm, bj = blockIdx.x / tbm;
    const int row0 = bi * BM, col0 = bj * BN;
    if (j == batch) {
        /* copy job: B_0 -> copy_dst over this block's region (rows of K) */
        for (int e = tid; e < BM * BN; e += NTH) {
            const int i = row0 + e % BM, q = col0 + e / BM;
            if (i < K && q < N) copy_dst[i + (size_t)q * ldb] = B0[i + (size_t)q * ldb];
        }
        return;
    }
    const double *A = A0 + (size_t)j * sA;
    const double *B = B0 + (size_t)j * sB;
    double *C = C0 + (size_t)j * sC;
    const int wm = (wave % (BM / WM)) * WM, wn = (wave / (BM / WM)) * WN;
    v4d acc[FM][FN];
#pragma unroll
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = zero4d();

    for (int kb = 0; kb < K; kb += BK) {
        /* A slice: op(A)(row0 + i, kb + p) */
        for (int e = tid; e < BK * BM; e += NTH) {
            int i, p;
            double v;
            if (TRANSA) {
                p = e % BK; i = e / BK;
                v = A[(kb + p) + (size_t)(row0 + i) * lda];
            } else {
                i = e % BM; p = e / BM;
                v = A[(row0 + i) + (size_t)(kb + p) * lda];
            }
            As[p][i] = v;
        }
        /* B slice: B(kb + p, col0 + q), K-contiguous */
        for (int e = tid; e < BK * BN; e += NTH) {
            const int p = e % BK, q = e / BK;
            Bs[p][q] = B[(kb + p) + (size_t)(col0 + q) * ldb];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double a[FM], bv[FN];
#pragma unroll
            for (int fm = 0; fm < FM; ++fm) a[fm] = As[kk + (lane >> 4)][wm + fm * 16 + (lane & 15)];
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) bv[fn] = Bs[kk + (lane >> 4)][wn + fn * 16 + (lane & 15)];
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < FN; ++fn) acc[fm][fn] = mfma_f64(bv[fn], a[fm], acc[fm][fn]);
        }
        __syncthreads();
    }
    for (int fm = 0; fm < FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int gi = row0 + wm + fm * 16 + (lane & 15);
                const int gj = col0 + wn + fn * 16 + (lane >> 4) + 4 * s;
                double *pc = C + gi + (size_t)gj * ldc;
                const double v = alpha * acc[fm][fn][s];
                *pc = (beta == 0.0) ? v : beta * (*pc) + v;
            }
}

template <int BM, int BN, int WM, int WN>
static void launch_gemm(int transA, int M, int N, int K, double alpha, const double *A, int lda, long long sA,
                        const double *B, int ldb, long long sB, double beta, double *C, int ldc, long long sC,
                        int batch, double *copy_dst, const int *info, hipStream_t s)
{
    constexpr int NTH = (BM / WM) * (BN / WN) * 64;
    dim3 grid((M / BM) * (N / BN), batch + (copy_dst ? 1 : 0));
    if (grid.y == 0) return;
    if (transA)
        hipLaunchKernelGGL((gemm_batched_kernel<BM, BN, WM, WN, 1>), grid, dim3(NTH), 0, s, M, N, K, alpha, A, lda,
                           sA, B, ldb, sB, beta, C, ldc, sC, batch, copy_dst, info);
    else
        hipLaunchKernelGGL((gemm_batched_kernel<BM, BN, WM, WN, 0>), grid, dim3(NTH), 0, s, M, N, K, alpha, A, lda,
                           sA, B, ldb, sB, beta, C, ldc, sC, batch, copy_dst, info);
}

void gemm_batched(int transA, int M, int N, int K, double alpha, const double *A, int lda, long long strideA,
                  const double *B, int ldb, long long strideB, double beta, double *C, int ldc, long long strideC,
                  int batch, double *copy_dst, const int *info, hipStream_t s)
{
    if (M % 64 == 0 && N % 64 == 0)
        launch_gemm<64, 64, 32, 32>(transA, M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc, strideC,
                                    batch, copy_dst, info, s);
    else if (M % 64 == 0 && N % 16 == 0)
        launch_gemm<64, 16, 16, 16>(transA, M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc, strideC,
                                    batch, copy_dst, info, s);
    else
        launch_gemm<16, 16, 16, 16>(transA, M, N, K, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc, strideC,
                                    batch, copy_dst, info, s);
}

} /* namespace k */
} /* namespace chx */
