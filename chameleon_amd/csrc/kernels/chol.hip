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

/* ================================================================== */
/* POTRF of one diagonal tile (one workgroup of 256 threads).
 *
 * Left-looking by 16-column blocks: block c is first updated with the
 * already factored columns (MFMA, operands streamed from L2), its 16 x 16
 * diagonal block is factored in LDS, inverted, and the rows below are
 * multiplied by the inverse transpose (MFMA).  The inverse of every 16 x 16
 * diagonal block is kept in `dinv` for the panel TRSM that follows.
 *
 * The strictly upper triangle of the tile is never written (it is not
 * referenced by a Lower factorisation and stays as the caller left it).
 * info: iinfo + (1-based column of the first non-positive pivot), written
 * only for pivots inside the valid part of the tile (nvalid columns).     */
__global__ __launch_bounds__(256) void potrf_diag_kernel(double *T, int ldt, double *dinv, int iinfo,
                                                          int nvalid, int *info)
{
    __shared__ double Ld[16][17];
    __shared__ double Li[16][17];
    __shared__ int s_fail;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (*info != 0) return;
    if (tid == 0) s_fail = 0;
    __syncthreads();

    for (int c = 0; c < ldt; c += 16) {
        const int nrb = (ldt - c) >> 4;
        /* (a) rows [c, ldt) x cols [c, c+16) -= T[., 0:c] T[c:c+16, 0:c]^T */
        if (c > 0) {
            for (int rb = wave; rb < nrb; rb += 4) {
                const int r0 = c + rb * 16;
                v4d acc0 = zero4d(), acc1 = zero4d();
                const double *pa = T + (r0 + (lane & 15)) + (size_t)(lane >> 4) * ldt;
                const double *pb = T + (c + (lane & 15)) + (size_t)(lane >> 4) * ldt;
                for (int p = 0; p < c; p += 8) {
                    double a0 = pa[(size_t)p * ldt], b0 = pb[(size_t)p * ldt];
                    double a1 = pa[(size_t)(p + 4) * ldt], b1 = pb[(size_t)(p + 4) * ldt];
                    acc0 = mfma_f64(b0, a0, acc0);
                    acc1 = mfma_f64(b1, a1, acc1);
                }
                /* lane holds (row r0 + lane%16, col c + lane/16 + 4 s) */
                for (int s = 0; s < 4; ++s) {
                    const int i = lane & 15, j = (lane >> 4) + 4 * s;
                    if (rb == 0 && i < j) continue;
                    T[(r0 + i) + (size_t)(c + j) * ldt] -= acc0[s] + acc1[s];
                }
            }
            __syncthreads();
        }
        /* (b) factor the 16 x 16 diagonal block in LDS */
        {
            const int i = tid & 15, j = tid >> 4;
            Ld[i][j] = (i >= j) ? T[(c + i) + (size_t)(c + j) * ldt] : 0.0;
        }
        __syncthreads();
        for (int j = 0; j < 16; ++j) {
            const double d = Ld[j][j];
            if (!(d > 0.0)) {
                if (tid == 0) s_fail = (c + j < nvalid) ? c + j + 1 : -1;
            }
            __syncthreads();
            if (s_fail) goto done;
            const double sd = sqrt(d);
            const int i = tid & 15, jj = tid >> 4;
            const double lij = (i > j) ? Ld[i][j] / sd : 0.0;
            const double ljj = (jj > j) ? Ld[jj][j] / sd : 0.0;
            __syncthreads();
            if (jj > j && i >= jj) Ld[i][jj] -= lij * ljj;
            if (jj == j) {
                if (i > j) Ld[i][j] = lij;
                else if (i == j) Ld[j][j] = sd;
            }
            __syncthreads();
        }
        /* (c) inverse of the lower triangular block, one column per thread */
        if (tid < 16) {
            const int j = tid;
            for (int i = 0; i < 16; ++i) {
                double s = (i == j) ? 1.0 : 0.0;
                for (int p = j; p < i; ++p) s -= Ld[i][p] * Li[p][j];
                Li[i][j] = (i < j) ? 0.0 : s / Ld[i][i];
            }
        }
        __syncthreads();
        {
            const int i = tid & 15, j = tid >> 4;
            if (i >= j) T[(c + i) + (size_t)(c + j) * ldt] = Ld[i][j];
            dinv[(size_t)(c >> 4) * 256 + i + j * 16] = Li[i][j];
        }
        /* (d) rows below: X = P Li^T, as D = Li P^T so the result lands in
         * the lane layout of coalesced column stores */
        for (int rb = 1 + wave; rb < nrb; rb += 4) {
            const int r0 = c + rb * 16;
            double pv[4];
            for (int kk = 0; kk < 4; ++kk)
                pv[kk] = T[(r0 + (lane & 15)) + (size_t)(c + 4 * kk + (lane >> 4)) * ldt];
            v4d acc = zero4d();
            for (int kk = 0; kk < 4; ++kk)
                acc = mfma_f64(Li[lane & 15][4 * kk + (lane >> 4)], pv[kk], acc);
            for (int s = 0; s < 4; ++s)
                T[(r0 + (lane & 15)) + (size_t)(c + (lane >> 4) + 4 * s) * ldt] = acc[s];
        }
        __syncthreads();
    }
done:
    if (tid == 0 && s_fail > 0) *info = iinfo + s_fail;
}

void potrf_diag(double *Akk, int ldt, double *dinv, int iinfo, int nvalid, int *info, hipStream_t s)
{
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(256), 0, s, Akk, ldt, dinv, iinfo, nvalid, info);
}

/* ================================================================== */
/* Right-side triangular solve X = P L^-T, in place, for a batch of tiles
 * (the TRSM(Right, Lower, Trans, NonUnit) of pzpotrf.c:99-105 applied to
 * every tile of panel column k at once).  One wavefront owns a 16-row strip
 * of one tile and walks its 16-column blocks left to right:
 *   R = P[:, c] - X[:, 0:c] L[c, 0:c]^T      (MFMA, K = c)
 *   X[:, c] = R Linv_c^T                     (MFMA with the 16x16 inverse)
 * Tile t uses L = L0 + t*lstride and the block inverses dinv0 + t*dstride,
 * so the same kernel inverts all diagonal tiles at once when applied to
 * identity tiles (X = I L^-T = Linv^T).                                  */
__global__ __launch_bounds__(256) void trsm_right_kernel(double *P0, int ntiles, long long pstride, int ldt,
                                                          const double *L0, long long lstride,
                                                          const double *dinv0, long long dstride,
                                                          const int *info)
{
    __shared__ double S[4][16][17];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (*info != 0) return;
    const int nrb = ldt >> 4;
    const int wid = blockIdx.x * 4 + wave;
    const int t = wid / nrb, rb = wid % nrb;
    if (t >= ntiles) return;
    double *X = P0 + (size_t)t * pstride;
    const double *L = L0 + (size_t)t * lstride;
    const double *dinv = dinv0 + (size_t)t * dstride;
    const int r0 = rb * 16;
    const int li = lane & 15, lk = lane >> 4;
    for (int c = 0; c < ldt; c += 16) {
        v4d acc0 = zero4d(), acc1 = zero4d();
        const double *pl = L + (c + li) + (size_t)lk * ldt;
        const double *px = X + (r0 + li) + (size_t)lk * ldt;
        for (int p = 0; p < c; p += 8) {
            double l0 = pl[(size_t)p * ldt], x0 = px[(size_t)p * ldt];
            double l1 = pl[(size_t)(p + 4) * ldt], x1 = px[(size_t)(p + 4) * ldt];
            acc0 = mfma_f64(l0, x0, acc0);
            acc1 = mfma_f64(l1, x1, acc1);
        }
        /* lane holds D[j][i] = sum_p L[c+j][p] X[r0+i][p], i = li, j = lk + 4s */
        for (int s = 0; s < 4; ++s) {
            const int j = lk + 4 * s;
            S[wave][li][j] = X[(r0 + li) + (size_t)(c + j) * ldt] - (acc0[s] + acc1[s]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const double *di = dinv + (size_t)(c >> 4) * 256;
        v4d y = zero4d();
        for (int kk = 0; kk < 4; ++kk) {
            const int q = 4 * kk + lk;
            y = mfma_f64(di[li + q * 16], S[wave][li][q], y);
        }
        for (int s = 0; s < 4; ++s)
            X[(r0 + li) + (size_t)(c + lk + 4 * s) * ldt] = y[s];
        __threadfence_block();
        __builtin_amdgcn_wave_barrier();
    }
}

void trsm_right_lower_trans_batched(double *P0, int ntiles, long long pstride, int ldt,
                                    const double *L0, long long lstride, const double *dinv0, long long dstride,
                                    const int *info, hipStream_t s)
{
    if (ntiles <= 0) return;
    const int waves = ntiles * (ldt >> 4);
    const int blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(trsm_right_kernel, dim3(blocks), dim3(256), 0, s, P0, ntiles, pstride, ldt,
                       L0, lstride, dinv0, dstride, info);
}

void trsm_right_lower_trans(double *panel, int ntiles, size_t tile_stride, int ldt, const double *Lkk,
                            const double *dinv, const int *info, hipStream_t s)
{
    trsm_right_lower_trans_batched(panel, ntiles, (long long)tile_stride, ldt, Lkk, 0, dinv, 0, info, s);
}

/* ================================================================== */
/* Inverses of the 16 x 16 diagonal blocks of every diagonal tile (for a
 * factor that was not produced by potrf_diag in this process).           */
__global__ __launch_bounds__(256) void dinv_kernel(const double *A, int nt, long long diag_stride, int ldt,
                                                    double *dinv)
{
    __shared__ double Ld[4][16][17];
    __shared__ double Li[4][16][17];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nrb = ldt >> 4;
    const int wid = blockIdx.x * 4 + wave;
    const int t = wid / nrb, c = (wid % nrb) * 16;
    const bool active = t < nt;
    if (active) {
        const double *T = A + (size_t)t * diag_stride;
        for (int e = lane; e < 256; e += 64) {
            const int i = e & 15, j = e >> 4;
            Ld[wave][i][j] = (i >= j) ? T[(c + i) + (size_t)(c + j) * ldt] : 0.0;
        }
    }
    __syncthreads();
    if (active && lane < 16) {
        const int j = lane;
        for (int i = 0; i < 16; ++i) {
            double s = (i == j) ? 1.0 : 0.0;
            for (int p = j; p < i; ++p) s -= Ld[wave][i][p] * Li[wave][p][j];
            Li[wave][i][j] = (i < j) ? 0.0 : s / Ld[wave][i][i];
        }
    }
    __syncthreads();
    if (active) {
        double *out = dinv + ((size_t)t * nrb + (c >> 4)) * 256;
        for (int e = lane; e < 256; e += 64) {
            const int i = e & 15, j = e >> 4;
            out[i + j * 16] = Li[wave][i][j];
        }
    }
}

void dinv_all(const double *A, int nt, size_t diag_stride, int ldt, double *dinv, hipStream_t s)
{
    const int waves = nt * (ldt >> 4);
    hipLaunchKernelGGL(dinv_kernel, dim3((waves + 3) / 4), dim3(256), 0, s, A, nt, (long long)diag_stride, ldt, dinv);
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
 * n <= m of either one column (n = n0, colmode) or the whole trailing
 * triangle n0 <= n <= m < NT.  Diagonal tiles (m == n) only write their
 * lower triangle: that is the SYRK of pzpotrf.c:112-117, the off-diagonal
 * pairs are its GEMMs (pzpotrf.c:123-129), all in one launch.
 *
 * Block tile BM x BN, K staged through LDS in BK-deep slices with the next
 * slice prefetched into registers while the current one is multiplied.
 * Each wavefront owns a WM x WN sub-tile of (WM/16)x(WN/16) MFMA blocks.
 * Operands are swapped (D = B A^T) so each lane's results are consecutive
 * rows of one column: stores of 16 lanes form one 128-byte segment.      */
template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64) void trailing_update_kernel(
    double *A, int lmt, int ldt, int k, int n0, int NT, int colmode, int total_blocks, const int *info)
{
    constexpr int NW = (BM / WM) * (BN / WN);
    constexpr int NTH = NW * 64;
    constexpr int BK = 16;
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LA = (BK * BM) / (2 * NTH);   /* v2d loads per thread for A */
    constexpr int LB = (BK * BN) / (2 * NTH);
    static_assert(LA >= 1 && LB >= 1, "tile too small for the thread count");
    __shared__ double As[BK][BM];
    __shared__ double Bs[BK][BN];

    if (*info != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bid = xcd_remap(blockIdx.x, total_blocks);
    const int tbm = ldt / BM, tbn = ldt / BN;
    const int bpt = tbm * tbn;
    const int pair = bid / bpt, b = bid % bpt;
    int m, n;
    if (colmode) {
        n = n0;
        m = n0 + pair;
    } else {
        /* pair -> (i, j), i >= j, row-major over the lower triangle */
        int i = (int)((sqrt(8.0 * (double)pair + 1.0) - 1.0) * 0.5);
        while ((long long)(i + 1) * (i + 2) / 2 <= pair) ++i;
        while ((long long)i * (i + 1) / 2 > pair) --i;
        const int j = pair - i * (i + 1) / 2;
        m = n0 + i;
        n = n0 + j;
    }
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
    gload_slice<BM, LA, NTH>(ra, Am, 0, ldt, tid);
    gload_slice<BN, LB, NTH>(rb, An, 0, ldt, tid);
    lstore_slice<BM, LA, NTH>(&As[0][0], ra, tid);
    lstore_slice<BN, LB, NTH>(&Bs[0][0], rb, tid);
    __syncthreads();
    for (int kb = 0; kb < ldt; kb += BK) {
        const bool more = kb + BK < ldt;
        if (more) {
            gload_slice<BM, LA, NTH>(ra, Am, kb + BK, ldt, tid);
            gload_slice<BN, LB, NTH>(rb, An, kb + BK, ldt, tid);
        }
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
        if (more) {
            lstore_slice<BM, LA, NTH>(&As[0][0], ra, tid);
            lstore_slice<BN, LB, NTH>(&Bs[0][0], rb, tid);
            __syncthreads();
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
                if (!diag || gi >= gj) C[gi + (size_t)gj * ldt] -= acc[fm][fn][s];
            }
}

template <int BM, int BN, int WM, int WN>
static void launch_update(double *A, int lmt, int ldt, int k, int n0, int NT, int colmode, int npairs,
                          const int *info, hipStream_t s)
{
    const int bpt = (ldt / BM) * (ldt / BN);
    const long long total = (long long)npairs * bpt;
    if (total <= 0) return;
    constexpr int NTH = (BM / WM) * (BN / WN) * 64;
    hipLaunchKernelGGL((trailing_update_kernel<BM, BN, WM, WN>), dim3((unsigned)total), dim3(NTH), 0, s, A, lmt,
                       ldt, k, n0, NT, colmode, (int)total, info);
}

void trailing_update(double *A, int lmt, int ldt, int k, int n0, int n1, int NT, const int *info, hipStream_t s)
{
    if (n0 >= NT || n1 <= n0) return;
    const bool colmode = (n1 == n0 + 1);
    const int S = NT - n0;
    const int npairs = colmode ? S : S * (S + 1) / 2;
    if (ldt % 128 == 0)
        launch_update<128, 128, 64, 64>(A, lmt, ldt, k, n0, NT, colmode, npairs, info, s);
    else if (ldt % 64 == 0)
        launch_update<64, 64, 32, 32>(A, lmt, ldt, k, n0, NT, colmode, npairs, info, s);
    else if (ldt % 32 == 0)
        launch_update<32, 32, 16, 16>(A, lmt, ldt, k, n0, NT, colmode, npairs, info, s);
    else
        launch_update<16, 16, 16, 16>(A, lmt, ldt, k, n0, NT, colmode, npairs, info, s);
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
