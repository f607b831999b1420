/*
 * solve.hip -- the two tiled triangular solves of dpotrs (reference
 * chameleon_pztrsm, compute/pztrsm.c:134-201: Left/Lower/NoTrans forward,
 * then Left/Lower/Trans backward) as one persistent cooperative kernel per
 * direction.
 *
 * The right-hand sides are held transposed during the solve (BT: nrhs x N,
 * element (j, p) at j + p*ldb) so that every MFMA operand load is a run of
 * 16 consecutive doubles.  With Tinv_k the inverse of diagonal tile k in
 * both orientations (TN = L_kk^-1, TT = L_kk^-T), step k of the forward
 * solve is
 *     phase 1:  W = L_kk^-1 B_k                      (B_k = rows of tile k)
 *     phase 2:  B_m -= L(m,k) W for m > k,  B_k = W
 * and of the backward solve (k descending)
 *     phase 1:  W = L_kk^-T B_k
 *     phase 2:  B_m -= L(k,m)^T W for m < k, B_k = W
 * Phases are separated by grid-wide barriers; work is cut in 16 x 16
 * output blocks with K = ldt, spread over all wavefronts of the grid.
 */
#include "../internal.h"
#include "mfma.h"

#include <hip/hip_cooperative_groups.h>

namespace cg = cooperative_groups;

namespace chx {
namespace k {

/* acc = sum_{p < K} opA(i0 + i, p) * Bv(p, j0 + j) for one 16 x 16 block.
 * opA element (i, p) at A[i * sa_i + p * sa_p]; Bv element (p, j) at
 * Bv[j + p * ldb].  Result slot s of lane l: row l/16 + 4 s, column l%16. */
__device__ __forceinline__ v4d block_dot(const double *A, long long sa_i, long long sa_p, const double *Bv, int ldb,
                                         int K, int lane)
{
    const int li = lane & 15, lk = lane >> 4;
    const double *pa = A + (long long)li * sa_i + (long long)lk * sa_p;
    const double *pb = Bv + li + (size_t)lk * ldb;
    v4d acc0 = zero4d(), acc1 = zero4d();
#pragma unroll 2
    for (int p = 0; p < K; p += 16) {
        double a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a[q] = pa[(long long)(p + 4 * q) * sa_p];
            b[q] = pb[(size_t)(p + 4 * q) * ldb];
        }
        acc0 = mfma_f64(a[0], b[0], acc0);
        acc1 = mfma_f64(a[1], b[1], acc1);
        acc0 = mfma_f64(a[2], b[2], acc0);
        acc1 = mfma_f64(a[3], b[3], acc1);
    }
    return acc0 + acc1;
}

struct SolveArgs {
    const double *A;     /* factor mirror, tile (m,n) at ((n*NT + m) * ldt*ldt) */
    const double *TN;    /* L_kk^-1, one ldt x ldt tile per k                  */
    const double *TT;    /* L_kk^-T                                            */
    double *BT;          /* right-hand sides, transposed: nrhs x (NT*ldt)       */
    double *W;           /* one solved block, transposed: nrhs x ldt            */
    int NT, ldt, nrhs;   /* nrhs: padded to a multiple of 16                    */
    int trans;           /* 0: forward L Y = B, 1: backward L^T X = Y           */
    const int *info;
};

__global__ __launch_bounds__(256) void solve_coop_kernel(SolveArgs a)
{
    cg::grid_group grid = cg::this_grid();
    if (*a.info != 0) return;   /* uniform across the grid: no barrier is left waiting */
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int gwave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * 4;
    const int ldt = a.ldt, NT = a.NT, nr = a.nrhs;
    const size_t ts = (size_t)ldt * ldt;
    const int ib = ldt / 16, jb = nr / 16;
    for (int kk = 0; kk < NT; ++kk) {
        const int k = a.trans ? NT - 1 - kk : kk;
        /* phase 1: W(i, j) = sum_p op(i, p) B_k(p, j), op = TN or TT of tile k
         * (both stored column-major: element (i, p) at i + p ldt) */
        {
            const double *op = (a.trans ? a.TT : a.TN) + (size_t)k * ts;
            const double *Bk = a.BT + (size_t)k * ldt * nr;
            for (int u = gwave; u < ib * jb; u += nwaves) {
                const int i0 = (u % ib) * 16, j0 = (u / ib) * 16;
                v4d acc = block_dot(op + i0, 1, ldt, Bk + j0, nr, ldt, lane);
#pragma unroll
                for (int s = 0; s < 4; ++s) a.W[(j0 + li) + (size_t)(i0 + lk + 4 * s) * nr] = acc[s];
            }
        }
        grid.sync();
        /* phase 2: B_m -= op(m) W for the blocks on the far side of k, B_k = W */
        {
            const int m0 = a.trans ? 0 : k + 1;
            const int cnt = a.trans ? k : NT - k - 1;
            const int units = cnt * ib * jb;
            for (int u = gwave; u < units + ib * jb; u += nwaves) {
                if (u >= units) {
                    /* copy W into block k */
                    const int v = u - units;
                    const int i0 = (v % ib) * 16, j0 = (v / ib) * 16;
                    double *Bk = a.BT + (size_t)k * ldt * nr;
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const size_t e = (j0 + li) + (size_t)(i0 + lk + 4 * s) * nr;
                        Bk[e] = a.W[e];
                    }
                    continue;
                }
                const int t = u / (ib * jb), r = u % (ib * jb);
                const int m = m0 + t;
                const int i0 = (r % ib) * 16, j0 = (r / ib) * 16;
                v4d acc;
                if (!a.trans) {
                    /* op = L(m,k), element (i, p) at i + p ldt */
                    const double *Amk = a.A + ((size_t)k * NT + m) * ts;
                    acc = block_dot(Amk + i0, 1, ldt, a.W + j0, nr, ldt, lane);
                } else {
                    /* op = L(k,m)^T, element (i, p) at p + i ldt */
                    const double *Akm = a.A + ((size_t)m * NT + k) * ts;
                    acc = block_dot(Akm + (size_t)i0 * ldt, ldt, 1, a.W + j0, nr, ldt, lane);
                }
                double *Bm = a.BT + (size_t)m * ldt * nr;
#pragma unroll
                for (int s = 0; s < 4; ++s) Bm[(j0 + li) + (size_t)(i0 + lk + 4 * s) * nr] -= acc[s];
            }
        }
        grid.sync();
    }
}

static int coop_blocks()
{
    static int cached = 0;
    if (cached) return cached;
    int dev = 0, ncu = 0, per = 0;
    CHX_HIP(hipGetDevice(&dev));
    CHX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    CHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, solve_coop_kernel, 256, 0));
    if (per < 1) per = 1;
    cached = ncu * (per < 2 ? per : 2);
    return cached;
}

void solve_coop(const double *A, const double *TN, const double *TT, double *BT, double *W, int NT, int ldt, int nrhs,
                int trans, const int *info, hipStream_t s)
{
    SolveArgs a;
    a.A = A; a.TN = TN; a.TT = TT; a.BT = BT; a.W = W;
    a.NT = NT; a.ldt = ldt; a.nrhs = nrhs; a.trans = trans; a.info = info;
    void *args[] = {&a};
    CHX_HIP(hipLaunchCooperativeKernel((const void *)solve_coop_kernel, dim3(coop_blocks()), dim3(256), args, 0, s));
}

/* ------------------------------------------------------------------ */
/* B mirror (NT x bnt tiles of ldt x wdt) <-> BT (nrhs_p x NT*ldt), with
 * nrhs_p = bnt * wdt. */
__global__ void rhs_transpose_kernel(double *Bm, double *BT, int NT, int ldt, int wdt, int bnt, int to_bt)
{
    __shared__ double S[32][33];
    const int nrp = bnt * wdt;
    const int rows = NT * ldt;
    const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;   /* rows of B, columns (rhs) of B */
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    auto baddr = [&](int r, int c) -> size_t {
        const int m = r / ldt, i = r % ldt, n = c / wdt, j = c % wdt;
        return ((size_t)n * NT + m) * ldt * wdt + i + (size_t)j * ldt;
    };
    if (to_bt) {
        for (int jj = ty; jj < 32; jj += 8) {
            const int r = r0 + tx, c = c0 + jj;
            S[tx][jj] = (r < rows && c < nrp) ? Bm[baddr(r, c)] : 0.0;
        }
        __syncthreads();
        for (int jj = ty; jj < 32; jj += 8) {
            const int c = c0 + tx, r = r0 + jj;
            if (r < rows && c < nrp) BT[c + (size_t)r * nrp] = S[jj][tx];
        }
    } else {
        for (int jj = ty; jj < 32; jj += 8) {
            const int c = c0 + tx, r = r0 + jj;
            S[jj][tx] = (r < rows && c < nrp) ? BT[c + (size_t)r * nrp] : 0.0;
        }
        __syncthreads();
        for (int jj = ty; jj < 32; jj += 8) {
            const int r = r0 + tx, c = c0 + jj;
            if (r < rows && c < nrp) Bm[baddr(r, c)] = S[tx][jj];
        }
    }
}

void rhs_transpose(double *Bm, double *BT, int NT, int ldt, int wdt, int bnt, int to_bt, hipStream_t s)
{
    const int rows = NT * ldt, nrp = bnt * wdt;
    dim3 grid((rows + 31) / 32, (nrp + 31) / 32);
    hipLaunchKernelGGL(rhs_transpose_kernel, grid, dim3(256), 0, s, Bm, BT, NT, ldt, wdt, bnt, to_bt);
}

/* Per-tile transpose of a batch of square tiles (TT -> TN). */
__global__ void tile_transpose_kernel(const double *src, double *dst, int ldt)
{
    __shared__ double S[32][33];
    const size_t ts = (size_t)ldt * ldt;
    const double *sp = src + (size_t)blockIdx.z * ts;
    double *dp = dst + (size_t)blockIdx.z * ts;
    const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int jj = ty; jj < 32; jj += 8) {
        const int r = r0 + tx, c = c0 + jj;
        S[tx][jj] = (r < ldt && c < ldt) ? sp[r + (size_t)c * ldt] : 0.0;
    }
    __syncthreads();
    for (int jj = ty; jj < 32; jj += 8) {
        const int r = c0 + tx, c = r0 + jj;
        if (r < ldt && c < ldt) dp[r + (size_t)c * ldt] = S[jj][tx];
    }
}

void tile_transpose(const double *src, double *dst, int ntiles, int ldt, hipStream_t s)
{
    if (ntiles <= 0) return;
    dim3 grid((ldt + 31) / 32, (ldt + 31) / 32, ntiles);
    hipLaunchKernelGGL(tile_transpose_kernel, grid, dim3(256), 0, s, src, dst, ldt);
}

} /* namespace k */
} /* namespace chx */
