/*
 * batch.hip -- many independent SPD systems on one device: generation and
 * the two triangular solves.  (The factorisations run on the job-queue
 * form of the POTRF kernel in panel.hip.)
 *
 * Storage of a batch: system j is an np x np column-major matrix (np = n
 * rounded up to 16, identity in the padding) with leading dimension ld,
 * its right-hand sides an np x nrhs column-major block (zero padding), and
 * the 64 x 64 inverses of its diagonal blocks, all at per-system offsets.
 */
#include "../internal.h"
#include "mfma.h"
#include "rng.h"

namespace chx {
namespace k {

/* ------------------------------------------------------------------ */
/* Generators.  Element (r, c), r, c < valid, of system j draws the LCG
 * value at position
 *     plgsy: (max(r,c) + base) + (min(r,c) + base) * bigM, + bump on the diagonal
 *     plrnt: (r + base) + (c + base) * bigM
 * (the element rule of coreblas/compute/core_zplgsy.c:38-112 and
 * core_zplrnt.c:36-54 for a matrix whose origin is at (base, base)).
 * One thread fills GRUN consecutive rows of one column of one system;
 * unit0 (count + 1 entries) is the prefix sum of the threads per system. */
#define GRUN 32

__global__ __launch_bounds__(256) void batch_gen_kernel(double *A, const BatchGenSys *sys, const long long *unit0,
                                                         int count, int sym)
{
    const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= unit0[count]) return;
    int lo = 0, hi = count - 1;
    while (lo < hi) {                    /* last system with unit0 <= u */
        const int mid = (lo + hi + 1) >> 1;
        if (unit0[mid] <= u) lo = mid; else hi = mid - 1;
    }
    const BatchGenSys g = sys[lo];
    const long long v = u - unit0[lo];
    const int runs = (g.rows + GRUN - 1) / GRUN;
    const int c = (int)(v / runs), r0 = (int)(v % runs) * GRUN;
    const int r1 = min(r0 + GRUN, g.rows);
    double *col = A + g.off + (size_t)c * g.ld;
    unsigned long long x = 0;
    int next = -1, mode = -1;            /* row at which x is valid, and its step kind */
    for (int r = r0; r < r1; ++r) {
        double val;
        if (r >= g.nvr || c >= g.nvc) {
            val = (sym && r == c && g.pad_one) ? 1.0 : 0.0;
        } else {
            unsigned long long pos;
            int m;                        /* 0: next row is pos + 1, 1: pos + bigM */
            if (!sym || r >= c) {
                pos = (unsigned long long)(r + g.base) + (unsigned long long)(c + g.base) * (unsigned long long)g.bigM;
                m = 0;
            } else {
                pos = (unsigned long long)(c + g.base) + (unsigned long long)(r + g.base) * (unsigned long long)g.bigM;
                m = 1;
            }
            if (next != r || mode != m) x = rnd_jump(pos, g.seed);
            val = rnd_val(x);
            if (sym && r == c) val += g.bump;
            x = m == 0 ? RND_A * x + RND_C : g.abig * x + g.cbig;
            next = r + 1;
            mode = m;
        }
        col[r] = val;
    }
}

void batch_gen(double *A, const BatchGenSys *sys, const long long *unit0, long long units, int count, int sym,
               hipStream_t s)
{
    if (units <= 0) return;
    hipLaunchKernelGGL(batch_gen_kernel, dim3((unsigned)((units + 255) / 256)), dim3(256), 0, s, A, sys, unit0, count,
                       sym);
}

/* ------------------------------------------------------------------ */
/* Forward and backward solve of every system, X = L^-T L^-1 B in place,
 * one workgroup of 4 wavefronts per system (taken from a queue), 64-row
 * blocks with the inverses of the diagonal blocks:
 *   forward  block J:  S = B_J - L(J, 0:J) X(0:J),          X_J = D_J S
 *   backward block J:  S = Y_J - L(J+1:, J)^T X(J+1:),      X_J = D_J^T S
 * Wave w owns rows 16w..16w+15 of the block; right-hand sides go in
 * chunks of 16 columns.  Systems whose factorisation failed are skipped. */
struct BatchSolveArgs {
    const double *A;
    const long long *aoff;
    const int *np, *ld;
    const double *dinv;
    const long long *doff;
    double *B;
    const long long *boff;
    const int *nrhs;
    const int *info;
    int count;
    int *queue;
};

__global__ __launch_bounds__(256) void batch_solve_kernel(BatchSolveArgs a)
{
    __shared__ double S[64][17];
    __shared__ int s_job;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_job = atomicAdd(a.queue, 1);
        __syncthreads();
        const int j = s_job;
        if (j >= a.count) break;
        if (a.info[j] != 0) continue;
        const double *L = a.A + a.aoff[j];
        const int np = a.np[j], ld = a.ld ? a.ld[j] : np;
        const double *Dv = a.dinv + a.doff[j];
        double *X = a.B + a.boff[j];
        const int ldb = np, nr = a.nrhs[j];
        const int nblk = (np + 63) / 64;
        for (int c0 = 0; c0 < nr; c0 += 16) {
            const int col = c0 + li;
            const bool cv = col < nr;
            const double *Xc = X + (size_t)(cv ? col : 0) * ldb;
            for (int dir = 0; dir < 2; ++dir) {
                for (int t = 0; t < nblk; ++t) {
                    const int J = dir == 0 ? t : nblk - 1 - t;
                    const int R0 = 64 * J, W = min(64, np - R0), ng = W >> 4;
                    const int r = R0 + 16 * wave;
                    if (wave < ng) {
                        v4d acc = zero4d();
                        const int pb = dir == 0 ? 0 : R0 + W, pe = dir == 0 ? R0 : np;
                        for (int p0 = pb; p0 < pe; p0 += 16) {
                            double av[4];
                            if (dir == 0) {
#pragma unroll
                                for (int q = 0; q < 4; ++q) av[q] = L[(r + li) + (size_t)(p0 + 4 * lk + q) * ld];
                            } else {
                                const v4d t4 = *(const v4d *)(L + (p0 + 4 * lk) + (size_t)(r + li) * ld);
#pragma unroll
                                for (int q = 0; q < 4; ++q) av[q] = t4[q];
                            }
                            v4d xb = zero4d();
                            if (cv) xb = *(const v4d *)(Xc + p0 + 4 * lk);
#pragma unroll
                            for (int q = 0; q < 4; ++q) acc = mfma_f64(av[q], xb[q], acc);
                        }
#pragma unroll
                        for (int e = 0; e < 4; ++e) S[16 * wave + lk + 4 * e][li] = (cv ? Xc[r + lk + 4 * e] : 0.0) - acc[e];
                    }
                    __syncthreads();
                    if (wave < ng) {
                        const double *D = Dv + (size_t)J * 4096;
                        v4d x = zero4d();
                        const int qb = dir == 0 ? 0 : 16 * wave, qe = dir == 0 ? 16 * wave + 16 : W;
                        for (int p0 = qb; p0 < qe; p0 += 16) {
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                const int p = p0 + 4 * lk + q;
                                const double dv = dir == 0 ? D[(16 * wave + li) + p * 64] : D[p + (16 * wave + li) * 64];
                                x = mfma_f64(dv, S[p][li], x);
                            }
                        }
                        if (cv) {
#pragma unroll
                            for (int e = 0; e < 4; ++e) X[(r + lk + 4 * e) + (size_t)col * ldb] = x[e];
                        }
                    }
                    __syncthreads();
                }
            }
        }
    }
}

void batch_solve(const double *A, const long long *aoff, const int *np, const int *ld, const double *dinv,
                 const long long *doff, double *B, const long long *boff, const int *nrhs, const int *info, int count,
                 int *queue, int grid, hipStream_t s)
{
    if (count <= 0) return;
    if (grid <= 0) {
        static int resident = 0;
        if (!resident) {
            int dev = 0, ncu = 0, per = 0;
            CHX_HIP(hipGetDevice(&dev));
            CHX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            CHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_solve_kernel, 256, 0));
            resident = ncu * (per > 0 ? per : 1);
        }
        grid = resident;
    }
    if (grid > count) grid = count;
    BatchSolveArgs a;
    a.A = A; a.aoff = aoff; a.np = np; a.ld = ld; a.dinv = dinv; a.doff = doff;
    a.B = B; a.boff = boff; a.nrhs = nrhs; a.info = info; a.count = count; a.queue = queue;
    hipLaunchKernelGGL(batch_solve_kernel, dim3(grid), dim3(256), 0, s, a);
}

/* Clears n ints (queue counters) inside the stream order, without a
 * memset node, so the sequence can be captured in a graph. */
__global__ void zero_ints_kernel(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void zero_ints(int *p, int n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}

/* Padded diagonal of every tile of a mirror seen as a batch of square
 * systems (tile (m,n) holds a v x v system, v = min(valid rows, valid
 * columns), padded to rup16(v)): mode 1 sets it to 1 (ready to factor),
 * mode 0 restores the mirror convention (1 in diagonal tiles, 0 in the
 * others).  One workgroup per tile. */
__global__ void tile_pad_diag_kernel(double *d, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N, int mode)
{
    const int t = blockIdx.x;
    const int m = t % mt, n = t / mt;
    const int vm = min(mb, M - m * mb), vn = min(nb, N - n * nb);
    const int v = min(vm, vn), np = min(min(ldt, wdt), (v + 15) & ~15);
    double *T = d + (size_t)t * ldt * wdt;
    const double val = (mode == 1 || m == n) ? 1.0 : 0.0;
    /* diagonal tiles keep the identity over their whole padding */
    const int end = (mode == 0 && m == n) ? min(ldt, wdt) : np;
    for (int i = v + threadIdx.x; i < end; i += blockDim.x) T[i + (size_t)i * ldt] = val;
}

void tile_pad_diag(double *d, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N, int mode, hipStream_t s)
{
    if (mt * nt <= 0) return;
    hipLaunchKernelGGL(tile_pad_diag_kernel, dim3(mt * nt), dim3(64), 0, s, d, mt, nt, ldt, wdt, mb, nb, M, N, mode);
}

} /* namespace k */
} /* namespace chx */
