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

#include <algorithm>
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

int batch_solve_resident()
{
    static int resident = 0;
    if (!resident) {
        int dev = 0, ncu = 0, per = 0;
        CHX_HIP(hipGetDevice(&dev));
        CHX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        CHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_solve_kernel, 256, 0));
        resident = ncu * (per > 0 ? per : 1);
    }
    return resident;
}

void batch_solve(const double *A, const long long *aoff, const int *np, const int *ld, const double *dinv,
                 const long long *doff, double *B, const long long *boff, const int *nrhs, const int *info, int count,
                 int *queue, int grid, hipStream_t s)
{
    if (count <= 0) return;
    if (grid <= 0) grid = batch_solve_resident();
    if (grid > count) grid = count;
    BatchSolveArgs a;
    a.A = A; a.aoff = aoff; a.np = np; a.ld = ld; a.dinv = dinv; a.doff = doff;
    a.B = B; a.boff = boff; a.nrhs = nrhs; a.info = info; a.count = count; a.queue = queue;
    hipLaunchKernelGGL(batch_solve_kernel, dim3(grid), dim3(256), 0, s, a);
}

/* ------------------------------------------------------------------ */
/* Refinement pieces of the batched mixed-precision solve.
 *
 * batch_residual_kernel: for every system j still refining (state[j] == 0)
 *   R = B - A X in FP64, A symmetric with its lower triangle stored
 *   (np x np, leading dimension np), then LAPACK dsposv's test per
 *   right-hand side c:  ||R_c||_inf <= ||X_c||_inf ||A||_inf eps sqrt(n);
 *   all pass -> state[j] = 1, else iters[j] += 1 and *unconv += 1.  The
 *   backward error max_c ||R_c|| / (||A|| ||X_c|| n eps) goes to berr[j].
 *   ||A||_inf (largest row sum of |A|) is formed in the same pass when
 *   anrm[j] < 0.
 * One workgroup of 4 wavefronts per system (queue).  A wavefront's unit is
 * a 16-row x 16-column block of A X, formed with FP64 MFMAs over k in
 * chunks of 16 (lane (li, lk) supplies k = p0 + 4 lk + q in product q):
 * chunks left of the block's diagonal read A as stored (16 consecutive
 * rows per load), chunks right of it read the stored lower triangle
 * transposed (each lane 4 consecutive k of one column: a 32-byte load),
 * the diagonal chunk element by element.  Each sum runs in a fixed order
 * and the maxima are order-free, so the result is deterministic. */
struct BatchResArgs {
    const double *A;
    const long long *aoff;
    const int *np, *n;
    const double *B, *X;
    double *R;
    const long long *boff;
    const int *nrhs;
    double *anrm;
    int *state, *iters, *unconv;
    double *berr;
    int count;
    int *queue;
};

#define RES_CB 256       /* right-hand-side columns per pass of a system */

__device__ __forceinline__ void lds_max_bits(double *p, double v)
{
    /* non-negative doubles order like their bits; a NaN (made positive)
     * compares above every number, so the maximum propagates it */
    unsigned long long u = (unsigned long long)__double_as_longlong(fabs(v));
    if (v != v) u = 0x7ff8000000000000ull;
    atomicMax(reinterpret_cast<unsigned long long *>(p), u);
}

__global__ __launch_bounds__(256) void batch_residual_kernel(BatchResArgs a)
{
    __shared__ double colr[RES_CB], colx[RES_CB];
    __shared__ double red[256];
    __shared__ int s_job;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const double eps = 0x1p-53;
    for (;;) {
        __syncthreads();
        if (tid == 0) s_job = atomicAdd(a.queue, 1);
        __syncthreads();
        const int j = s_job;
        if (j >= a.count) break;
        if (a.state[j] != 0) continue;
        const double *A = a.A + a.aoff[j];
        const int np = a.np[j], n = a.n[j], nr = a.nrhs[j];
        const double *B = a.B + a.boff[j], *X = a.X + a.boff[j];
        double *R = a.R + a.boff[j];
        const bool want_norm = a.anrm[j] < 0.0;
        const int ngr = (n + 15) / 16;             /* 16-row groups; 16 * ngr <= np */
        double an = want_norm ? 0.0 : a.anrm[j];
        double rowmax = 0.0, worst = 0.0;
        bool ok = true;
        for (int cb = 0; cb < nr; cb += RES_CB) {
            const int ncb = min(RES_CB, nr - cb), nfr = (ncb + 15) / 16;
            const bool norm_pass = want_norm && cb == 0;
            for (int e = tid; e < RES_CB; e += 256) colr[e] = colx[e] = 0.0;
            __syncthreads();
            for (int u = wave; u < ngr * nfr; u += 4) {
                const int g = u % ngr, f = u / ngr;
                const int r0 = 16 * g, i = r0 + li;        /* A row this lane supplies */
                const int cl = 16 * f + li;                /* column of the pass, B operand */
                const bool cv = cl < ncb;
                const double *Xc = X + (size_t)(cb + (cv ? cl : 0)) * np;
                v4d acc = zero4d();
                double rs = 0.0;
                for (int p0 = 0; p0 < 16 * ngr; p0 += 16) {
                    double av[4];
                    if (p0 + 16 <= r0) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) av[q] = A[i + (size_t)(p0 + 4 * lk + q) * np];
                    } else if (p0 >= r0 + 16) {
                        const v4d t4 = *(const v4d *)(A + (p0 + 4 * lk) + (size_t)i * np);
#pragma unroll
                        for (int q = 0; q < 4; ++q) av[q] = t4[q];
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int k = p0 + 4 * lk + q;
                            av[q] = i >= k ? A[i + (size_t)k * np] : A[k + (size_t)i * np];
                        }
                    }
                    v4d xb = zero4d();
                    if (cv) xb = *(const v4d *)(Xc + p0 + 4 * lk);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int k = p0 + 4 * lk + q;
                        if (k >= n) av[q] = 0.0;
                        if (norm_pass && f == 0) rs += fabs(av[q]);
                        acc = mfma_f64(av[q], xb[q], acc);
                    }
                }
                if (norm_pass && f == 0) {
                    /* row i's |A| sum: the four k quarters meet */
                    rs += __shfl_xor(rs, 16);
                    rs += __shfl_xor(rs, 32);
                    if (lk == 0 && i < n) rowmax = nan_max(rowmax, rs);
                }
                /* slot s: row r0 + lk + 4 s, column cl */
                double rm = 0.0, xm = 0.0;
#pragma unroll
                for (int s2 = 0; s2 < 4; ++s2) {
                    const int row = r0 + lk + 4 * s2;
                    if (cv && row < n) {
                        const size_t o = row + (size_t)(cb + cl) * np;
                        const double r = B[o] - acc[s2];
                        R[o] = r;
                        rm = nan_max(rm, fabs(r));
                        xm = nan_max(xm, fabs(X[o]));
                    }
                }
                if (cv) {
                    lds_max_bits(&colr[cl], rm);
                    lds_max_bits(&colx[cl], xm);
                }
            }
            if (norm_pass) {
                __syncthreads();
                red[tid] = rowmax;
                __syncthreads();
                for (int w = 128; w > 0; w >>= 1) {
                    if (tid < w) red[tid] = nan_max(red[tid], red[tid + w]);
                    __syncthreads();
                }
                an = red[0];
                if (tid == 0) a.anrm[j] = an;
            }
            __syncthreads();
            for (int c = 0; c < ncb; ++c) {
                if (!(colr[c] <= colx[c] * an * eps * sqrt((double)n))) ok = false;
                const double d = an * colx[c] * n * eps;
                worst = nan_max(worst, d > 0.0 ? colr[c] / d : (colr[c] > 0.0 ? INFINITY : 0.0));
            }
            __syncthreads();
        }
        if (tid == 0) {
            a.berr[j] = worst;
            if (ok) {
                a.state[j] = 1;
            } else {
                a.iters[j] += 1;
                atomicAdd(a.unconv, 1);
            }
        }
    }
}

/* dst = src on the part of every system the factorisation reads: in each
 * 64-column block, rows from the block's first column down (the lower
 * triangle and the diagonal 64 x 64 blocks whole).  Workgroup per system,
 * consecutive threads on consecutive rows. */
__global__ __launch_bounds__(256) void batch_copy_lower_kernel(const double *src, double *dst, const long long *aoff,
                                                               const int *np, int count)
{
    for (int j = blockIdx.x; j < count; j += gridDim.x) {
        const int n = np[j];
        const double *S = src + aoff[j];
        double *D = dst + aoff[j];
        for (int c0 = 0; c0 < n; c0 += 64) {
            const int h = n - c0, w = min(64, h);
            const int tot = h * w;
#pragma unroll 4
            for (int e = threadIdx.x; e < tot; e += 256) {
                const int c = c0 + e / h, r = c0 + e % h;
                D[r + (size_t)c * n] = S[r + (size_t)c * n];
            }
        }
    }
}

void batch_copy_lower(const double *src, double *dst, const long long *aoff, const int *np, int count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(batch_copy_lower_kernel, dim3(std::min(count, 8192)), dim3(256), 0, s, src, dst, aoff, np,
                       count);
}

void batch_residual(const double *A, const long long *aoff, const int *np, const int *n, const double *B,
                    const double *X, double *R, const long long *boff, const int *nrhs, double *anrm, int *state,
                    int *iters, int *unconv, double *berr, int count, int *queue, hipStream_t s)
{
    if (count <= 0) return;
    BatchResArgs a;
    a.A = A; a.aoff = aoff; a.np = np; a.n = n; a.B = B; a.X = X; a.R = R; a.boff = boff; a.nrhs = nrhs;
    a.anrm = anrm; a.state = state; a.iters = iters; a.unconv = unconv; a.berr = berr; a.count = count;
    a.queue = queue;
    int grid = std::min(count, 2048);
    hipLaunchKernelGGL(batch_residual_kernel, dim3(grid), dim3(256), 0, s, a);
}

/* X += R for every system with state[j] == 0 (`add`), or X = B for every
 * system with state[j] != 1 (copy: the systems the FP64 fallback takes). */
__global__ void batch_update_kernel(double *X, const double *R, const long long *boff, const int *np,
                                    const int *nrhs, const int *state, int count, int add)
{
    const int j = blockIdx.x;
    if (j >= count) return;
    if (add ? state[j] != 0 : state[j] == 1) return;
    const size_t len = (size_t)np[j] * nrhs[j];
    double *x = X + boff[j];
    const double *r = R + boff[j];
    for (size_t e = threadIdx.x; e < len; e += blockDim.x) x[e] = add ? x[e] + r[e] : r[e];
}

void batch_update(double *X, const double *R, const long long *boff, const int *np, const int *nrhs,
                  const int *state, int count, int add, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(batch_update_kernel, dim3(count), dim3(256), 0, s, X, R, boff, np, nrhs, state, count, add);
}

/* state[j] = (info[j] != 0) ? 2 : 0, iters[j] = 0, anrm[j] = -1 */
__global__ void batch_refine_init_kernel(const int *info, int *state, int *iters, double *anrm, int count)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    state[j] = info[j] != 0 ? 2 : 0;
    iters[j] = 0;
    anrm[j] = -1.0;
}

void batch_refine_init(const int *info, int *state, int *iters, double *anrm, int count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(batch_refine_init_kernel, dim3((count + 255) / 256), dim3(256), 0, s, info, state, iters, anrm,
                       count);
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
