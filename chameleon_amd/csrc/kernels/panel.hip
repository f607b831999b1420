/*
 * panel.hip -- the critical-path kernels of the tiled Cholesky on gfx950:
 * POTRF of a diagonal tile (the tile kernel behind INSERT_TASK_dpotrf,
 * coreblas/compute/core_zpotrf.c:75-81, called at compute/pzpotrf.c:90-94)
 * and the panel TRSM (INSERT_TASK_dtrsm(Right, Lower, Trans, NonUnit),
 * pzpotrf.c:99-105), plus the 64 x 64 diagonal-block inverses they share.
 *
 * Both work on 64-column blocks.  POTRF factors a 64-column block of the
 * diagonal tile with a 16-column inner loop (16 x 16 blocks factored in
 * registers by one wavefront), then builds the explicit inverse of the
 * 64 x 64 diagonal block from the 16 x 16 inverses; TRSM then needs one
 * 64-column MFMA product per block instead of a 16-column substitution
 * chain.  The inverses are stored as 64 x 64 column-major blocks, ldt/64
 * (rounded up) per diagonal tile.
 */
#include "../internal.h"
#include "mfma.h"

namespace chx {
namespace k {

#define OB 64       /* outer block width */
#define OBP 65      /* padded LDS row */

/* ------------------------------------------------------------------ */
/* acc[f] += X[0:16, 0:K] * Y[16f:16f+16, 0:K]^T for f < NF, one wavefront;
 * X, Y column-major with leading dimension ld, K a multiple of 16, loads
 * pipelined one 16-deep slice ahead.  Formed as D = Y X^T: lane l holds
 * (row l%16, col 16f + l/16 + 4s) in slot s of acc[f].                    */
template <int NF>
__device__ __forceinline__ void wave_gemm_nt(const double *X, const double *Y, int ld, int K, int lane, v4d *acc)
{
    const int li = lane & 15, lk = lane >> 4;
    if (K <= 0) return;
    const double *px = X + li + (size_t)lk * ld;
    const double *py = Y + li + (size_t)lk * ld;
    double x[4], y[NF][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        x[q] = px[(size_t)(4 * q) * ld];
#pragma unroll
        for (int f = 0; f < NF; ++f) y[f][q] = py[16 * f + (size_t)(4 * q) * ld];
    }
    for (int p = 0; p < K; p += 16) {
        double xn[4], yn[NF][4];
        const bool more = p + 16 < K;
        if (more) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                xn[q] = px[(size_t)(p + 16 + 4 * q) * ld];
#pragma unroll
                for (int f = 0; f < NF; ++f) yn[f][q] = py[16 * f + (size_t)(p + 16 + 4 * q) * ld];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int f = 0; f < NF; ++f) acc[f] = mfma_f64(y[f][q], x[q], acc[f]);
        if (more) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                x[q] = xn[q];
#pragma unroll
                for (int f = 0; f < NF; ++f) y[f][q] = yn[f][q];
            }
        }
    }
}

__device__ __forceinline__ void wave_gemm_nt_dyn(int nf, const double *X, const double *Y, int ld, int K, int lane,
                                                 v4d *acc)
{
    switch (nf) {
    case 1: wave_gemm_nt<1>(X, Y, ld, K, lane, acc); break;
    case 2: wave_gemm_nt<2>(X, Y, ld, K, lane, acc); break;
    case 3: wave_gemm_nt<3>(X, Y, ld, K, lane, acc); break;
    default: wave_gemm_nt<4>(X, Y, ld, K, lane, acc); break;
    }
}

/* 1/sqrt(d) to full double precision: hardware estimate refined by two
 * Newton-Raphson steps (quadratic convergence from ~2^-22).  NaN and
 * non-positive d propagate as NaN or are caught by the caller's pivot test. */
__device__ __forceinline__ double rsqrt_f64(double d)
{
    double r = __builtin_amdgcn_rsq(d);
    r = r * (1.5 - 0.5 * d * r * r);
    r = r * (1.5 - 0.5 * d * r * r);
    return r;
}

/* Cholesky of a 16 x 16 block by one wavefront: lane i (of each 16-lane
 * group; the four groups compute the same thing) holds row i in a[];
 * pivots and column entries are read from lanes 0-15 with readlane (a
 * scalar broadcast, no LDS round trip).  On return a[] holds row i of L
 * and rd = 1 / L(i,i).  `fail` is the 1-based index of the first
 * non-positive pivot (uniform), 0 if none. */
__device__ __forceinline__ void block16_potrf(double (&a)[16], int lane, int &fail, double &rd)
{
    const int i = lane & 15;
    fail = 0;
    rd = 0.0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const double d = readlane_f64(a[j], j);
        if (fail == 0 && !(d > 0.0)) fail = j + 1;
        const double r = rsqrt_f64(d);
        const double lij = a[j] * r;                      /* l_ij for i > j, sqrt(d) for i == j */
        if (i == j) rd = r;
        if (i >= j) a[j] = lij;
#pragma unroll
        for (int k = j + 1; k < 16; ++k) {
            const double lkj = readlane_f64(lij, k);
            if (i > j && k <= i) a[k] -= lij * lkj;
        }
    }
}

/* One 16-column step of a 64-row column panel by one wavefront: lanes
 * 0-15 hold the rows of the 16 x 16 diagonal block, lanes 16-63 the rows
 * below it, a[] their 16 columns.  The right-looking Cholesky steps that
 * factor the diagonal block (pivots and column entries read from lanes
 * 0-15) turn the rows below into X = P L16^-T at the same time.  Entries
 * above the diagonal of lanes 0-15 end up meaningless (never read).  rd =
 * 1 / L(l,l) on lanes l < 16; `fail` as in block16_potrf. */
__device__ __forceinline__ void block16_panel(double (&a)[16], int lane, int &fail, double &rd)
{
    fail = 0;
    rd = 0.0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const double d = readlane_f64(a[j], j);
        if (fail == 0 && !(d > 0.0)) fail = j + 1;
        const double r = rsqrt_f64(d);
        const double lij = a[j] * r;
        if (lane == j) rd = r;
        a[j] = lij;
#pragma unroll
        for (int k = j + 1; k < 16; ++k) a[k] -= lij * readlane_f64(lij, k);
    }
}

/* Column (lane % 16) of the inverse of the lower triangular 16 x 16 block
 * L(r,p) = Ls[r * lds + p], by forward substitution, with the reciprocal
 * diagonal rdiag[r] = 1 / L(r,r). */
__device__ __forceinline__ void block16_inverse_col(const double *Ls, int lds, const double *rdiag, int lane,
                                                    double (&col)[16])
{
    const int jc = lane & 15;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        double s = (r == jc) ? 1.0 : 0.0;
#pragma unroll
        for (int p = 0; p < r; ++p) s -= Ls[r * lds + p] * col[p];
        col[r] = (r < jc) ? 0.0 : s * rdiag[r];
    }
}

/* Inverse Lv of the lower triangular W x W block Ld (W = 16 nfb), given
 * the inverses Lb[I] of its 16 x 16 diagonal blocks, by block rows:
 *   Lv(I,I) = Lb(I),  Lv(I,J) = -Lb(I) sum_{J<=P<I} Ld(I,P) Lv(P,J).
 * Called by the whole workgroup (barriers inside); wavefront J computes
 * block column J. */
__device__ void assemble_inverse(const double (*Ld)[OBP], const double (*Lb)[16][17], double (*Lv)[OBP], int nfb,
                                 int wave, int lane)
{
    const int li = lane & 15, lk = lane >> 4;
    /* diagonal blocks and zero upper blocks */
    for (int e = threadIdx.x; e < OB * OB; e += blockDim.x) {
        const int r = e / OB, c = e % OB;
        const int I = r >> 4, J = c >> 4;
        Lv[r][c] = (I == J && I < nfb) ? Lb[I][r & 15][c & 15] : 0.0;
    }
    __syncthreads();
    for (int I = 1; I < nfb; ++I) {
        const int J = wave;
        if (J < I) {
            v4d S = zero4d();
            for (int P = J; P < I; ++P) {
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
                    S = mfma_f64(Ld[16 * I + li][16 * P + 4 * kk + lk], Lv[16 * P + 4 * kk + lk][16 * J + li], S);
            }
            /* slot kk of S holds S(4kk + lk, li): the B-operand layout */
            v4d R = zero4d();
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) R = mfma_f64(Lb[I][li][4 * kk + lk], S[kk], R);
#pragma unroll
            for (int s = 0; s < 4; ++s) Lv[16 * I + lk + 4 * s][16 * J + li] = -R[s];
        }
        __syncthreads();
    }
}

/* ================================================================== */
/* POTRF of one matrix held in HBM per job, one workgroup of 8 wavefronts
 * per job: the diagonal tile of the tiled factorisation (one job), or
 * every system of a batch (jobs taken from a queue).
 * For each 64-column block C (left-looking):
 *   (a) rows [C, np) of the block -= T[., 0:C] T[C:C+64, 0:C]^T   (MFMA)
 *   (b) factor the 64 x 64 diagonal block in LDS by 16-column steps
 *   (c) build its inverse, store it in dinv
 *   (d) rows below: X = P Linv^T                                  (MFMA)
 * The strictly upper triangle of the matrix is never written.  A failed
 * job reports the 1-based column of its first non-positive pivot, only
 * for pivots inside its valid part (nvalid columns).                    */

struct PotrfJobs {
    double *A;               /* job j: matrix at A + off[j] (off null: A)               */
    const long long *off;
    const int *np;           /* order (multiple of 16); null: np0                        */
    const int *ld;           /* leading dimension; null: ld0, or np when ld0 = 0         */
    const int *nvalid;       /* valid columns; null: nvalid0                             */
    double *dinv;            /* 64 x 64 inverses: job j at dinv + doff[j] (null: dinv)   */
    const long long *doff;
    int np0, ld0, nvalid0;
    int count;               /* number of jobs                                           */
    int *queue;              /* job counter (batches); null: job = blockIdx.x            */
    int *info;               /* single job: *info = iinfo + column on failure;
                              * batches: info[j] = column or 0                           */
    int iinfo, batch;
};

/* Phase (a) of potrf_jobs_kernel (below) for a full 64-column block C: the
 * same products in the same order as its general loop, without the
 * run-time tests of that loop (block width, last slice, rows past the
 * matrix) around the slice loads; the last slice is peeled off.  A wave
 * whose rows lie past the matrix reads row li instead and discards it. */
template <int POTRF_WAVES, bool LP>
__device__ __forceinline__ void potrf_update_full(double *T, int ldt, int C, int r0, bool active,
                                                  double (*Ys)[16 * OB], int tid, int li, int lk)
{
    constexpr int POTRF_THREADS = POTRF_WAVES * 64, YH = 512 / POTRF_THREADS, YS = POTRF_THREADS / 32;
    v4d acc[4] = {zero4d(), zero4d(), zero4d(), zero4d()};
    v4f accf[4] = {zero4f(), zero4f(), zero4f(), zero4f()};
    const double *px = T + (active ? r0 + li : li) + (size_t)lk * ldt;
    const int yq = tid >> 5, yr = (tid & 31) * 2;
    const double *py = T + (C + yr) + (size_t)yq * ldt;
    auto products = [&](int cur, const double (&x)[4]) {
        if (!active) return;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                if constexpr (LP)
                    accf[f] = mfma_f32((float)Ys[cur][(4 * q + lk) * OB + 16 * f + li], (float)x[q], accf[f]);
                else
                    acc[f] = mfma_f64(Ys[cur][(4 * q + lk) * OB + 16 * f + li], x[q], acc[f]);
            }
    };
    v2d yreg[YH];
    double x[4];
#pragma unroll
    for (int h = 0; h < YH; ++h) yreg[h] = *reinterpret_cast<const v2d *>(py + (size_t)(h * YS) * ldt);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double v = px[(size_t)(4 * q) * ldt];
        x[q] = active ? v : 0.0;
    }
#pragma unroll
    for (int h = 0; h < YH; ++h) *reinterpret_cast<v2d *>(&Ys[0][(yq + h * YS) * OB + yr]) = yreg[h];
    __syncthreads();
    int p = 0;
    for (; p + 16 < C; p += 16) {
        const int cur = (p >> 4) & 1;
        double xn[4];
#pragma unroll
        for (int h = 0; h < YH; ++h)
            yreg[h] = *reinterpret_cast<const v2d *>(py + (size_t)(p + 16 + h * YS) * ldt);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double v = px[(size_t)(p + 16 + 4 * q) * ldt];
            xn[q] = active ? v : 0.0;
        }
        products(cur, x);
#pragma unroll
        for (int h = 0; h < YH; ++h) *reinterpret_cast<v2d *>(&Ys[cur ^ 1][(yq + h * YS) * OB + yr]) = yreg[h];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = xn[q];
        __syncthreads();
    }
    products((p >> 4) & 1, x);
    __syncthreads();
    if (active) {
        /* the 16 old values first, all in flight (a load-subtract-store per
         * element waited for each load in turn); the block lies inside the
         * matrix, so the upper-triangle ones are read and left as they are */
        double old[4][4];
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int c = C + 16 * f + (LP ? 4 * lk + s : lk + 4 * s);
                old[f][s] = T[(r0 + li) + (size_t)c * ldt];
            }
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int r = r0 + li;
                const int c = C + 16 * f + (LP ? 4 * lk + s : lk + 4 * s);
                if (r >= c) T[r + (size_t)c * ldt] = old[f][s] - (LP ? (double)accf[f][s] : acc[f][s]);
            }
    }
}

template <int POTRF_WAVES, bool LP = false>
__global__ __launch_bounds__(POTRF_WAVES * 64, 8 / POTRF_WAVES) void potrf_jobs_kernel(PotrfJobs J)
{
    constexpr int POTRF_THREADS = POTRF_WAVES * 64;
    constexpr int YH = 512 / POTRF_THREADS;     /* Y-slice columns staged per thread */
    __shared__ double Ld[OB][OBP];
    /* Lv (the 64 x 64 inverse, phases (c)-(d)) and Ys (the staged Y slices
     * of phase (a), 2 x 64 rows x 16 k) are never live at the same time */
    __shared__ double LvYs[OB * OBP];
    __shared__ double Lb[4][16][17];
    __shared__ double Rd[16];
    __shared__ int s_fail, s_job;
    double (*Lv)[OBP] = reinterpret_cast<double (*)[OBP]>(LvYs);
    double (*Ys)[16 * OB] = reinterpret_cast<double (*)[16 * OB]>(LvYs);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    if (!J.batch && *J.info != 0) return;
    /* the single-job form is on the critical path of the factorisation:
     * its waves win issue arbitration against co-resident update waves */
    __builtin_amdgcn_s_setprio(2);
  for (int pass = 0;; ++pass) {
    if (tid == 0) s_job = J.queue ? atomicAdd(J.queue, 1) : (pass == 0 ? (int)blockIdx.x : J.count);
    __syncthreads();
    const int job = s_job;
    if (job >= J.count) break;
    double *T = J.A + (J.off ? J.off[job] : 0);
    double *dinv = J.dinv + (J.doff ? J.doff[job] : 0);
    const int np = J.np ? J.np[job] : J.np0;
    const int ldt = J.ld ? J.ld[job] : (J.ld0 ? J.ld0 : np);
    const int nvalid = J.nvalid ? J.nvalid[job] : J.nvalid0;
    if (tid == 0) s_fail = 0;
    __syncthreads();

    for (int C = 0; C < np; C += OB) {
        const int W = min(OB, np - C), nfb = W >> 4;
        /* (a) rows [C, np) x cols [C, C+W) -= T[., 0:C] T[C:C+W, 0:C]^T.
         * The 16-deep slices of the W shared rows (Y) are staged in LDS,
         * double-buffered, so one CU's L2 bandwidth carries each slice
         * once; every wavefront streams its own 16 rows (X). */
        if (C > 0) {
            const int nrb = (np - C) >> 4;
            for (int rb0 = 0; rb0 < nrb; rb0 += POTRF_WAVES) {
                const int rb = rb0 + wave;
                const bool active = rb < nrb;
                const int r0 = C + rb * 16;
                if (W == OB) {
                    potrf_update_full<POTRF_WAVES, LP>(T, ldt, C, r0, active, Ys, tid, li, lk);
                    continue;
                }
                v4d acc[4] = {zero4d(), zero4d(), zero4d(), zero4d()};
                /* LP: the same products with FP32 operands and accumulation
                 * (v_mfma_f32_16x16x4_f32, result slot s at column 4 lk + s) */
                v4f accf[4] = {zero4f(), zero4f(), zero4f(), zero4f()};
                const double *px = T + (r0 + li) + (size_t)lk * ldt;
                /* Y slice element (row r < 64, k q < 16) -> Ys[q * 64 + r];
                 * thread tid moves rows 2*(tid%32)..+1 of columns
                 * tid/32 + h * (threads/32), h < YH */
                const int yq = tid >> 5, yr = (tid & 31) * 2;
                const bool yv = yr < W;
                v2d yreg[YH];
#pragma unroll
                for (int h = 0; h < YH; ++h) {
                    yreg[h] = v2d{0.0, 0.0};
                    if (yv) yreg[h] = *reinterpret_cast<const v2d *>(T + (C + yr) + (size_t)(yq + h * (POTRF_THREADS / 32)) * ldt);
                }
                double x[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) x[q] = active ? px[(size_t)(4 * q) * ldt] : 0.0;
#pragma unroll
                for (int h = 0; h < YH; ++h)
                    *reinterpret_cast<v2d *>(&Ys[0][(yq + h * (POTRF_THREADS / 32)) * OB + yr]) = yreg[h];
                __syncthreads();
                for (int p = 0; p < C; p += 16) {
                    const int cur = (p >> 4) & 1;
                    const bool more = p + 16 < C;
                    double xn[4];
                    if (more) {
#pragma unroll
                        for (int h = 0; h < YH; ++h)
                            if (yv)
                                yreg[h] = *reinterpret_cast<const v2d *>(
                                    T + (C + yr) + (size_t)(p + 16 + yq + h * (POTRF_THREADS / 32)) * ldt);
#pragma unroll
                        for (int q = 0; q < 4; ++q) xn[q] = active ? px[(size_t)(p + 16 + 4 * q) * ldt] : 0.0;
                    }
                    if (active) {
#pragma unroll
                        for (int q = 0; q < 4; ++q)
#pragma unroll
                            for (int f = 0; f < 4; ++f)
                                if (f < nfb) {
                                    if constexpr (LP)
                                        accf[f] = mfma_f32((float)Ys[cur][(4 * q + lk) * OB + 16 * f + li], (float)x[q],
                                                           accf[f]);
                                    else
                                        acc[f] = mfma_f64(Ys[cur][(4 * q + lk) * OB + 16 * f + li], x[q], acc[f]);
                                }
                    }
                    if (more) {
#pragma unroll
                        for (int h = 0; h < YH; ++h)
                            *reinterpret_cast<v2d *>(&Ys[cur ^ 1][(yq + h * (POTRF_THREADS / 32)) * OB + yr]) = yreg[h];
#pragma unroll
                        for (int q = 0; q < 4; ++q) x[q] = xn[q];
                    }
                    __syncthreads();
                }
                if (active) {
                    for (int f = 0; f < nfb; ++f)
#pragma unroll
                        for (int s = 0; s < 4; ++s) {
                            const int r = r0 + li;
                            const int c = C + 16 * f + (LP ? 4 * lk + s : lk + 4 * s);
                            if (r >= c) T[r + (size_t)c * ldt] -= LP ? (double)accf[f][s] : acc[f][s];
                        }
                }
            }
            __syncthreads();
        }
        /* (b) diagonal block into LDS (lower part; zeros above) */
        for (int e = tid; e < OB * OB; e += POTRF_THREADS) {
            const int r = e % OB, c = e / OB;
            Ld[r][c] = (r < W && c < W && r >= c) ? T[(C + r) + (size_t)(C + c) * ldt] : 0.0;
        }
        __syncthreads();
        for (int cb = 0; cb < nfb; ++cb) {
            const int c0 = 16 * cb;
            /* (b1) columns [c0, c0+16) of rows [c0, W) -= Ld[., 0:c0] Ld[c0.., 0:c0]^T */
            if (cb > 0) {
                const int rb = wave;
                if (rb < nfb - cb) {
                    const int r0 = c0 + 16 * rb;
                    v4d acc = zero4d();
                    for (int p = 0; p < c0; p += 4)
                        acc = mfma_f64(Ld[c0 + li][p + lk], Ld[r0 + li][p + lk], acc);
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const int c = lk + 4 * s;
                        if (rb > 0 || li >= c) Ld[r0 + li][c0 + c] -= acc[s];
                    }
                }
                __syncthreads();
            }
            /* (b2) factor and invert the 16 x 16 block */
            if (wave == 0) {
                double a[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) a[q] = (q <= li) ? Ld[c0 + li][c0 + q] : 0.0;
                int fail;
                double rd;
                block16_potrf(a, lane, fail, rd);
                if (lane < 16) {
#pragma unroll
                    for (int q = 0; q < 16; ++q) Ld[c0 + li][c0 + q] = (q <= li) ? a[q] : 0.0;
                    Rd[li] = rd;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                double col[16];
                block16_inverse_col(&Ld[c0][c0], OBP, Rd, lane, col);
                if (lane < 16) {
#pragma unroll
                    for (int q = 0; q < 16; ++q) Lb[cb][q][li] = col[q];
                }
                if (lane == 0 && fail) s_fail = (C + c0 + fail - 1 < nvalid) ? C + c0 + fail : -1;
            }
            __syncthreads();
            if (s_fail) break;
            /* (b3) rows [c0+16, W): X = P Lb^T */
            {
                const int rb = wave;
                if (rb < nfb - cb - 1) {
                    const int r0 = c0 + 16 + 16 * rb;
                    double pv[4];
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) pv[kk] = Ld[r0 + li][c0 + 4 * kk + lk];
                    v4d acc = zero4d();
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) acc = mfma_f64(Lb[cb][li][4 * kk + lk], pv[kk], acc);
#pragma unroll
                    for (int s = 0; s < 4; ++s) Ld[r0 + li][c0 + lk + 4 * s] = acc[s];
                }
            }
            __syncthreads();
        }
        if (s_fail) break;
        /* (c) inverse of the 64 x 64 diagonal block */
        assemble_inverse(Ld, Lb, Lv, nfb, wave, lane);
        /* factor back into the tile, inverse into dinv (column-major 64 x 64) */
        double *dv = dinv + (size_t)(C / OB) * OB * OB;
        for (int e = tid; e < OB * OB; e += POTRF_THREADS) {
            const int r = e % OB, c = e / OB;
            if (r < W && c < W && r >= c) T[(C + r) + (size_t)(C + c) * ldt] = Ld[r][c];
            dv[e] = Lv[r][c];
        }
        /* (d) rows [C+W, np): X = P Lv^T, D = Lv P^T */
        {
            const int nrb = (np - C - W) >> 4;
            for (int rb = wave; rb < nrb; rb += POTRF_WAVES) {
                const int r0 = C + W + rb * 16;
                double pv[16];
                if (W == OB) {
                    /* full block: the 16 loads with no run-time test between
                     * them (in flight together), the same products */
#pragma unroll
                    for (int kq = 0; kq < 16; ++kq) pv[kq] = T[(r0 + li) + (size_t)(C + 4 * kq + lk) * ldt];
                    v4d acc[4] = {zero4d(), zero4d(), zero4d(), zero4d()};
#pragma unroll
                    for (int f = 0; f < 4; ++f)
#pragma unroll
                        for (int kq = 0; kq < 4 * (f + 1); ++kq)
                            acc[f] = mfma_f64(Lv[16 * f + li][4 * kq + lk], pv[kq], acc[f]);
#pragma unroll
                    for (int f = 0; f < 4; ++f)
#pragma unroll
                        for (int s = 0; s < 4; ++s) T[(r0 + li) + (size_t)(C + 16 * f + lk + 4 * s) * ldt] = acc[f][s];
                    continue;
                }
#pragma unroll
                for (int kq = 0; kq < 16; ++kq)
                    pv[kq] = (4 * kq < W) ? T[(r0 + li) + (size_t)(C + 4 * kq + lk) * ldt] : 0.0;
                v4d acc[4] = {zero4d(), zero4d(), zero4d(), zero4d()};
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    if (f < nfb) {
#pragma unroll
                        for (int kq = 0; kq < 16; ++kq)
                            if (kq < 4 * (f + 1)) acc[f] = mfma_f64(Lv[16 * f + li][4 * kq + lk], pv[kq], acc[f]);
                    }
                }
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    if (f < nfb) {
#pragma unroll
                        for (int s = 0; s < 4; ++s) T[(r0 + li) + (size_t)(C + 16 * f + lk + 4 * s) * ldt] = acc[f][s];
                    }
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (J.batch) J.info[job] = s_fail > 0 ? s_fail : 0;
        else if (s_fail > 0) *J.info = J.iinfo + s_fail;
    }
    __syncthreads();
  }
}

void potrf_diag(double *Akk, int ldt, double *dinv, int iinfo, int nvalid, int *info, int waves, hipStream_t s)
{
    PotrfJobs J = {};
    J.A = Akk; J.dinv = dinv;
    J.np0 = ldt; J.ld0 = ldt; J.nvalid0 = nvalid;
    J.count = 1; J.info = info; J.iinfo = iinfo; J.batch = 0;
    if (waves == 8) hipLaunchKernelGGL(potrf_jobs_kernel<8>, dim3(1), dim3(512), 0, s, J);
    else hipLaunchKernelGGL(potrf_jobs_kernel<4>, dim3(1), dim3(256), 0, s, J);
}

/* Batched form: `count` matrices, job j of order np[j] (multiple of 16)
 * at A + off[j] with leading dimension ld[j] (null: np[j]), inverses at
 * dinv + doff[j]; info[j] = first failing column or 0.  `queue` is a
 * zeroed device counter; jobs are taken in index order by `grid`
 * persistent workgroups (grid <= 0: as many as are resident at once; sort
 * the jobs largest first for balance). */
template <int W, bool LP>
static int potrf_jobs_resident()
{
    static int resident = 0;
    if (!resident) {
        int dev = 0, ncu = 0, per = 0;
        CHX_HIP(hipGetDevice(&dev));
        CHX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        CHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, potrf_jobs_kernel<W, LP>, W * 64, 0));
        resident = ncu * (per > 0 ? per : 1);
    }
    return resident;
}

template <int W, bool LP>
static void potrf_jobs_launch(const PotrfJobs &J, int grid, hipStream_t s)
{
    if (grid <= 0) grid = potrf_jobs_resident<W, LP>();
    if (grid > J.count) grid = J.count;
    hipLaunchKernelGGL((potrf_jobs_kernel<W, LP>), dim3(grid), dim3(W * 64), 0, s, J);
}

void potrf_jobs(double *A, const long long *off, const int *np, const int *ld, const int *nvalid, double *dinv,
                const long long *doff, int count, int *queue, int *info, int grid, hipStream_t s, int waves, bool lp)
{
    PotrfJobs J = {};
    J.A = A; J.off = off; J.np = np; J.ld = ld; J.nvalid = nvalid; J.dinv = dinv; J.doff = doff;
    J.count = count; J.queue = queue; J.info = info; J.batch = 1;
    if (count <= 0) return;
    if (waves == 8) {
        if (lp) potrf_jobs_launch<8, true>(J, grid, s);
        else potrf_jobs_launch<8, false>(J, grid, s);
    } else {
        if (lp) potrf_jobs_launch<4, true>(J, grid, s);
        else potrf_jobs_launch<4, false>(J, grid, s);
    }
}

void potrf_jobs_prepare()
{
    (void)potrf_jobs_resident<4, false>();
    (void)potrf_jobs_resident<8, false>();
    (void)potrf_jobs_resident<4, true>();
    (void)potrf_jobs_resident<8, true>();
}

/* ================================================================== */
/* 64 x 64 diagonal-block inverses of every diagonal tile of a factor that
 * was not produced by potrf_diag in this process.  One workgroup (4
 * wavefronts) per (tile, 64-column block). */
__global__ __launch_bounds__(256) void dinv64_kernel(const double *A, int nt, long long diag_stride, int ldt,
                                                      double *dinv)
{
    __shared__ double Ld[OB][OBP];
    __shared__ double Lv[OB][OBP];
    __shared__ double Lb[4][16][17];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15;
    const int nob = (ldt + OB - 1) / OB;
    const int t = blockIdx.x / nob, C = (blockIdx.x % nob) * OB;
    if (t >= nt) return;
    const int W = min(OB, ldt - C), nfb = W >> 4;
    const double *T = A + (size_t)t * diag_stride;
    for (int e = tid; e < OB * OB; e += 256) {
        const int r = e % OB, c = e / OB;
        Ld[r][c] = (r < W && c < W && r >= c) ? T[(C + r) + (size_t)(C + c) * ldt] : 0.0;
    }
    __syncthreads();
    if (wave < nfb) {
        double col[16], rdiag[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) rdiag[r] = 1.0 / Ld[16 * wave + r][16 * wave + r];
        block16_inverse_col(&Ld[16 * wave][16 * wave], OBP, rdiag, lane, col);
        if (lane < 16) {
#pragma unroll
            for (int q = 0; q < 16; ++q) Lb[wave][q][li] = col[q];
        }
    }
    __syncthreads();
    assemble_inverse(Ld, Lb, Lv, nfb, wave, lane);
    double *dv = dinv + ((size_t)t * nob + C / OB) * OB * OB;
    for (int e = tid; e < OB * OB; e += 256) dv[e] = Lv[e % OB][e / OB];
}

void dinv_all(const double *A, int nt, size_t diag_stride, int ldt, double *dinv, hipStream_t s)
{
    const int nob = (ldt + OB - 1) / OB;
    hipLaunchKernelGGL(dinv64_kernel, dim3(nt * nob), dim3(256), 0, s, A, nt, (long long)diag_stride, ldt, dinv);
}

/* ================================================================== */
/* Right-side triangular solve X = P L^-T, in place, for a batch of tiles.
 * One wavefront owns a 16-row strip of one tile and walks its 64-column
 * blocks left to right:
 *   R = P[:, C:C+64] - X[:, 0:C] L[C:C+64, 0:C]^T    (MFMA, K = C)
 *   X[:, C:C+64] = R Linv_C^T                         (MFMA, K = 64)
 * Tile t uses L0 + t*lstride and the block inverses dinv0 + t*dstride, so
 * applied to identity tiles the kernel forms Linv^T for every diagonal
 * tile at once.                                                          */
__global__ __launch_bounds__(256) void trsm_right_kernel(double *P0, int ntiles, long long pstride, int ldt,
                                                          const double *L0, long long lstride,
                                                          const double *dinv0, long long dstride,
                                                          const int *info)
{
    __shared__ double S[4][16][OBP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    if (*info != 0) return;
    __builtin_amdgcn_s_setprio(2);   /* critical path, see potrf_diag_kernel */
    const int nrb = ldt >> 4;
    const int wid = blockIdx.x * 4 + wave;
    const int t = wid / nrb, rb = wid % nrb;
    if (t >= ntiles) return;
    double *X = P0 + (size_t)t * pstride;
    const double *L = L0 + (size_t)t * lstride;
    const double *dinv = dinv0 + (size_t)t * dstride;
    const int r0 = rb * 16;
    for (int C = 0; C < ldt; C += OB) {
        const int W = min(OB, ldt - C), nfb = W >> 4;
        v4d acc[4] = {zero4d(), zero4d(), zero4d(), zero4d()};
        wave_gemm_nt_dyn(nfb, X + r0, L + C, ldt, C, lane, acc);
        for (int f = 0; f < nfb; ++f)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int c = 16 * f + lk + 4 * s;
                S[wave][li][c] = X[(r0 + li) + (size_t)(C + c) * ldt] - acc[f][s];
            }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const double *Lv = dinv + (size_t)(C / OB) * OB * OB;   /* column-major 64 x 64 */
        v4d y[4] = {zero4d(), zero4d(), zero4d(), zero4d()};
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            if (f < nfb) {
#pragma unroll
                for (int kq = 0; kq < 16; ++kq)
                    if (kq < 4 * (f + 1))
                        y[f] = mfma_f64(Lv[(16 * f + li) + (size_t)(4 * kq + lk) * OB], S[wave][li][4 * kq + lk], y[f]);
            }
        }
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            if (f < nfb) {
#pragma unroll
                for (int s = 0; s < 4; ++s) X[(r0 + li) + (size_t)(C + 16 * f + lk + 4 * s) * ldt] = y[f][s];
            }
        }
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
/* Fused panel factorisation of step k: POTRF of the diagonal tile
 * (pzpotrf.c:90-94) and the TRSM of the S tiles below it (pzpotrf.c:99-105)
 * in one launch, spread over 64-row units so that the panel is not bounded
 * by one compute unit.
 *
 * Units (one workgroup of 4 wavefronts each, numbered in the order the
 * workgroups start, through a ticket counter):
 *   r < ND      owner of row block r of the diagonal tile (ND = ldt / 64)
 *   r >= ND     one 64-row block of a tile below
 * Every unit walks the 64-column blocks C of the panel left-looking:
 *   R = A(u, C) - sum_{p<C} X(u, p) X(C, p)^T      X(C, .) published by owner C
 *   X(u, C) = R Linv_C^T                           Linv_C published by owner C
 * and owner r, after its blocks C < r, factors its diagonal block
 *   L(r, r) L(r, r)^T = A(r, r) - sum_{p<r} X(r, p) X(r, p)^T
 * and publishes L(r, r)'s inverse.  The SYRK part over p < r-1 is formed
 * before the wait for Linv_{r-1}, so the chain per column block is one
 * 64 x 64 product, one rank-64 update and the 64 x 64 factorisation.
 *
 * Publication: flag F[r * ND + C] set to this launch's epoch once owner r
 * stored X(r, C) (C < r) or L(r, r) and Linv_r (C = r), after a
 * device-scope release.  Published blocks are read with device-scope
 * loads (past the non-coherent L2 of other XCDs).  A unit only waits for
 * flags of owners, and owners only for owners with smaller tickets, which
 * have started: the waits end.  A wait also ends when the info flag is
 * set (a failed pivot) or after a bounded number of polls.             */

#define PF_SPIN_LIMIT (1 << 22)

struct PanelArgs {
    double *T;           /* diagonal tile                                         */
    double *P;           /* first tile below it; tile t at P + t * ts             */
    long long ts;
    int S, ldt, ND;
    double *dinv;        /* Linv_C at dinv + C * 4096, column-major 64 x 64        */
    int *flags;          /* ND x ND publication flags                              */
    unsigned *ctr;       /* [0] tickets taken, [1] workgroups finished (both zero
                          * between launches: the last workgroup clears them)      */
    int epoch;
    int *info;
    int iinfo, nvalid;
};

/* Buffer accesses relative to a tile: one per-lane offset register, the
 * uniform part of each address in a scalar offset.  aux 16 sets sc1, a
 * device-scope (coherent) access, the form of a relaxed device-scope
 * atomic load. */
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ __forceinline__ rsrc_t pf_rsrc(const double *p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(p), 0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ double pf_ld(rsrc_t r, int vo, int so)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}
__device__ __forceinline__ double pf_ld_dev(rsrc_t r, int vo, int so)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 16));
}
__device__ __forceinline__ void pf_st(double v, rsrc_t r, int vo, int so)
{
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, vo, so, 0);
}

/* Wait for flag idx of this launch; false if the factorisation failed
 * meanwhile (or the poll limit was reached).  Uniform over the workgroup. */
__device__ bool pf_wait(const PanelArgs &a, int idx, int tid, int *s_bad)
{
    if (tid == 0) {
        int bad = 0, n = 0;
        while (__hip_atomic_load(&a.flags[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != a.epoch) {
            if (__hip_atomic_load(a.info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { bad = 1; break; }
            if (++n > PF_SPIN_LIMIT) {
                atomicCAS(a.info, 0, PANEL_STALLED);
                bad = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        *s_bad = bad;
    }
    __syncthreads();
    const int bad = *s_bad;
    __syncthreads();
    return bad == 0;
}

__device__ __forceinline__ void pf_publish(const PanelArgs &a, int idx, int tid)
{
    __syncthreads();
    if (tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&a.flags[idx], a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

/* acc[f] += X[16w + i, 0:64n] Y[16f + j, 0:64n]^T (wave w, lane (i, j) of
 * the result layout).  X: the unit's own rows (plain loads); Y: 64 rows
 * read with device-scope loads and staged in LDS once per 64-column block;
 * the next block is loaded while the current one is multiplied.  */
__device__ void pf_kloop(const double *X, const double *Y, int ldt, int n, double *Ys, int tid, int wave, int lane,
                         v4d (&acc)[4])
{
    const int li = lane & 15, lk = lane >> 4;
    const rsrc_t rx = pf_rsrc(X), ry = pf_rsrc(Y);
    const int vx = (16 * wave + li + lk * ldt) * 8;            /* X(16w + li, lk)          */
    const int vy = ((tid & 63) + (tid >> 6) * ldt) * 8;        /* Y(tid % 64, tid / 64)    */
    const int c4 = 4 * ldt * 8;                                 /* four columns, in bytes   */
    double x[16], yst[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) x[q] = pf_ld(rx, vx, q * c4);
#pragma unroll
    for (int h = 0; h < 16; ++h) yst[h] = pf_ld_dev(ry, vy, h * c4);
    for (int p = 0; p < n; ++p) {
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 16; ++h) Ys[tid + 256 * h] = yst[h];
        __syncthreads();
        const bool more = p + 1 < n;
        const int o = (p + 1) * 16 * c4;
        if (more) {
#pragma unroll
            for (int h = 0; h < 16; ++h) yst[h] = pf_ld_dev(ry, vy, o + h * c4);
        }
        /* x[q] is consumed by its four products, then refilled with the
         * next block's value */
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const double xq = x[q];
            if (more) x[q] = pf_ld(rx, vx, o + q * c4);
#pragma unroll
            for (int f = 0; f < 4; ++f) acc[f] = mfma_f64(Ys[(4 * q + lk) * 64 + 16 * f + li], xq, acc[f]);
        }
    }
}

/* Factor the 64 x 64 block in Ld (lower part, zeros above) in place and
 * build its inverse in Lv, by 16-column steps: (1) columns [c0, c0+16)
 * of rows [c0, 64) take the earlier steps' products (MFMA), (2) one
 * wavefront factors that 16-column panel in registers (block16_panel:
 * the diagonal block and, in the same steps, the rows below).  The
 * 16 x 16 inverses are only needed for the 64 x 64 inverse, so the four
 * are formed after the loop, one per wavefront, off the column-to-column
 * chain.  Rd holds 1 / L(i,i) of all 64 columns.  Returns the 1-based
 * column of the first non-positive pivot, 0 if none (uniform). */
__device__ int pf_factor64(double (*Ld)[OBP], double (*Lb)[16][17], double (*Lv)[OBP], double *Rd, int *s_fail,
                           int wave, int lane)
{
    const int li = lane & 15, lk = lane >> 4;
    const int tid = threadIdx.x;
    if (tid == 0) *s_fail = 0;
    __syncthreads();
    for (int cb = 0; cb < 4; ++cb) {
        const int c0 = 16 * cb;
        if (cb > 0) {
            const int rb = wave;
            if (rb < 4 - cb) {
                const int r0 = c0 + 16 * rb;
                v4d acc = zero4d();
                for (int p = 0; p < c0; p += 4) acc = mfma_f64(Ld[c0 + li][p + lk], Ld[r0 + li][p + lk], acc);
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int c = lk + 4 * s;
                    if (rb > 0 || li >= c) Ld[r0 + li][c0 + c] -= acc[s];
                }
            }
            __syncthreads();
        }
        if (wave == 0) {
            /* lane (g, i): row c0 + 16 g + i of the panel, g = 0 the diagonal block */
            const int row = c0 + 16 * lk + li;
            const bool valid = row < OB;
            double a[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) a[q] = (valid && (lk > 0 || q <= li)) ? Ld[row][c0 + q] : 0.0;
            int fail;
            double rd;
            block16_panel(a, lane, fail, rd);
            if (lk == 0) {
#pragma unroll
                for (int q = 0; q < 16; ++q) Ld[row][c0 + q] = (q <= li) ? a[q] : 0.0;
                Rd[c0 + li] = rd;
            } else if (valid) {
#pragma unroll
                for (int q = 0; q < 16; ++q) Ld[row][c0 + q] = a[q];
            }
            if (lane == 0 && fail && *s_fail == 0) *s_fail = c0 + fail;
        }
        __syncthreads();
        if (*s_fail) return *s_fail;
    }
    /* the four 16 x 16 inverses, one per wavefront */
    {
        double col[16];
        block16_inverse_col(&Ld[16 * wave][16 * wave], OBP, &Rd[16 * wave], lane, col);
        if (lane < 16) {
#pragma unroll
            for (int q = 0; q < 16; ++q) Lb[wave][q][li] = col[q];
        }
    }
    __syncthreads();
    assemble_inverse(Ld, Lb, Lv, 4, wave, lane);
    return 0;
}

__device__ __forceinline__ void panel_unit(const PanelArgs &a, int t, double *Ys, double (*Ld)[OBP],
                                           double (*Lb)[16][17], double *Rd, int *s_int)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int ND = a.ND, ldt = a.ldt;
    const bool owner = t < ND;
    const int r = owner ? t : (t - ND) % ND;
    double *R = owner ? a.T + 64 * r : a.P + (size_t)((t - ND) / ND) * a.ts + 64 * r;
    const int nsteps = owner ? r : ND;
    const int rowo = 16 * wave + li;
    const rsrc_t rR = pf_rsrc(R);
    const int vr = (rowo + lk * ldt) * 8;          /* R(rowo, lk) */
    /* owner r keeps A(r, r) - sum_{p<=C} X(r, p) X(r, p)^T of its diagonal
     * block in Ld ([row][col]): each X block just published is folded in
     * while the next one is awaited */
    for (int C = 0; C < nsteps; ++C) {
        v4d acc[4] = {zero4d(), zero4d(), zero4d(), zero4d()};
        if (C > 0) {
            if (!pf_wait(a, C * ND + C - 1, tid, s_int)) return;
            pf_kloop(R, a.T + 64 * C, ldt, C, Ys, tid, wave, lane, acc);
        }
        double av[4][4];
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int s = 0; s < 4; ++s) av[f][s] = pf_ld(rR, vr, (64 * C + 16 * f + 4 * s) * ldt * 8) - acc[f][s];
        if (!pf_wait(a, C * ND + C, tid, s_int)) return;
        const rsrc_t rL = pf_rsrc(a.dinv + (size_t)C * 4096);
        double lv[16];
#pragma unroll
        for (int h = 0; h < 16; ++h) lv[h] = pf_ld_dev(rL, tid * 8, h * 2048);
#pragma unroll
        for (int h = 0; h < 16; ++h) Ys[tid + 256 * h] = lv[h];
        __syncthreads();
        v4d xo[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            xo[f] = zero4d();
#pragma unroll
            for (int kq = 0; kq < 16; ++kq)
                if (kq < 4 * (f + 1)) xo[f] = mfma_f64(Ys[(4 * kq + lk) * 64 + 16 * f + li], av[kq >> 2][kq & 3], xo[f]);
        }
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int s = 0; s < 4; ++s) pf_st(xo[f][s], rR, vr, (64 * C + 16 * f + 4 * s) * ldt * 8);
        if (owner) {
            pf_publish(a, r * ND + C, tid);
            /* at the first column block the diagonal block starts from A(r, r):
             * its 16 values per lane requested together, ahead of the
             * products (a load per value behind a run-time test put a
             * wait after each one) */
            double base0[4][4];
            if (C == 0) {
#pragma unroll
                for (int f = 0; f < 4; ++f)
#pragma unroll
                    for (int s = 0; s < 4; ++s) base0[f][s] = pf_ld(rR, vr, (64 * r + 16 * f + 4 * s) * ldt * 8);
            }
            /* rank-64 term X(r, C) X(r, C)^T of the diagonal block: the block
             * through LDS ([k][row]) as the shared operand */
#pragma unroll
            for (int f = 0; f < 4; ++f)
#pragma unroll
                for (int s = 0; s < 4; ++s) Ys[(16 * f + lk + 4 * s) * 64 + rowo] = xo[f][s];
            __syncthreads();
            v4d a2[4] = {zero4d(), zero4d(), zero4d(), zero4d()};
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const double xq = xo[q >> 2][q & 3];
#pragma unroll
                for (int f = 0; f < 4; ++f) a2[f] = mfma_f64(Ys[(4 * q + lk) * 64 + 16 * f + li], xq, a2[f]);
            }
#pragma unroll
            for (int f = 0; f < 4; ++f)
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int c = 16 * f + lk + 4 * s;
                    const double base = C == 0 ? base0[f][s] : Ld[rowo][c];
                    Ld[rowo][c] = base - a2[f][s];
                }
        }
    }
    if (!owner) return;

    /* diagonal block of owner r: A(r, r) - sum_p X(r, p) X(r, p)^T, lower part */
    double (*Lv)[OBP] = reinterpret_cast<double (*)[OBP]>(Ys);
    __syncthreads();
    if (r == 0) {
        /* owner 0 starts the chain from A(0, 0): all 16 loads in flight */
        double v0[4][4];
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int s = 0; s < 4; ++s) v0[f][s] = pf_ld(rR, vr, (16 * f + 4 * s) * ldt * 8);
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int c = 16 * f + lk + 4 * s;
                Ld[rowo][c] = (rowo >= c) ? v0[f][s] : 0.0;
            }
    } else {
#pragma unroll
        for (int f = 0; f < 4; ++f)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int c = 16 * f + lk + 4 * s;
                if (rowo < c) Ld[rowo][c] = 0.0;
            }
    }
    __syncthreads();
    const int fail = pf_factor64(Ld, Lb, Lv, Rd, s_int, wave, lane);
    if (fail) {
        const int col = 64 * r + fail;
        if (col - 1 < a.nvalid) {
            if (tid == 0 && __hip_atomic_load(a.info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
                __hip_atomic_store(a.info, a.iinfo + col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
    }
    double *Tr = a.T + 64 * r + (size_t)(64 * r) * ldt;
    double *Dv = a.dinv + (size_t)r * 4096;
    for (int e = tid; e < OB * OB; e += 256) {
        const int rr = e % OB, cc = e / OB;
        if (rr >= cc) Tr[rr + (size_t)cc * ldt] = Ld[rr][cc];
        Dv[e] = Lv[rr][cc];
    }
    pf_publish(a, r * ND + r, tid);
}

__global__ __launch_bounds__(256, 2) void panel_factor_kernel(PanelArgs a)
{
    __shared__ double Ys[OB * OBP];       /* Y blocks, Linv_C, the inverse being built */
    __shared__ double Ld[OB][OBP];        /* diagonal block, the last X block          */
    __shared__ double Lb[4][16][17];
    __shared__ double Rd[64];
    __shared__ int s_int, s_ticket, s_info;
    const int tid = threadIdx.x;
    __builtin_amdgcn_s_setprio(2);        /* critical path: win issue next to update waves */
    if (tid == 0) {
        s_ticket = (int)atomicAdd(&a.ctr[0], 1u);
        s_info = __hip_atomic_load(a.info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    /* a ticket past the grid would mean the counters were not cleared */
    if (s_ticket >= a.ND * (a.S + 1)) {
        if (tid == 0) atomicCAS(a.info, 0, PANEL_STALLED);
    } else if (s_info == 0) {
        panel_unit(a, s_ticket, Ys, Ld, Lb, Rd, &s_int);
    }
    __syncthreads();
    if (tid == 0 && atomicAdd(&a.ctr[1], 1u) == gridDim.x - 1) {
        __hip_atomic_store(&a.ctr[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&a.ctr[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

void panel_factor(double *T, double *P, long long ts, int S, int ldt, double *dinv, int *flags, unsigned *ctr,
                  int epoch, int *info, int iinfo, int nvalid, hipStream_t s)
{
    PanelArgs a;
    a.T = T; a.P = P; a.ts = ts; a.S = S; a.ldt = ldt; a.ND = ldt / 64;
    a.dinv = dinv; a.flags = flags; a.ctr = ctr; a.epoch = epoch;
    a.info = info; a.iinfo = iinfo; a.nvalid = nvalid;
    const int grid = a.ND * (S + 1);
    hipLaunchKernelGGL(panel_factor_kernel, dim3(grid), dim3(256), 0, s, a);
}

} /* namespace k */
} /* namespace chx */
