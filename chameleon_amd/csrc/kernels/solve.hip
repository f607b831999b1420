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
/* Strip solve without grid-wide barriers.
 *
 * The rows are cut in strips of R rows (R = 64 or 128, dividing ldt).
 * Strip s is owned by one workgroup of R/16 wavefronts, wave w holding
 * rows 16w .. 16w+15 of the strip and all NC*16 columns of this launch's
 * right-hand-side chunk.  The workgroup keeps acc = B_s in registers,
 * subtracts the contribution of every strip on the near side of it as soon
 * as that strip's solution is published, then applies the inverse of its
 * diagonal R x R block and publishes X_s:
 *   forward  (L Y = B):    acc -= L(s, j) X_j     for j = 0 .. s-1
 *                          X_s  = D_s acc           D_s = L(s,s)^-1
 *   backward (L^T X = Y):  acc -= L(j, s)^T X_j   for j = S-1 .. s+1
 *                          X_s  = D_s^T acc
 * The strips form a chain (s needs s-1): the time per link bounds the
 * solve, so everything of a link that does not depend on the previous
 * strip -- the rows of D_s and the factor block against the previous
 * strip -- is loaded into registers before the wait.
 *
 * Publication is a per-strip flag set to this launch's epoch after a
 * device-scope release; readers poll the flags and acquire.  A flag of
 * strip j is only set after strip j consumed every strip before it, so
 * seeing flag j also makes all earlier strips visible.  Strips are dealt
 * to workgroups in chain order (t = blockIdx.x, + gridDim.x, ...) and the
 * grid is launched cooperatively, so every strip a workgroup waits for is
 * owned by a resident workgroup that is either on it or on an earlier
 * strip: the waits always end.
 *
 * The MFMA k index of lane (li, lk) in product q is p0 + 4*lk + q, which
 * turns the transposed operand reads of the backward solve into 32-byte
 * vector loads.                                                          */
struct StripArgs {
    const double *A;      /* factor mirror (Lower orientation)                       */
    const double *D;      /* inverses of the diagonal R x R blocks, element (r, p) of
                           * strip s at D[(s/q)*dts + (s%q)*doff + r + p*ldd]         */
    long long dts, doff;
    int ldd;
    int dtrans;           /* D holds the transposes L(s,s)^-T instead                 */
    double *BT;           /* right-hand sides, element (c, r) at c + r * nrp         */
    int *flags;           /* one per strip                                            */
    int NT, ldt, nrp, c0; /* c0: first column of this launch's chunk                  */
    int S;                /* number of strips                                         */
    int trans, epoch;
    const int *info;
};

/* operand of lane (li, lk) for k-chunk p0 of a 16-row group: element
 * (row0 + li, p0 + 4 lk + q), q < 4, of a matrix whose element (r, p) is at
 * M[r + p*ld] (vec = false) or M[p + r*ld] (vec = true) */
__device__ __forceinline__ void load_op4(const double *M, int ld, bool vec, int row0, int p0, int li, int lk,
                                         double out[4])
{
    if (vec) {
        const v4d v = *(const v4d *)(M + (p0 + 4 * lk) + (size_t)(row0 + li) * ld);
#pragma unroll
        for (int q = 0; q < 4; ++q) out[q] = v[q];
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) out[q] = M[(row0 + li) + (size_t)(p0 + 4 * lk + q) * ld];
    }
}

template <int R, int NC>
__global__ __launch_bounds__(R * 4) void strip_solve_kernel(StripArgs a)
{
    constexpr int NCOL = 16 * NC, NPC = R / 16;
    __shared__ double Xs[R * NCOL];     /* acc of the strip, input of the diagonal product */
    __shared__ double Xc[R * NCOL];     /* solution of the chain predecessor               */
    if (*a.info != 0) return;     /* uniform over the grid: nobody is left waiting */
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int q = a.ldt / R, nrp = a.nrp, ldt = a.ldt;
    const size_t ts = (size_t)ldt * ldt;
    const int wr = wave * 16;
    const bool fwd = a.trans == 0;
    /* operand addressing of the factor blocks: forward reads L(s, j)
     * column-wise, backward reads L(j, s)^T as row vectors */
    const bool lvec = !fwd;
    const bool dvec = fwd ? a.dtrans != 0 : a.dtrans == 0;
    for (int t = blockIdx.x; t < a.S; t += gridDim.x) {
        const int s = fwd ? t : a.S - 1 - t;
        const int tm = s / q, rs = (s % q) * R;
        double *Bs = a.BT + a.c0 + (size_t)s * R * nrp;
        auto block_of = [&](int j) -> const double * {
            const int tj = j / q, rj = (j % q) * R;
            return fwd ? a.A + ((size_t)tj * a.NT + tm) * ts + rs + (size_t)rj * ldt
                       : a.A + ((size_t)tm * a.NT + tj) * ts + rj + (size_t)rs * ldt;
        };

        /* registers filled before any wait: the right-hand side, this
         * wave's rows of D_s, and the factor block against the chain
         * predecessor */
        v4d acc[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[c][e] = Bs[(c * 16 + li) + (size_t)(wr + lk + 4 * e) * nrp];
        const double *Dg = a.D + (size_t)tm * a.dts + (size_t)(s % q) * a.doff;
        const int pc0 = fwd ? 0 : wave, pc1 = fwd ? wave + 1 : NPC;   /* nonzero chunks of row group `wave` */
        double dreg[NPC][4];
#pragma unroll
        for (int pc = 0; pc < NPC; ++pc) {
            if (pc >= pc0 && pc < pc1) load_op4(Dg, a.ldd, dvec, wr, pc * 16, li, lk, dreg[pc]);
            else {
#pragma unroll
                for (int qq = 0; qq < 4; ++qq) dreg[pc][qq] = 0.0;
            }
        }
        const int jc = fwd ? s - 1 : s + 1;          /* chain predecessor */
        const bool has_jc = fwd ? s > 0 : s < a.S - 1;
        double lreg[NPC][4];
        if (has_jc) {
            const double *L = block_of(jc);
#pragma unroll
            for (int pc = 0; pc < NPC; ++pc) load_op4(L, ldt, lvec, wr, pc * 16, li, lk, lreg[pc]);
        }

        int ready = fwd ? -1 : a.S;
        auto wait_for = [&](int j) {
            if (fwd ? j <= ready : j >= ready) return;
            /* the lanes poll the flags of the next 64 strips; published
             * strips form a prefix of the chain, so the run of set flags
             * starting at j is all that may be consumed, behind one acquire */
            for (;;) {
                const int jj = fwd ? j + lane : j - lane;
                const bool inr = fwd ? jj < s : jj > s;
                const bool set =
                    inr && __hip_atomic_load(&a.flags[jj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == a.epoch;
                const unsigned long long m = __ballot(set);
                if (m & 1ull) {
                    const int run = (~m == 0ull) ? 64 : __builtin_ctzll(~m);
                    ready = fwd ? j + run - 1 : j - run + 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        };

        /* strips before the chain predecessor */
        const int jstep = fwd ? 1 : -1;
        for (int j = fwd ? 0 : a.S - 1; j != jc && has_jc; j += jstep) {
            wait_for(j);
            const double *Xj = a.BT + a.c0 + (size_t)j * R * nrp;
            const double *L = block_of(j);
#pragma unroll 2
            for (int pc = 0; pc < NPC; ++pc) {
                double av[4], bv[NC][4];
                load_op4(L, ldt, lvec, wr, pc * 16, li, lk, av);
#pragma unroll
                for (int qq = 0; qq < 4; ++qq)
#pragma unroll
                    for (int c = 0; c < NC; ++c) bv[c][qq] = Xj[(c * 16 + li) + (size_t)(pc * 16 + 4 * lk + qq) * nrp];
#pragma unroll
                for (int qq = 0; qq < 4; ++qq)
#pragma unroll
                    for (int c = 0; c < NC; ++c) acc[c] = mfma_f64(-av[qq], bv[c][qq], acc[c]);
            }
        }
        /* the chain predecessor: its factor block is already in registers,
         * its solution is staged once in LDS for all waves */
        if (has_jc) {
            wait_for(jc);
            const double *Xj = a.BT + a.c0 + (size_t)jc * R * nrp;
            for (int e = threadIdx.x; e < R * NCOL / 2; e += R * 4) {
                const int r = (2 * e) / NCOL, c = (2 * e) % NCOL;
                *(v2d *)&Xc[r * NCOL + c] = *(const v2d *)&Xj[c + (size_t)r * nrp];
            }
            __syncthreads();
#pragma unroll
            for (int pc = 0; pc < NPC; ++pc)
#pragma unroll
                for (int qq = 0; qq < 4; ++qq)
#pragma unroll
                    for (int c = 0; c < NC; ++c)
                        acc[c] = mfma_f64(-lreg[pc][qq], Xc[(pc * 16 + 4 * lk + qq) * NCOL + c * 16 + li], acc[c]);
        }

        /* diagonal block: X_s = D_s acc (forward) or D_s^T acc (backward) */
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) Xs[(wr + lk + 4 * e) * NCOL + c * 16 + li] = acc[c][e];
        __syncthreads();
        v4d x[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) x[c] = zero4d();
#pragma unroll
        for (int pc = 0; pc < NPC; ++pc) {
            if (pc < pc0 || pc >= pc1) continue;
#pragma unroll
            for (int qq = 0; qq < 4; ++qq)
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    x[c] = mfma_f64(dreg[pc][qq], Xs[(pc * 16 + 4 * lk + qq) * NCOL + c * 16 + li], x[c]);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) Bs[(c * 16 + li) + (size_t)(wr + lk + 4 * e) * nrp] = x[c][e];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(&a.flags[s], a.epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int R, int NC>
static void strip_launch(const StripArgs &a, hipStream_t s)
{
    static int per_cu = -1, ncu = 0;
    if (per_cu < 0) {
        int dev = 0;
        CHX_HIP(hipGetDevice(&dev));
        CHX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        CHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, strip_solve_kernel<R, NC>, R * 4, 0));
        if (per_cu < 1) per_cu = 1;
    }
    int grid = ncu * per_cu;
    if (grid > a.S) grid = a.S;
    StripArgs x = a;
    void *args[] = {&x};
    CHX_HIP(hipLaunchCooperativeKernel((const void *)strip_solve_kernel<R, NC>, dim3(grid), dim3(R * 4), args, 0, s));
}

int solve_strips(const double *A, const double *D, long long dts, long long doff, int ldd, int dtrans, int R,
                 double *BT, int NT, int ldt, int nrp, int trans, int *flags, int epoch, const int *info, hipStream_t s)
{
    StripArgs a;
    a.A = A; a.D = D; a.dts = dts; a.doff = doff; a.ldd = ldd; a.dtrans = dtrans;
    a.BT = BT; a.flags = flags;
    a.NT = NT; a.ldt = ldt; a.nrp = nrp;
    a.S = NT * ldt / R;
    a.trans = trans;
    a.info = info;
    int used = 0;
    for (int c0 = 0; c0 < nrp; c0 += 64) {
        const int nc = (nrp - c0 >= 64) ? 4 : (nrp - c0) / 16;
        a.c0 = c0;
        a.epoch = epoch + used++;
        if (R == 128) {
            switch (nc) {
            case 1: strip_launch<128, 1>(a, s); break;
            case 2: strip_launch<128, 2>(a, s); break;
            case 3: strip_launch<128, 3>(a, s); break;
            default: strip_launch<128, 4>(a, s); break;
            }
        } else {
            switch (nc) {
            case 1: strip_launch<64, 1>(a, s); break;
            case 2: strip_launch<64, 2>(a, s); break;
            case 3: strip_launch<64, 3>(a, s); break;
            default: strip_launch<64, 4>(a, s); break;
            }
        }
    }
    return used;
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
