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
 * The rows are cut in strips of 64 (64 divides ldt).  Strip s is owned by
 * one workgroup of 4 wavefronts, wave w holding rows 16w .. 16w+15 of the
 * strip and all NC*16 columns of this launch's right-hand-side chunk.  The
 * workgroup keeps acc = B_s in registers, subtracts the contribution of
 * every strip on the near side of it as soon as that strip's solution is
 * published, then applies the inverse of its 64 x 64 diagonal block and
 * publishes X_s:
 *   forward  (L Y = B):    acc -= L(s, j) X_j     for j = 0 .. s-1
 *                          X_s  = D_s acc           D_s = L(s,s)^-1
 *   backward (L^T X = Y):  acc -= L(j, s)^T X_j   for j = S-1 .. s+1
 *                          X_s  = D_s^T acc
 * A strip far down the chain consumes hundreds of published blocks, so
 * the block loop is a pipeline: while block j is multiplied, the factor
 * rows of block j+1 are loaded into registers (before waiting for its
 * publication: the factor does not depend on it) and its solution block
 * into staging registers, which go to LDS (one copy for the 4 waves)
 * after the block is done.  D_s goes to LDS before any wait.
 *
 * Publication is a per-strip flag set to this launch's epoch after a
 * device-scope release (the strip's stores written back from L2).  Readers
 * poll the flags with device-scope loads and read published blocks with
 * device-scope loads as well, which see memory past the non-coherent L2
 * lines of other XCDs; a device-scope acquire fence would invalidate the
 * reader's whole L2 instead, ~6 us per chain link (measured).  A flag of
 * strip j is only set after strip j consumed every strip before it, so
 * seeing flag j also covers all earlier strips.  Strips are dealt
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
    const double *D;      /* 64 x 64 inverses of the diagonal blocks, element (r, p) of
                           * strip s at D[(s/q)*dts + (s%q)*doff + r + p*ldd]         */
    long long dts, doff;
    int ldd;
    double *BT;           /* right-hand sides, element (c, r) at c + r * nrp         */
    int *flags;           /* one per strip and column chunk: chunk c at flags + c*S   */
    unsigned *ticket;     /* task counter of the launch (cleared before it)           */
    int NT, ldt, nrp, c0; /* c0: first column of this launch's chunks                 */
    int nchunk;           /* column chunks of 16*NC columns solved side by side       */
    int S;                /* number of strips                                         */
    int s0;               /* forward: first strip solved (the rows above are solved and
                           * already subtracted from the right-hand sides)           */
    int trans, epoch;
    const int *info;
    /* null, or the predecessor products E_s (strip_products) of this
     * direction: strip s's 64 x 64 E_s at E + s*4096, element (r, p) at
     * r + p*64 */
    const double *E;
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

#define SR 64            /* strip rows */
/* polls of one flag wait before it gives up: about a second at ~0.3 us per
 * poll, far beyond any chain (a full direction at N = 32768 is ~3 ms) */
#define STRIP_SPIN_LIMIT (1 << 22)
template <int NC, bool PE, bool FWD>
__global__ __launch_bounds__(256, 2) void strip_solve_kernel(StripArgs a)
{
    constexpr int NCOL = 16 * NC, NPC = SR / 16;
    constexpr int XV = SR * NCOL / (2 * 256);     /* 16-byte pieces of a block staged per thread */
    /* the solution block being consumed; after the block loop, the acc of
     * the strip (input of the diagonal product) */
    __shared__ double Xb[SR * NCOL];
    __shared__ double Ds[SR * (SR + 1)];          /* D_s, element (r, p) at r + p*65 */
    double *Xs = Xb;
    if (*a.info != 0) return;     /* uniform over the grid: nobody is left waiting */
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int q = a.ldt / SR, nrp = a.nrp, ldt = a.ldt;
    const size_t ts = (size_t)ldt * ldt;
    const int wr = wave * 16;
    /* the direction is a template parameter: with a run-time choice
     * between the two operand loads the compiler put a wait for all
     * outstanding loads before each 16-column piece of a factor block */
    constexpr bool fwd = FWD;
    constexpr bool lvec = !FWD;   /* backward reads L(j, s)^T as row vectors */
    /* task t = (strip in chain order) * nchunk + chunk: the chunks are
     * independent chains, interleaved so that the tasks a task waits for
     * always have smaller indices.  Tasks are taken from a ticket counter
     * (cleared before the launch), so every task a workgroup waits for
     * belongs to a workgroup that has started: the chain moves whatever
     * else occupies the device, no co-residency of the grid is needed. */
    __shared__ int s_task;
    unsigned *const ticket = a.ticket;
    const int ntask = (a.S - a.s0) * a.nchunk;
    for (;;) {
        __syncthreads();
        if (tid == 0) s_task = (int)atomicAdd(ticket, 1u);
        __syncthreads();
        const int t = s_task;
        if (t >= ntask) break;
        const int si = t / a.nchunk, cc = t - si * a.nchunk;
        const int s = fwd ? a.s0 + si : a.S - 1 - si;
        const int c0 = a.c0 + cc * NCOL;
        int *const flags = a.flags + (size_t)cc * a.S;
        const int tm = s / q, rs = (s % q) * SR;
        double *Bs = a.BT + c0 + (size_t)s * SR * nrp;
        auto block_of = [&](int j) -> const double * {
            const int tj = j / q, rj = (j % q) * SR;
            return fwd ? a.A + ((size_t)tj * a.NT + tm) * ts + rs + (size_t)rj * ldt
                       : a.A + ((size_t)tm * a.NT + tj) * ts + rj + (size_t)rs * ldt;
        };
        auto xblock = [&](int j) -> const double * { return a.BT + c0 + (size_t)j * SR * nrp; };

        v4d acc[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[c][e] = Bs[(c * 16 + li) + (size_t)(wr + lk + 4 * e) * nrp];
        /* D_s into LDS before any wait (read by the diagonal product) */
        const double *Dg = a.D + (size_t)tm * a.dts + (size_t)(s % q) * a.doff;
        {
            /* all 16 loads in flight before the first LDS store */
            double dv[SR * SR / 256];
#pragma unroll
            for (int h = 0; h < SR * SR / 256; ++h) {
                const int e = tid + 256 * h;
                dv[h] = Dg[(e % SR) + (size_t)(e / SR) * a.ldd];
            }
#pragma unroll
            for (int h = 0; h < SR * SR / 256; ++h) {
                const int e = tid + 256 * h;
                Ds[(e % SR) + (e / SR) * (SR + 1)] = dv[h];
            }
        }
        const int pc0 = fwd ? 0 : wave, pc1 = fwd ? wave + 1 : NPC;   /* nonzero chunks of row group `wave` */

        int ready = fwd ? a.s0 - 1 : a.S;
        auto wait_for = [&](int j) {
            if (fwd ? j <= ready : j >= ready) return;
            /* the lanes poll the flags of the next 64 strips; published
             * strips form a prefix of the chain, so the run of set flags
             * starting at j is all that may be consumed, behind one acquire.
             * The wait is bounded: when the info flag turns nonzero (the
             * factorisation failed, or another wait ran out) or the poll
             * limit is reached (PANEL_STALLED recorded in info), the wave
             * stops waiting and goes on with whatever the blocks hold, so
             * every workgroup still reaches its barriers and the grid
             * drains; the host reports the info value, not the result. */
            for (int polls = 0;; ++polls) {
                const int jj = fwd ? j + lane : j - lane;
                const bool inr = fwd ? jj < s : jj > s;
                const bool set =
                    inr && __hip_atomic_load(&flags[jj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == a.epoch;
                const unsigned long long m = __ballot(set);
                if (m & 1ull) {
                    const int run = (~m == 0ull) ? 64 : __builtin_ctzll(~m);
                    ready = fwd ? j + run - 1 : j - run + 1;
                    break;
                }
                if ((polls & 63) == 63) {
                    const int inf = __hip_atomic_load(a.info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (inf == 0 && polls >= STRIP_SPIN_LIMIT && lane == 0)
                        atomicCAS((int *)a.info, 0, PANEL_STALLED);
                    if (inf != 0 || polls >= STRIP_SPIN_LIMIT) {
                        ready = fwd ? a.S : -1;      /* no further waits in this task */
                        break;
                    }
                }
                __builtin_amdgcn_s_sleep(1);
            }
            /* no device-scope acquire here: it would invalidate the whole
             * L2 of this XCD (about 6 us per chain link, measured); the
             * published blocks are read with device-scope (coherent) loads
             * instead, like the flags, after the flag was seen */
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        };
        auto stage_load = [&](int j, v2d (&xv)[XV]) {
            const double *Xj = xblock(j);
#pragma unroll
            for (int h = 0; h < XV; ++h) {
                const int e = 2 * (tid + 256 * h), r = e / NCOL, c = e % NCOL;
                const unsigned long long *src = (const unsigned long long *)&Xj[c + (size_t)r * nrp];
                const unsigned long long u0 = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned long long u1 = __hip_atomic_load(src + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                xv[h] = v2d{__longlong_as_double((long long)u0), __longlong_as_double((long long)u1)};
            }
        };
        auto stage_store = [&](const v2d (&xv)[XV]) {
#pragma unroll
            for (int h = 0; h < XV; ++h) *(v2d *)&Xb[2 * (tid + 256 * h)] = xv[h];
        };

        /* blocks j in chain order; the last one is the chain predecessor.
         * Block j is multiplied before the wait for block j+1, so a strip
         * at the front of the chain has its partial sum ready when its
         * predecessor publishes.  With E the predecessor is left out here:
         * X_s = D_s acc' - E_s X_pred, where acc' lacks the predecessor's
         * block and E_s = D_s L(s, s-1) (forward) or D_s^T L(s+1, s)^T
         * (backward), so the diagonal product runs before the predecessor
         * publishes and a chain link holds one product instead of two. */
        const int nall = fwd ? s - a.s0 : a.S - 1 - s;
        const bool use_e = PE && nall > 0;
        const int nblk = use_e ? nall - 1 : nall;
        if (nblk > 0) {
            double lcur[NPC][4], lnext[NPC][4];
            v2d xv[XV];
            int j = fwd ? a.s0 : a.S - 1;
            {
                const double *L = block_of(j);
#pragma unroll
                for (int pc = 0; pc < NPC; ++pc) load_op4(L, ldt, lvec, wr, pc * 16, li, lk, lcur[pc]);
                wait_for(j);
                stage_load(j, xv);
                stage_store(xv);
                __syncthreads();
            }
            for (int b = 0; b < nblk; ++b) {
                const bool more = b + 1 < nblk;
                const int jn = fwd ? j + 1 : j - 1;
                if (more) {
                    const double *L = block_of(jn);
#pragma unroll
                    for (int pc = 0; pc < NPC; ++pc) load_op4(L, ldt, lvec, wr, pc * 16, li, lk, lnext[pc]);
                }
#pragma unroll
                for (int pc = 0; pc < NPC; ++pc)
#pragma unroll
                    for (int qq = 0; qq < 4; ++qq)
#pragma unroll
                        for (int c = 0; c < NC; ++c)
                            acc[c] = mfma_f64(-lcur[pc][qq], Xb[(pc * 16 + 4 * lk + qq) * NCOL + c * 16 + li], acc[c]);
                if (more) {
                    wait_for(jn);
                    stage_load(jn, xv);
                }
                __syncthreads();                 /* every wave is done with Xb */
                if (more) {
                    stage_store(xv);
#pragma unroll
                    for (int pc = 0; pc < NPC; ++pc)
#pragma unroll
                        for (int qq = 0; qq < 4; ++qq) lcur[pc][qq] = lnext[pc][qq];
                }
                __syncthreads();                 /* the next block is in Xb */
                j = jn;
            }
        }

        /* the predecessor product's operand, requested before the
         * diagonal step (its latency hides behind it) */
        double ev[NPC][4];
        if constexpr (PE) {
            const double *Eg = a.E + (size_t)s * SR * SR;
#pragma unroll
            for (int pc = 0; pc < NPC; ++pc) load_op4(Eg, SR, false, wr, pc * 16, li, lk, ev[pc]);
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
        for (int pc = pc0; pc < pc1; ++pc) {
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) {
                const int p = pc * 16 + 4 * lk + qq, r = wr + li;
                const double dv = fwd ? Ds[r + p * (SR + 1)] : Ds[p + r * (SR + 1)];
#pragma unroll
                for (int c = 0; c < NC; ++c) x[c] = mfma_f64(dv, Xs[p * NCOL + c * 16 + li], x[c]);
            }
        }
        if (PE && use_e) {
            /* X_s -= E_s X_pred once the predecessor is published */
            const int jp = fwd ? s - 1 : s + 1;
            v2d xv[XV];
            wait_for(jp);
            stage_load(jp, xv);
            __syncthreads();                     /* every wave is done with Xs */
            stage_store(xv);
            __syncthreads();
#pragma unroll
            for (int pc = 0; pc < NPC; ++pc)
#pragma unroll
                for (int qq = 0; qq < 4; ++qq)
#pragma unroll
                    for (int c = 0; c < NC; ++c)
                        x[c] = mfma_f64(-ev[pc][qq], Xb[(pc * 16 + 4 * lk + qq) * NCOL + c * 16 + li], x[c]);
        }
        {
            /* publication by completion: X_s is written with device-scope
             * (coherent, sc1) stores, which reach the memory side past this
             * XCD's L2 by themselves; once every wave has waited for its
             * stores to complete (vmcnt 0) and the workgroup has met at the
             * barrier, the flag store follows them.  A release fence would
             * write back the whole L2 for the same guarantee (measured
             * slower: dpotrs at N = 32768 3.90 against 4.13 ms, NRHS 1). */
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    __hip_atomic_store(&Bs[(c * 16 + li) + (size_t)(wr + lk + 4 * e) * nrp], x[c][e], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_s_waitcnt(0x0F70);   /* vmcnt(0), expcnt and lgkmcnt not waited for */
            __syncthreads();
            if (tid == 0) __hip_atomic_store(&flags[s], a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <int NC, bool PE, bool FWD>
static void strip_launch_dir(const StripArgs &a, hipStream_t s)
{
    static int per_cu = -1, ncu = 0;
    if (per_cu < 0) {
        int dev = 0;
        CHX_HIP(hipGetDevice(&dev));
        CHX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        CHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, strip_solve_kernel<NC, PE, FWD>, 256, 0));
        if (per_cu < 1) per_cu = 1;
    }
    int grid = ncu * per_cu;
    if (grid > (a.S - a.s0) * a.nchunk) grid = (a.S - a.s0) * a.nchunk;
    zero_ints(reinterpret_cast<int *>(a.ticket), 1, s);
    hipLaunchKernelGGL((strip_solve_kernel<NC, PE, FWD>), dim3(grid), dim3(256), 0, s, a);
}

template <int NC, bool PE>
static void strip_launch(const StripArgs &a, hipStream_t s)
{
    if (a.trans) strip_launch_dir<NC, PE, false>(a, s);
    else strip_launch_dir<NC, PE, true>(a, s);
}

/* ------------------------------------------------------------------ */
/* Narrow right-hand sides (at most 8 columns): strips of 128 rows solved
 * with FP64 FMAs on the vector ALUs.
 *
 * An FP64 MFMA computes 16 columns whether or not they are in use, and on
 * this device the FP64 matrix and vector peaks are the same, so for one
 * to eight columns the vector ALUs do the chain's arithmetic 2-16x sooner
 * and strips twice as tall halve the number of chain links (each link is
 * a few memory round trips: flag seen, solution block read, solution
 * stored and flag set).  The strip's diagonal step uses the 128 x 128
 * inverse of its diagonal block (D128, built by dinv128 from the 64 x 64
 * inverses): X_s = D128_s acc (forward) or D128_s^T acc (backward).
 *
 * Workgroup of 512 threads per strip: thread (r, kq) = (tid % 128,
 * tid / 128) owns row r of the strip and the k range [32 kq, 32 kq + 32)
 * of every 128 x 128 factor block; it accumulates its partial products
 * over all blocks in registers, and the four partial sums of a row meet
 * once, in LDS, before the diagonal step.  Factor rows of the next block
 * are loaded before the wait for the current one's publication (the
 * factor does not depend on it); wave 0 alone polls the flags and reads
 * the published block into LDS (double-buffered, one barrier per block).
 * Tasks come from a ticket counter in chain order and publication is the
 * same as in strip_solve_kernel (device-scope stores, store completion,
 * barrier, flag = epoch), so the same argument shows every wait ends.   */
#define NSR 128

struct NarrowArgs {
    const double *A;      /* factor mirror (Lower orientation)                          */
    const double *D128;   /* 128 x 128 inverse of the diagonal block of strip s at
                           * D128 + s * 16384, element (r, p) at r + p * 128            */
    double *BT;           /* right-hand sides, element (c, r) at c + r * nrp            */
    int *flags;           /* one per strip                                              */
    unsigned *ticket;     /* task counter of the launch (cleared before it)              */
    int NT, ldt, nrp;
    int S, s0;            /* strips of 128 rows; forward: first strip solved            */
    int trans, epoch;
    const int *info;
};

template <int NCV, bool FWD>
__global__ __launch_bounds__(512, 1) void strip_narrow_kernel(NarrowArgs a)
{
    /* backward: one factor block at a time goes through LDS (rows of TLD
     * doubles: 16-byte reads of 8 consecutive rows hit distinct banks), and
     * the diagonal step's buffers reuse that space once the blocks are done */
    constexpr int TLD = NSR + 2;
    constexpr int PSZ = 5 * NSR * NCV;
    constexpr int SMEM = FWD ? PSZ : (NSR * TLD > PSZ ? NSR * TLD : PSZ);
    __shared__ double Xs[2][NSR * NCV];   /* published blocks, element (p, c) at p*NCV + c */
    __shared__ double smem[SMEM];
    double *const T = smem;               /* backward: op(r, p) at T[r*TLD + p]               */
    double (*const P)[NSR * NCV] = reinterpret_cast<double (*)[NSR * NCV]>(smem);   /* partial sums */
    double *const Acc = smem + 4 * NSR * NCV;                                        /* diagonal input */
    __shared__ int s_task;
    if (*a.info != 0) return;     /* uniform over the grid at launch */
    /* wave and k range are uniform over a wavefront: scalar registers, so
     * the uniform part of every load address is scalar too */
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), kq = wave >> 1;
    const int r = tid & (NSR - 1);
    const int ldt = a.ldt, nrp = a.nrp, q = ldt / NSR;
    const size_t ts = (size_t)ldt * ldt;
    constexpr bool fwd = FWD;
    const int ntask = a.S - a.s0;
    for (;;) {
        __syncthreads();
        if (tid == 0) s_task = (int)atomicAdd(a.ticket, 1u);
        __syncthreads();
        const int t = s_task;
        if (t >= ntask) break;
        const int s = fwd ? a.s0 + t : a.S - 1 - t;
        const int tm = s / q, rs = (s % q) * NSR;
        double *Bs = a.BT + (size_t)s * NSR * nrp;
        auto block_of = [&](int j) -> const double * {
            const int tj = j / q, rj = (j % q) * NSR;
            return fwd ? a.A + ((size_t)tj * a.NT + tm) * ts + rs + (size_t)rj * ldt
                       : a.A + ((size_t)tm * a.NT + tj) * ts + rj + (size_t)rs * ldt;
        };
        /* forward: op(r, p) = L(s, j)(r, p) of block j for p in this
         * thread's k range (lanes on consecutive r: coalesced) */
        auto load_l = [&](int j, double (&l)[32]) {
            const double *b = block_of(j);
#pragma unroll
            for (int i = 0; i < 32; ++i) l[i] = b[r + (size_t)(32 * kq + i) * ldt];
        };
        /* backward: op(r, p) = L(j, s)(p, r) is contiguous in p; wave w
         * reads rows r = 16 w .. 16 w + 15, one 1 KB row per instruction */
        auto load_g = [&](int j, v2d (&g)[16]) {
            const double *b = block_of(j);
#pragma unroll
            for (int i = 0; i < 16; ++i) g[i] = *(const v2d *)(b + 2 * lane + (size_t)(16 * wave + i) * ldt);
        };
        double bval[NCV];
        if (kq == 0) {
#pragma unroll
            for (int c = 0; c < NCV; ++c) bval[c] = Bs[c + (size_t)r * nrp];
        }
        /* row r (forward) or column r (backward) of the inverse, k range kq */
        double dreg[32];
        {
            const double *Dg = a.D128 + (size_t)s * NSR * NSR;
            if (fwd) {
#pragma unroll
                for (int i = 0; i < 32; ++i) dreg[i] = Dg[r + (size_t)(32 * kq + i) * NSR];
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const v4d v = *(const v4d *)(Dg + 32 * kq + 4 * i + (size_t)r * NSR);
#pragma unroll
                    for (int e = 0; e < 4; ++e) dreg[4 * i + e] = v[e];
                }
            }
        }
        double part[NCV];
#pragma unroll
        for (int c = 0; c < NCV; ++c) part[c] = 0.0;

        int ready = fwd ? a.s0 - 1 : a.S;
        /* wave 0 only: wait until strip j is published (bounded, as in
         * strip_solve_kernel), then copy its solution block into Xs[buf] */
        auto fetch = [&](int j, int buf) {
            if (fwd ? j > ready : j < ready) {
                for (int polls = 0;; ++polls) {
                    const int jj = fwd ? j + lane : j - lane;
                    const bool inr = fwd ? jj < s : jj > s;
                    const bool set = inr && __hip_atomic_load(&a.flags[jj], __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT) == a.epoch;
                    const unsigned long long m = __ballot(set);
                    if (m & 1ull) {
                        const int run = (~m == 0ull) ? 64 : __builtin_ctzll(~m);
                        ready = fwd ? j + run - 1 : j - run + 1;
                        break;
                    }
                    if ((polls & 63) == 63) {
                        const int inf = __hip_atomic_load(a.info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (inf == 0 && polls >= STRIP_SPIN_LIMIT && lane == 0)
                            atomicCAS((int *)a.info, 0, PANEL_STALLED);
                        if (inf != 0 || polls >= STRIP_SPIN_LIMIT) {
                            ready = fwd ? a.S : -1;
                            break;
                        }
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
            const double *Xj = a.BT + (size_t)j * NSR * nrp;
            double xv[2][NCV];
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int c = 0; c < NCV; ++c) {
                    const unsigned long long u = __hip_atomic_load(
                        (const unsigned long long *)&Xj[c + (size_t)(lane + 64 * h) * nrp], __ATOMIC_RELAXED,
                        __HIP_MEMORY_SCOPE_AGENT);
                    xv[h][c] = __longlong_as_double((long long)u);
                }
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int c = 0; c < NCV; ++c) Xs[buf][(lane + 64 * h) * NCV + c] = xv[h][c];
        };

        /* block j's factor rows are requested before the wait for its
         * solution, so their latency hides behind the wait; one block in
         * flight per workgroup (128 KB) is enough to keep HBM busy */
        const int nblk = fwd ? s - a.s0 : a.S - 1 - s;
        int j = fwd ? a.s0 : a.S - 1;
        if constexpr (FWD) {
            for (int b = 0; b < nblk; ++b) {
                double l[32];
                load_l(j, l);
                /* Xs[b&1] was last read in iteration b-2, before the barrier
                 * of iteration b-1 */
                if (wave == 0) fetch(j, b & 1);
                __syncthreads();
                const double *xb = Xs[b & 1];
#pragma unroll
                for (int i = 0; i < 32; ++i) {
                    const int p = 32 * kq + i;
#pragma unroll
                    for (int c = 0; c < NCV; ++c) part[c] = fma(l[i], xb[p * NCV + c], part[c]);
                }
                ++j;
            }
        } else {
            /* block j in registers (coalesced) -> T -> row reads; the next
             * block's loads are issued before this block's products */
            v2d g[16];
            if (nblk > 0) load_g(j, g);
            for (int b = 0; b < nblk; ++b) {
                if (wave == 0) fetch(j, b & 1);
                __syncthreads();                 /* T no longer read; Xs[b&1] ready */
#pragma unroll
                for (int i = 0; i < 16; ++i) *(v2d *)&T[(16 * wave + i) * TLD + 2 * lane] = g[i];
                __syncthreads();                 /* T holds block j */
                if (b + 1 < nblk) load_g(j - 1, g);
                const double *xb = Xs[b & 1];
                const double *tr = T + r * TLD + 32 * kq;
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const v2d t = *(const v2d *)(tr + 2 * i);
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int p = 32 * kq + 2 * i + e;
#pragma unroll
                        for (int c = 0; c < NCV; ++c) part[c] = fma(t[e], xb[p * NCV + c], part[c]);
                    }
                }
                --j;
            }
            __syncthreads();                     /* T no longer read: P and Acc reuse it */
        }

        /* acc = B_s - sum of the four partial products, then the diagonal
         * step; P and Acc were last read before the previous barriers */
#pragma unroll
        for (int c = 0; c < NCV; ++c) P[kq][r * NCV + c] = part[c];
        __syncthreads();
        if (kq == 0) {
#pragma unroll
            for (int c = 0; c < NCV; ++c) {
                const int e = r * NCV + c;
                Acc[e] = bval[c] - (((P[0][e] + P[1][e]) + P[2][e]) + P[3][e]);
            }
        }
        __syncthreads();
        double xp[NCV];
#pragma unroll
        for (int c = 0; c < NCV; ++c) xp[c] = 0.0;
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            const int p = 32 * kq + i;
#pragma unroll
            for (int c = 0; c < NCV; ++c) xp[c] = fma(dreg[i], Acc[p * NCV + c], xp[c]);
        }
#pragma unroll
        for (int c = 0; c < NCV; ++c) P[kq][r * NCV + c] = xp[c];
        __syncthreads();
        if (kq == 0) {
#pragma unroll
            for (int c = 0; c < NCV; ++c) {
                const int e = r * NCV + c;
                const double x = ((P[0][e] + P[1][e]) + P[2][e]) + P[3][e];
                __hip_atomic_store(&Bs[c + (size_t)r * nrp], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        /* publication by completion, as in strip_solve_kernel */
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&a.flags[s], a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int NCV>
static void narrow_launch(const NarrowArgs &a, hipStream_t s)
{
    static int per_cu = -1, ncu = 0;
    if (per_cu < 0) {
        int dev = 0, pf = 0, pb = 0;
        CHX_HIP(hipGetDevice(&dev));
        CHX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        CHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pf, strip_narrow_kernel<NCV, true>, 512, 0));
        CHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pb, strip_narrow_kernel<NCV, false>, 512, 0));
        per_cu = pf < pb ? pf : pb;
        if (per_cu < 1) per_cu = 1;
    }
    int grid = ncu * per_cu;
    if (grid > a.S - a.s0) grid = a.S - a.s0;
    zero_ints(reinterpret_cast<int *>(a.ticket), 1, s);
    if (a.trans)
        hipLaunchKernelGGL((strip_narrow_kernel<NCV, false>), dim3(grid), dim3(512), 0, s, a);
    else
        hipLaunchKernelGGL((strip_narrow_kernel<NCV, true>), dim3(grid), dim3(512), 0, s, a);
}

/* D128 of strip s (rows 128 s .. 128 s + 127, inside diagonal tile
 * tm = s / (ldt/128)) from the 64 x 64 inverses Da, Db of its two halves
 * and the factor block L_ba between them:
 *     [ L_aa   0   ]^-1   [ Da               0  ]
 *     [ L_ba  L_bb ]    = [ -Db (L_ba Da)    Db ]
 * The inverses' upper triangles are taken as zero.  One workgroup per
 * strip, each thread a 4 x 4 block of the two 64 x 64 products. */
__global__ __launch_bounds__(256) void dinv128_kernel(const double *A, const double *D, long long dts,
                                                      long long doff, int ldd, int NT, int ldt, double *D128)
{
    __shared__ double X[64 * 65], Y[64 * 65];
    const int s = blockIdx.x, tid = threadIdx.x;
    const int q = ldt / NSR, tm = s / q, rs = (s % q) * NSR;
    const int qa = ldt / 64, sa = 2 * s, sb = 2 * s + 1;
    const size_t ts = (size_t)ldt * ldt;
    const double *Da = D + (size_t)(sa / qa) * dts + (size_t)(sa % qa) * doff;
    const double *Db = D + (size_t)(sb / qa) * dts + (size_t)(sb % qa) * doff;
    const double *Lba = A + ((size_t)tm * NT + tm) * ts + (rs + 64) + (size_t)rs * ldt;
    double *O = D128 + (size_t)s * NSR * NSR;
    for (int e = tid; e < 4096; e += 256) {
        const int i = e & 63, k = e >> 6;
        X[i + k * 65] = Lba[i + (size_t)k * ldt];
        Y[i + k * 65] = i >= k ? Da[i + (size_t)k * ldd] : 0.0;
    }
    __syncthreads();
    const int i0 = (tid & 15) * 4, j0 = (tid >> 4) * 4;
    double g[4][4] = {};
    for (int k = 0; k < 64; ++k) {
        double x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            x[u] = X[(i0 + u) + k * 65];
            y[u] = Y[k + (j0 + u) * 65];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) g[u][v] = fma(x[u], y[v], g[u][v]);
    }
    /* top half of the result: [Da 0] */
    for (int e = tid; e < 64 * NSR; e += 256) {
        const int i = e & 63, j = e >> 6;
        O[i + (size_t)j * NSR] = j < 64 ? Y[i + j * 65] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) X[(i0 + u) + (j0 + v) * 65] = g[u][v];
    for (int e = tid; e < 4096; e += 256) {
        const int i = e & 63, k = e >> 6;
        Y[i + k * 65] = i >= k ? Db[i + (size_t)k * ldd] : 0.0;
    }
    __syncthreads();
    double c[4][4] = {};
    for (int k = 0; k < 64; ++k) {
        double y[4], x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            y[u] = Y[(i0 + u) + k * 65];
            x[u] = X[k + (j0 + u) * 65];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) c[u][v] = fma(y[u], x[v], c[u][v]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) O[(64 + i0 + u) + (size_t)(j0 + v) * NSR] = -c[u][v];
    for (int e = tid; e < 4096; e += 256) {
        const int i = e & 63, j = e >> 6;
        O[(64 + i) + (size_t)(64 + j) * NSR] = Y[i + j * 65];
    }
}

bool narrow_eligible(int ldt, int ncols) { return ldt % NSR == 0 && ncols >= 1 && ncols <= 8; }

void dinv128(const double *A, const double *D, long long dts, long long doff, int ldd, int NT, int ldt,
             double *D128, hipStream_t s)
{
    const int S = NT * ldt / NSR;
    if (S <= 0) return;
    hipLaunchKernelGGL(dinv128_kernel, dim3(S), dim3(256), 0, s, A, D, dts, doff, ldd, NT, ldt, D128);
}

/* Predecessor products of the 64-row strips (strip s = rows 64 s ..):
 *   forward  Ef_s = D_s L(s, s-1)          for s >= 1
 *   backward Eb_s = (L(s+1, s) D_s)^T      for s <= S-2
 * with D_s the 64 x 64 inverse of the diagonal block (upper triangle taken
 * as zero), each 64 x 64 at E + s*4096, element (r, p) at r + p*64.  One
 * workgroup per strip, each thread a 4 x 4 block of the products. */
__global__ __launch_bounds__(256) void strip_products_kernel(const double *A, const double *D, long long dts,
                                                             long long doff, int ldd, int NT, int ldt, double *Ef,
                                                             double *Eb)
{
    __shared__ double Dl[64 * 65], Lb[64 * 65];
    const int s = blockIdx.x, tid = threadIdx.x, S = NT * ldt / SR;
    const int q = ldt / SR;
    const size_t ts = (size_t)ldt * ldt;
    const double *Dg = D + (size_t)(s / q) * dts + (size_t)(s % q) * doff;
    auto block = [&](int row_strip, int col_strip) -> const double * {
        const int tr = row_strip / q, rr = (row_strip % q) * SR, tc = col_strip / q, rc = (col_strip % q) * SR;
        return A + ((size_t)tc * NT + tr) * ts + rr + (size_t)rc * ldt;   /* element (i, k) at i + k*ldt */
    };
    for (int e = tid; e < 4096; e += 256) {
        const int i = e & 63, k = e >> 6;
        Dl[i + k * 65] = i >= k ? Dg[i + (size_t)k * ldd] : 0.0;
    }
    const int i0 = (tid & 15) * 4, j0 = (tid >> 4) * 4;
    if (s >= 1) {
        const double *L = block(s, s - 1);
        for (int e = tid; e < 4096; e += 256) {
            const int i = e & 63, k = e >> 6;
            Lb[i + k * 65] = L[i + (size_t)k * ldt];
        }
        __syncthreads();
        double g[4][4] = {};
        for (int k = 0; k < 64; ++k) {
            double d[4], l[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                d[u] = Dl[(i0 + u) + k * 65];
                l[u] = Lb[k + (j0 + u) * 65];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) g[u][v] = fma(d[u], l[v], g[u][v]);
        }
        double *O = Ef + (size_t)s * 4096;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) O[(i0 + u) + (j0 + v) * 64] = g[u][v];
    }
    if (s <= S - 2) {
        const double *L = block(s + 1, s);
        __syncthreads();                  /* Lb no longer read */
        for (int e = tid; e < 4096; e += 256) {
            const int i = e & 63, k = e >> 6;
            Lb[i + k * 65] = L[i + (size_t)k * ldt];
        }
        __syncthreads();
        /* G = L(s+1, s) D_s; Eb = G^T: Eb(r, p) = G(p, r) */
        double g[4][4] = {};
        for (int k = 0; k < 64; ++k) {
            double l[4], d[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                l[u] = Lb[(i0 + u) + k * 65];
                d[u] = Dl[k + (j0 + u) * 65];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) g[u][v] = fma(l[u], d[v], g[u][v]);
        }
        double *O = Eb + (size_t)s * 4096;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) O[(j0 + v) + (i0 + u) * 64] = g[u][v];
    }
}

void strip_products(const double *A, const double *D, long long dts, long long doff, int ldd, int NT, int ldt,
                    double *Ef, double *Eb, hipStream_t s)
{
    const int S = NT * ldt / SR;
    if (S <= 0) return;
    hipLaunchKernelGGL(strip_products_kernel, dim3(S), dim3(256), 0, s, A, D, dts, doff, ldd, NT, ldt, Ef, Eb);
}

int solve_strips(const double *A, const double *D, long long dts, long long doff, int ldd, double *BT, int NT,
                 int ldt, int nrp, int trans, int *flags, int epoch, unsigned *ticket, const int *info, hipStream_t s,
                 int s0, const double *D128, int ncols, const double *E)
{
    if (D128 && narrow_eligible(ldt, ncols)) {
        NarrowArgs na;
        na.A = A; na.D128 = D128; na.BT = BT; na.flags = flags; na.ticket = ticket;
        na.NT = NT; na.ldt = ldt; na.nrp = nrp;
        na.S = NT * ldt / NSR;
        na.s0 = trans ? 0 : s0 / 2;     /* s0 counts 64-row strips; ldt / 64 is even here */
        if (na.s0 >= na.S) return 0;
        na.trans = trans;
        na.epoch = epoch;
        na.info = info;
        if (ncols == 1) narrow_launch<1>(na, s);
        else if (ncols == 2) narrow_launch<2>(na, s);
        else if (ncols <= 4) narrow_launch<4>(na, s);
        else narrow_launch<8>(na, s);
        return 1;
    }
    StripArgs a;
    a.A = A; a.D = D; a.dts = dts; a.doff = doff; a.ldd = ldd;
    a.BT = BT; a.flags = flags; a.ticket = ticket;
    a.NT = NT; a.ldt = ldt; a.nrp = nrp;
    a.S = NT * ldt / SR;
    a.s0 = trans ? 0 : s0;
    if (a.s0 >= a.S) return 0;
    a.trans = trans;
    a.info = info;
    a.E = E;
    /* groups of up to 64 columns per launch; a full group is split into
     * two chunks of 32 columns that run as separate chains side by side
     * (shorter work per chain link).  Each chunk fetches the factor rows
     * past the L2 itself: at N = 32768, NRHS 64 two chunks fetch ~10.5 GB
     * per direction for a 4.4 GB factor, and still beat one 64-column
     * chain (3.2 against 4.0 ms) or four 16-column ones (3.9 ms). */
    const int snc = 2;
    int used = 0;
    for (int c0 = 0; c0 < nrp; c0 += 64) {
        const int cols = (nrp - c0 >= 64) ? 64 : nrp - c0;
        a.c0 = c0;
        a.epoch = epoch + used++;
        int nc = cols / 16;
        a.nchunk = 1;
        if (cols == 64 && snc < 4) {
            nc = snc;
            a.nchunk = 4 / snc;
        }
        /* the predecessor products for chunks of 16 and 32 columns (wider
         * chunks would not fit their registers) */
        if (nc <= 2 && E) {
            if (nc == 1) strip_launch<1, true>(a, s);
            else strip_launch<2, true>(a, s);
        } else {
            a.E = nullptr;
            switch (nc) {
            case 1: strip_launch<1, false>(a, s); break;
            case 2: strip_launch<2, false>(a, s); break;
            case 3: strip_launch<3, false>(a, s); break;
            default: strip_launch<4, false>(a, s); break;
            }
            a.E = E;
        }
    }
    return used;
}

/* ------------------------------------------------------------------ */
/* Forward substitution of dposv beside the factorisation, in the
 * right-looking order of Chameleon's forward pztrsm (compute/pztrsm.c:
 * 134-160: TRSM of B(k) by L(k,k), then GEMM of every B(m), m > k, by
 * L(m,k)).  Both run as soon as panel k is factored: L(k,k), the tiles
 * below it and the 64 x 64 inverses of its diagonal blocks (D, written by
 * the panel kernel, element (r, p) of block b at D[b*4096 + r + p*64], zero
 * above the diagonal) are final then.  BT holds the right-hand sides
 * transposed (element (c, r) at c + r*nrp), solved in place.
 *
 * fwd_tile_kernel: one workgroup per 16*NC columns walks the 64-row
 * strips b of tile k in order: acc = B_b - sum_{j<b} L(b,j) X_j over the
 * strips of this tile (the earlier tiles were subtracted by fwd_update),
 * then X_b = D_b acc.  Each wave owns 16 rows of the strip; the MFMA
 * lane layout is the one of strip_solve_kernel (k index p0 + 4 lk + q).   */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const double *p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(p), 0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, int vo, int so)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}

template <int NC>
__global__ __launch_bounds__(256) void fwd_tile_kernel(const double *A, const double *D, double *BT, int k, int NT,
                                                       int ldt, int nrp, const int *info)
{
    constexpr int NCOL = 16 * NC;
    /* few registers and little LDS on purpose: the launch runs beside the
     * bulk updates and has to fit a CU next to them */
    __shared__ double Xs[64 * NCOL];
    if (*info != 0) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4, wr = wave * 16;
    const size_t ts = (size_t)ldt * ldt;
    const double *L = A + ((size_t)k * NT + k) * ts;
    double *B = BT + blockIdx.x * NCOL + (size_t)k * ldt * nrp;
    /* buffer loads: a 32-bit offset per lane and a scalar one per load,
     * instead of a 64-bit address per load */
    const __amdgpu_buffer_rsrc_t rl = buf_rsrc(L), rb = buf_rsrc(B);
    for (int b = 0; b < ldt / 64; ++b) {
        const int vl = ((64 * b + wr + li) + 4 * lk * ldt) * 8, vb = (li + 4 * lk * nrp) * 8;
        v4d acc[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[c][e] = B[(c * 16 + li) + (size_t)(64 * b + wr + lk + 4 * e) * nrp];
        /* D_b operands of this wave's rows, requested before the products
         * (the blocks above the diagonal are zero and not multiplied) */
        const double *Db = D + (size_t)b * 4096;
        double dv[4][4];
#pragma unroll
        for (int pc = 0; pc < 4; ++pc) load_op4(Db, 64, false, wr, pc * 16, li, lk, dv[pc]);
        /* acc -= L(b, j) X_j over the earlier strips: per strip j all its
         * operands (the factor block and the solution rows X_j, written by
         * this workgroup before the barrier that ended strip j) requested
         * together, then the 16 products (operand loads inside the product
         * loop made every product wait for its own load) */
        for (int j = 0; j < b; ++j) {
            double l[4][4], xo[4][4][NC];
#pragma unroll
            for (int pc = 0; pc < 4; ++pc)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    /* L(64 b + wr + li, 64 j + pc 16 + 4 lk + q) and
                     * X(64 j + pc 16 + 4 lk + q, c 16 + li) */
                    const int kk = 64 * j + pc * 16 + q;
                    l[pc][q] = buf_ld(rl, vl, kk * ldt * 8);
#pragma unroll
                    for (int c = 0; c < NC; ++c) xo[pc][q][c] = buf_ld(rb, vb + c * 128, kk * nrp * 8);
                }
#pragma unroll
            for (int pc = 0; pc < 4; ++pc)
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int c = 0; c < NC; ++c) acc[c] = mfma_f64(-l[pc][q], xo[pc][q][c], acc[c]);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) Xs[(wr + lk + 4 * e) * NCOL + c * 16 + li] = acc[c][e];
        __syncthreads();
        v4d x[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) x[c] = zero4d();
#pragma unroll
        for (int pc = 0; pc < 4; ++pc) {
            if (pc > wave) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int pp = pc * 16 + 4 * lk + q;
#pragma unroll
                for (int c = 0; c < NC; ++c) x[c] = mfma_f64(dv[pc][q], Xs[pp * NCOL + c * 16 + li], x[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) B[(c * 16 + li) + (size_t)(64 * b + wr + lk + 4 * e) * nrp] = x[c][e];
        /* X_b is read by the later strips of this tile; Xs is refilled */
        __syncthreads();
    }
}

/* fwd_update_kernel: B(rows of tiles m > k) -= L(m,k) X_k.  Workgroup
 * (x, y): 64 rows from row (k+1)*ldt + 64x, columns 16*NC*y ..; each wave
 * 16 rows, the whole K = ldt in registers-only MFMA steps (L(m,k) is read
 * once, X_k from the cache). */
template <int NC>
__global__ __launch_bounds__(256) void fwd_update_kernel(const double *A, double *BT, int k, int NT, int ldt, int nrp,
                                                         const int *info)
{
    constexpr int NCOL = 16 * NC;
    if (*info != 0) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4, wr = wave * 16;
    const int R0 = (k + 1) * ldt + blockIdx.x * 64;
    const int m = R0 / ldt, rr = R0 - m * ldt;
    const size_t ts = (size_t)ldt * ldt;
    const double *L = A + ((size_t)k * NT + m) * ts;
    const double *X = BT + blockIdx.y * NCOL + (size_t)k * ldt * nrp;
    double *Bo = BT + blockIdx.y * NCOL + (size_t)R0 * nrp;
    v4d acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[c][e] = Bo[(c * 16 + li) + (size_t)(wr + lk + 4 * e) * nrp];
#pragma unroll 2
    for (int pc = 0; pc < ldt / 16; ++pc) {
        double l[4];
        load_op4(L, ldt, false, rr + wr, pc * 16, li, lk, l);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const size_t row = (size_t)(pc * 16 + 4 * lk + q) * nrp;
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[c] = mfma_f64(-l[q], X[(c * 16 + li) + row], acc[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) Bo[(c * 16 + li) + (size_t)(wr + lk + 4 * e) * nrp] = acc[c][e];
}

void fwd_tile(const double *A, const double *D, double *BT, int k, int NT, int ldt, int nrp, const int *info,
              hipStream_t s)
{
    /* 16-column chains: the most workgroups on this short latency chain */
    if (!fwd_tile_fits(ldt, nrp))   /* callers keep such shapes off this path */
        ::chx::fatal_hip(hipErrorInvalidValue, "fwd_tile: offsets past 32 bits", __FILE__, __LINE__);
    hipLaunchKernelGGL(fwd_tile_kernel<1>, dim3(nrp / 16), dim3(256), 0, s, A, D, BT, k, NT, ldt, nrp, info);
}

void fwd_update(const double *A, double *BT, int k, int NT, int ldt, int nrp, const int *info, hipStream_t s)
{
    const int rows = (NT - k - 1) * ldt;
    if (rows <= 0) return;
    if (nrp % 64 == 0)
        hipLaunchKernelGGL(fwd_update_kernel<4>, dim3(rows / 64, nrp / 64), dim3(256), 0, s, A, BT, k, NT, ldt, nrp,
                           info);
    else if (nrp % 32 == 0)
        hipLaunchKernelGGL(fwd_update_kernel<2>, dim3(rows / 64, nrp / 32), dim3(256), 0, s, A, BT, k, NT, ldt, nrp,
                           info);
    else
        hipLaunchKernelGGL(fwd_update_kernel<1>, dim3(rows / 64, nrp / 16), dim3(256), 0, s, A, BT, k, NT, ldt, nrp,
                           info);
}

/* ------------------------------------------------------------------ */
/* B mirror (NT x bnt tiles of ldt x wdt) <-> BT (nrhs_p x NT*ldt), with
 * nrhs_p = bnt * wdt. */
__global__ void rhs_transpose_kernel(double *Bm, double *BT, int NT, int ldt, int wdt, int bnt, int to_bt,
                                     const int *info)
{
    __shared__ double S[32][33];
    if (info && *info != 0) return;   /* failed factorisation: B keeps its values */
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

void rhs_transpose(double *Bm, double *BT, int NT, int ldt, int wdt, int bnt, int to_bt, hipStream_t s,
                   const int *info)
{
    const int rows = NT * ldt, nrp = bnt * wdt;
    dim3 grid((rows + 31) / 32, (nrp + 31) / 32);
    hipLaunchKernelGGL(rhs_transpose_kernel, grid, dim3(256), 0, s, Bm, BT, NT, ldt, wdt, bnt, to_bt, info);
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
