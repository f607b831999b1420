/*
 * cholesky.cpp -- the single-GPU scheduler of the tiled Cholesky
 * factor-and-solve path and its tile-layout entry points.
 *
 * Factorisation (Chameleon: chameleon_pzpotrf, compute/pzpotrf.c:84-193).
 * For each panel k the reference submits POTRF(k,k), TRSM(m,k) for m > k,
 * SYRK(n,n) and GEMM(m,n) for k < n <= m, and a runtime (StarPU, OpenMP)
 * orders them by data dependencies.  Here the same right-looking order is
 * issued as launches on two HIP streams of default priority:
 *
 *   critical stream : the panel of step k (POTRF of A(k,k) and TRSM of the
 *                     tiles below in one panel_factor launch) and the
 *                     lookahead columns (trailing_update on ncol columns)
 *   update stream   : the bulk of the trailing update (trailing_update on
 *                     the triangle right of the lookahead columns)
 *
 * By default the steps go in pairs (k, k+1) with two lookahead columns and
 * one bulk launch that applies both panels (potrf_run); events order the
 * streams, so the next panels are factored while the bulk update still
 * runs (the role of the task priorities 2NT-2k[-m][-n] in
 * pzpotrf.c:90-123).  CHAMELEON_LOOKAHEAD = 0 issues everything on one
 * stream.  The Upper variant runs on the transposed device mirror (U = L^T).
 *
 * Solves (chameleon_pztrsm, compute/pztrsm.c:134-201): with tiles a
 * multiple of 64, the 64 x 64 diagonal-block inverses are formed once and
 * one strip-solve launch per direction walks the 64-row strips
 * (kernels/solve.hip); other tile sizes multiply by tile inverses and
 * update the blocks below (forward) or above (backward) in one strided
 * batched GEMM launch per step.
 */
#include "internal.h"
#include "chameleon_amd.h"

#include <algorithm>
#include <cstring>

using namespace chx;

namespace chx {

void potrf_run(const PotrfPlan &p)
{
    Context *c = ctx();
    Lane *L = p.lane ? p.lane : main_lane();
    hipStream_t sm = L->sm, su = L->su;
    auto event = [L](size_t i) { return lane_event(L, i); };
    const int NT = p.NT, ldt = p.ldt;
    const size_t ts = (size_t)ldt * ldt;
    /* explicit inverses of the 64 x 64 diagonal blocks, per diagonal tile */
    const size_t dper = (size_t)((ldt + 63) / 64) * 64 * 64;
    double *dinv = p.dinv ? p.dinv : (double *)workspace((size_t)NT * dper * sizeof(double));
    auto tile = [&](int m, int n) { return p.A + ((size_t)n * NT + m) * ts; };
    /* algorithmic flops (include/chameleon/flops.h) of each launch, for
     * the kernel statistics */
    auto tsz = [&](int k) { return (double)((k == NT - 1) ? p.nvalid_last : p.nb); };
    const double nbd = p.nb;
    auto update_flops = [&](int k, int n0, int n1) {
        double f = 0.0;
        for (int n = n0; n < n1; ++n) {
            f += nbd * tsz(n) * (tsz(n) + 1.0);
            for (int m = n + 1; m < NT; ++m) f += 2.0 * tsz(m) * tsz(n) * nbd;
        }
        (void)k;
        return f;
    };
    /* paired bulk updates: the trailing matrix right of the two lookahead
     * columns takes the panels of steps k and k+1 in one pass */
    const bool pairs = c->pairs && c->lookahead > 0 && su && NT >= 4 && (p.lp_mode < 0 || p.P[3] != nullptr);
    auto pbuf = [&](int k) { return pairs ? p.P[k % 4] : p.P[k & 1]; };
    auto pbuf2 = [&](int k) { return pairs ? p.P2[k % 4] : p.P2[k & 1]; };
    auto update = [&](int k, int n0, int n1, hipStream_t s, int kp = 1) {
        n1 = std::min(n1, NT);
        if (n0 >= NT || n1 <= n0) return;
        /* with lookahead, the critical stream only updates the lookahead
         * columns; they are timed apart from the bulk update */
        const int cls = (s == sm && c->lookahead > 0 && su) ? CHAMAMD_KERNEL_COLUPD : CHAMAMD_KERNEL_UPDATE;
        timed(cls, kp * update_flops(k, n0, n1), s, [&] {
            if (p.lp_mode < 0) {
                k::trailing_update(p.A, NT, ldt, k, n0, n1, NT, p.d_info, s, kp);
            } else if (p.lp_mode == CHAMAMD_PREC_ADAPTIVE) {
                /* every tile in the precision the map gives it: three
                 * launches over the same tiles, each skipping the others' */
                k::trailing_update(p.A, NT, ldt, k, n0, n1, NT, p.d_info, s, kp, p.pmap);
                k::trailing_update_lp(p.A, pbuf(k), kp == 2 ? pbuf(k + 1) : nullptr, CHAMAMD_PREC_FP32, NT, ldt, n0,
                                      n1, NT, p.d_info, s, p.pmap);
                k::trailing_update_lp(p.A, pbuf2(k), kp == 2 ? pbuf2(k + 1) : nullptr, CHAMAMD_PREC_BF16, NT, ldt, n0,
                                      n1, NT, p.d_info, s, p.pmap);
            } else {
                k::trailing_update_lp(p.A, pbuf(k), kp == 2 ? pbuf(k + 1) : nullptr, p.lp_mode, NT, ldt, n0, n1, NT,
                                      p.d_info, s);
            }
        });
    };
    /* POTRF(k) runs next to the bulk update of step k-1 (columns k+1..).
     * While that grid fills the device, a compute unit only frees one
     * update workgroup's worth of registers at a time, which the 4-wave
     * POTRF fits and the 8-wave one does not; once the bulk grid is smaller
     * than the device, the faster 8-wave POTRF gets a compute unit of its
     * own. */
    int ncu = 0;
    CHX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    const int bpt = std::max(1, (ldt / 128) * (ldt / 128));     /* update workgroups per tile */
    auto potrf_waves = [&](int k) {
        if (k == 0 || c->lookahead <= 0) return 8;
        const long long S = NT - k - 1;                           /* columns of bulk update k-1 */
        return (S * (S + 1) / 2) * bpt >= 2LL * ncu ? 4 : 8;
    };
    const bool fused = c->panel && ldt % 64 == 0;
    /* the tiles below the diagonal of panel k in the precision of the
     * trailing products */
    auto demote = [&](int k) {
        if (p.lp_mode == CHAMAMD_PREC_FP8) {
            k::panel_demote_fp8(p.A, NT, ldt, k, k + 1, NT, pbuf(k), p.d_info, sm);
        } else if (p.lp_mode == CHAMAMD_PREC_ADAPTIVE) {
            k::panel_demote(p.A, NT, ldt, k, k + 1, NT, pbuf(k), CHAMAMD_PREC_FP32, sm);
            k::panel_demote(p.A, NT, ldt, k, k + 1, NT, pbuf2(k), CHAMAMD_PREC_BF16, sm);
        } else {
            k::panel_demote(p.A, NT, ldt, k, k + 1, NT, pbuf(k), p.lp_mode, sm);
        }
    };
    auto panel = [&](int k) {
        const int nvalid = (k == NT - 1) ? p.nvalid_last : p.nb;
        double *dk = dinv + (size_t)k * dper;
        const double nk = tsz(k);
        if (fused) {
            /* POTRF(k) and the TRSM of the tiles below in one launch */
            double f = nk * (nk * nk / 3.0 + nk / 2.0 + 1.0 / 6.0);
            for (int m = k + 1; m < NT; ++m) f += tsz(m) * nbd * nbd;
            const int ND = ldt / 64;
            int epoch = 0;
            int *flags = lane_flags(L, ND * ND, 1, &epoch);
            unsigned *ctr = lane_counters(L);
            timed(CHAMAMD_KERNEL_PANEL, f, sm, [&] {
                k::panel_factor(tile(k, k), tile(std::min(k + 1, NT - 1), k), (long long)ts, NT - k - 1, ldt, dk,
                                flags, ctr, epoch, p.d_info, p.nb * k, nvalid, sm);
            });
            if (k + 1 < NT && p.lp_mode >= 0) demote(k);
            return;
        }
        const int waves = potrf_waves(k);
        timed(CHAMAMD_KERNEL_POTRF, nk * (nk * nk / 3.0 + nk / 2.0 + 1.0 / 6.0), sm,
              [&] { k::potrf_diag(tile(k, k), ldt, dk, p.nb * k, nvalid, p.d_info, waves, sm); });
        if (k + 1 < NT) {
            double f = 0.0;
            for (int m = k + 1; m < NT; ++m) f += tsz(m) * nbd * nbd;
            timed(CHAMAMD_KERNEL_TRSM, f, sm, [&] {
                k::trsm_right_lower_trans(tile(k + 1, k), NT - k - 1, ts, ldt, tile(k, k), dk, p.d_info, sm);
            });
            if (p.lp_mode >= 0) demote(k);
        }
    };
    if (c->lookahead <= 0 || !su || NT <= 2) {
        for (int k = 0; k < NT; ++k) {
            panel(k);
            xfer_stream_columns(p.A, k + 1, sm);
            update(k, k + 1, NT, sm);
        }
        return;
    }
    if (pairs) {
        /* Steps in pairs (k, k+1), two lookahead columns.  At the start of
         * a pair, panel k is factored, column k+1 holds every update
         * before step k, and the columns from k+2 on hold them once the
         * paired bulk update of (k-2, k-1) is done.
         *   main:   column k+1 by panel k, POTRF+TRSM(k+1);
         *           [wait bulk(k-2,k-1)] columns k+2, k+3 by panels k, k+1
         *           (one two-column launch);
         *           POTRF+TRSM(k+2)
         *   update: bulk(k, k+1) on the columns from k+4 (after panel k+1)
         * Events: 2k = panel k+1 done, 2k+1 = bulk(k, k+1) done. */
        CHX_HIP(hipEventRecord(event(0), sm));
        CHX_HIP(hipStreamWaitEvent(su, event(0), 0));
        panel(0);
        int last_bulk = -1;
        for (int k = 0; k + 1 < NT; k += 2) {
            update(k, k + 1, k + 2, sm);
            panel(k + 1);
            if (k + 2 >= NT) break;
            CHX_HIP(hipEventRecord(event(2 * k), sm));
            CHX_HIP(hipStreamWaitEvent(su, event(2 * k), 0));
            update(k, k + 4, NT, su, 2);
            CHX_HIP(hipEventRecord(event(2 * k + 1), su));
            if (p.after_panels) p.after_panels(p.after_arg, k + 1);
            xfer_stream_columns(p.A, k + 2, sm);        /* tile columns 0..k+1 are final */
            if (last_bulk >= 0) CHX_HIP(hipStreamWaitEvent(sm, event(2 * last_bulk + 1), 0));
            last_bulk = k;
            update(k, k + 2, k + 4, sm, 2);
            panel(k + 2);
        }
        if (last_bulk >= 0) CHX_HIP(hipStreamWaitEvent(sm, event(2 * last_bulk + 1), 0));
        return;
    }
    /* events: P(k) = 2k (panel k done), R(k) = 2k+1 (rest update of step k
     * done) */
    /* the update stream must not start before earlier work on the main stream */
    CHX_HIP(hipEventRecord(event(0), sm));
    CHX_HIP(hipStreamWaitEvent(su, event(0), 0));
    panel(0);
    CHX_HIP(hipEventRecord(event(2 * 0), sm));
    for (int k = 0; k < NT - 1; ++k) {
        /* update stream: columns k+2.. with panel k */
        CHX_HIP(hipStreamWaitEvent(su, event(2 * k), 0));
        if (k + 2 < NT) update(k, k + 2, NT, su);
        CHX_HIP(hipEventRecord(event(2 * k + 1), su));
        /* critical stream: column k+1, then panel k+1 */
        if (p.after_panels) p.after_panels(p.after_arg, k);
        xfer_stream_columns(p.A, k + 1, sm);            /* tile columns 0..k are final */
        if (k >= 1) CHX_HIP(hipStreamWaitEvent(sm, event(2 * (k - 1) + 1), 0));
        update(k, k + 1, k + 2, sm);
        panel(k + 1);
        CHX_HIP(hipEventRecord(event(2 * (k + 1)), sm));
    }
    /* join: the last rest update ends before anything else on the main stream */
    CHX_HIP(hipStreamWaitEvent(sm, event(2 * (NT - 2) + 1), 0));
}

/* Workspace of the solves.  With ldt a multiple of 64 the strip solve runs
 * (kernels/solve.hip) with the 64 x 64 inverses of the diagonal blocks
 * (dinv).  Other tile sizes use the tile-step cooperative solve, which
 * needs the tile inverses in both orientations (TT = L_kk^-T, TN = L_kk^-1)
 * and one solved block W.  BT holds the transposed right-hand sides
 * (nrp x NT*ldt). */

SolvePlan solve_plan(double *A, int NT, int ldt, int nrp, int *d_info, int ncols)
{
    Context *c = ctx();
    hipStream_t sm = c->s_main;
    const size_t ts = (size_t)ldt * ldt;
    const size_t dper = (size_t)((ldt + 63) / 64) * 64 * 64;
    SolvePlan p;
    p.R = (ldt % 64 == 0) ? 64 : 0;
    p.ncols = ncols;
    const bool narrow = k::narrow_eligible(ldt, ncols);
    const size_t ntile_inv = p.R == 64 ? 0 : 2;
    /* the narrow solve's 128 x 128 strip inverses, or the MFMA strip
     * solve's predecessor products (both directions) */
    const size_t n128 = narrow ? (size_t)NT * ldt * 128 : 0;
    const size_t npe = (!narrow && p.R == 64) ? 2 * (size_t)NT * ldt * 64 : 0;
    const size_t need = (ntile_inv * NT * ts + (size_t)nrp * NT * ldt + (size_t)nrp * ldt + (size_t)NT * dper + n128 +
                         npe) * sizeof(double);
    double *ws = (double *)workspace(need);
    p.BT = ws;
    p.W = p.BT + (size_t)nrp * NT * ldt;
    p.dinv = p.W + (size_t)nrp * ldt;
    p.d128 = narrow ? p.dinv + (size_t)NT * dper : nullptr;
    p.ef = npe ? p.dinv + (size_t)NT * dper + n128 : nullptr;
    p.eb = npe ? p.ef + (size_t)NT * ldt * 64 : nullptr;
    p.TT = ntile_inv ? p.dinv + (size_t)NT * dper + n128 + npe : nullptr;
    p.TN = ntile_inv ? p.TT + (size_t)NT * ts : nullptr;
    k::dinv_all(A, NT, (size_t)(NT + 1) * ts, ldt, p.dinv, sm);
    if (narrow) k::dinv128(A, p.dinv, (long long)dper, 64LL * 64, 64, NT, ldt, p.d128, sm);
    if (npe) k::strip_products(A, p.dinv, (long long)dper, 64LL * 64, 64, NT, ldt, p.ef, p.eb, sm);
    if (p.R == 64) return p;
    k::set_identity(p.TT, NT, ldt, sm);
    k::trsm_right_lower_trans_batched(p.TT, NT, (long long)ts, ldt, A, (long long)(NT + 1) * ts, p.dinv,
                                      (long long)dper, d_info, sm);
    k::tile_transpose(p.TT, p.TN, NT, ldt, sm);
    return p;
}

/* Forward and backward solve of the transposed right-hand sides BT
 * (nrp x NT*ldt) in place, Lower factor in A. */
static void solve_pair(const SolvePlan &p, double *A, int NT, int ldt, double *BT, int nrp, int *d_info)
{
    hipStream_t sm = ctx()->s_main;
    if (p.R == 0) {
        k::solve_coop(A, p.TN, p.TT, BT, p.W, NT, ldt, nrp, 0, d_info, sm);
        k::solve_coop(A, p.TN, p.TT, BT, p.W, NT, ldt, nrp, 1, d_info, sm);
        return;
    }
    const int S = NT * ldt / 64;
    const int chunks = (nrp + 63) / 64;
    int epoch = 0;
    int *flags = strip_flags(4 * S, 2 * chunks, &epoch);    /* up to 4 column chunks side by side */
    const size_t dper = (size_t)((ldt + 63) / 64) * 64 * 64;
    /* 64 x 64 inverse of block b of diagonal tile t at dinv + (t*nob + b)*4096, ld 64 */
    for (int dir = 0; dir < 2; ++dir)
        k::solve_strips(A, p.dinv, (long long)dper, 64LL * 64, 64, BT, NT, ldt, nrp, dir, flags, epoch + dir * chunks,
                        panel_counters() + 2, d_info, sm, 0, p.d128, p.ncols, dir ? p.eb : p.ef);
}

/* In-place solve of transposed right-hand sides BT (nrp x NT*ldt) with the
 * Lower factor in A. */
void solve_bt(const SolvePlan &p, double *A, int NT, int ldt, double *BT, int nrp, int *d_info)
{
    solve_pair(p, A, NT, ldt, BT, nrp, d_info);
}

void solve_strips_lane(Lane *L, const double *A, const double *dinv, int NT, int ldt, double *BT, int nrp,
                       const int *d_info, const double *d128, int ncols, const double *ef, const double *eb)
{
    const int S = NT * ldt / 64, chunks = (nrp + 63) / 64;
    int epoch = 0;
    int *flags = lane_flags(L, 4 * S, 2 * chunks, &epoch);
    const size_t dper = (size_t)(ldt / 64) * 64 * 64;
    for (int dir = 0; dir < 2; ++dir)
        k::solve_strips(A, dinv, (long long)dper, 64LL * 64, 64, BT, NT, ldt, nrp, dir, flags, epoch + dir * chunks,
                        lane_counters(L) + 2, d_info, L->sm, 0, d128, ncols, dir ? eb : ef);
}

/* B <- A^-1 B with the Lower factor in A: forward then backward solve on
 * the transposed right-hand sides. */
static void solve_run(const SolvePlan &p, double *A, int NT, int ldt, double *B, int Bnt, int wdt, int *d_info,
                      double flops)
{
    Context *c = ctx();
    hipStream_t sm = c->s_main;
    const int nrp = Bnt * wdt;
    timed(CHAMAMD_KERNEL_SOLVE, flops, sm, [&] {
        k::rhs_transpose(B, p.BT, NT, ldt, wdt, Bnt, 1, sm);
        solve_pair(p, A, NT, ldt, p.BT, nrp, d_info);
        k::rhs_transpose(B, p.BT, NT, ldt, wdt, Bnt, 0, sm);
    });
}

static int check_posv_descs(const char *fn, CHAM_desc_t *A, CHAM_desc_t *B, RUNTIME_sequence_t *sequence,
                            RUNTIME_request_t *request)
{
    if (!ctx()) {
        error(fn, "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (!sequence) {
        error(fn, "NULL sequence");
        return CHAMELEON_ERR_UNALLOCATED;
    }
    if (!request) {
        error(fn, "NULL request");
        return CHAMELEON_ERR_UNALLOCATED;
    }
    if (sequence->status == CHAMELEON_SUCCESS) request->status = CHAMELEON_SUCCESS;
    else return request_fail(sequence, request, CHAMELEON_ERR_SEQUENCE_FLUSHED);
    if (desc_check(A) != CHAMELEON_SUCCESS) {
        error(fn, "invalid first descriptor");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (B && desc_check(B) != CHAMELEON_SUCCESS) {
        error(fn, "invalid second descriptor");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (A->nb != A->mb || (B && B->nb != B->mb)) {
        error(fn, "only square tiles supported");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (cham_get_flttype(A->dtyp) != ChamRealDouble || (B && cham_get_flttype(B->dtyp) != ChamRealDouble)) {
        error(fn, "descriptors must hold real double precision values");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (A->m != A->n) {
        error(fn, "matrix A must be square");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (B && (B->m != A->n || B->mb != A->mb)) {
        error(fn, "B must have as many rows (and the same tiling) as A");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    return CHAMELEON_SUCCESS;
}

int potrf_async(cham_uplo_t uplo, CHAM_desc_t *A, RUNTIME_sequence_t *sequence)
{
    Context *c = ctx();
    if (A->m == 0) return CHAMELEON_SUCCESS;
    mirror_acquire(A, uplo == ChamUpper ? MIRROR_TRANS : MIRROR_NORMAL, true, c->s_main, uplo == ChamLower);
    Mirror *M = mirror_of(A);
    if (c->stream_flush == A && uplo == ChamLower) xfer_stream_begin(A, M);
    SeqState *st = seq_state(sequence);
    PotrfPlan p;
    p.A = (double *)M->d;
    p.NT = M->mt;
    p.ldt = M->ldt;
    p.nb = A->nb;
    p.nvalid_last = A->m - (M->mt - 1) * A->mb;
    p.d_info = st->d_info;
    potrf_run(p);
    st->pending_info = true;
    mirror_mark_written(A);
    A->sync = 1;
    return CHAMELEON_SUCCESS;
}

int potrs_async(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, RUNTIME_sequence_t *sequence)
{
    Context *c = ctx();
    if (A->m == 0 || B->n == 0) return CHAMELEON_SUCCESS;
    mirror_acquire(A, uplo == ChamUpper ? MIRROR_TRANS : MIRROR_NORMAL, true, c->s_main, uplo == ChamLower);
    mirror_acquire(B, MIRROR_NORMAL, false, c->s_main);
    Mirror *MA = mirror_of(A), *MB = mirror_of(B);
    SeqState *st = seq_state(sequence);
    if (MB->ldt != MA->ldt) {
        error("CHAMELEON_dpotrs", "B tiling does not match A");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    const int nrp = MB->nt * MB->wdt;
    SolvePlan plan = solve_plan((double *)MA->d, MA->mt, MA->ldt, nrp, st->d_info, B->n);
    solve_run(plan, (double *)MA->d, MA->mt, MA->ldt, (double *)MB->d, MB->nt, MB->wdt, st->d_info,
              2.0 * (double)A->m * (double)A->m * (double)B->n);
    /* a solve wait that ran out is reported through the info flag */
    st->pending_info = true;
    mirror_mark_written(B);
    B->sync = 1;
    return CHAMELEON_SUCCESS;
}

/* dposv with the forward substitution beside the factorisation.
 * Chameleon's zposv (compute/zposv.c) submits pzpotrf and the two pztrsm
 * to one sequence, and its runtime starts each TRSM/GEMM task of the
 * forward solve as soon as the factor tiles it reads are final.  Here the
 * forward step of tile row k (fwd_tile + fwd_update, with the panel's own
 * 64 x 64 block inverses) is issued on the main stream where it would
 * otherwise wait for a bulk update, as long as those updates are long
 * enough to hide it; the last tile rows, where the factorisation is a
 * chain of short launches, are solved forward by one strip solve from
 * their first strip, then the backward strip solve runs.  Used with the
 * fused panel, lookahead and tiles a multiple of 64; otherwise dposv is
 * dpotrf followed by dpotrs. */
struct FwdSolve {
    double *A, *dinv, *BT;
    size_t dper;
    int NT, ldt, nrp, nrhs, nb;
    hipStream_t s;              /* the factorisation's critical stream */
    int next;                   /* first tile row not yet issued      */
    int last;                   /* last tile row issued beside potrf  */
    int *d_info;
};

static void fwd_after_panels(void *arg, int k)
{
    FwdSolve *f = (FwdSolve *)arg;
    hipStream_t sm = f->s;
    for (; f->next <= k && f->next <= f->last; ++f->next) {
        const int t = f->next;
        const double nb = f->nb, below = (double)(f->NT - t - 1) * nb;
        timed(CHAMAMD_KERNEL_SOLVE, (nb * nb + 2.0 * below * nb) * f->nrhs, sm, [&] {
            k::fwd_tile(f->A, f->dinv + (size_t)t * f->dper, f->BT, t, f->NT, f->ldt, f->nrp, f->d_info, sm);
            k::fwd_update(f->A, f->BT, t, f->NT, f->ldt, f->nrp, f->d_info, sm);
        });
    }
}

/* The last tile row whose forward step hides behind the bulk updates: the
 * bulk update next to tile row t has about S = NT - t trailing columns
 * and takes 2 nb^3 S (S + 1) flops; at the update rate `rate` (flop/s; FP64
 * on MI355X ~60e12) it must last 3 ms or more.  The forward steps of a pair
 * of tile rows take 0.5-1 ms of the main stream next to the bulk, and the
 * pairs before the factorisation turns into a chain of short launches need
 * that margin (measured at N = 32768, nb = 512 in FP64: a tail of 25-30
 * tile rows is best, 17 gains nothing). */
static int fwd_overlap_last(int NT, int nb, double rate)
{
    const double nb3 = (double)nb * nb * nb;
    int t = NT - 1;
    while (t >= 0) {
        const double S = NT - t;
        if (2.0 * nb3 * S * (S + 1.0) / rate >= 3e-3) break;
        --t;
    }
    return t;
}

bool potrf_solve_eligible(int ldt)
{
    Context *c = ctx();
    return c->panel && c->lookahead > 0 && ldt % 64 == 0;
}

void potrf_solve_run(PotrfPlan p, double *BT, int nrp, int nrhs, double rate)
{
    Context *c = ctx();
    Lane *L = p.lane ? p.lane : main_lane();
    hipStream_t sm = L->sm;
    const int NT = p.NT, ldt = p.ldt;
    FwdSolve f;
    f.dper = (size_t)(ldt / 64) * 64 * 64;
    /* where potrf_run puts the block inverses */
    f.dinv = p.dinv ? p.dinv : (double *)workspace((size_t)NT * f.dper * sizeof(double));
    p.dinv = f.dinv;
    f.s = sm;
    f.A = p.A;
    f.BT = BT;
    f.NT = NT; f.ldt = ldt; f.nrp = nrp; f.nrhs = nrhs; f.nb = p.nb;
    f.next = 0;
    f.last = c->fwd_tail >= 0 ? NT - 1 - c->fwd_tail : fwd_overlap_last(NT, p.nb, rate);
    if (!k::fwd_tile_fits(ldt, nrp)) f.last = -1;    /* all rows by the strip solve below */
    f.d_info = p.d_info;
    p.after_panels = fwd_after_panels;
    p.after_arg = &f;
    potrf_run(p);

    /* the rows not solved beside the factorisation, then backward */
    const int S = NT * ldt / 64, chunks = (nrp + 63) / 64, s0 = f.next * (ldt / 64);
    int epoch = 0;
    int *flags = lane_flags(L, 4 * S, 2 * chunks, &epoch);
    const double n = (double)(NT - 1) * p.nb + p.nvalid_last, rows_left = n - (double)f.next * p.nb;
    /* with few right-hand sides: the 128 x 128 strip inverses, from the
     * panels' 64 x 64 ones (all final once potrf_run has issued) */
    const bool narrow = p.d128 && k::narrow_eligible(ldt, nrhs);
    if (narrow) k::dinv128(f.A, f.dinv, (long long)f.dper, 64LL * 64, 64, NT, ldt, p.d128, sm);
    const double *d128 = narrow ? p.d128 : nullptr;
    /* otherwise the predecessor products of the MFMA strip solve */
    double *ef = (!narrow && p.epred) ? p.epred : nullptr, *eb = ef ? ef + (size_t)NT * ldt * 64 : nullptr;
    if (ef) k::strip_products(f.A, f.dinv, (long long)f.dper, 64LL * 64, 64, NT, ldt, ef, eb, sm);
    timed(CHAMAMD_KERNEL_SOLVE, (rows_left * rows_left + n * n) * nrhs, sm, [&] {
        k::solve_strips(f.A, f.dinv, (long long)f.dper, 64LL * 64, 64, f.BT, NT, ldt, nrp, 0, flags, epoch,
                        lane_counters(L) + 2, p.d_info, sm, s0, d128, nrhs, ef);
        k::solve_strips(f.A, f.dinv, (long long)f.dper, 64LL * 64, 64, f.BT, NT, ldt, nrp, 1, flags,
                        epoch + chunks, lane_counters(L) + 2, p.d_info, sm, 0, d128, nrhs, eb);
    });
}

int posv_async(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, RUNTIME_sequence_t *sequence)
{
    Context *c = ctx();
    hipStream_t sm = c->s_main;
    if (A->m == 0) return CHAMELEON_SUCCESS;
    if (B->n == 0) return potrf_async(uplo, A, sequence);
    mirror_acquire(A, uplo == ChamUpper ? MIRROR_TRANS : MIRROR_NORMAL, true, sm, uplo == ChamLower);
    mirror_acquire(B, MIRROR_NORMAL, false, sm);
    Mirror *MA = mirror_of(A), *MB = mirror_of(B);
    if (c->stream_flush == A && uplo == ChamLower) xfer_stream_begin(A, MA);
    if (!(potrf_solve_eligible(MA->ldt) && MB->ldt == MA->ldt)) {
        potrf_async(uplo, A, sequence);
        /* the solve kernels return at once when the device info flag is
         * set, so a failed factorisation costs no host round trip here */
        return potrs_async(uplo, A, B, sequence);
    }
    SeqState *st = seq_state(sequence);
    const int NT = MA->mt, ldt = MA->ldt, nrp = MB->nt * MB->wdt;
    const size_t dper = (size_t)(ldt / 64) * 64 * 64;
    /* one workspace: the block inverses of the factorisation (where
     * potrf_run puts them), the transposed right-hand sides, and with few
     * right-hand sides the 128 x 128 strip inverses of the solves */
    const size_t n128 = k::narrow_eligible(ldt, B->n) ? (size_t)NT * ldt * 128 : 0;
    const size_t npe = n128 ? 0 : 2 * (size_t)NT * ldt * 64;      /* predecessor products */
    double *ws = (double *)workspace(((size_t)NT * dper + (size_t)nrp * NT * ldt + n128 + npe) * sizeof(double));
    double *BT = ws + (size_t)NT * dper;
    k::rhs_transpose((double *)MB->d, BT, NT, ldt, MB->wdt, MB->nt, 1, sm);

    PotrfPlan p;
    p.A = (double *)MA->d;
    p.NT = NT;
    p.ldt = ldt;
    p.nb = A->nb;
    p.nvalid_last = A->m - (NT - 1) * A->mb;
    p.d_info = st->d_info;
    p.d128 = n128 ? BT + (size_t)nrp * NT * ldt : nullptr;
    p.epred = npe ? BT + (size_t)nrp * NT * ldt : nullptr;
    potrf_solve_run(p, BT, nrp, B->n, 60e12);
    st->pending_info = true;
    mirror_mark_written(A);
    A->sync = 1;
    /* a failed factorisation leaves B as it was */
    k::rhs_transpose((double *)MB->d, BT, NT, ldt, MB->wdt, MB->nt, 0, sm, st->d_info);
    mirror_mark_written(B);
    B->sync = 1;
    return CHAMELEON_SUCCESS;
}

} /* namespace chx */

extern "C" {

int CHAMAMD_Set_Forward_Tail(int tiles)
{
    if (!ctx()) {
        error("CHAMAMD_Set_Forward_Tail", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (tiles < -1) {
        error("CHAMAMD_Set_Forward_Tail", "illegal tile count");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    ctx()->fwd_tail = tiles;
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_dpotrf_Tile_Async(cham_uplo_t uplo, CHAM_desc_t *A, RUNTIME_sequence_t *sequence,
                                RUNTIME_request_t *request)
{
    int rc = check_posv_descs("CHAMELEON_dpotrf_Tile", A, nullptr, sequence, request);
    if (rc != CHAMELEON_SUCCESS) return rc;
    if (uplo != ChamUpper && uplo != ChamLower) {
        error("CHAMELEON_dpotrf_Tile", "illegal value of uplo");
        return request_fail(sequence, request, -1);
    }
    return potrf_async(uplo, A, sequence);
}

int CHAMELEON_dpotrs_Tile_Async(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, RUNTIME_sequence_t *sequence,
                                RUNTIME_request_t *request)
{
    int rc = check_posv_descs("CHAMELEON_dpotrs_Tile", A, B, sequence, request);
    if (rc != CHAMELEON_SUCCESS) return rc;
    if (uplo != ChamUpper && uplo != ChamLower) {
        error("CHAMELEON_dpotrs_Tile", "illegal value of uplo");
        return request_fail(sequence, request, -1);
    }
    rc = potrs_async(uplo, A, B, sequence);
    return rc == CHAMELEON_SUCCESS ? rc : request_fail(sequence, request, rc);
}

int CHAMELEON_dposv_Tile_Async(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, RUNTIME_sequence_t *sequence,
                               RUNTIME_request_t *request)
{
    int rc = check_posv_descs("CHAMELEON_dposv_Tile", A, B, sequence, request);
    if (rc != CHAMELEON_SUCCESS) return rc;
    if (uplo != ChamUpper && uplo != ChamLower) {
        error("CHAMELEON_dposv_Tile", "illegal value of uplo");
        return request_fail(sequence, request, -1);
    }
    rc = posv_async(uplo, A, B, sequence);
    return rc == CHAMELEON_SUCCESS ? rc : request_fail(sequence, request, rc);
}

/* stream_A: the call factors A (Lower, column-major storage): its factor
 * columns are downloaded while the factorisation runs, the flush below
 * only waits for the rest */
static int sync_wrap(int (*fn)(RUNTIME_sequence_t *, RUNTIME_request_t *, void *), void *arg, CHAM_desc_t *A,
                     CHAM_desc_t *B, const char *name, bool stream_A = false)
{
    if (!ctx()) {
        error(name, "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    RUNTIME_sequence_t *sequence = nullptr;
    RUNTIME_request_t request = RUNTIME_REQUEST_INITIALIZER;
    CHAMELEON_Sequence_Create(&sequence);
    if (stream_A && A) ctx()->stream_flush = A;
    fn(sequence, &request, arg);
    ctx()->stream_flush = nullptr;
    if (A) CHAMELEON_Desc_Flush(A, sequence);
    if (B) CHAMELEON_Desc_Flush(B, sequence);
    int status = sequence_wait_internal(sequence);
    CHAMELEON_Sequence_Destroy(sequence);
    return status;
}

struct PosvArgs { cham_uplo_t uplo; CHAM_desc_t *A, *B; };

int CHAMELEON_dposv_Tile(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B)
{
    PosvArgs a = {uplo, A, B};
    return sync_wrap([](RUNTIME_sequence_t *s, RUNTIME_request_t *r, void *p) {
        PosvArgs *x = (PosvArgs *)p;
        return CHAMELEON_dposv_Tile_Async(x->uplo, x->A, x->B, s, r);
    }, &a, A, B, "CHAMELEON_dposv_Tile", uplo == ChamLower);
}

int CHAMELEON_dpotrf_Tile(cham_uplo_t uplo, CHAM_desc_t *A)
{
    PosvArgs a = {uplo, A, nullptr};
    return sync_wrap([](RUNTIME_sequence_t *s, RUNTIME_request_t *r, void *p) {
        PosvArgs *x = (PosvArgs *)p;
        return CHAMELEON_dpotrf_Tile_Async(x->uplo, x->A, s, r);
    }, &a, A, nullptr, "CHAMELEON_dpotrf_Tile", uplo == ChamLower);
}

int CHAMELEON_dpotrs_Tile(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B)
{
    PosvArgs a = {uplo, A, B};
    return sync_wrap([](RUNTIME_sequence_t *s, RUNTIME_request_t *r, void *p) {
        PosvArgs *x = (PosvArgs *)p;
        return CHAMELEON_dpotrs_Tile_Async(x->uplo, x->A, x->B, s, r);
    }, &a, nullptr, B, "CHAMELEON_dpotrs_Tile");
}

} /* extern "C" */
