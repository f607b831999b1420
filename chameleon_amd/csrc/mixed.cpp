/*
 * mixed.cpp -- mixed-precision Cholesky solve with FP64 iterative
 * refinement, and the FP64 backward-error report.
 *
 * The semantics are LAPACK dsposv's (the routine Chameleon declares as
 * CHAMELEON_zcposv / dsposv but does not build, include/chameleon/
 * chameleon_zc.h:34,42):
 *   - A is factored in reduced precision; X0 = A_lp^-1 B;
 *   - R = B - A X in FP64; column j has converged when
 *       ||R_j||_inf <= ||X_j||_inf * ||A||_inf * eps * sqrt(N);
 *   - otherwise X += A_lp^-1 R, at most ITERMAX = 30 times;
 *   - with a refinement target t > 0 (CHAMAMD_Set_Refinement_Target) the
 *     test is instead max_j ||R_j||_inf / (||A||_inf ||X_j||_inf N eps) <= t,
 *     so that a solve can be driven to the FP64 solve's backward error;
 *   - on a reduced-precision factorisation failure (ITER = -3) or no
 *     convergence (ITER = -31) the plain FP64 factorisation is used and A
 *     holds its factor on return; otherwise A is unchanged.
 * ITER >= 0 is the number of refinement steps taken.
 *
 * Reduced precision here means reduced-precision *products*: the working
 * copy of A stays in FP64 storage, its panels are factored in FP64, and the
 * trailing updates multiply FP32, BF16, split-BF16 or FP8 (E4M3, a
 * power-of-two scale per tile) operands on the matching MFMA units with
 * FP32 accumulation (kernels/mixed.hip).  By default the working copy is
 * first scaled symmetrically to a unit diagonal (S A S, S = diag(A)^-1/2),
 * and solves with its factor are scaled back.
 */
#include "internal.h"
#include "chameleon_amd.h"
#include "chameleon/chameleon_ds.h"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace chx;

namespace chx {
int potrf_async(cham_uplo_t uplo, CHAM_desc_t *A, RUNTIME_sequence_t *sequence);
int potrs_async(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, RUNTIME_sequence_t *sequence);
}

namespace {

const int ITERMAX = 30;

double norm_inf_sym(CHAM_desc_t *A)
{
    /* the mirror is in Lower orientation (possibly transposed storage of an
     * Upper matrix): the symmetric inf-norm of its lower triangle */
    Context *c = ctx();
    Mirror *M = mirror_of(A);
    const size_t n = (size_t)A->m + A->n + 2;
    double *work = (double *)workspace(n * sizeof(double));
    CHX_HIP(hipMemsetAsync(work, 0, n * sizeof(double), c->s_main));
    k::norm_partials((const double *)M->d, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, ChamLower, 0,
                     work, work + A->n, work + A->n + A->m, c->s_main);
    std::vector<double> h(A->m);
    CHX_HIP(hipMemcpyAsync(h.data(), work + A->n, A->m * sizeof(double), hipMemcpyDeviceToHost, c->s_main));
    CHX_HIP(hipStreamSynchronize(c->s_main));
    double r = 0.0;
    for (double v : h) r = nan_max(r, v);
    return r;
}

/* column maxima of X and R (transposed, nrp x rows) on the host, one
 * launch and one round trip; dwork holds 2 nrp doubles */
void colmax_pair_host(const double *XT, const double *RT, int nrp, int rows, double *dwork, std::vector<double> &xn,
                      std::vector<double> &rn)
{
    hipStream_t s = ctx()->s_main;
    k::zero_ints((int *)dwork, 4 * nrp, s);
    k::colmax2(XT, RT, nrp, nrp, rows, dwork, s);
    std::vector<double> h(2 * (size_t)nrp);
    CHX_HIP(hipMemcpyAsync(h.data(), dwork, 2 * (size_t)nrp * sizeof(double), hipMemcpyDeviceToHost, s));
    CHX_HIP(hipStreamSynchronize(s));
    xn.assign(h.begin(), h.begin() + nrp);
    rn.assign(h.begin() + nrp, h.end());
}

/* Per-tile precision of the trailing products, Chameleon's rule for
 * demoting tiles (compute/pzhered.c:176-288, runtime/starpu/codelets/
 * codelet_zgered.c:62-163): with gnorm = ||A||_F of the symmetric matrix
 * and threshold = u_high gnorm / NT (u_high: the requested accuracy,
 * 1e-12 unless set, compute/zhered.c:170-176), tile (m,n) takes the
 * lowest precision whose unit roundoff u_low still has
 * ||A_mn||_F < threshold / u_low: BF16 (u_low = 2^-8, the half-width type
 * with MFMA on CDNA4, where the reference's CUDA path uses FP16 with
 * u_low = 1e-4), FP32 (u_low = 2^-24, slamch), else FP64.  Diagonal tiles
 * use threshold_diag = threshold while u_high < 2^-24, else 2^-24 gnorm/NT
 * (pzhered.c:249).  The GEMM/SYRK of an output tile then runs in that
 * tile's precision, as the reference's dtype-dispatched gemm does
 * (runtime/starpu/codelets/codelet_gemm.c:35-120: the C tile's type). */
void tile_precision_map(const double *Aw, int NT, int ldt, double *tss, unsigned char *dmap, hipStream_t s)
{
    Context *c = ctx();
    k::tile_sumsq_lower(Aw, NT, ldt, tss, s);
    std::vector<double> ss((size_t)NT * NT, 0.0);
    CHX_HIP(hipMemcpyAsync(ss.data(), tss, ss.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    CHX_HIP(hipStreamSynchronize(s));
    double g2 = 0.0;
    for (int n = 0; n < NT; ++n)
        for (int m = n; m < NT; ++m) g2 += (m == n ? 1.0 : 2.0) * ss[m + (size_t)n * NT];
    const double gnorm = std::sqrt(g2);
    const double u_high = c->tile_accuracy, u32 = std::ldexp(1.0, -24), u16 = std::ldexp(1.0, -8);
    const double threshold = u_high * gnorm / NT;
    const double threshold_diag = u_high < u32 ? threshold : u32 * gnorm / NT;
    std::vector<unsigned char> map((size_t)NT * NT, 0);
    long long cnt[3] = {0, 0, 0};
    for (int n = 0; n < NT; ++n)
        for (int m = n; m < NT; ++m) {
            const double t = (m == n) ? threshold_diag : threshold;
            const double ln = std::sqrt(ss[m + (size_t)n * NT]);
            const unsigned char pr = ln < t / u16 ? 2 : (ln < t / u32 ? 1 : 0);
            map[m + (size_t)n * NT] = pr;
            ++cnt[pr];
        }
    for (int i = 0; i < 3; ++i) c->tile_counts[i] = cnt[i];
    CHX_HIP(hipMemcpyAsync(dmap, map.data(), map.size(), hipMemcpyHostToDevice, s));
    CHX_HIP(hipStreamSynchronize(s));
}

int check_descs(const char *fn, CHAM_desc_t *A, CHAM_desc_t *B, CHAM_desc_t *X)
{
    if (!ctx()) {
        error(fn, "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    CHAM_desc_t *d[3] = {A, B, X};
    for (int i = 0; i < 3; ++i) {
        if (desc_check(d[i]) != CHAMELEON_SUCCESS || cham_get_flttype(d[i]->dtyp) != ChamRealDouble) {
            error(fn, "invalid descriptor");
            return CHAMELEON_ERR_ILLEGAL_VALUE;
        }
        if (d[i]->mb != d[i]->nb) {
            error(fn, "only square tiles supported");
            return CHAMELEON_ERR_ILLEGAL_VALUE;
        }
    }
    if (A->m != A->n || B->m != A->m || X->m != A->m || B->n != X->n || B->mb != A->mb || X->mb != A->mb) {
        error(fn, "inconsistent dimensions");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    return CHAMELEON_SUCCESS;
}

} /* namespace */

namespace chx {

/* Backward error max_j ||B_j - A X_j||_inf / (||A||_inf ||X_j||_inf N eps)
 * on the device, A symmetric given by its `uplo` triangle. */
int backward_error(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *X, CHAM_desc_t *B, double *berr)
{
    Context *c = ctx();
    hipStream_t s = c->s_main;
    mirror_acquire(A, uplo == ChamUpper ? MIRROR_TRANS : MIRROR_NORMAL, true, s, uplo == ChamLower);
    mirror_acquire(X, MIRROR_NORMAL, false, s);
    mirror_acquire(B, MIRROR_NORMAL, false, s);
    Mirror *MA = mirror_of(A), *MX = mirror_of(X), *MB = mirror_of(B);
    const int NT = MA->mt, ldt = MA->ldt, nrp = MX->nt * MX->wdt, rows = NT * ldt;
    const double anrm = norm_inf_sym(A);
    const size_t vb = (size_t)nrp * rows;
    const size_t pv = k::residual_scratch(NT, ldt, nrp);
    double *w = (double *)workspace2((3 * vb + 2 * (size_t)nrp + pv) * sizeof(double));
    double *XT = w, *BT = w + vb, *RT = w + 2 * vb, *nw = w + 3 * vb, *part = nw + 2 * (size_t)nrp;
    k::rhs_transpose((double *)MX->d, XT, NT, ldt, MX->wdt, MX->nt, 1, s);
    k::rhs_transpose((double *)MB->d, BT, NT, ldt, MB->wdt, MB->nt, 1, s);
    k::residual((const double *)MA->d, NT, ldt, XT, BT, RT, nrp, part, s);
    std::vector<double> xn, rn;
    colmax_pair_host(XT, RT, nrp, rows, nw, xn, rn);
    const double eps = std::ldexp(1.0, -53);
    double worst = 0.0;
    for (int j = 0; j < X->n; ++j) {
        const int pj = (j / X->nb) * MX->wdt + (j % X->nb);   /* padded column index */
        const double d = anrm * xn[pj] * A->m * eps;
        worst = nan_max(worst, d > 0 ? rn[pj] / d : (rn[pj] > 0 ? INFINITY : 0.0));
    }
    *berr = worst;
    return CHAMELEON_SUCCESS;
}

int mixed_posv(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, CHAM_desc_t *X, int prec, int *iter, double *berr)
{
    Context *c = ctx();
    hipStream_t s = c->s_main;
    const int orient = uplo == ChamUpper ? MIRROR_TRANS : MIRROR_NORMAL;
    mirror_acquire(A, orient, true, s, uplo == ChamLower);
    mirror_acquire(B, MIRROR_NORMAL, false, s);
    mirror_prepare_overwrite(X, MIRROR_NORMAL, false);
    Mirror *MA = mirror_of(A), *MB = mirror_of(B), *MX = mirror_of(X);
    const int NT = MA->mt, ldt = MA->ldt, N = A->m;
    const int nrp = MB->nt * MB->wdt, rows = NT * ldt;
    const size_t ts = (size_t)ldt * ldt;
    const double eps = std::ldexp(1.0, -53);
    int it_count = 0;
    double worst = 0.0;
    bool converged = false, lp_ok = (ldt % 128 == 0);

    RUNTIME_sequence_t *seq = nullptr;
    CHAMELEON_Sequence_Create(&seq);
    SeqState *st = seq_state(seq);

    if (lp_ok) {
        /* workspace2: working factor | panels (4) | BT | XT | RT | norms | residual pieces
         * (the pieces hold the norm partials first) */
        const size_t abytes = MA->bytes;
        const size_t pbytes = (size_t)NT * ts * 4;
        const size_t vb = (size_t)nrp * rows;
        const size_t pv = std::max(k::residual_scratch(NT, ldt, nrp), k::copy_lower_norm_scratch(NT, ldt));
        /* per-tile precision: a second set of panel buffers (BF16) and the map */
        const bool adaptive = prec == CHAMAMD_PREC_ADAPTIVE;
        const int npanel = adaptive ? 8 : 4;
        const size_t mapbytes = adaptive ? (((size_t)NT * NT + 255) & ~(size_t)255) : 0;
        /* at the end, for the first solve (beside the factorisation): with
         * few right-hand sides the 128 x 128 strip inverses, else the
         * predecessor products of the MFMA strip solve */
        const size_t n128 = k::narrow_eligible(ldt, X->n) ? (size_t)NT * ldt * 128 : 2 * (size_t)NT * ldt * 64;
        char *w = (char *)workspace2(abytes + npanel * pbytes + mapbytes +
                                     (3 * vb + 2 * (size_t)nrp + pv + rows + (size_t)NT * NT + n128) * sizeof(double));
        double *Aw = (double *)w;
        unsigned char *dmap = (unsigned char *)(w + abytes + npanel * pbytes);
        double *BT = (double *)(w + abytes + npanel * pbytes + mapbytes), *XT = BT + vb, *RT = XT + vb, *nw = RT + vb,
               *part = nw + 2 * (size_t)nrp;
        double *sc = part + pv;       /* diagonal scaling of the working copy */
        double *tss = sc + rows;      /* tile norms of the per-tile precision choice */
        /* the factorisation reads and writes the lower tiles only: copy
         * those and form ||A||_inf in the same pass (row sums into RT) */
        /* the working copy becomes S A S, S = diag(A)^-1/2 (a solve with
         * it is X = S (S A S)^-1 S B): the scales from A's diagonal first,
         * then applied by the copy itself */
        const bool scaled = c->mixed_scaling != 0;
        if (scaled) k::diag_rsqrt((const double *)MA->d, NT, ldt, sc, s);
        double anrm = 0.0;
        if (k::copy_lower_norm((const double *)MA->d, Aw, NT, ldt, A->mb, N, part, RT, s, scaled ? sc : nullptr) == 0) {
            std::vector<double> h(N);
            CHX_HIP(hipMemcpyAsync(h.data(), RT, N * sizeof(double), hipMemcpyDeviceToHost, s));
            CHX_HIP(hipStreamSynchronize(s));
            for (double v : h) anrm = nan_max(anrm, v);
        } else {
            anrm = norm_inf_sym(A);
            CHX_HIP(hipMemcpyAsync(Aw, MA->d, abytes, hipMemcpyDeviceToDevice, s));
            if (scaled) k::diag_scaling(Aw, NT, ldt, sc, s);
        }
        const double cte = anrm * eps * std::sqrt((double)N);
        if (adaptive) tile_precision_map(Aw, NT, ldt, tss, dmap, s);

        k::rhs_transpose((double *)MB->d, BT, NT, ldt, MB->wdt, MB->nt, 1, s);
        CHX_HIP(hipMemcpyAsync(XT, BT, vb * sizeof(double), hipMemcpyDeviceToDevice, s));
        PotrfPlan p;
        p.A = Aw;
        p.NT = NT;
        p.ldt = ldt;
        p.nb = A->nb;
        p.nvalid_last = N - (NT - 1) * A->mb;
        p.d_info = st->d_info;
        p.lp_mode = prec;
        if (k::narrow_eligible(ldt, X->n))
            p.d128 = tss + (size_t)NT * NT;
        else
            p.epred = tss + (size_t)NT * NT;     /* predecessor products of the first solve */
        for (int i = 0; i < 4; ++i) p.P[i] = w + abytes + i * pbytes;
        if (adaptive) {
            for (int i = 0; i < 4; ++i) p.P2[i] = w + abytes + (4 + i) * pbytes;
            p.pmap = dmap;
        }
        /* X0 = A_lp^-1 B: with the fused panel the forward solve runs beside
         * the factorisation, sized for the rate of this precision's bulk
         * update (measured on MI355X, FP64-equivalent flop/s) */
        const bool overlap = potrf_solve_eligible(ldt);
        if (scaled) k::scale_bt(XT, nrp, rows, sc, s);
        if (overlap) {
            /* FP32, BF16, BF16x3, FP8, per-tile (sized for its FP64 tiles) */
            static const double rate[5] = {98e12, 377e12, 205e12, 400e12, 60e12};
            potrf_solve_run(p, XT, nrp, X->n, rate[prec]);
        } else {
            potrf_run(p);
        }
        st->pending_info = true;
        const int info = sequence_wait_internal(seq);
        if (info != 0 || *st->h_info == LP_RANGE) {
            it_count = -3;
        } else {
            SolvePlan sp = solve_plan(Aw, NT, ldt, nrp, st->d_info, X->n);
            if (!overlap) solve_bt(sp, Aw, NT, ldt, XT, nrp, st->d_info);
            if (scaled) k::scale_bt(XT, nrp, rows, sc, s);
            for (int it = 0; it <= ITERMAX; ++it) {
                k::residual((const double *)MA->d, NT, ldt, XT, BT, RT, nrp, part, s);
                std::vector<double> xn, rn;
                colmax_pair_host(XT, RT, nrp, rows, nw, xn, rn);
                bool ok = true;
                worst = 0.0;
                for (int j = 0; j < X->n; ++j) {
                    const int pj = (j / X->nb) * MX->wdt + (j % X->nb);
                    if (!(rn[pj] <= xn[pj] * cte)) ok = false;
                    const double d = anrm * xn[pj] * N * eps;
                    worst = nan_max(worst, d > 0 ? rn[pj] / d : (rn[pj] > 0 ? INFINITY : 0.0));
                }
                /* with a target, refinement goes on (converged or not by the
                 * test above) until every column's backward error is within it */
                if (c->ir_target > 0.0) ok = (worst <= c->ir_target);
                if (ok) {
                    converged = true;
                    it_count = it;
                    break;
                }
                if (it == ITERMAX) break;
                if (scaled) k::scale_bt(RT, nrp, rows, sc, s);
                solve_bt(sp, Aw, NT, ldt, RT, nrp, st->d_info);
                if (scaled) k::scale_bt(RT, nrp, rows, sc, s);
                k::axpy(XT, RT, vb, s);
            }
            /* a strip-solve wait that ran out leaves its code in the info
             * flag: the refinement result is then not trusted */
            st->pending_info = true;
            if (sequence_wait_internal(seq) != CHAMELEON_SUCCESS) converged = false;
            if (converged) {
                k::rhs_transpose((double *)MX->d, XT, NT, ldt, MX->wdt, MX->nt, 0, s);
                CHX_HIP(hipStreamSynchronize(s));
            } else {
                it_count = -(ITERMAX + 1);
            }
        }
    } else {
        it_count = -1;
    }
    int status = CHAMELEON_SUCCESS;
    if (!converged) {
        /* plain FP64 factorisation and solve; A holds its factor */
        CHAMELEON_Sequence_Destroy(seq);
        CHAMELEON_Sequence_Create(&seq);
        CHX_HIP(hipMemcpyAsync(MX->d, MB->d, MB->bytes, hipMemcpyDeviceToDevice, s));
        MX->state = MB->state;
        MX->diag_identity = false;
        potrf_async(uplo, A, seq);
        potrs_async(uplo, A, X, seq);
        status = sequence_wait_internal(seq);
        worst = -1.0;
    }
    CHAMELEON_Sequence_Destroy(seq);
    mirror_mark_written(X);
    X->sync = 1;
    if (iter) *iter = it_count;
    if (berr) *berr = worst;
    return status;
}

} /* namespace chx */

extern "C" {

static int mixed_checked(const char *fn, cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, CHAM_desc_t *X,
                         int prec, int *iter, double *berr)
{
    int rc = check_descs(fn, A, B, X);
    if (rc != CHAMELEON_SUCCESS) return rc;
    if (uplo != ChamUpper && uplo != ChamLower) {
        error(fn, "illegal value of uplo");
        return -1;
    }
    if (prec < CHAMAMD_PREC_FP32 || prec > CHAMAMD_PREC_ADAPTIVE) {
        error(fn, "illegal value of prec");
        return -5;
    }
    if (A->m == 0 || B->n == 0) {
        if (iter) *iter = 0;
        return CHAMELEON_SUCCESS;
    }
    return mixed_posv(uplo, A, B, X, prec, iter, berr);
}

int CHAMAMD_dposv_mixed_Tile(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, CHAM_desc_t *X, int prec, int *iter,
                             double *berr)
{
    int rc = mixed_checked("CHAMAMD_dposv_mixed_Tile", uplo, A, B, X, prec, iter, berr);
    if (rc < 0 || !ctx()) return rc;
    CHAMELEON_Desc_Flush(X, nullptr);
    CHAMELEON_Desc_Flush(A, nullptr);
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    return rc;
}

int CHAMAMD_dposv_mixed_Tile_Async(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, CHAM_desc_t *X, int prec,
                                   int *iter, double *berr, RUNTIME_sequence_t *sequence,
                                   RUNTIME_request_t *request)
{
    const char *fn = "CHAMAMD_dposv_mixed_Tile_Async";
    if (!ctx()) {
        error(fn, "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (!sequence || !request) {
        error(fn, "NULL sequence or request");
        return CHAMELEON_ERR_UNALLOCATED;
    }
    if (sequence->status != CHAMELEON_SUCCESS) return request_fail(sequence, request, CHAMELEON_ERR_SEQUENCE_FLUSHED);
    request->status = CHAMELEON_SUCCESS;
    int rc = mixed_checked(fn, uplo, A, B, X, prec, iter, berr);
    return rc == CHAMELEON_SUCCESS ? rc : request_fail(sequence, request, rc);
}

int CHAMAMD_Set_Tile_Accuracy(double u_high)
{
    if (!ctx()) {
        error("CHAMAMD_Set_Tile_Accuracy", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (!(u_high > 0.0)) {
        error("CHAMAMD_Set_Tile_Accuracy", "the accuracy must be > 0");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    ctx()->tile_accuracy = u_high;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Mixed_Tile_Counts(long long *counts)
{
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    if (!counts) return CHAMELEON_ERR_ILLEGAL_VALUE;
    for (int i = 0; i < 3; ++i) counts[i] = ctx()->tile_counts[i];
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Set_Mixed_Scaling(int on)
{
    if (!ctx()) {
        error("CHAMAMD_Set_Mixed_Scaling", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    ctx()->mixed_scaling = on ? 1 : 0;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Set_Refinement_Target(double berr)
{
    if (!ctx()) {
        error("CHAMAMD_Set_Refinement_Target", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (!(berr >= 0.0)) {
        error("CHAMAMD_Set_Refinement_Target", "the target must be >= 0");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    ctx()->ir_target = berr;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Backward_Error(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *X, CHAM_desc_t *B, double *berr)
{
    int rc = check_descs("CHAMAMD_Backward_Error", A, B, X);
    if (rc != CHAMELEON_SUCCESS) return rc;
    return backward_error(uplo, A, X, B, berr);
}

int CHAMELEON_dsposv_Tile(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, CHAM_desc_t *X, int *ITER)
{
    return CHAMAMD_dposv_mixed_Tile(uplo, A, B, X, CHAMAMD_PREC_FP32, ITER, nullptr);
}

int CHAMELEON_dsposv(cham_uplo_t uplo, int N, int NRHS, double *A, int LDA, double *B, int LDB, double *X, int LDX,
                     int *ITER)
{
    if (!ctx()) {
        error("CHAMELEON_dsposv", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (uplo != ChamUpper && uplo != ChamLower) { error("CHAMELEON_dsposv", "illegal value of uplo"); return -1; }
    if (N < 0) { error("CHAMELEON_dsposv", "illegal value of N"); return -2; }
    if (NRHS < 0) { error("CHAMELEON_dsposv", "illegal value of NRHS"); return -3; }
    if (LDA < (N > 1 ? N : 1)) { error("CHAMELEON_dsposv", "illegal value of LDA"); return -5; }
    if (LDB < (N > 1 ? N : 1)) { error("CHAMELEON_dsposv", "illegal value of LDB"); return -7; }
    if (LDX < (N > 1 ? N : 1)) { error("CHAMELEON_dsposv", "illegal value of LDX"); return -9; }
    if (N == 0 || NRHS == 0) {
        if (ITER) *ITER = 0;
        return CHAMELEON_SUCCESS;
    }
    const int nb = ctx()->nb;
    CHAM_desc_t *dA = nullptr, *dB = nullptr, *dX = nullptr;
    CHAMELEON_Desc_Create(&dA, A, ChamRealDouble, nb, nb, nb * nb, LDA, N, 0, 0, N, N, 1, 1);
    CHAMELEON_Desc_Create(&dB, B, ChamRealDouble, nb, nb, nb * nb, LDB, NRHS, 0, 0, N, NRHS, 1, 1);
    CHAMELEON_Desc_Create(&dX, X, ChamRealDouble, nb, nb, nb * nb, LDX, NRHS, 0, 0, N, NRHS, 1, 1);
    int rc = (dA && dB && dX) ? CHAMELEON_dsposv_Tile(uplo, dA, dB, dX, ITER) : CHAMELEON_ERR_ILLEGAL_VALUE;
    if (dA) CHAMELEON_Desc_Destroy(&dA);
    if (dB) CHAMELEON_Desc_Destroy(&dB);
    if (dX) CHAMELEON_Desc_Destroy(&dX);
    return rc;
}

} /* extern "C" */
