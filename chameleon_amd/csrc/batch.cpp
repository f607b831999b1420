/*
 * batch.cpp -- many independent SPD systems on one device.
 *
 * Two front doors:
 *
 *  - Chameleon's tile batch (compute/zpotrf_batch.c:89-118,
 *    compute/zplghe_batch.c:103-139): every tile of a descriptor is an
 *    independent tempnn x tempnn matrix (tempnn = width of its tile
 *    column).  CHAMELEON_dplgsy_batch_Tile generates tile (m,n) with the
 *    element rule of CORE_dplgsy at origin (n*mt + m)*mb*nb and bigM = 0,
 *    CHAMELEON_dpotrf_batch_Tile factors every tile in place.
 *
 *  - Uneven batches (no counterpart in the reference): systems of any
 *    orders n_i with nrhs_i right-hand sides in one device-resident
 *    CHAMAMD_batch_t.  Systems are stored largest first (the order the
 *    persistent kernels take them from their queues), each padded to a
 *    multiple of 16 with identity, so one kernel shape serves all sizes.
 *
 * Factorisations use the job-queue form of the diagonal-tile POTRF kernel
 * (one workgroup per system), solves the batched strip solve of
 * kernels/batch.hip.  A system is computed by one workgroup in a fixed
 * order of operations, so its result does not depend on the rest of the
 * batch: solved alone or in any mix it is bitwise the same.
 */
#include "internal.h"
#include "chameleon_amd.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

using namespace chx;

struct CHAMAMD_batch_s {
    int count = 0;
    std::vector<int> slot;                     /* user index -> storage slot          */
    std::vector<int> user;                     /* slot -> user index                  */
    std::vector<int> n, np, nrhs;              /* per slot                            */
    std::vector<long long> aoff, boff, doff;   /* per slot, in elements               */
    size_t a_elems = 0, b_elems = 0, d_elems = 0;
    double *dA = nullptr, *dB = nullptr, *dD = nullptr;
    int *d_np = nullptr, *d_n = nullptr, *d_nrhs = nullptr, *d_info = nullptr, *d_queue = nullptr;
    long long *d_aoff = nullptr, *d_boff = nullptr, *d_doff = nullptr;
};

namespace {

int rup16(int x) { return (x + 15) & ~15; }

/* the LCG map of `n` steps: x -> A x + C (coreblas/include/coreblas/random.h) */
void lcg_pow(unsigned long long n, unsigned long long *A, unsigned long long *C)
{
    unsigned long long a = 6364136223846793005ULL, c = 1ULL, ra = 1ULL, rc = 0ULL;
    while (n) {
        if (n & 1ULL) { rc = a * rc + c; ra = a * ra; }
        c = c * (a + 1ULL);
        a = a * a;
        n >>= 1;
    }
    *A = ra;
    *C = rc;
}

template <typename T>
T *upload(const std::vector<T> &v)
{
    T *d = nullptr;
    if (v.empty()) return nullptr;
    CHX_HIP(hipMalloc(&d, v.size() * sizeof(T)));
    CHX_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

/* Generate systems described by `sys` (device-side list built here). */
void run_gen(double *base, std::vector<BatchGenSys> &sys, int sym, hipStream_t s)
{
    std::vector<long long> unit0(sys.size() + 1, 0);
    for (size_t i = 0; i < sys.size(); ++i) {
        const long long runs = (sys[i].rows + 31) / 32;
        unit0[i + 1] = unit0[i] + runs * sys[i].cols;
    }
    BatchGenSys *d_sys = upload(sys);
    long long *d_u = upload(unit0);
    k::batch_gen(base, d_sys, d_u, unit0.back(), (int)sys.size(), sym, s);
    CHX_HIP(hipStreamSynchronize(s));
    CHX_HIP(hipFree(d_sys));
    CHX_HIP(hipFree(d_u));
}

BatchGenSys gen_sys(long long off, int ld, int rows, int cols, int nvr, int nvc, long long base, int bigM,
                    unsigned long long seed, double bump)
{
    BatchGenSys g;
    g.off = off;
    g.base = base;
    g.seed = seed;
    lcg_pow((unsigned long long)bigM, &g.abig, &g.cbig);
    g.bump = bump;
    g.ld = ld;
    g.rows = rows;
    g.cols = cols;
    g.nvr = nvr;
    g.nvc = nvc;
    g.bigM = bigM;
    g.pad_one = 1;
    return g;
}

bool check_batch(const char *fn, const CHAMAMD_batch_t *b)
{
    if (!ctx()) {
        error(fn, "CHAMELEON not initialized");
        return false;
    }
    if (!b) {
        error(fn, "NULL batch");
        return false;
    }
    return true;
}

int batch_potrf(CHAMAMD_batch_t *b, hipStream_t s)
{
    k::zero_ints(b->d_queue, 1, s);
    k::potrf_jobs(b->dA, b->d_aoff, b->d_np, nullptr, b->d_n, b->dD, b->d_doff, b->count, b->d_queue, b->d_info, 0,
                  s);
    return CHAMELEON_SUCCESS;
}

int batch_potrs(CHAMAMD_batch_t *b, hipStream_t s)
{
    k::zero_ints(b->d_queue + 1, 1, s);
    k::batch_solve(b->dA, b->d_aoff, b->d_np, nullptr, b->dD, b->d_doff, b->dB, b->d_boff, b->d_nrhs, b->d_info,
                   b->count, b->d_queue + 1, 0, s);
    return CHAMELEON_SUCCESS;
}

/* first failing system (in user order) of the last factorisation, info of it */
int batch_status(CHAMAMD_batch_t *b, int *info_user)
{
    std::vector<int> h(b->count);
    CHX_HIP(hipMemcpy(h.data(), b->d_info, b->count * sizeof(int), hipMemcpyDeviceToHost));
    int status = 0;
    for (int i = 0; i < b->count; ++i) {
        const int v = h[b->slot[i]];
        if (info_user) info_user[i] = v;
        if (v != 0 && status == 0) status = v;
    }
    return status;
}

} /* namespace */

extern "C" {

int CHAMAMD_Batch_Create(CHAMAMD_batch_t **batch, int count, const int *n, const int *nrhs)
{
    if (!ctx()) {
        error("CHAMAMD_Batch_Create", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (!batch || count < 0 || (count > 0 && (!n || !nrhs))) {
        error("CHAMAMD_Batch_Create", "illegal arguments");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    for (int i = 0; i < count; ++i)
        if (n[i] < 1 || nrhs[i] < 0) {
            error("CHAMAMD_Batch_Create", "every system needs n >= 1 and nrhs >= 0");
            return CHAMELEON_ERR_ILLEGAL_VALUE;
        }
    CHAMAMD_batch_t *b = new CHAMAMD_batch_t();
    b->count = count;
    b->user.resize(count);
    std::iota(b->user.begin(), b->user.end(), 0);
    /* largest first; ties keep user order */
    std::stable_sort(b->user.begin(), b->user.end(), [&](int x, int y) { return n[x] > n[y]; });
    b->slot.resize(count);
    b->n.resize(count);
    b->np.resize(count);
    b->nrhs.resize(count);
    b->aoff.resize(count);
    b->boff.resize(count);
    b->doff.resize(count);
    for (int sl = 0; sl < count; ++sl) {
        const int u = b->user[sl];
        b->slot[u] = sl;
        b->n[sl] = n[u];
        b->np[sl] = rup16(n[u]);
        b->nrhs[sl] = nrhs[u];
        b->aoff[sl] = (long long)b->a_elems;
        b->boff[sl] = (long long)b->b_elems;
        b->doff[sl] = (long long)b->d_elems;
        b->a_elems += (size_t)b->np[sl] * b->np[sl];
        b->b_elems += (size_t)b->np[sl] * nrhs[u];
        b->d_elems += (size_t)((b->np[sl] + 63) / 64) * 4096;
    }
    if (count > 0) {
        CHX_HIP(hipMalloc(&b->dA, b->a_elems * sizeof(double)));
        CHX_HIP(hipMalloc(&b->dB, (b->b_elems ? b->b_elems : 1) * sizeof(double)));
        CHX_HIP(hipMalloc(&b->dD, b->d_elems * sizeof(double)));
        CHX_HIP(hipMemset(b->dB, 0, (b->b_elems ? b->b_elems : 1) * sizeof(double)));
        b->d_n = upload(b->n);
        b->d_np = upload(b->np);
        b->d_nrhs = upload(b->nrhs);
        b->d_aoff = upload(b->aoff);
        b->d_boff = upload(b->boff);
        b->d_doff = upload(b->doff);
        CHX_HIP(hipMalloc(&b->d_info, count * sizeof(int)));
        CHX_HIP(hipMemset(b->d_info, 0, count * sizeof(int)));
        CHX_HIP(hipMalloc(&b->d_queue, 2 * sizeof(int)));
        /* identity systems until the caller provides matrices */
        std::vector<BatchGenSys> sys;
        for (int sl = 0; sl < count; ++sl)
            sys.push_back(gen_sys(b->aoff[sl], b->np[sl], b->np[sl], b->np[sl], 0, 0, 0, 0, 0, 0.0));
        run_gen(b->dA, sys, 1, ctx()->s_main);
    }
    *batch = b;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_Destroy(CHAMAMD_batch_t **batch)
{
    if (!batch || !*batch) return CHAMELEON_ERR_UNALLOCATED;
    CHAMAMD_batch_t *b = *batch;
    if (ctx()) CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    void *ptrs[] = {b->dA, b->dB, b->dD, b->d_np, b->d_n, b->d_nrhs, b->d_info, b->d_queue, b->d_aoff, b->d_boff,
                    b->d_doff};
    for (void *p : ptrs)
        if (p) CHX_HIP(hipFree(p));
    delete b;
    *batch = nullptr;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_SetA(CHAMAMD_batch_t *b, int i, cham_uplo_t uplo, const double *A, int lda)
{
    if (!check_batch("CHAMAMD_Batch_SetA", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (i < 0 || i >= b->count) return -2;
    if (uplo != ChamUpper && uplo != ChamLower && uplo != ChamUpperLower) return -3;
    const int sl = b->slot[i], n = b->n[sl], np = b->np[sl];
    if (!A) return -4;
    if (lda < n) return -5;
    /* lower storage with identity padding; Upper input is transposed */
    std::vector<double> h((size_t)np * np, 0.0);
    for (int c = 0; c < np; ++c)
        for (int r = c; r < np; ++r) {
            double v;
            if (r >= n || c >= n) v = (r == c) ? 1.0 : 0.0;
            else v = (uplo == ChamUpper) ? A[c + (size_t)r * lda] : A[r + (size_t)c * lda];
            h[r + (size_t)c * np] = v;
        }
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    CHX_HIP(hipMemcpy(b->dA + b->aoff[sl], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_GetA(CHAMAMD_batch_t *b, int i, cham_uplo_t uplo, double *A, int lda)
{
    if (!check_batch("CHAMAMD_Batch_GetA", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (i < 0 || i >= b->count) return -2;
    if (uplo != ChamUpper && uplo != ChamLower) return -3;
    const int sl = b->slot[i], n = b->n[sl], np = b->np[sl];
    if (!A) return -4;
    if (lda < n) return -5;
    std::vector<double> h((size_t)np * np);
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    CHX_HIP(hipMemcpy(h.data(), b->dA + b->aoff[sl], h.size() * sizeof(double), hipMemcpyDeviceToHost));
    /* only the uplo triangle (diagonal included) of A is written */
    for (int c = 0; c < n; ++c)
        for (int r = c; r < n; ++r) {
            if (uplo == ChamLower) A[r + (size_t)c * lda] = h[r + (size_t)c * np];
            else A[c + (size_t)r * lda] = h[r + (size_t)c * np];
        }
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_SetB(CHAMAMD_batch_t *b, int i, const double *B, int ldb)
{
    if (!check_batch("CHAMAMD_Batch_SetB", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (i < 0 || i >= b->count) return -2;
    const int sl = b->slot[i], n = b->n[sl], np = b->np[sl], nr = b->nrhs[sl];
    if (!B && nr > 0) return -3;
    if (ldb < n) return -4;
    if (nr == 0) return CHAMELEON_SUCCESS;
    std::vector<double> h((size_t)np * nr, 0.0);
    for (int c = 0; c < nr; ++c)
        for (int r = 0; r < n; ++r) h[r + (size_t)c * np] = B[r + (size_t)c * ldb];
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    CHX_HIP(hipMemcpy(b->dB + b->boff[sl], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_GetB(CHAMAMD_batch_t *b, int i, double *B, int ldb)
{
    if (!check_batch("CHAMAMD_Batch_GetB", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (i < 0 || i >= b->count) return -2;
    const int sl = b->slot[i], n = b->n[sl], np = b->np[sl], nr = b->nrhs[sl];
    if (!B && nr > 0) return -3;
    if (ldb < n) return -4;
    if (nr == 0) return CHAMELEON_SUCCESS;
    std::vector<double> h((size_t)np * nr);
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    CHX_HIP(hipMemcpy(h.data(), b->dB + b->boff[sl], h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int c = 0; c < nr; ++c)
        for (int r = 0; r < n; ++r) B[r + (size_t)c * ldb] = h[r + (size_t)c * np];
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_dplgsy_batch(CHAMAMD_batch_t *b, unsigned long long seed)
{
    if (!check_batch("CHAMAMD_dplgsy_batch", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    std::vector<BatchGenSys> sys;
    for (int sl = 0; sl < b->count; ++sl) {
        const int n = b->n[sl], np = b->np[sl];
        sys.push_back(gen_sys(b->aoff[sl], np, np, np, n, n, 0, n, seed + (unsigned long long)b->user[sl], (double)n));
    }
    if (!sys.empty()) run_gen(b->dA, sys, 1, ctx()->s_main);
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_dplrnt_batch(CHAMAMD_batch_t *b, unsigned long long seed)
{
    if (!check_batch("CHAMAMD_dplrnt_batch", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    std::vector<BatchGenSys> sys;
    for (int sl = 0; sl < b->count; ++sl) {
        const int n = b->n[sl], np = b->np[sl], nr = b->nrhs[sl];
        if (nr == 0) continue;
        sys.push_back(gen_sys(b->boff[sl], np, np, nr, n, nr, 0, n, seed + (unsigned long long)b->user[sl], 0.0));
    }
    if (!sys.empty()) run_gen(b->dB, sys, 0, ctx()->s_main);
    return CHAMELEON_SUCCESS;
}

static int batch_async(const char *fn, CHAMAMD_batch_t *b, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request,
                       bool factor, bool solve)
{
    if (!check_batch(fn, b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (!sequence || !request) {
        error(fn, "NULL sequence or request");
        return CHAMELEON_ERR_UNALLOCATED;
    }
    if (sequence->status != CHAMELEON_SUCCESS) return request_fail(sequence, request, CHAMELEON_ERR_SEQUENCE_FLUSHED);
    request->status = CHAMELEON_SUCCESS;
    if (b->count == 0) return CHAMELEON_SUCCESS;
    hipStream_t s = ctx()->s_main;
    if (factor) batch_potrf(b, s);
    if (solve) batch_potrs(b, s);
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_dpotrf_batch_Async(CHAMAMD_batch_t *b, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request)
{
    return batch_async("CHAMAMD_dpotrf_batch_Async", b, sequence, request, true, false);
}

int CHAMAMD_dpotrs_batch_Async(CHAMAMD_batch_t *b, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request)
{
    return batch_async("CHAMAMD_dpotrs_batch_Async", b, sequence, request, false, true);
}

int CHAMAMD_dposv_batch_Async(CHAMAMD_batch_t *b, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request)
{
    return batch_async("CHAMAMD_dposv_batch_Async", b, sequence, request, true, true);
}

int CHAMAMD_Batch_Info(CHAMAMD_batch_t *b, int *info)
{
    if (!check_batch("CHAMAMD_Batch_Info", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (b->count == 0) return CHAMELEON_SUCCESS;
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    return batch_status(b, info);
}

int CHAMAMD_dposv_batch(CHAMAMD_batch_t *b, int *info)
{
    if (!check_batch("CHAMAMD_dposv_batch", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (b->count == 0) return CHAMELEON_SUCCESS;
    hipStream_t s = ctx()->s_main;
    batch_potrf(b, s);
    batch_potrs(b, s);
    CHX_HIP(hipStreamSynchronize(s));
    return batch_status(b, info);
}

int CHAMAMD_dpotrf_batch(CHAMAMD_batch_t *b, int *info)
{
    if (!check_batch("CHAMAMD_dpotrf_batch", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (b->count == 0) return CHAMELEON_SUCCESS;
    hipStream_t s = ctx()->s_main;
    batch_potrf(b, s);
    CHX_HIP(hipStreamSynchronize(s));
    return batch_status(b, info);
}

int CHAMAMD_dpotrs_batch(CHAMAMD_batch_t *b)
{
    if (!check_batch("CHAMAMD_dpotrs_batch", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (b->count == 0) return CHAMELEON_SUCCESS;
    hipStream_t s = ctx()->s_main;
    batch_potrs(b, s);
    CHX_HIP(hipStreamSynchronize(s));
    return CHAMELEON_SUCCESS;
}

/* ------------------------------------------------- Chameleon tile batch */

int CHAMELEON_dplgsy_batch_Tile(double bump, CHAM_desc_t *A, unsigned long long int seed)
{
    const char *fn = "CHAMELEON_dplgsy_batch_Tile";
    if (!ctx()) {
        error(fn, "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (desc_check(A) != CHAMELEON_SUCCESS || cham_get_flttype(A->dtyp) != ChamRealDouble) {
        error(fn, "invalid descriptor");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    hipStream_t s = ctx()->s_main;
    mirror_prepare_overwrite(A, MIRROR_NORMAL, false);
    Mirror *M = mirror_of(A);
    std::vector<BatchGenSys> sys;
    for (int n = 0; n < A->nt; ++n)
        for (int m = 0; m < A->mt; ++m) {
            const int vm = std::min(A->mb, A->m - m * A->mb), vn = std::min(A->nb, A->n - n * A->nb);
            const int v = std::min(vm, vn);
            /* zplghe_batch.c:56-57: origin (n*mt + m)*mb*nb, bigM = 0 */
            const long long base = ((long long)n * A->mt + m) * (long long)A->mb * A->nb;
            BatchGenSys g = gen_sys((long long)((size_t)n * M->mt + m) * M->tile_elems, M->ldt, M->ldt, M->wdt, v, v,
                                    base, 0, seed, bump);
            g.pad_one = 0;
            sys.push_back(g);
        }
    run_gen((double *)M->d, sys, 1, s);
    /* the mirror convention: identity on the padded diagonal of diagonal tiles */
    k::tile_pad_diag((double *)M->d, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, 0, s);
    M->diag_identity = true;
    mirror_mark_written(A);
    CHAMELEON_Desc_Flush(A, nullptr);
    CHX_HIP(hipStreamSynchronize(s));
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_dpotrf_batch_Tile(cham_uplo_t uplo, CHAM_desc_t *A)
{
    const char *fn = "CHAMELEON_dpotrf_batch_Tile";
    if (!ctx()) {
        error(fn, "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (desc_check(A) != CHAMELEON_SUCCESS || cham_get_flttype(A->dtyp) != ChamRealDouble) {
        error(fn, "invalid descriptor");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if (uplo != ChamUpper && uplo != ChamLower) {
        error(fn, "illegal value of uplo");
        return -1;
    }
    hipStream_t s = ctx()->s_main;
    mirror_acquire(A, MIRROR_NORMAL, true, s);
    Mirror *M = mirror_of(A);
    if (uplo == ChamUpper && M->ldt != M->wdt) {
        error(fn, "Upper needs square padded tiles");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    const int ntiles = M->mt * M->nt;
    std::vector<long long> off(ntiles), doff(ntiles);
    std::vector<int> np(ntiles), nvalid(ntiles);
    double *T = (double *)M->d;
    if (uplo == ChamUpper) {
        /* factor the transposed tiles: the upper triangle becomes the lower
         * one, the untouched lower triangle is carried along and restored */
        T = (double *)workspace2((size_t)ntiles * M->tile_elems * sizeof(double));
        k::tile_transpose((const double *)M->d, T, ntiles, M->ldt, s);
    }
    long long dsum = 0;
    for (int n = 0; n < A->nt; ++n)
        for (int m = 0; m < A->mt; ++m) {
            const int t = n * A->mt + m;
            const int vm = std::min(A->mb, A->m - m * A->mb), vn = std::min(A->nb, A->n - n * A->nb);
            const int v = std::min(vm, vn);
            off[t] = (long long)t * (long long)M->tile_elems;
            nvalid[t] = v;
            np[t] = std::min(std::min(M->ldt, M->wdt), rup16(v));
            doff[t] = dsum;
            dsum += (long long)((np[t] + 63) / 64) * 4096;
        }
    long long *d_off = upload(off), *d_doff = upload(doff);
    int *d_np = upload(np), *d_nv = upload(nvalid);
    int *d_info = nullptr, *d_q = nullptr;
    CHX_HIP(hipMalloc(&d_info, ntiles * sizeof(int)));
    CHX_HIP(hipMalloc(&d_q, sizeof(int)));
    double *dinv = (double *)workspace((size_t)dsum * sizeof(double));
    std::vector<int> ldv(ntiles, M->ldt);
    int *d_ld = upload(ldv);
    k::tile_pad_diag(T, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, 1, s);
    k::zero_ints(d_q, 1, s);
    k::potrf_jobs(T, d_off, d_np, d_ld, d_nv, dinv, d_doff, ntiles, d_q, d_info, 0, s);
    k::tile_pad_diag(T, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, 0, s);
    if (uplo == ChamUpper) k::tile_transpose(T, (double *)M->d, ntiles, M->ldt, s);
    std::vector<int> info(ntiles);
    CHX_HIP(hipMemcpyAsync(info.data(), d_info, ntiles * sizeof(int), hipMemcpyDeviceToHost, s));
    CHX_HIP(hipStreamSynchronize(s));
    void *ptrs[] = {d_off, d_doff, d_np, d_nv, d_info, d_q, d_ld};
    for (void *p : ptrs) CHX_HIP(hipFree(p));
    mirror_mark_written(A);
    CHAMELEON_Desc_Flush(A, nullptr);
    CHX_HIP(hipStreamSynchronize(s));
    /* the reference only prints failing tiles (zpotrf_batch.c:46-48); here
     * the first failing tile's column is returned as well */
    for (int t = 0; t < ntiles; ++t)
        if (info[t] > 0) {
            char msg[96];
            snprintf(msg, sizeof msg, "tile %d of the batch is not positive definite (info = %d)", t, info[t]);
            warning(fn, msg);
            return info[t];
        }
    return CHAMELEON_SUCCESS;
}

} /* extern "C" */
