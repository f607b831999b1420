/*
 * aux.cpp -- generators (dplgsy, dplrnt), norms (dlange, dlansy) and tile
 * copy (dlacpy) on device mirrors, with their tile, asynchronous and
 * LAPACK-layout entry points (reference compute/zplgsy.c, zplrnt.c,
 * zlange.c, zlansy.c, zlacpy.c).
 */
#include "internal.h"

#include <cmath>
#include <vector>

#include "chameleon_amd.h"

using namespace chx;

namespace {

int precheck(const char *fn, CHAM_desc_t *A, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request)
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
        error(fn, "invalid descriptor");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (cham_get_flttype(A->dtyp) != ChamRealDouble) {
        error(fn, "descriptor must hold real double precision values");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    return CHAMELEON_SUCCESS;
}

/* run `body` in a fresh sequence, flush `A` back to the host, wait */
template <typename F>
int run_sync(const char *fn, CHAM_desc_t *A, F body)
{
    if (!ctx()) {
        error(fn, "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    RUNTIME_sequence_t *seq = nullptr;
    RUNTIME_request_t req = RUNTIME_REQUEST_INITIALIZER;
    CHAMELEON_Sequence_Create(&seq);
    body(seq, &req);
    if (A) CHAMELEON_Desc_Flush(A, seq);
    int status = sequence_wait_internal(seq);
    CHAMELEON_Sequence_Destroy(seq);
    return status;
}

/* LAPACK-layout wrapper: an in-place column-major descriptor over the
 * user's array (ChamInPlace translation, control/compute_z.h:471-507). */
CHAM_desc_t *lap_desc(double *A, int LDA, int M, int N)
{
    CHAM_desc_t *d = nullptr;
    const int nb = ctx()->nb;
    if (CHAMELEON_Desc_Create(&d, A, ChamRealDouble, nb, nb, nb * nb, LDA, N, 0, 0, M, N, 1, 1) != CHAMELEON_SUCCESS)
        return nullptr;
    return d;
}

double host_norm(CHAM_desc_t *A, cham_normtype_t norm, int uplo)
{
    Context *c = ctx();
    hipStream_t s = c->s_main;
    Mirror *M = mirror_of(A);
    if (!M || M->state == MIRROR_NONE) {
        mirror_acquire(A, MIRROR_NORMAL, false, s);
        M = mirror_of(A);
    } else if (M->lower_only && uplo != ChamLower) {
        mirror_complete(A, s);
    }
    const int trans = (M->state == MIRROR_TRANS);
    const size_t n = (size_t)A->m + A->n + 2;
    double *work = (double *)workspace(n * sizeof(double));
    CHX_HIP(hipMemsetAsync(work, 0, n * sizeof(double), s));
    double *colsum = work, *rowsum = work + A->n, *w = work + A->n + A->m;
    k::norm_partials((const double *)M->d, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, uplo, trans,
                     colsum, rowsum, w, s);
    std::vector<double> h(n);
    CHX_HIP(hipMemcpyAsync(h.data(), work, n * sizeof(double), hipMemcpyDeviceToHost, s));
    CHX_HIP(hipStreamSynchronize(s));
    double r = 0.0;
    switch (norm) {
    case ChamOneNorm:
        for (int j = 0; j < A->n; ++j) r = nan_max(r, h[j]);
        break;
    case ChamInfNorm:
        for (int i = 0; i < A->m; ++i) r = nan_max(r, h[A->n + i]);
        break;
    case ChamMaxNorm:
        r = h[A->n + A->m];
        break;
    case ChamFrobeniusNorm:
        r = std::sqrt(h[A->n + A->m + 1]);
        break;
    default:
        error("CHAMELEON_dlange", "illegal value of norm");
        return -1.0;
    }
    return r;
}

} /* namespace */

extern "C" {

/* ------------------------------------------------------------ dplgsy */
int CHAMELEON_dplgsy_Tile_Async(double bump, cham_uplo_t uplo, CHAM_desc_t *A, unsigned long long int seed,
                                RUNTIME_sequence_t *sequence, RUNTIME_request_t *request)
{
    int rc = precheck("CHAMELEON_dplgsy_Tile", A, sequence, request);
    if (rc != CHAMELEON_SUCCESS) return rc;
    if (A->mb != A->nb) {
        error("CHAMELEON_dplgsy_Tile", "only square tiles supported");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (A->m == 0 || A->n == 0) return CHAMELEON_SUCCESS;
    Context *c = ctx();
    const bool square = (A->m == A->n);
    if (uplo == ChamUpperLower) mirror_prepare_overwrite(A, MIRROR_NORMAL, square);
    else mirror_acquire(A, MIRROR_NORMAL, square, c->s_main);
    Mirror *M = mirror_of(A);
    k::plgsy((double *)M->d, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, bump, (int)uplo, A->m, A->i, A->j,
             seed, 0, square, c->s_main);
    mirror_mark_written(A);
    A->sync = 1;
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_dplgsy_Tile(double bump, cham_uplo_t uplo, CHAM_desc_t *A, unsigned long long int seed)
{
    return run_sync("CHAMELEON_dplgsy_Tile", A, [&](RUNTIME_sequence_t *s, RUNTIME_request_t *r) {
        return CHAMELEON_dplgsy_Tile_Async(bump, uplo, A, seed, s, r);
    });
}

int CHAMELEON_dplgsy(double bump, cham_uplo_t uplo, int N, double *A, int LDA, unsigned long long int seed)
{
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    if (N < 0) { error("CHAMELEON_dplgsy", "illegal value of N"); return -2; }
    if (LDA < (N > 1 ? N : 1)) { error("CHAMELEON_dplgsy", "illegal value of LDA"); return -4; }
    if (N == 0) return CHAMELEON_SUCCESS;
    CHAM_desc_t *d = lap_desc(A, LDA, N, N);
    if (!d) return CHAMELEON_ERR_ILLEGAL_VALUE;
    int rc = CHAMELEON_dplgsy_Tile(bump, uplo, d, seed);
    CHAMELEON_Desc_Destroy(&d);
    return rc;
}

/* ------------------------------------------------------------ dplrnt */
int CHAMELEON_dplrnt_Tile_Async(CHAM_desc_t *A, unsigned long long int seed, RUNTIME_sequence_t *sequence,
                                RUNTIME_request_t *request)
{
    int rc = precheck("CHAMELEON_dplrnt_Tile", A, sequence, request);
    if (rc != CHAMELEON_SUCCESS) return rc;
    if (A->m == 0 || A->n == 0) return CHAMELEON_SUCCESS;
    Context *c = ctx();
    mirror_prepare_overwrite(A, MIRROR_NORMAL, false);
    Mirror *M = mirror_of(A);
    k::plrnt((double *)M->d, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, A->m, A->i, A->j, seed, 0,
             c->s_main);
    mirror_mark_written(A);
    A->sync = 1;
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_dplrnt_Tile(CHAM_desc_t *A, unsigned long long int seed)
{
    return run_sync("CHAMELEON_dplrnt_Tile", A, [&](RUNTIME_sequence_t *s, RUNTIME_request_t *r) {
        return CHAMELEON_dplrnt_Tile_Async(A, seed, s, r);
    });
}

int CHAMELEON_dplrnt(int M, int N, double *A, int LDA, unsigned long long int seed)
{
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    if (M < 0) { error("CHAMELEON_dplrnt", "illegal value of M"); return -1; }
    if (N < 0) { error("CHAMELEON_dplrnt", "illegal value of N"); return -2; }
    if (LDA < (M > 1 ? M : 1)) { error("CHAMELEON_dplrnt", "illegal value of LDA"); return -4; }
    if (M == 0 || N == 0) return CHAMELEON_SUCCESS;
    CHAM_desc_t *d = lap_desc(A, LDA, M, N);
    if (!d) return CHAMELEON_ERR_ILLEGAL_VALUE;
    int rc = CHAMELEON_dplrnt_Tile(d, seed);
    CHAMELEON_Desc_Destroy(&d);
    return rc;
}

/* ------------------------------------------------------------- norms */
int CHAMELEON_dlange_Tile_Async(cham_normtype_t norm, CHAM_desc_t *A, double *value, RUNTIME_sequence_t *sequence,
                                RUNTIME_request_t *request)
{
    int rc = precheck("CHAMELEON_dlange_Tile", A, sequence, request);
    if (rc != CHAMELEON_SUCCESS) return rc;
    if (norm != ChamOneNorm && norm != ChamInfNorm && norm != ChamMaxNorm && norm != ChamFrobeniusNorm) {
        error("CHAMELEON_dlange_Tile", "illegal value of norm");
        return request_fail(sequence, request, -1);
    }
    *value = (A->m == 0 || A->n == 0) ? 0.0 : host_norm(A, norm, ChamUpperLower);
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_dlansy_Tile_Async(cham_normtype_t norm, cham_uplo_t uplo, CHAM_desc_t *A, double *value,
                                RUNTIME_sequence_t *sequence, RUNTIME_request_t *request)
{
    int rc = precheck("CHAMELEON_dlansy_Tile", A, sequence, request);
    if (rc != CHAMELEON_SUCCESS) return rc;
    if (norm != ChamOneNorm && norm != ChamInfNorm && norm != ChamMaxNorm && norm != ChamFrobeniusNorm) {
        error("CHAMELEON_dlansy_Tile", "illegal value of norm");
        return request_fail(sequence, request, -1);
    }
    if (uplo != ChamUpper && uplo != ChamLower) {
        error("CHAMELEON_dlansy_Tile", "illegal value of uplo");
        return request_fail(sequence, request, -2);
    }
    *value = (A->m == 0) ? 0.0 : host_norm(A, norm, uplo);
    return CHAMELEON_SUCCESS;
}

double CHAMELEON_dlange_Tile(cham_normtype_t norm, CHAM_desc_t *A)
{
    double v = -1.0;
    int rc = run_sync("CHAMELEON_dlange_Tile", nullptr, [&](RUNTIME_sequence_t *s, RUNTIME_request_t *r) {
        return CHAMELEON_dlange_Tile_Async(norm, A, &v, s, r);
    });
    return rc == CHAMELEON_SUCCESS ? v : (double)rc;
}

double CHAMELEON_dlansy_Tile(cham_normtype_t norm, cham_uplo_t uplo, CHAM_desc_t *A)
{
    double v = -1.0;
    int rc = run_sync("CHAMELEON_dlansy_Tile", nullptr, [&](RUNTIME_sequence_t *s, RUNTIME_request_t *r) {
        return CHAMELEON_dlansy_Tile_Async(norm, uplo, A, &v, s, r);
    });
    return rc == CHAMELEON_SUCCESS ? v : (double)rc;
}

double CHAMELEON_dlange(cham_normtype_t norm, int M, int N, double *A, int LDA)
{
    if (!ctx()) return (double)CHAMELEON_ERR_NOT_INITIALIZED;
    if (M < 0 || N < 0 || LDA < (M > 1 ? M : 1)) return (double)CHAMELEON_ERR_ILLEGAL_VALUE;
    if (M == 0 || N == 0) return 0.0;
    CHAM_desc_t *d = lap_desc(A, LDA, M, N);
    double v = CHAMELEON_dlange_Tile(norm, d);
    CHAMELEON_Desc_Destroy(&d);
    return v;
}

double CHAMELEON_dlansy(cham_normtype_t norm, cham_uplo_t uplo, int N, double *A, int LDA)
{
    if (!ctx()) return (double)CHAMELEON_ERR_NOT_INITIALIZED;
    if (N < 0 || LDA < (N > 1 ? N : 1)) return (double)CHAMELEON_ERR_ILLEGAL_VALUE;
    if (N == 0) return 0.0;
    CHAM_desc_t *d = lap_desc(A, LDA, N, N);
    double v = CHAMELEON_dlansy_Tile(norm, uplo, d);
    CHAMELEON_Desc_Destroy(&d);
    return v;
}

/* ------------------------------------------------------------- lacpy */
int CHAMELEON_dlacpy_Tile_Async(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, RUNTIME_sequence_t *sequence,
                                RUNTIME_request_t *request)
{
    int rc = precheck("CHAMELEON_dlacpy_Tile", A, sequence, request);
    if (rc != CHAMELEON_SUCCESS) return rc;
    rc = precheck("CHAMELEON_dlacpy_Tile", B, sequence, request);
    if (rc != CHAMELEON_SUCCESS) return rc;
    if (A->m != B->m || A->n != B->n || A->mb != B->mb || A->nb != B->nb) {
        error("CHAMELEON_dlacpy_Tile", "A and B must have the same size and tiling");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (A->m == 0 || A->n == 0) return CHAMELEON_SUCCESS;
    Context *c = ctx();
    mirror_acquire(A, MIRROR_NORMAL, false, c->s_main);
    if (uplo == ChamUpperLower) mirror_prepare_overwrite(B, MIRROR_NORMAL, false);
    else mirror_acquire(B, MIRROR_NORMAL, false, c->s_main);
    Mirror *MA = mirror_of(A), *MB = mirror_of(B);
    k::lacpy_tiles((const double *)MA->d, (double *)MB->d, MA->mt, MA->nt, MA->ldt, MA->wdt, A->mb, A->nb, A->m, A->n,
                   (int)uplo, c->s_main);
    MB->diag_identity = MA->diag_identity;
    mirror_mark_written(B);
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_dlacpy_Tile(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B)
{
    return run_sync("CHAMELEON_dlacpy_Tile", B, [&](RUNTIME_sequence_t *s, RUNTIME_request_t *r) {
        return CHAMELEON_dlacpy_Tile_Async(uplo, A, B, s, r);
    });
}

} /* extern "C" */
