/*
 * lapack.cpp -- LAPACK-layout entry points of the Cholesky path.
 *
 * CHAMELEON_dposv / dpotrf / dpotrs (compute/zposv.c:93-172,
 * compute/zpotrf.c:81-145, compute/zpotrs.c:78-152): argument checks with
 * the reference's -i return codes, then the column-major array is wrapped
 * in place as a CM descriptor with the context tile size (the default
 * ChamInPlace translation, control/compute_z.h:471-507), the tile algorithm
 * runs on the device mirror, and the results are copied back into the
 * user's array before returning.
 *
 * CHAMELEON_lapacke_d* and the Fortran chameleon_lapack_d*_
 * (lapack_api/src/lapack_z{posv,potrf,potrs}.c) convert the layout and
 * uplo arguments and forward here.
 */
#include "internal.h"
#include "chameleon/chameleon_dlapack.h"

using namespace chx;

namespace {

CHAM_desc_t *wrap(double *A, int LDA, int M, int N)
{
    CHAM_desc_t *d = nullptr;
    const int nb = ctx()->nb;
    if (CHAMELEON_Desc_Create(&d, A, ChamRealDouble, nb, nb, nb * nb, LDA, N, 0, 0, M, N, 1, 1) != CHAMELEON_SUCCESS)
        return nullptr;
    return d;
}

int uplo_of(char c)
{
    if (c == 'U' || c == 'u') return ChamUpper;
    if (c == 'L' || c == 'l') return ChamLower;
    fprintf(stderr, "CHAMELEON ERROR: %s(): %s\n", "chameleon_blastocblas_uplo", "illegal value of BLAS uplo parameter");
    return CHAMELEON_ERR_ILLEGAL_VALUE;
}

} /* namespace */

extern "C" {

int CHAMELEON_dposv(cham_uplo_t uplo, int N, int NRHS, double *A, int LDA, double *B, int LDB)
{
    if (!ctx()) {
        error("CHAMELEON_dposv", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (uplo != ChamUpper && uplo != ChamLower) { error("CHAMELEON_dposv", "illegal value of uplo"); return -1; }
    if (N < 0) { error("CHAMELEON_dposv", "illegal value of N"); return -2; }
    if (NRHS < 0) { error("CHAMELEON_dposv", "illegal value of NRHS"); return -3; }
    if (LDA < (N > 1 ? N : 1)) { error("CHAMELEON_dposv", "illegal value of LDA"); return -5; }
    if (LDB < (N > 1 ? N : 1)) { error("CHAMELEON_dposv", "illegal value of LDB"); return -7; }
    if (N == 0 || NRHS == 0) return CHAMELEON_SUCCESS;
    CHAM_desc_t *dA = wrap(A, LDA, N, N), *dB = wrap(B, LDB, N, NRHS);
    int rc = (dA && dB) ? CHAMELEON_dposv_Tile(uplo, dA, dB) : CHAMELEON_ERR_ILLEGAL_VALUE;
    if (dA) CHAMELEON_Desc_Destroy(&dA);
    if (dB) CHAMELEON_Desc_Destroy(&dB);
    return rc;
}

int CHAMELEON_dpotrf(cham_uplo_t uplo, int N, double *A, int LDA)
{
    if (!ctx()) {
        error("CHAMELEON_dpotrf", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (uplo != ChamUpper && uplo != ChamLower) { error("CHAMELEON_dpotrf", "illegal value of uplo"); return -1; }
    if (N < 0) { error("CHAMELEON_dpotrf", "illegal value of N"); return -2; }
    if (LDA < (N > 1 ? N : 1)) { error("CHAMELEON_dpotrf", "illegal value of LDA"); return -4; }
    if (N == 0) return CHAMELEON_SUCCESS;
    CHAM_desc_t *dA = wrap(A, LDA, N, N);
    int rc = dA ? CHAMELEON_dpotrf_Tile(uplo, dA) : CHAMELEON_ERR_ILLEGAL_VALUE;
    if (dA) CHAMELEON_Desc_Destroy(&dA);
    return rc;
}

int CHAMELEON_dpotrs(cham_uplo_t uplo, int N, int NRHS, double *A, int LDA, double *B, int LDB)
{
    if (!ctx()) {
        error("CHAMELEON_dpotrs", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (uplo != ChamUpper && uplo != ChamLower) { error("CHAMELEON_dpotrs", "illegal value of uplo"); return -1; }
    if (N < 0) { error("CHAMELEON_dpotrs", "illegal value of N"); return -2; }
    if (NRHS < 0) { error("CHAMELEON_dpotrs", "illegal value of NRHS"); return -3; }
    if (LDA < (N > 1 ? N : 1)) { error("CHAMELEON_dpotrs", "illegal value of LDA"); return -5; }
    if (LDB < (N > 1 ? N : 1)) { error("CHAMELEON_dpotrs", "illegal value of LDB"); return -7; }
    if (N == 0 || NRHS == 0) return CHAMELEON_SUCCESS;
    CHAM_desc_t *dA = wrap(A, LDA, N, N), *dB = wrap(B, LDB, N, NRHS);
    int rc = (dA && dB) ? CHAMELEON_dpotrs_Tile(uplo, dA, dB) : CHAMELEON_ERR_ILLEGAL_VALUE;
    if (dA) CHAMELEON_Desc_Destroy(&dA);
    if (dB) CHAMELEON_Desc_Destroy(&dB);
    return rc;
}

int CHAMELEON_lapacke_dposv(int matrix_layout, char uplo, int N, int NRHS, double *A, int lda, double *B, int ldb)
{
    if (matrix_layout != CHAMELEON_CblasColMajor) {
        fprintf(stderr, "CHAMELEON ERROR: %s(): %s\n", "CHAMELEON_lapacke_dposv", "illegal value of matrix_layout");
        return -1;
    }
    return CHAMELEON_dposv((cham_uplo_t)uplo_of(uplo), N, NRHS, A, lda, B, ldb);
}

int CHAMELEON_lapacke_dpotrf(int matrix_layout, char uplo, int N, double *A, int lda)
{
    if (matrix_layout != CHAMELEON_CblasColMajor) {
        fprintf(stderr, "CHAMELEON ERROR: %s(): %s\n", "CHAMELEON_lapacke_dpotrf", "illegal value of matrix_layout");
        return -1;
    }
    return CHAMELEON_dpotrf((cham_uplo_t)uplo_of(uplo), N, A, lda);
}

int CHAMELEON_lapacke_dpotrs(int matrix_layout, char uplo, int N, int NRHS, const double *A, int lda, double *B,
                             int ldb)
{
    if (matrix_layout != CHAMELEON_CblasColMajor) {
        fprintf(stderr, "CHAMELEON ERROR: %s(): %s\n", "CHAMELEON_lapacke_dpotrs", "illegal value of matrix_layout");
        return -1;
    }
    return CHAMELEON_dpotrs((cham_uplo_t)uplo_of(uplo), N, NRHS, const_cast<double *>(A), lda, B, ldb);
}

void chameleon_lapack_dposv_(const char *uplo, const int *n, const int *nrhs, double *a, const int *lda, double *b,
                             const int *ldb, int *info)
{
    *info = CHAMELEON_lapacke_dposv(CHAMELEON_CblasColMajor, *uplo, *n, *nrhs, a, *lda, b, *ldb);
}

void chameleon_lapack_dpotrf_(const char *uplo, const int *n, double *a, const int *lda, int *info)
{
    *info = CHAMELEON_lapacke_dpotrf(CHAMELEON_CblasColMajor, *uplo, *n, a, *lda);
}

void chameleon_lapack_dpotrs_(const char *uplo, const int *n, const int *nrhs, const double *a, const int *lda,
                              double *b, const int *ldb, int *info)
{
    *info = CHAMELEON_lapacke_dpotrs(CHAMELEON_CblasColMajor, *uplo, *n, *nrhs, a, *lda, b, *ldb);
}

} /* extern "C" */
