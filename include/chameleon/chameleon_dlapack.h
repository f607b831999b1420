/*
 * chameleon/chameleon_dlapack.h -- LAPACK/LAPACKE-compatible front door of
 * the Cholesky path (reference lapack_api/include/chameleon/
 * chameleon_zlapack.h:105-117 and lapack_api/src/lapack_z{posv,potrf,potrs}.c).
 * Only column-major storage is accepted: any other matrix_layout returns -1.
 * The Fortran entry points take every argument by address and follow the
 * lower-case + underscore name mangling of gfortran.
 */
#ifndef CHAMELEON_AMD_CHAMELEON_DLAPACK_H
#define CHAMELEON_AMD_CHAMELEON_DLAPACK_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef CHAMELEON_CBLAS_ORDER_DEFINED
#define CHAMELEON_CBLAS_ORDER_DEFINED
enum { CHAMELEON_CblasRowMajor = 101, CHAMELEON_CblasColMajor = 102 };
#endif

int CHAMELEON_lapacke_dposv (int matrix_layout, char uplo, int N, int NRHS,
                             double *A, int lda, double *B, int ldb);
int CHAMELEON_lapacke_dpotrf(int matrix_layout, char uplo, int N, double *A, int lda);
int CHAMELEON_lapacke_dpotrs(int matrix_layout, char uplo, int N, int NRHS,
                             const double *A, int lda, double *B, int ldb);

void chameleon_lapack_dposv_ (const char *uplo, const int *n, const int *nrhs,
                              double *a, const int *lda, double *b, const int *ldb, int *info);
void chameleon_lapack_dpotrf_(const char *uplo, const int *n, double *a, const int *lda, int *info);
void chameleon_lapack_dpotrs_(const char *uplo, const int *n, const int *nrhs,
                              const double *a, const int *lda, double *b, const int *ldb, int *info);

#ifdef __cplusplus
}
#endif

#endif
