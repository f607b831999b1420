/*
 * chameleon/chameleon_ds.h -- mixed-precision (double/single) Cholesky solve
 * with iterative refinement.  The reference declares this routine for the
 * z/c pair only as a commented-out prototype (include/chameleon/
 * chameleon_zc.h:34 CHAMELEON_zcposv, :42 CHAMELEON_zcposv_Tile); this is
 * its d/s instantiation with LAPACK dsposv semantics:
 *   ITER >= 0 : number of refinement steps, A unchanged;
 *   ITER = -1 : refinement not used (tile size not a multiple of 128);
 *   ITER = -3 : the reduced-precision factorisation failed;
 *   ITER = -31: no convergence in 30 steps;
 *   ITER < 0  : the FP64 factorisation was used and A holds its factor.
 */
#ifndef CHAMELEON_AMD_CHAMELEON_DS_H
#define CHAMELEON_AMD_CHAMELEON_DS_H

#include "chameleon/struct.h"

#ifdef __cplusplus
extern "C" {
#endif

int CHAMELEON_dsposv(cham_uplo_t uplo, int N, int NRHS, double *A, int LDA, double *B, int LDB,
                     double *X, int LDX, int *ITER);
int CHAMELEON_dsposv_Tile(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, CHAM_desc_t *X, int *ITER);

#ifdef __cplusplus
}
#endif

#endif
