/*
 * chameleon_amd.h -- entry points of libchameleon_amd.so that have no
 * counterpart in the reference Chameleon API.
 */
#ifndef CHAMELEON_AMD_EXTRAS_H
#define CHAMELEON_AMD_EXTRAS_H

#include "chameleon.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel classes timed while CHAMELEON_GENERATE_STATS is enabled (the
 * reference's GENERATE_STATS switches on StarPU per-kernel statistics,
 * runtime/starpu/control/runtime_profiling.c:120-150; here each launch of
 * the class is bracketed by HIP events on its own stream). */
enum {
    CHAMAMD_KERNEL_UPDATE = 0,   /* trailing SYRK+GEMM update of the factorisation */
    CHAMAMD_KERNEL_POTRF  = 1,   /* diagonal tile POTRF                             */
    CHAMAMD_KERNEL_TRSM   = 2,   /* panel TRSM                                      */
    CHAMAMD_KERNEL_SOLVE  = 3,   /* triangular-solve GEMMs (forward + backward)     */
    CHAMAMD_KERNEL_NCLASS = 4
};

/* Accumulated launches, summed device time (ms) and algorithmic flops
 * (Chameleon's flop model, include/chameleon/flops.h) of one kernel class
 * since the last reset.  Results of a sequence are folded in when it is
 * waited on.  Returns CHAMELEON_SUCCESS or CHAMELEON_ERR_ILLEGAL_VALUE. */
int CHAMAMD_Kernel_Stats(int kernel_class, long long *launches, double *total_ms, double *flops);
int CHAMAMD_Kernel_Stats_Reset(void);

/* Products used by the trailing updates of the mixed-precision solve (the
 * matrix itself stays in FP64 storage, panels are factored in FP64). */
enum {
    CHAMAMD_PREC_FP32   = 0,   /* FP32 operands, v_mfma_f32_16x16x4_f32                      */
    CHAMAMD_PREC_BF16   = 1,   /* BF16 operands, v_mfma_f32_16x16x32_bf16                    */
    CHAMAMD_PREC_BF16X3 = 2    /* operands split hi+lo in BF16, 3 products (~2^-16 accuracy) */
};

/* Mixed-precision Cholesky solve X = A^-1 B with FP64 iterative refinement
 * and FP64 fallback (LAPACK dsposv semantics, see chameleon/chameleon_ds.h).
 * `iter` receives the ITER code, `berr` the final backward error
 * max_j ||B_j - A X_j||_inf / (||A||_inf ||X_j||_inf N eps) (-1 when the
 * FP64 fallback was taken).  B is not modified. */
int CHAMAMD_dposv_mixed_Tile(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, CHAM_desc_t *X, int prec,
                             int *iter, double *berr);

/* The same without the final flush of A and X to their host storage: the
 * results stay in the device tile mirrors (read them with
 * CHAMELEON_Desc_Flush or CHAMELEON_Desc2Lap).  The refinement loop needs
 * host-side convergence decisions, so the call returns once the solve is
 * complete; errors are recorded in `sequence`/`request` like the other
 * _Tile_Async routines. */
int CHAMAMD_dposv_mixed_Tile_Async(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B, CHAM_desc_t *X, int prec,
                                   int *iter, double *berr, RUNTIME_sequence_t *sequence,
                                   RUNTIME_request_t *request);

/* Symmetric positive definite matrix with eigenvalues kappa^(-i/(N-1)),
 * i < N, generated on the device: A = Q diag(lambda) Q^T with Q the
 * product of `nrefl` random Householder reflectors (seeded).  Conditioned
 * test matrices in the role of the reference's dlatms generator
 * (compute/pzlatms.c, mode 3 spectrum of core_zlatm1.c), at O(nrefl N^2)
 * cost.  The full matrix (both triangles) is written. */
int CHAMAMD_dplgsy_cond_Tile(double kappa, int nrefl, CHAM_desc_t *A, unsigned long long int seed);

/* The same backward error for any X, A symmetric given by its `uplo`
 * triangle, computed on the device in FP64. */
int CHAMAMD_Backward_Error(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *X, CHAM_desc_t *B, double *berr);

#ifdef __cplusplus
}
#endif

#endif
