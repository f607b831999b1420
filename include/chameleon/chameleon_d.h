/*
 * chameleon/chameleon_d.h -- double-precision routines of the Cholesky
 * factor-and-solve path.  Prototypes are the "d" instantiation of the
 * reference's include/chameleon/chameleon_z.h (LAPACK layout :73-95, tile
 * layout :154-179, asynchronous tile layout :234-259).
 */
#ifndef CHAMELEON_AMD_CHAMELEON_D_H
#define CHAMELEON_AMD_CHAMELEON_D_H

#include "chameleon/constants.h"
#include "chameleon/struct.h"
#include "chameleon/runtime_struct.h"

#ifdef __cplusplus
extern "C" {
#endif

/* LAPACK layout: column-major host arrays */
int    CHAMELEON_dposv (cham_uplo_t uplo, int N, int NRHS, double *A, int LDA, double *B, int LDB);
int    CHAMELEON_dpotrf(cham_uplo_t uplo, int N, double *A, int LDA);
int    CHAMELEON_dpotrs(cham_uplo_t uplo, int N, int NRHS, double *A, int LDA, double *B, int LDB);
int    CHAMELEON_dplgsy(double bump, cham_uplo_t uplo, int N, double *A, int LDA, unsigned long long int seed);
int    CHAMELEON_dplrnt(int M, int N, double *A, int LDA, unsigned long long int seed);
double CHAMELEON_dlange(cham_normtype_t norm, int M, int N, double *A, int LDA);
double CHAMELEON_dlansy(cham_normtype_t norm, cham_uplo_t uplo, int N, double *A, int LDA);

/* Tile layout, synchronous */
int    CHAMELEON_dposv_Tile (cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B);
int    CHAMELEON_dpotrf_Tile(cham_uplo_t uplo, CHAM_desc_t *A);
int    CHAMELEON_dpotrs_Tile(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B);
int    CHAMELEON_dplgsy_Tile(double bump, cham_uplo_t uplo, CHAM_desc_t *A, unsigned long long int seed);
int    CHAMELEON_dplrnt_Tile(CHAM_desc_t *A, unsigned long long int seed);
double CHAMELEON_dlange_Tile(cham_normtype_t norm, CHAM_desc_t *A);
double CHAMELEON_dlansy_Tile(cham_normtype_t norm, cham_uplo_t uplo, CHAM_desc_t *A);
int    CHAMELEON_dlacpy_Tile(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B);

/* Tile layout, asynchronous */
int CHAMELEON_dposv_Tile_Async (cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B,
                                RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);
int CHAMELEON_dpotrf_Tile_Async(cham_uplo_t uplo, CHAM_desc_t *A,
                                RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);
int CHAMELEON_dpotrs_Tile_Async(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B,
                                RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);
int CHAMELEON_dplgsy_Tile_Async(double bump, cham_uplo_t uplo, CHAM_desc_t *A, unsigned long long int seed,
                                RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);
int CHAMELEON_dplrnt_Tile_Async(CHAM_desc_t *A, unsigned long long int seed,
                                RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);
int CHAMELEON_dlange_Tile_Async(cham_normtype_t norm, CHAM_desc_t *A, double *value,
                                RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);
int CHAMELEON_dlansy_Tile_Async(cham_normtype_t norm, cham_uplo_t uplo, CHAM_desc_t *A, double *value,
                                RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);
int CHAMELEON_dlacpy_Tile_Async(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B,
                                RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);

/* Tile batches: every tile of A is an independent matrix of the order of
 * its tile column.  dplgsy_batch: compute/zplghe_batch.c:103 (tile (m,n)
 * generated at origin (n*mt + m)*mb*nb with bigM = 0); dpotrf_batch:
 * compute/zpotrf_batch.c:89 (each tile factored in place; returns the
 * first failing tile's column where the reference only prints it). */
int CHAMELEON_dplgsy_batch_Tile(double bump, CHAM_desc_t *A, unsigned long long int seed);
int CHAMELEON_dpotrf_batch_Tile(cham_uplo_t uplo, CHAM_desc_t *A);

#ifdef __cplusplus
}
#endif

#endif
