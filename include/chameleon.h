/*
 * chameleon.h -- public entry of the MI355X-native Chameleon tiled Cholesky
 * library (libchameleon_amd.so).
 *
 * The names, argument order and return codes are those of the reference
 * Chameleon API (include/chameleon.h:119-205, :279-290), so code written for
 * Chameleon's dpotrf/dpotrs/dposv path compiles and links against this
 * library unchanged.  Computation runs on one AMD Instinct MI355X (gfx950);
 * descriptors keep a device mirror of their tiles (see chameleon/struct.h,
 * field schedopt) that is synchronised with host memory by
 * CHAMELEON_Desc_Flush / CHAMELEON_Desc_Acquire, or at the end of every
 * synchronous call.
 */
#ifndef CHAMELEON_AMD_CHAMELEON_H
#define CHAMELEON_AMD_CHAMELEON_H

#include <sys/types.h>
#include "chameleon/constants.h"
#include "chameleon/struct.h"
#include "chameleon/runtime_struct.h"
#include "chameleon/flops.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- lifecycle (reference include/chameleon.h:119-125, :279-290) ---- */
int CHAMELEON_Version(int *ver_major, int *ver_minor, int *ver_micro);
int CHAMELEON_Initialized(void);
int __chameleon_init(int nworkers, int ncudas);
int __chameleon_initpar(int nworkers, int ncudas, int nthreads_per_worker);
int __chameleon_finalize(void);
int CHAMELEON_Pause(void);
int CHAMELEON_Resume(void);
int CHAMELEON_GetThreadNbr(void);
int CHAMELEON_Comm_size(void);
int CHAMELEON_Comm_rank(void);

#define CHAMELEON_Init(_nworkers_, _ncudas_) __chameleon_init((_nworkers_), (_ncudas_))
#define CHAMELEON_InitPar(_nworkers_, _ncudas_, _nthreads_per_worker_) \
    __chameleon_initpar((_nworkers_), (_ncudas_), (_nthreads_per_worker_))
#define CHAMELEON_Finalize() __chameleon_finalize()

/* ---- options (reference include/chameleon.h:193-199) ---- */
int CHAMELEON_Enable(int option);
int CHAMELEON_Disable(int option);
int CHAMELEON_Set(int param, int value);
int CHAMELEON_Get(int param, int *value);

/* ---- descriptors (reference include/chameleon.h:141-176) ---- */
ssize_t CHAMELEON_Element_Size(cham_flttype_t type);
int CHAMELEON_Desc_Create(CHAM_desc_t **desc, void *mat, cham_flttype_t dtyp,
                          int mb, int nb, int bsiz, int lm, int ln,
                          int i, int j, int m, int n, int p, int q);
int CHAMELEON_Desc_Create_User(CHAM_desc_t **desc, void *mat, cham_flttype_t dtyp,
                               int mb, int nb, int bsiz, int lm, int ln,
                               int i, int j, int m, int n, int p, int q,
                               blkaddr_fct_t get_blkaddr, blkldd_fct_t get_blkldd,
                               blkrankof_fct_t get_rankof, void *get_rankof_arg);
CHAM_desc_t *CHAMELEON_Desc_Copy(const CHAM_desc_t *descin, void *mat);
CHAM_desc_t *CHAMELEON_Desc_SubMatrix(CHAM_desc_t *descA, int i, int j, int m, int n);
int CHAMELEON_Desc_Destroy(CHAM_desc_t **desc);
int CHAMELEON_Desc_Acquire(const CHAM_desc_t *desc);
int CHAMELEON_Desc_Release(const CHAM_desc_t *desc);
int CHAMELEON_Desc_Flush(CHAM_desc_t *desc, const RUNTIME_sequence_t *sequence);
int CHAMELEON_Lap2Desc(cham_uplo_t uplo, void *A, int LDA, CHAM_desc_t *descA);
int CHAMELEON_Desc2Lap(cham_uplo_t uplo, CHAM_desc_t *descA, void *A, int LDA);

/* ---- sequences (reference include/chameleon.h:202-205) ---- */
int CHAMELEON_Sequence_Create(RUNTIME_sequence_t **sequence);
int CHAMELEON_Sequence_Destroy(RUNTIME_sequence_t *sequence);
int CHAMELEON_Sequence_Wait(RUNTIME_sequence_t *sequence);
int CHAMELEON_Sequence_Flush(RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);

#ifdef __cplusplus
}
#endif

#include "chameleon/chameleon_d.h"

#endif /* CHAMELEON_AMD_CHAMELEON_H */
