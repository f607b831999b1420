/*
 * chameleon/constants.h -- enumerations, option codes and return codes of the
 * Chameleon C API, restricted to what the tiled Cholesky path uses.
 *
 * Every numeric value below is part of the API contract: user code written
 * against Chameleon passes these integers through unchanged, so they carry
 * the same values as the reference (include/chameleon/constants.h:41-193,
 * :348-445 of Chameleon 1.4).
 */
#ifndef CHAMELEON_AMD_CONSTANTS_H
#define CHAMELEON_AMD_CONSTANTS_H

/* Storage precision of a descriptor: low 2 bits = width, bits 2-3 = kind,
 * bit 4 = mixed-precision marker (reference constants.h:41-92). */
typedef enum chameleon_arithmetic_e {
    Cham8 = 0, ChamHalf = 1, ChamSingle = 2, ChamDouble = 3
} cham_arithmetic_t;

typedef enum chameleon_ftype_e {
    ChamInt = 0, ChamReal = 1, ChamComplex = 2
} cham_ftype_t;

#define CHAM_ARITHMETIC_MASK        (0x3)
#define CHAM_FTYPE_MASK             (0x3 << 2)
#define CHAM_MIXED_MASK             (0x1 << 4)
#define cham_get_arith(t)           ((t) & CHAM_ARITHMETIC_MASK)
#define cham_get_ftype(t)           (((t) & CHAM_FTYPE_MASK) >> 2)
#define cham_get_flttype(t)         ((t) & (CHAM_FTYPE_MASK | CHAM_ARITHMETIC_MASK))
#define cham_is_mixed(t)            ((t) & CHAM_MIXED_MASK)
#define CHAMELEON_FLTTYPE(kind, w)  (((kind) << 2) | (w))

typedef enum chameleon_flttype_e {
    ChamByte               = CHAMELEON_FLTTYPE(ChamInt, Cham8),
    ChamInteger16          = CHAMELEON_FLTTYPE(ChamInt, ChamHalf),
    ChamInteger            = CHAMELEON_FLTTYPE(ChamInt, ChamSingle),
    ChamInteger32          = CHAMELEON_FLTTYPE(ChamInt, ChamSingle),
    ChamInteger64          = CHAMELEON_FLTTYPE(ChamInt, ChamDouble),
    ChamRealHalf           = CHAMELEON_FLTTYPE(ChamReal, ChamHalf),
    ChamRealFloat          = CHAMELEON_FLTTYPE(ChamReal, ChamSingle),
    ChamRealDouble         = CHAMELEON_FLTTYPE(ChamReal, ChamDouble),
    ChamComplexHalf        = CHAMELEON_FLTTYPE(ChamComplex, ChamHalf),
    ChamComplexFloat       = CHAMELEON_FLTTYPE(ChamComplex, ChamSingle),
    ChamComplexDouble      = CHAMELEON_FLTTYPE(ChamComplex, ChamDouble),
    ChamRealDoubleMixed    = ChamRealDouble | CHAM_MIXED_MASK,
    ChamComplexDoubleMixed = ChamComplexDouble | CHAM_MIXED_MASK
} cham_flttype_t;

#define ChamRealSingle    ChamRealFloat
#define ChamComplexSingle ChamComplexFloat

typedef enum chameleon_dim_e { DIM_m = 0, DIM_n = 1 } cham_dim_t;

typedef enum chameleon_storage_e {
    ChamCM = 101, ChamRM = 102, ChamCCRB = 103, ChamCRRB = 104, ChamRCRB = 105, ChamRRRB = 106
} cham_storage_t;

typedef enum chameleon_trans_e {
    ChamNoTrans = 111, ChamTrans = 112, ChamConjTrans = 113, ChamConjNoTrans = 114
} cham_trans_t;

typedef enum chameleon_uplo_e {
    ChamUpper = 121, ChamLower = 122, ChamUpperLower = 123
} cham_uplo_t;

typedef enum chameleon_diag_e { ChamNonUnit = 131, ChamUnit = 132 } cham_diag_t;

typedef enum chameleon_side_e { ChamLeft = 141, ChamRight = 142 } cham_side_t;

typedef enum chameleon_normtype_e {
    ChamOneNorm = 171, ChamTwoNorm = 173, ChamFrobeniusNorm = 174, ChamInfNorm = 175, ChamMaxNorm = 177
} cham_normtype_t;

typedef enum chameleon_mtxtype_e {
    ChamGeneral = 231, ChamSymmetric = 232, ChamHermitian = 233, ChamTriangular = 234
} cham_mtxtype_t;

/* distributions of the random matrix generators (reference
 * include/chameleon/constants.h:211-215) */
typedef enum chameleon_dist_e {
    ChamDistUniform = 201, ChamDistSymmetric = 202, ChamDistNormal = 203
} cham_dist_t;

typedef enum chameleon_access_e { ChamR = 1, ChamW = 2, ChamRW = 3 } cham_access_t;

typedef enum chameleon_translation_e { ChamInPlace = 1, ChamOutOfPlace = 2 } cham_translation_t;

typedef enum chameleon_desc_acccess_e {
    ChamDescInput = 1, ChamDescOutput = 2, ChamDescInout = 3
} cham_desc_access_t;

#define CHAMELEON_FALSE 0
#define CHAMELEON_TRUE  1

/* CHAMELEON_Enable / CHAMELEON_Disable options */
#define CHAMELEON_WARNINGS            1
#define CHAMELEON_ERRORS              2
#define CHAMELEON_AUTOTUNING          3
#define CHAMELEON_DAG                 4
#define CHAMELEON_GENERATE_TRACE      5
#define CHAMELEON_PROFILING_MODE      CHAMELEON_GENERATE_TRACE
#define CHAMELEON_GENERATE_STATS      6
#define CHAMELEON_KERNELPROFILE_MODE  CHAMELEON_GENERATE_STATS
#define CHAMELEON_PARALLEL_KERNEL     7
#define CHAMELEON_BOUND               8
#define CHAMELEON_PROGRESS            9
#define CHAMELEON_GEMM3M             10
#define CHAMELEON_GENERIC            11

/* CHAMELEON_Set / CHAMELEON_Get parameters */
#define CHAMELEON_CONCURRENCY       1
#define CHAMELEON_TILE_SIZE         2
#define CHAMELEON_INNER_BLOCK_SIZE  3
#define CHAMELEON_HOUSEHOLDER_MODE  5
#define CHAMELEON_HOUSEHOLDER_SIZE  6
#define CHAMELEON_TRANSLATION_MODE  7
#define CHAMELEON_LOOKAHEAD         8
#define CHAMELEON_RUNTIME           9
#define CHAMELEON_DAG_COLOR_MODE   10

/* Special values of the `mat` argument of CHAMELEON_Desc_Create */
#define CHAMELEON_MAT_ALLOC_GLOBAL NULL
#define CHAMELEON_MAT_ALLOC_TILE   ((void*)-1)
#define CHAMELEON_MAT_OOC          ((void*)-2)

/* Return codes */
#define CHAMELEON_SUCCESS                 0
#define CHAMELEON_ERR_NOT_INITIALIZED  -101
#define CHAMELEON_ERR_REINITIALIZED    -102
#define CHAMELEON_ERR_NOT_SUPPORTED    -103
#define CHAMELEON_ERR_ILLEGAL_VALUE    -104
#define CHAMELEON_ERR_NOT_FOUND        -105
#define CHAMELEON_ERR_OUT_OF_RESOURCES -106
#define CHAMELEON_ERR_INTERNAL_LIMIT   -107
#define CHAMELEON_ERR_UNALLOCATED      -108
#define CHAMELEON_ERR_FILESYSTEM       -109
#define CHAMELEON_ERR_UNEXPECTED       -110
#define CHAMELEON_ERR_SEQUENCE_FLUSHED -111

#endif /* CHAMELEON_AMD_CONSTANTS_H */
