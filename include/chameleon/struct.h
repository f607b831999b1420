/*
 * chameleon/struct.h -- tile and descriptor records of the Chameleon C API.
 *
 * The records keep the reference's field list, order and types
 * (include/chameleon/struct.h:49-56 for CHAM_tile_t, :152-243 for
 * CHAM_desc_t, default build without CHAMELEON_KERNELS_TRACE), so a caller
 * compiled against Chameleon's own header reads every field at the same
 * offset.  The multi-process fields (data distribution, MPI tag layout)
 * are present but hold their single-process values: this library drives
 * one GPU from one process.  The tile address maps (column-major "CM" for
 * a user pointer, tile-contiguous "CCRB" for library-owned memory) are
 * kept exactly, because they decide where each element of a user's matrix
 * lives.
 */
#ifndef CHAMELEON_AMD_STRUCT_H
#define CHAMELEON_AMD_STRUCT_H

#include <stddef.h>
#include <stdint.h>
#include "chameleon/constants.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CHAMELEON_TILE_FULLRANK (1 << 0)
#define CHAMELEON_TILE_DESC     (1 << 1)
#define CHAMELEON_TILE_HMAT     (1 << 2)

#define CHAMELEON_MAX_DIMENSION 10

typedef struct chameleon_tile_s {
    void          *mat;      /* host address of the tile's first element       */
    int            rank;     /* owning process (always 0 here)                 */
    int            m, n;     /* tile extent                                    */
    int            ld;       /* leading dimension of the tile in host memory   */
    int8_t         format;   /* CHAMELEON_TILE_FULLRANK                        */
    cham_flttype_t flttype;  /* element type                                   */
} CHAM_tile_t;

struct chameleon_desc_s;
typedef struct chameleon_desc_s CHAM_desc_t;

typedef void*        (*blkaddr_fct_t)  (const CHAM_desc_t*, int, int);
typedef int          (*blkldd_fct_t)   (const CHAM_desc_t*, int);
typedef int          (*blkdim_fct_t)   (const CHAM_desc_t*, int, cham_dim_t, int);
typedef int          (*blkrankof_fct_t)(const CHAM_desc_t*, int, int);
typedef int          (*datadist_access_fct_t)(const CHAM_desc_t*, int, ...);
typedef CHAM_tile_t* (*blktile_fct_t)  (const CHAM_desc_t*, int, int);

/* Data distribution of a multi-process descriptor (reference
 * include/chameleon/struct.h:122-126); always NULL here. */
typedef struct cham_data_dist_s {
    datadist_access_fct_t get_distrib;
    int distrib_array_size;
    int distrib[CHAMELEON_MAX_DIMENSION];
} cham_data_dist_t;

/*
 * Tile matrix descriptor.  Library-allocated matrices (CCRB) hold, in this
 * order, the full mb x nb tiles (A11), the partial last tile row (A21), the
 * partial last tile column (A12) and the corner tile (A22).
 */
struct chameleon_desc_s {
    char           *name;
    blktile_fct_t   get_blktile;
    blkaddr_fct_t   get_blkaddr;
    blkldd_fct_t    get_blkldd;
    blkdim_fct_t    get_blkdim;
    blkrankof_fct_t get_rankof;
    blkrankof_fct_t get_rankof_init;     /* rank function given at creation          */
    void           *get_rankof_init_arg; /* its argument                              */

    CHAM_tile_t    *tiles;   /* lmt x lnt tile records, column-major      */
    void           *mat;     /* host matrix base (NULL if none)            */
    size_t          A21, A12, A22;
    cham_storage_t  styp;
    cham_flttype_t  dtyp;
    int mb, nb, bsiz;

    int i, j;                /* submatrix origin (elements)               */
    int m, n;                /* submatrix extent (elements)               */
    int lm, ln;              /* full matrix extent (elements)             */
    int mt, nt;              /* submatrix extent (tiles)                  */
    int lmt, lnt;            /* full matrix extent (tiles)                */

    cham_data_dist_t *data_dist;   /* NULL: one process                   */
    int dist_it, dist_jt;          /* 0: tile grid origin                 */
    int dist_mstride, dist_nstride;/* 1: unit strides                     */
    int llm, lln, llm1, lln1, llmt, llnt;

    int id;
    int occurences;
    int sync;
    int use_mat;
    int alloc_mat;
    int register_mat;
    int myrank;
    int ooc;
    int runtime_view;

    /* MPI tag layout of the reference's distributed runtime: unused, 0 */
    int64_t      mpitag;
    int64_t      mpitag_size;
    int64_t      mpitag_tile_span;
    int64_t      mpitag_tree_span;
    int          mpitag_ld;
    unsigned int mpitag_owner : 1;
    unsigned int mpitag_use   : 1;
    unsigned int mpitag_dist  : 1;

    void *schedopt;          /* device mirror owned by the runtime         */
};

#ifdef __cplusplus
}
#endif

#endif /* CHAMELEON_AMD_STRUCT_H */
