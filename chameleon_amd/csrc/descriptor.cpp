/*
 * descriptor.cpp -- tile matrix descriptors.
 *
 * Host address maps are those of the reference so that a user's matrix is
 * read and written where Chameleon would read and write it:
 *   CM   (user pointer): tile (m,n) at mat + lm*nb*n + mb*m, ld = lm
 *        (control/descriptor.h:238-252, descriptor_helpers.c:497-503);
 *   CCRB (library-allocated): full tiles contiguous, mb*nb each, tile
 *        column-major; the partial last tile row, partial last tile column
 *        and corner tile stored after the full block
 *        (control/descriptor_helpers.c:450-477, :573-577,
 *        control/descriptor.c:373-375).
 * Tile extents follow chameleon_getblkdim (include/chameleon/
 * descriptor_helpers.h:125-185).
 */
#include "internal.h"

#include <cstring>
#include <new>

using namespace chx;

extern "C" {

ssize_t CHAMELEON_Element_Size(cham_flttype_t type)
{
    switch (cham_get_flttype(type)) {
    case ChamByte: return 1;
    case ChamInteger16: return 2;
    case ChamInteger32: return 4;
    case ChamInteger64: return 8;
    case ChamRealHalf: return 2;
    case ChamRealFloat: return 4;
    case ChamRealDouble: return 8;
    case ChamComplexHalf: return 4;
    case ChamComplexFloat: return 8;
    case ChamComplexDouble: return 16;
    default:
        error("CHAMELEON_Element_Size", "undefined type");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
}

static void *getaddr_ccrb(const CHAM_desc_t *A, int m, int n)
{
    const size_t mm = (size_t)m + A->i / A->mb;
    const size_t nn = (size_t)n + A->j / A->nb;
    const size_t es = (size_t)CHAMELEON_Element_Size(A->dtyp);
    size_t off;
    if (mm < (size_t)A->llm1) {
        if (nn < (size_t)A->lln1) off = (size_t)A->bsiz * (mm + (size_t)A->llm1 * nn);
        else off = A->A12 + (size_t)A->mb * (A->lln % A->nb) * mm;
    } else {
        if (nn < (size_t)A->lln1) off = A->A21 + (size_t)(A->llm % A->mb) * A->nb * nn;
        else off = A->A22;
    }
    return (char *)A->mat + off * es;
}

static int getblkldd_ccrb(const CHAM_desc_t *A, int m)
{
    const int mm = m + A->i / A->mb;
    return ((mm + 1) == A->lmt && (A->lm % A->mb) != 0) ? A->lm % A->mb : A->mb;
}

static void *getaddr_cm(const CHAM_desc_t *A, int m, int n)
{
    const size_t mm = (size_t)m + A->i / A->mb;
    const size_t nn = (size_t)n + A->j / A->nb;
    const size_t es = (size_t)CHAMELEON_Element_Size(A->dtyp);
    const size_t off = (size_t)A->llm * A->nb * nn + (size_t)A->mb * mm;
    return (char *)A->mat + off * es;
}

static int getblkldd_cm(const CHAM_desc_t *A, int m)
{
    (void)m;
    return A->llm;
}

static int getblkdim(const CHAM_desc_t *A, int k, cham_dim_t dim, int len)
{
    const int b = (dim == DIM_m) ? A->mb : A->nb;
    return ((k + 1) * b > len) ? len - k * b : b;
}

static int getrankof(const CHAM_desc_t *A, int m, int n)
{
    (void)A; (void)m; (void)n;
    return 0;
}

static CHAM_tile_t *gettile(const CHAM_desc_t *A, int m, int n)
{
    const size_t mm = (size_t)m + A->i / A->mb;
    const size_t nn = (size_t)n + A->j / A->nb;
    return A->tiles + mm + nn * (size_t)A->lmt;
}

static int ceil_div(int a, int b) { return (a + b - 1) / b; }

static int desc_init(CHAM_desc_t *d, void *mat, cham_flttype_t dtyp, int mb, int nb, int bsiz, int lm, int ln,
                     int i, int j, int m, int n, blkaddr_fct_t get_blkaddr, blkldd_fct_t get_blkldd,
                     blkrankof_fct_t get_rankof, void *get_rankof_arg)
{
    static int nbdesc = 0;
    memset(d, 0, sizeof(*d));
    d->get_blktile = gettile;
    d->get_blkdim = getblkdim;
    d->get_rankof = get_rankof ? get_rankof : getrankof;
    d->get_rankof_init = d->get_rankof;
    d->get_rankof_init_arg = get_rankof_arg;
    d->data_dist = nullptr;
    d->dist_mstride = d->dist_nstride = 1;
    d->dtyp = dtyp;
    d->mb = mb; d->nb = nb; d->bsiz = bsiz;
    d->i = i; d->j = j; d->m = m; d->n = n;
    d->lm = lm; d->ln = ln;
    d->mt = ceil_div(m, mb); d->nt = ceil_div(n, nb);
    d->lmt = ceil_div(lm, mb); d->lnt = ceil_div(ln, nb);
    d->llm = lm; d->lln = ln;
    d->llmt = d->lmt; d->llnt = d->lnt;
    d->llm1 = lm / mb; d->lln1 = ln / nb;
    d->A21 = (size_t)(d->llm - d->llm % mb) * (size_t)(d->lln - d->lln % nb);
    d->A12 = (size_t)(d->llm % mb) * (size_t)(d->lln - d->lln % nb) + d->A21;
    d->A22 = (size_t)(d->llm - d->llm % mb) * (size_t)(d->lln % nb) + d->A12;
    d->id = nbdesc++;
    d->occurences = 1;
    d->myrank = 0;

    if (mat == CHAMELEON_MAT_OOC) {
        error("CHAMELEON_Desc_Create", "out-of-core descriptors are not supported");
        return CHAMELEON_ERR_NOT_SUPPORTED;
    }
    if (mat == CHAMELEON_MAT_ALLOC_GLOBAL || mat == CHAMELEON_MAT_ALLOC_TILE) {
        const size_t bytes = (size_t)lm * ln * (size_t)CHAMELEON_Element_Size(dtyp);
        d->mat = malloc(bytes > 0 ? bytes : 1);
        if (!d->mat) {
            error("CHAMELEON_Desc_Create", "malloc() failed");
            return CHAMELEON_ERR_OUT_OF_RESOURCES;
        }
        d->alloc_mat = 1;
        d->use_mat = 1;
        d->get_blkaddr = get_blkaddr ? get_blkaddr : getaddr_ccrb;
        d->get_blkldd = get_blkldd ? get_blkldd : getblkldd_ccrb;
        d->styp = ChamCCRB;
    } else {
        d->mat = mat;
        d->use_mat = 1;
        d->get_blkaddr = get_blkaddr ? get_blkaddr : getaddr_cm;
        d->get_blkldd = get_blkldd ? get_blkldd : getblkldd_cm;
        d->styp = (d->get_blkaddr == getaddr_cm) ? ChamCM : ChamCCRB;
    }
    /* tile records over the full matrix */
    d->tiles = (CHAM_tile_t *)malloc(sizeof(CHAM_tile_t) * (size_t)(d->lmt > 0 ? d->lmt : 1) * (d->lnt > 0 ? d->lnt : 1));
    const int si = d->i, sj = d->j;
    d->i = 0; d->j = 0;
    for (int jj = 0; jj < d->lnt; ++jj)
        for (int ii = 0; ii < d->lmt; ++ii) {
            CHAM_tile_t *t = d->tiles + ii + (size_t)jj * d->lmt;
            t->format = CHAMELEON_TILE_FULLRANK;
            t->flttype = (cham_flttype_t)cham_get_flttype(dtyp);
            t->rank = 0;
            t->m = (ii == d->lmt - 1) ? lm - ii * mb : mb;
            t->n = (jj == d->lnt - 1) ? ln - jj * nb : nb;
            t->mat = d->get_blkaddr(d, ii, jj);
            t->ld = d->get_blkldd(d, ii);
        }
    d->i = si; d->j = sj;
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Desc_Create_User(CHAM_desc_t **descptr, void *mat, cham_flttype_t dtyp, int mb, int nb, int bsiz,
                               int lm, int ln, int i, int j, int m, int n, int p, int q,
                               blkaddr_fct_t get_blkaddr, blkldd_fct_t get_blkldd, blkrankof_fct_t get_rankof,
                               void *get_rankof_arg)
{
    if (!ctx()) {
        error("CHAMELEON_Desc_Create", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (!descptr) {
        error("CHAMELEON_Desc_Create", "NULL descriptor");
        return CHAMELEON_ERR_UNALLOCATED;
    }
    if (p * q != 1) {
        error("CHAMELEON_Desc_Create", "only a 1x1 process grid is supported (one GPU, one process)");
        return CHAMELEON_ERR_NOT_SUPPORTED;
    }
    CHAM_desc_t *d = (CHAM_desc_t *)calloc(1, sizeof(CHAM_desc_t));
    if (!d) return CHAMELEON_ERR_OUT_OF_RESOURCES;
    int rc = desc_init(d, mat, dtyp, mb, nb, bsiz, lm, ln, i, j, m, n, get_blkaddr, get_blkldd, get_rankof,
                       get_rankof_arg);
    if (rc == CHAMELEON_SUCCESS) rc = desc_check(d);
    if (rc != CHAMELEON_SUCCESS) {
        if (d->alloc_mat) free(d->mat);
        free(d->tiles);
        free(d);
        *descptr = nullptr;
        return rc;
    }
    *descptr = d;
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Desc_Create(CHAM_desc_t **desc, void *mat, cham_flttype_t dtyp, int mb, int nb, int bsiz, int lm,
                          int ln, int i, int j, int m, int n, int p, int q)
{
    return CHAMELEON_Desc_Create_User(desc, mat, dtyp, mb, nb, bsiz, lm, ln, i, j, m, n, p, q, nullptr, nullptr,
                                      nullptr, nullptr);
}

CHAM_desc_t *CHAMELEON_Desc_Copy(const CHAM_desc_t *in, void *mat)
{
    CHAM_desc_t *d = nullptr;
    CHAMELEON_Desc_Create(&d, mat, in->dtyp, in->mb, in->nb, in->bsiz, in->lm, in->ln, in->i, in->j, in->m, in->n,
                          1, 1);
    return d;
}

CHAM_desc_t *CHAMELEON_Desc_SubMatrix(CHAM_desc_t *A, int i, int j, int m, int n)
{
    if ((A->i + i) % A->mb != 0 || (A->j + j) % A->nb != 0) {
        error("CHAMELEON_Desc_SubMatrix", "submatrix must start on a tile boundary");
        return nullptr;
    }
    /* the submatrix shares host memory: bring the parent home first */
    if (ctx()) {
        mirror_flush(A, ctx()->s_main);
        CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    }
    CHAM_desc_t *d = (CHAM_desc_t *)malloc(sizeof(CHAM_desc_t));
    *d = *A;
    d->i = A->i + i;
    d->j = A->j + j;
    d->m = m;
    d->n = n;
    d->mt = ceil_div(m, A->mb);
    d->nt = ceil_div(n, A->nb);
    d->alloc_mat = 0;
    d->occurences = 2;
    d->name = nullptr;
    d->schedopt = nullptr;
    d->tiles = A->tiles;
    return d;
}

int CHAMELEON_Desc_Destroy(CHAM_desc_t **desc)
{
    if (!ctx()) {
        error("CHAMELEON_Desc_Destroy", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (!desc || !*desc) {
        error("CHAMELEON_Desc_Destroy", "attempting to destroy a NULL descriptor");
        return CHAMELEON_ERR_UNALLOCATED;
    }
    CHAM_desc_t *d = *desc;
    mirror_release(d);
    if (d->occurences == 1) {
        if (d->alloc_mat) free(d->mat);
        free(d->tiles);
    }
    free(d->name);
    free(d);
    *desc = nullptr;
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Desc_Acquire(const CHAM_desc_t *desc)
{
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    CHAM_desc_t *d = const_cast<CHAM_desc_t *>(desc);
    int rc = mirror_flush(d, ctx()->s_main);
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    return rc;
}

int CHAMELEON_Desc_Release(const CHAM_desc_t *desc)
{
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    /* the host copy may have been modified between Acquire and Release:
     * it becomes the reference copy and the device mirror is dropped */
    Mirror *M = mirror_of(desc);
    if (M) M->state = MIRROR_NONE;
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Desc_Flush(CHAM_desc_t *desc, const RUNTIME_sequence_t *sequence)
{
    (void)sequence;
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    if (!desc) return CHAMELEON_ERR_UNALLOCATED;
    return mirror_flush(desc, ctx()->s_main);
}

static void lap_copy(cham_uplo_t uplo, CHAM_desc_t *d, double *A, int LDA, bool to_desc)
{
    for (int n = 0; n < d->nt; ++n) {
        const int tn = d->get_blkdim(d, n, DIM_n, d->n);
        for (int m = 0; m < d->mt; ++m) {
            const int tm = d->get_blkdim(d, m, DIM_m, d->m);
            double *t = (double *)d->get_blkaddr(d, m, n);
            const int ld = d->get_blkldd(d, m);
            for (int jj = 0; jj < tn; ++jj)
                for (int ii = 0; ii < tm; ++ii) {
                    const int r = m * d->mb + ii, c = n * d->nb + jj;
                    if (uplo == ChamLower && r < c) continue;
                    if (uplo == ChamUpper && r > c) continue;
                    double *a = A + r + (size_t)c * LDA;
                    if (to_desc) t[ii + (size_t)jj * ld] = *a;
                    else *a = t[ii + (size_t)jj * ld];
                }
        }
    }
}

int CHAMELEON_Lap2Desc(cham_uplo_t uplo, void *A, int LDA, CHAM_desc_t *descA)
{
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    if (cham_get_flttype(descA->dtyp) != ChamRealDouble) return CHAMELEON_ERR_NOT_SUPPORTED;
    if (LDA < descA->m) return CHAMELEON_ERR_ILLEGAL_VALUE;
    /* a partial copy needs the other part's current values on the host */
    mirror_flush(descA, ctx()->s_main);
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    lap_copy(uplo, descA, (double *)A, LDA, true);
    Mirror *M = mirror_of(descA);
    if (M) M->state = MIRROR_NONE;
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Desc2Lap(cham_uplo_t uplo, CHAM_desc_t *descA, void *A, int LDA)
{
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    if (cham_get_flttype(descA->dtyp) != ChamRealDouble) return CHAMELEON_ERR_NOT_SUPPORTED;
    if (LDA < descA->m) return CHAMELEON_ERR_ILLEGAL_VALUE;
    mirror_flush(descA, ctx()->s_main);
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    lap_copy(uplo, descA, (double *)A, LDA, false);
    return CHAMELEON_SUCCESS;
}

} /* extern "C" */

namespace chx {

/* descriptor validity, same conditions and codes as the reference
 * chameleon_desc_check (control/descriptor.c:627-675) */
int desc_check(const CHAM_desc_t *d)
{
    if (d == nullptr) {
        error("chameleon_desc_check", "NULL descriptor");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (d->mat == nullptr && d->use_mat == 1) {
        error("chameleon_desc_check", "NULL matrix pointer");
        return CHAMELEON_ERR_UNALLOCATED;
    }
    const int ft = cham_get_flttype(d->dtyp);
    if (ft != ChamInteger && ft != ChamRealHalf && ft != ChamRealFloat && ft != ChamRealDouble &&
        ft != ChamComplexHalf && ft != ChamComplexFloat && ft != ChamComplexDouble) {
        error("chameleon_desc_check", "invalid matrix type");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if (d->mb <= 0 || d->nb <= 0) {
        error("chameleon_desc_check", "negative tile dimension");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if (d->bsiz < d->mb * d->nb) {
        error("chameleon_desc_check", "tile memory size smaller than the product of dimensions");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if (d->lm <= 0 || d->ln <= 0) {
        error("chameleon_desc_check", "negative matrix dimension");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if (d->lm < d->m || d->ln < d->n) {
        error("chameleon_desc_check", "matrix dimensions larger than leading dimensions");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if ((d->i > 0 && d->i >= d->lm) || (d->j > 0 && d->j >= d->ln)) {
        error("chameleon_desc_check", "beginning of the matrix out of scope");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if (d->i + d->m > d->lm || d->j + d->n > d->ln) {
        error("chameleon_desc_check", "submatrix out of scope");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    return CHAMELEON_SUCCESS;
}

} /* namespace chx */

bool chx::desc_is_cm(const CHAM_desc_t *A) { return A->get_blkaddr == getaddr_cm; }
