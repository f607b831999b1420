/*
 * cham_oracle.c -- CPU restatement of Chameleon's tiled Cholesky path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark call it to produce
 * the reference answer and the reference timing.  Nothing in the product
 * library (chameleon_amd/) links, loads or calls it.
 *
 * What it restates, function by function:
 *   orc_rnd64_jump / orc_dlaran      coreblas/include/coreblas/random.h:32-73
 *   orc_dplgsy_tile                  coreblas/compute/core_zplgsy.c:38-112 (d, NBELEM=1)
 *   orc_dplrnt_tile                  coreblas/compute/core_zplrnt.c:36-54
 *   orc_dplgsy / orc_dplrnt          compute/pzplgsy.c:31-83, compute/pzplrnt.c (tile loops,
 *                                    bigM = A->m, tiles outside uplo skipped)
 *   tile kernels potrf/trsm/syrk/gemm  coreblas/compute/core_z{potrf,trsm,syrk,gemm}.c
 *                                    (LAPACK/BLAS semantics; either the textbook loops below or,
 *                                    when orc_set_blas() was given a BLAS, that BLAS)
 *   orc_pdpotrf                      compute/pzpotrf.c:84-193 (right-looking task loops)
 *   orc_pdtrsm_left                  compute/pztrsm.c:60-201 (the four Left branches)
 *   orc_dpotrf/dpotrs/dposv          compute/zpotrf.c, zpotrs.c, zposv.c (CM in-place layout)
 *   orc_dplgsy_batch                 compute/zplghe_batch.c:36-66 (tile origin (n*mt+m)*mb*nb, bigM 0)
 *   orc_dpotrf_batch                 compute/zpotrf_batch.c:28-49 (every tile factored alone)
 * Tasks are OpenMP tasks with depend() clauses on the tiles they read and
 * write, the scheme of runtime/openmp/codelets/codelet_z{potrf,trsm,syrk,gemm}.c.
 *
 * Storage: a column-major matrix with leading dimension lda cut into nb x nb
 * tiles, tile (m,n) at A + nb*n*lda + nb*m (control/descriptor.h:238-252).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

enum { Upper = 121, Lower = 122, UpperLower = 123, NoTrans = 111, Trans = 112,
       NonUnit = 131, Unit = 132, Left = 141, Right = 142 };

/* ------------------------------------------------------------------ RNG */
#define RND_A 6364136223846793005ULL
#define RND_C 1ULL

/* state after n steps of x <- a x + c from `seed`, by binary powering of
 * the affine map (random.h:32-52) */
uint64_t orc_rnd64_jump(uint64_t n, uint64_t seed)
{
    uint64_t a = RND_A, c = RND_C, x = seed;
    while (n) {
        if (n & 1) x = a * x + c;
        c = c * (a + 1);
        a = a * a;
        n >>= 1;
    }
    return x;
}

/* value drawn from state x, then x advances one step (random.h:63-70):
 * 0.5 - x * 2^-64 in double precision. */
static inline double orc_dlaran(uint64_t *x)
{
    double v = 0.5 - (double)(*x) * 5.4210108624275222e-20;
    *x = RND_A * (*x) + RND_C;
    return v;
}

double orc_dlaran_at(uint64_t pos, uint64_t seed)
{
    uint64_t x = orc_rnd64_jump(pos, seed);
    return orc_dlaran(&x);
}

/* ----------------------------------------------------------- generators */
/* Symmetric random tile: element (r,c) of the big matrix, r >= c, takes the
 * stream position r + c*bigM; the upper half mirrors it; the diagonal gets
 * `bump` added.  (m0,n0) is the tile's origin in the big matrix. */
void orc_dplgsy_tile(double bump, int m, int n, double *A, int lda,
                     int bigM, int m0, int n0, uint64_t seed)
{
    int64_t i, j;
    if (m0 == n0) {
        int mn = m < n ? m : n;
        for (j = 0; j < mn; j++) {
            uint64_t x = orc_rnd64_jump((uint64_t)(m0 + j) + (uint64_t)(n0 + j) * (uint64_t)bigM, seed);
            A[j * lda + j] = orc_dlaran(&x) + bump;
            for (i = j + 1; i < m; i++) A[j * lda + i] = orc_dlaran(&x);
        }
        for (i = 0; i < mn; i++) {
            uint64_t x = orc_rnd64_jump((uint64_t)(m0 + i + 1) + (uint64_t)(n0 + i) * (uint64_t)bigM, seed);
            for (j = i + 1; j < n; j++) A[j * lda + i] = orc_dlaran(&x);
        }
    } else if (m0 > n0) {
        for (j = 0; j < n; j++) {
            uint64_t x = orc_rnd64_jump((uint64_t)m0 + (uint64_t)(n0 + j) * (uint64_t)bigM, seed);
            for (i = 0; i < m; i++) A[j * lda + i] = orc_dlaran(&x);
        }
    } else {
        for (i = 0; i < m; i++) {
            uint64_t x = orc_rnd64_jump((uint64_t)n0 + (uint64_t)(m0 + i) * (uint64_t)bigM, seed);
            for (j = 0; j < n; j++) A[j * lda + i] = orc_dlaran(&x);
        }
    }
}

/* General random tile: element (r,c) takes stream position r + c*bigM. */
void orc_dplrnt_tile(int m, int n, double *A, int lda, int bigM, int m0, int n0, uint64_t seed)
{
    for (int64_t j = 0; j < n; j++) {
        uint64_t x = orc_rnd64_jump((uint64_t)m0 + (uint64_t)(n0 + j) * (uint64_t)bigM, seed);
        for (int64_t i = 0; i < m; i++) A[j * lda + i] = orc_dlaran(&x);
    }
}

static inline int tdim(int k, int nb, int n) { return (k + 1) * nb > n ? n - k * nb : nb; }

/* Whole-matrix generators driven tile by tile like chameleon_pzplgsy:
 * tiles entirely outside `uplo` are left untouched. */
void orc_dplgsy(double bump, int uplo, int N, double *A, int lda, int nb, uint64_t seed)
{
    int nt = (N + nb - 1) / nb;
    /* tiles are independent (every element's stream position is fixed):
     * tile columns are generated in parallel */
    #pragma omp parallel for schedule(dynamic)
    for (int n = 0; n < nt; n++) {
        int nn = tdim(n, nb, N), n0 = n * nb;
        for (int m = 0; m < nt; m++) {
            int mm = tdim(m, nb, N), m0 = m * nb;
            if (uplo == Lower && n0 >= m0 + mm) continue;
            if (uplo == Upper && m0 >= n0 + nn) continue;
            orc_dplgsy_tile(bump, mm, nn, A + (size_t)n0 * lda + m0, lda, N, m0, n0, seed);
        }
    }
}

void orc_dplrnt(int M, int N, double *A, int lda, int nb, uint64_t seed)
{
    int mt = (M + nb - 1) / nb, nt = (N + nb - 1) / nb;
    #pragma omp parallel for schedule(dynamic)
    for (int n = 0; n < nt; n++)
        for (int m = 0; m < mt; m++)
            orc_dplrnt_tile(tdim(m, nb, M), tdim(n, nb, N),
                            A + (size_t)n * nb * lda + (size_t)m * nb, lda, M, m * nb, n * nb, seed);
}

/* --------------------------------------------------------- tile kernels */
typedef void (*f_dpotrf)(const char*, const int*, double*, const int*, int*, size_t);
typedef void (*f_dtrsm)(const char*, const char*, const char*, const char*, const int*, const int*,
                        const double*, const double*, const int*, double*, const int*,
                        size_t, size_t, size_t, size_t);
typedef void (*f_dsyrk)(const char*, const char*, const int*, const int*, const double*,
                        const double*, const int*, const double*, double*, const int*, size_t, size_t);
typedef void (*f_dgemm)(const char*, const char*, const int*, const int*, const int*, const double*,
                        const double*, const int*, const double*, const int*, const double*,
                        double*, const int*, size_t, size_t);
typedef void (*f_setthreads)(int);

static f_dpotrf b_potrf;
static f_dtrsm  b_trsm;
static f_dsyrk  b_syrk;
static f_dgemm  b_gemm;

/* Install a Fortran-interface BLAS/LAPACK (e.g. OpenBLAS) for the tile
 * kernels, as coreblas does with CBLAS/LAPACKE; NULL restores the loops. */
void orc_set_blas(void *potrf, void *trsm, void *syrk, void *gemm, void *setthreads)
{
    b_potrf = (f_dpotrf)potrf; b_trsm = (f_dtrsm)trsm;
    b_syrk = (f_dsyrk)syrk;    b_gemm = (f_dgemm)gemm;
    if (setthreads) ((f_setthreads)setthreads)(1);
}

#define AT(M, ld, i, j) (M)[(size_t)(j) * (ld) + (i)]

/* Cholesky of one n x n tile (LAPACK dpotrf): returns 0 or the 1-based
 * column at which a non-positive pivot appeared. */
int orc_dpotrf_tile(int uplo, int n, double *A, int lda)
{
    if (b_potrf) {
        int info = 0; char u = uplo == Lower ? 'L' : 'U';
        b_potrf(&u, &n, A, &lda, &info, 1);
        return info;
    }
    for (int j = 0; j < n; j++) {
        double d = AT(A, lda, j, j);
        for (int p = 0; p < j; p++) {
            double v = uplo == Lower ? AT(A, lda, j, p) : AT(A, lda, p, j);
            d -= v * v;
        }
        if (!(d > 0.0)) return j + 1;
        d = sqrt(d);
        AT(A, lda, j, j) = d;
        for (int i = j + 1; i < n; i++) {
            double s;
            if (uplo == Lower) {
                s = AT(A, lda, i, j);
                for (int p = 0; p < j; p++) s -= AT(A, lda, i, p) * AT(A, lda, j, p);
                AT(A, lda, i, j) = s / d;
            } else {
                s = AT(A, lda, j, i);
                for (int p = 0; p < j; p++) s -= AT(A, lda, p, i) * AT(A, lda, p, j);
                AT(A, lda, j, i) = s / d;
            }
        }
    }
    return 0;
}

/* op(T)(r,c) of a triangular tile, with zeros outside the triangle */
static inline double tri_get(const double *T, int ldt, int uplo, int trans, int diag, int r, int c)
{
    int rr = trans == NoTrans ? r : c, cc = trans == NoTrans ? c : r;
    if (rr == cc) return diag == Unit ? 1.0 : AT(T, ldt, rr, cc);
    if (uplo == Lower ? rr < cc : rr > cc) return 0.0;
    return AT(T, ldt, rr, cc);
}

/* B <- alpha op(T)^-1 B (Left) or alpha B op(T)^-1 (Right), T triangular */
void orc_dtrsm_tile(int side, int uplo, int trans, int diag, int m, int n, double alpha,
                    const double *T, int ldt, double *B, int ldb)
{
    if (b_trsm) {
        char s = side == Left ? 'L' : 'R', u = uplo == Lower ? 'L' : 'U';
        char t = trans == NoTrans ? 'N' : 'T', d = diag == Unit ? 'U' : 'N';
        b_trsm(&s, &u, &t, &d, &m, &n, &alpha, T, &ldt, B, &ldb, 1, 1, 1, 1);
        return;
    }
    /* effective op(T) is lower triangular when (Lower,NoTrans) or (Upper,Trans) */
    int lower_eff = (uplo == Lower) == (trans == NoTrans);
    if (side == Left) {
        for (int c = 0; c < n; c++) {
            for (int i = 0; i < m; i++) AT(B, ldb, i, c) *= alpha;
            if (lower_eff) {
                for (int i = 0; i < m; i++) {
                    double s = AT(B, ldb, i, c);
                    for (int p = 0; p < i; p++) s -= tri_get(T, ldt, uplo, trans, diag, i, p) * AT(B, ldb, p, c);
                    AT(B, ldb, i, c) = s / tri_get(T, ldt, uplo, trans, diag, i, i);
                }
            } else {
                for (int i = m - 1; i >= 0; i--) {
                    double s = AT(B, ldb, i, c);
                    for (int p = i + 1; p < m; p++) s -= tri_get(T, ldt, uplo, trans, diag, i, p) * AT(B, ldb, p, c);
                    AT(B, ldb, i, c) = s / tri_get(T, ldt, uplo, trans, diag, i, i);
                }
            }
        }
    } else {
        /* X op(T) = alpha B, row by row: x_r op(T) = b_r */
        for (int r = 0; r < m; r++) {
            for (int j = 0; j < n; j++) AT(B, ldb, r, j) *= alpha;
            if (lower_eff) {
                for (int j = n - 1; j >= 0; j--) {
                    double s = AT(B, ldb, r, j);
                    for (int p = j + 1; p < n; p++) s -= AT(B, ldb, r, p) * tri_get(T, ldt, uplo, trans, diag, p, j);
                    AT(B, ldb, r, j) = s / tri_get(T, ldt, uplo, trans, diag, j, j);
                }
            } else {
                for (int j = 0; j < n; j++) {
                    double s = AT(B, ldb, r, j);
                    for (int p = 0; p < j; p++) s -= AT(B, ldb, r, p) * tri_get(T, ldt, uplo, trans, diag, p, j);
                    AT(B, ldb, r, j) = s / tri_get(T, ldt, uplo, trans, diag, j, j);
                }
            }
        }
    }
}

/* C <- alpha op(A) op(A)^T + beta C on the `uplo` triangle only;
 * op(A) is n x k (NoTrans) or A^T with A k x n (Trans). */
void orc_dsyrk_tile(int uplo, int trans, int n, int k, double alpha, const double *A, int lda,
                    double beta, double *C, int ldc)
{
    if (b_syrk) {
        char u = uplo == Lower ? 'L' : 'U', t = trans == NoTrans ? 'N' : 'T';
        b_syrk(&u, &t, &n, &k, &alpha, A, &lda, &beta, C, &ldc, 1, 1);
        return;
    }
    for (int j = 0; j < n; j++) {
        int i0 = uplo == Lower ? j : 0, i1 = uplo == Lower ? n : j + 1;
        for (int i = i0; i < i1; i++) {
            double s = 0.0;
            for (int p = 0; p < k; p++)
                s += trans == NoTrans ? AT(A, lda, i, p) * AT(A, lda, j, p) : AT(A, lda, p, i) * AT(A, lda, p, j);
            AT(C, ldc, i, j) = alpha * s + beta * AT(C, ldc, i, j);
        }
    }
}

/* C <- alpha op(A) op(B) + beta C */
void orc_dgemm_tile(int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
                    const double *B, int ldb, double beta, double *C, int ldc)
{
    if (b_gemm) {
        char a = ta == NoTrans ? 'N' : 'T', b = tb == NoTrans ? 'N' : 'T';
        b_gemm(&a, &b, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc, 1, 1);
        return;
    }
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double s = 0.0;
            for (int p = 0; p < k; p++) {
                double a = ta == NoTrans ? AT(A, lda, i, p) : AT(A, lda, p, i);
                double b = tb == NoTrans ? AT(B, ldb, p, j) : AT(B, ldb, j, p);
                s += a * b;
            }
            AT(C, ldc, i, j) = alpha * s + beta * AT(C, ldc, i, j);
        }
}

/* ------------------------------------------------------ tiled algorithms */
typedef struct { double *mat; int ld, nb, m, n, mt, nt; } orc_desc;

static inline double *TILE(const orc_desc *d, int m, int n)
{
    return d->mat + (size_t)d->ld * d->nb * n + (size_t)d->nb * m;
}
static inline int DM(const orc_desc *d, int m) { return tdim(m, d->nb, d->m); }
static inline int DN(const orc_desc *d, int n) { return tdim(n, d->nb, d->n); }

static void set_info(int *info, int v)
{
    #pragma omp critical(orc_info)
    if (*info == 0 || v < *info) *info = v;
}

/* Right-looking tiled Cholesky (pzpotrf.c:84-193).  Returns 0, or the
 * global 1-based order of the first leading minor found not positive
 * definite (tile potrf info offset by nb*k, pzpotrf.c:91-93). */
static int pdpotrf(int uplo, orc_desc *A)
{
    int info = 0;
    #pragma omp parallel
    #pragma omp single
    {
        int nt = A->nt, nb = A->nb;
        for (int k = 0; k < nt; k++) {
            int kk = DN(A, k);
            double *Akk = TILE(A, k, k);
            #pragma omp task depend(inout: Akk[0]) firstprivate(Akk, kk, k) shared(info)
            {
                int ii = orc_dpotrf_tile(uplo, kk, Akk, A->ld);
                if (ii) set_info(&info, nb * k + ii);
            }
            for (int m = k + 1; m < nt; m++) {
                int mm = DM(A, m);
                if (uplo == Lower) {
                    double *Amk = TILE(A, m, k);
                    #pragma omp task depend(in: Akk[0]) depend(inout: Amk[0]) firstprivate(Akk, Amk, mm)
                    orc_dtrsm_tile(Right, Lower, Trans, NonUnit, mm, nb, 1.0, Akk, A->ld, Amk, A->ld);
                } else {
                    double *Akm = TILE(A, k, m);
                    #pragma omp task depend(in: Akk[0]) depend(inout: Akm[0]) firstprivate(Akk, Akm, mm)
                    orc_dtrsm_tile(Left, Upper, Trans, NonUnit, nb, mm, 1.0, Akk, A->ld, Akm, A->ld);
                }
            }
            for (int n = k + 1; n < nt; n++) {
                int nn = DN(A, n);
                double *Ann = TILE(A, n, n);
                if (uplo == Lower) {
                    double *Ank = TILE(A, n, k);
                    #pragma omp task depend(in: Ank[0]) depend(inout: Ann[0]) firstprivate(Ank, Ann, nn)
                    orc_dsyrk_tile(Lower, NoTrans, nn, nb, -1.0, Ank, A->ld, 1.0, Ann, A->ld);
                    for (int m = n + 1; m < nt; m++) {
                        int mm = DM(A, m);
                        double *Amk = TILE(A, m, k), *Amn = TILE(A, m, n);
                        #pragma omp task depend(in: Amk[0], Ank[0]) depend(inout: Amn[0]) firstprivate(Amk, Ank, Amn, mm, nn)
                        orc_dgemm_tile(NoTrans, Trans, mm, nn, nb, -1.0, Amk, A->ld, Ank, A->ld, 1.0, Amn, A->ld);
                    }
                } else {
                    double *Akn = TILE(A, k, n);
                    #pragma omp task depend(in: Akn[0]) depend(inout: Ann[0]) firstprivate(Akn, Ann, nn)
                    orc_dsyrk_tile(Upper, Trans, nn, nb, -1.0, Akn, A->ld, 1.0, Ann, A->ld);
                    for (int m = n + 1; m < nt; m++) {
                        int mm = DN(A, m);
                        double *Akm = TILE(A, k, m), *Anm = TILE(A, n, m);
                        #pragma omp task depend(in: Akn[0], Akm[0]) depend(inout: Anm[0]) firstprivate(Akn, Akm, Anm, mm, nn)
                        orc_dgemm_tile(Trans, NoTrans, nn, mm, nb, -1.0, Akn, A->ld, Akm, A->ld, 1.0, Anm, A->ld);
                    }
                }
            }
        }
    }
    return info;
}

/* Left tiled triangular solve B <- op(A)^-1 B with A's `uplo` triangle
 * (pztrsm.c:60-201), alpha = 1 as potrs/posv use it. */
static void pdtrsm_left(int uplo, int trans, orc_desc *A, orc_desc *B)
{
    #pragma omp parallel
    #pragma omp single
    {
        int mt = B->mt, nt = B->nt;
        /* forward when op(A) is lower triangular, backward otherwise */
        int forward = (uplo == Lower) == (trans == NoTrans);
        for (int kk = 0; kk < mt; kk++) {
            int k = forward ? kk : mt - 1 - kk;
            int km = DM(B, k);
            double *Akk = TILE(A, k, k);
            for (int n = 0; n < nt; n++) {
                int nn = DN(B, n);
                double *Bkn = TILE(B, k, n);
                #pragma omp task depend(in: Akk[0]) depend(inout: Bkn[0]) firstprivate(Akk, Bkn, km, nn)
                orc_dtrsm_tile(Left, uplo, trans, NonUnit, km, nn, 1.0, Akk, A->ld, Bkn, B->ld);
            }
            for (int mm_ = kk + 1; mm_ < mt; mm_++) {
                int m = forward ? mm_ : mt - 1 - mm_;
                int mm = DM(B, m);
                /* op(A)(m,k): stored tile is A(m,k) or A(k,m) */
                double *Aop = trans == NoTrans ? TILE(A, m, k) : TILE(A, k, m);
                int ta = trans;
                for (int n = 0; n < nt; n++) {
                    int nn = DN(B, n);
                    double *Bkn = TILE(B, k, n), *Bmn = TILE(B, m, n);
                    #pragma omp task depend(in: Aop[0], Bkn[0]) depend(inout: Bmn[0]) firstprivate(Aop, Bkn, Bmn, mm, nn, km, ta)
                    orc_dgemm_tile(ta, NoTrans, mm, nn, km, -1.0, Aop, A->ld, Bkn, B->ld, 1.0, Bmn, B->ld);
                }
            }
        }
    }
}

static void mkdesc(orc_desc *d, double *mat, int ld, int nb, int m, int n)
{
    d->mat = mat; d->ld = ld; d->nb = nb; d->m = m; d->n = n;
    d->mt = (m + nb - 1) / nb; d->nt = (n + nb - 1) / nb;
}

static void set_threads(int nthreads) { if (nthreads > 0) omp_set_num_threads(nthreads); }

/* CHAMELEON_dpotrf on a column-major array, in place (zpotrf.c:81-145) */
int orc_dpotrf(int uplo, int N, double *A, int lda, int nb, int nthreads)
{
    if (uplo != Upper && uplo != Lower) return -1;
    if (N < 0) return -2;
    if (lda < (N > 1 ? N : 1)) return -4;
    if (N == 0) return 0;
    set_threads(nthreads);
    orc_desc dA; mkdesc(&dA, A, lda, nb, N, N);
    return pdpotrf(uplo, &dA);
}

/* CHAMELEON_dpotrs (zpotrs.c:78-152): two tiled triangular solves */
int orc_dpotrs(int uplo, int N, int NRHS, double *A, int lda, double *B, int ldb, int nb, int nthreads)
{
    if (uplo != Upper && uplo != Lower) return -1;
    if (N < 0) return -2;
    if (NRHS < 0) return -3;
    if (lda < (N > 1 ? N : 1)) return -5;
    if (ldb < (N > 1 ? N : 1)) return -7;
    if (N == 0 || NRHS == 0) return 0;
    set_threads(nthreads);
    orc_desc dA, dB; mkdesc(&dA, A, lda, nb, N, N); mkdesc(&dB, B, ldb, nb, N, NRHS);
    pdtrsm_left(uplo, uplo == Upper ? Trans : NoTrans, &dA, &dB);
    pdtrsm_left(uplo, uplo == Upper ? NoTrans : Trans, &dA, &dB);
    return 0;
}

/* CHAMELEON_dposv (zposv.c:93-172 and :276-333) */
int orc_dposv(int uplo, int N, int NRHS, double *A, int lda, double *B, int ldb, int nb, int nthreads)
{
    if (uplo != Upper && uplo != Lower) return -1;
    if (N < 0) return -2;
    if (NRHS < 0) return -3;
    if (lda < (N > 1 ? N : 1)) return -5;
    if (ldb < (N > 1 ? N : 1)) return -7;
    if (N == 0 || NRHS == 0) return 0;
    set_threads(nthreads);
    orc_desc dA, dB; mkdesc(&dA, A, lda, nb, N, N); mkdesc(&dB, B, ldb, nb, N, NRHS);
    int info = pdpotrf(uplo, &dA);
    if (info) return info;
    pdtrsm_left(uplo, uplo == Upper ? Trans : NoTrans, &dA, &dB);
    pdtrsm_left(uplo, uplo == Upper ? NoTrans : Trans, &dA, &dB);
    return 0;
}

/* ------------------------------------------------------------ batches */
/* CHAMELEON_zplghe_batch_Tile (compute/zplghe_batch.c:36-66, d instance):
 * every tile (m, n) of an M x N matrix cut in mb x nb tiles is an
 * independent tempnn x tempnn matrix generated with origin
 * m0 = n0 = (n*mt + m)*mb*nb and bigM = 0.  CM layout: tile (m, n) at
 * A + n*nb*lda + m*mb. */
void orc_dplgsy_batch(double bump, int M, int N, int mb, int nb, double *A, int lda, uint64_t seed)
{
    int mt = (M + mb - 1) / mb, nt = (N + nb - 1) / nb;
    for (int n = 0; n < nt; n++)
        for (int m = 0; m < mt; m++) {
            int tempnn = tdim(n, nb, N);
            int m0 = (n * mt + m) * (mb * nb);
            orc_dplgsy_tile(bump, tempnn, tempnn, A + (size_t)n * nb * lda + (size_t)m * mb, lda, 0, m0, m0, seed);
        }
}

/* CHAMELEON_zpotrf_batch_Tile (compute/zpotrf_batch.c:28-49): every tile
 * factored in place by the tile kernel; info[n*mt + m] receives each
 * tile's result (the reference prints failures and returns 0). */
void orc_dpotrf_batch(int uplo, int M, int N, int mb, int nb, double *A, int lda, int *info)
{
    int mt = (M + mb - 1) / mb, nt = (N + nb - 1) / nb;
    for (int n = 0; n < nt; n++)
        for (int m = 0; m < mt; m++) {
            int tempnn = tdim(n, nb, N);
            info[n * mt + m] = orc_dpotrf_tile(uplo, tempnn, A + (size_t)n * nb * lda + (size_t)m * mb, lda);
        }
}

double orc_wtime(void) { return omp_get_wtime(); }
