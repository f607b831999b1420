/*
 * latms.cpp -- conditioned SPD test matrices on the device:
 * A = Q diag(lambda) Q^T with Q the orthogonal factor of the QR
 * factorisation of a seeded random N x N matrix G.
 *
 * This is the construction of the reference's chameleon_pzlatms
 * (compute/pzlatms.c:194-239: U <- plrnt, U <- qr(U), A <- U D, A <- A U^T)
 * with the mode-3 spectrum of core_zlatm1.c:172-180, lambda_i = alpha^i,
 * alpha = kappa^(-1/(N-1)).  G is either Chameleon's uniform dplrnt matrix
 * with the seed advanced by 2N as pzlatms.c:139 does (ChamDistSymmetric:
 * the reference's own choice) or a matrix of N(0,1) entries
 * (ChamDistNormal, whose Q is Haar distributed).
 *
 * The QR factor is computed as shifted CholeskyQR3 (Fukaya, Kannan,
 * Nakatsukasa, Zhang, Yamamoto, SIAM J. Sci. Comput. 42 (2020) A477):
 *   pass 1:  C = G^T G + s I,  C = R^T R (dpotrf),  G <- G R^-1
 *   pass 2, 3: the same without the shift,
 * s = 11 (N^2 + N(N+1)) eps ||G||_F^2.  The Q of the QR factorisation is
 * unique up to the signs of its columns, which cancel in Q D Q^T, so A is
 * the matrix pzlatms defines whatever QR algorithm forms Q.  Every step
 * runs on the library's own kernels: the tile product of kernels/
 * latms.hip for G^T G and Q D Q^T, the tiled Cholesky of the dpotrf path,
 * and the panel TRSM for G R^-1 (right-looking over tile columns).
 *
 * Q depends only on (N, tiling, seed, distribution); it is kept on the
 * device after a call so that matrices of other condition numbers from the
 * same seed cost one tile product each.
 */
#include "internal.h"
#include "chameleon_amd.h"

#include <cmath>
#include <vector>

using namespace chx;

namespace {

struct LatmsCache {
    int N = 0, mb = 0, ldt = 0, NT = 0, dist = 0;
    unsigned long long seed = 0;
    double *Q = nullptr;
};

LatmsCache *cache() { return (LatmsCache *)ctx()->latms_cache; }

/* The orthogonal factor for (N, mb, seed, dist) in the cache, computed if
 * needed.  Returns CHAMELEON_SUCCESS or an error code. */
int latms_q(int N, int mb, int ldt, int NT, int dist, unsigned long long seed)
{
    Context *c = ctx();
    LatmsCache *lc = cache();
    if (lc && lc->Q && lc->N == N && lc->mb == mb && lc->ldt == ldt && lc->seed == seed && lc->dist == dist)
        return CHAMELEON_SUCCESS;
    latms_release();
    hipStream_t s = c->s_main;
    const size_t ts = (size_t)ldt * ldt, elems = ts * NT * NT;
    const size_t dper = (size_t)((ldt + 63) / 64) * 64 * 64;
    double *Q = nullptr, *T = nullptr, *C = nullptr, *dinv = nullptr, *part = nullptr;
    int *info = nullptr;
    if (hipMalloc(&Q, elems * sizeof(double)) != hipSuccess || hipMalloc(&T, elems * sizeof(double)) != hipSuccess ||
        hipMalloc(&C, elems * sizeof(double)) != hipSuccess ||
        hipMalloc(&dinv, (size_t)NT * dper * sizeof(double)) != hipSuccess ||
        hipMalloc(&part, (size_t)NT * NT * sizeof(double)) != hipSuccess || hipMalloc(&info, sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        if (Q) (void)hipFree(Q);
        if (T) (void)hipFree(T);
        if (C) (void)hipFree(C);
        if (dinv) (void)hipFree(dinv);
        if (part) (void)hipFree(part);
        error("CHAMAMD_dplgsy_cond_Tile", "cannot allocate the generator's workspace (3 N^2 doubles)");
        return CHAMELEON_ERR_OUT_OF_RESOURCES;
    }
    CHX_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
    if (dist == ChamDistNormal) {
        k::gauss(Q, NT, ldt, mb, N, seed, s);
    } else {
        /* pzlatms.c:139: the seed moves 2 min(M,N) states on, then
         * U = plrnt(seed) with bigM = M (pzlatms.c:197) */
        unsigned long long a = 6364136223846793005ULL, cc = 1ULL, x = seed;
        for (unsigned long long e = 2ULL * (unsigned long long)N; e; e >>= 1) {
            if (e & 1ULL) x = a * x + cc;
            cc = cc * (a + 1ULL);
            a = a * a;
        }
        k::plrnt(Q, NT, NT, ldt, ldt, mb, mb, N, N, N, 0, 0, x, 0, s);
        k::fix_padding(Q, NT, NT, ldt, ldt, mb, mb, N, N, 8, true, s);
    }
    /* shift of the first pass from ||G||_F^2 (an upper bound of ||G||_2^2),
     * summed over tiles in a fixed order */
    k::sumsq_tiles(Q, NT * NT, ldt, part, s);
    std::vector<double> hp((size_t)NT * NT);
    CHX_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    CHX_HIP(hipStreamSynchronize(s));
    double fro2 = 0.0;
    for (double v : hp) fro2 += v;
    fro2 -= (double)(NT * ldt - N);                       /* the padded identity entries */
    const double eps = std::ldexp(1.0, -53);
    const double shift = 11.0 * ((double)N * N + (double)N * (N + 1.0)) * eps * fro2;

    int rc = CHAMELEON_SUCCESS;
    for (int pass = 0; pass < 3 && rc == CHAMELEON_SUCCESS; ++pass) {
        /* C = G^T G (lower tiles): C(m,n) = sum_k G^T(m,k) G^T(n,k)^T */
        k::transpose_full(Q, T, NT, ldt, s);
        k::tile_gemm_nt(T, NT, T, NT, C, NT, ldt, 0, NT, 0, NT, 0, NT, 1, 1.0, 0.0, s);
        if (pass == 0) k::add_diag(C, NT, ldt, mb, N, shift, s);
        /* C = L L^T, L = R^T */
        PotrfPlan p;
        p.A = C;
        p.NT = NT;
        p.ldt = ldt;
        p.nb = mb;
        p.nvalid_last = N - (NT - 1) * mb;
        p.d_info = info;
        potrf_run(p);
        int hinfo = 0;
        CHX_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
        CHX_HIP(hipStreamSynchronize(s));
        if (hinfo != 0) {
            error("CHAMAMD_dplgsy_cond_Tile", "CholeskyQR: G^T G not positive definite");
            rc = CHAMELEON_ERR_UNEXPECTED;
            break;
        }
        /* G <- G L^-T, right-looking over tile columns j:
         *   G(:, j) <- G(:, j) L(j,j)^-T;  G(:, j') -= G(:, j) L(j',j)^T, j' > j */
        k::dinv_all(C, NT, (size_t)(NT + 1) * ts, ldt, dinv, s);
        for (int j = 0; j < NT; ++j) {
            k::trsm_right_lower_trans(Q + (size_t)j * NT * ts, NT, ts, ldt, C + ((size_t)j * NT + j) * ts,
                                      dinv + (size_t)j * dper, info, s);
            if (j + 1 < NT)
                k::tile_gemm_nt(Q, NT, C, NT, Q, NT, ldt, j, j + 1, 0, NT, j + 1, NT - j - 1, 0, -1.0, 1.0, s);
        }
    }
    CHX_HIP(hipStreamSynchronize(s));
    (void)hipFree(T);
    (void)hipFree(C);
    (void)hipFree(dinv);
    (void)hipFree(part);
    (void)hipFree(info);
    if (rc != CHAMELEON_SUCCESS) {
        (void)hipFree(Q);
        return rc;
    }
    LatmsCache *n = new LatmsCache();
    n->N = N; n->mb = mb; n->ldt = ldt; n->NT = NT; n->dist = dist; n->seed = seed; n->Q = Q;
    c->latms_cache = n;
    return CHAMELEON_SUCCESS;
}

} /* namespace */

namespace chx {

void latms_release()
{
    Context *c = ctx();
    if (!c || !c->latms_cache) return;
    LatmsCache *lc = (LatmsCache *)c->latms_cache;
    if (lc->Q) {
        CHX_HIP(hipStreamSynchronize(c->s_main));
        CHX_HIP(hipFree(lc->Q));
    }
    delete lc;
    c->latms_cache = nullptr;
}

} /* namespace chx */

extern "C" {

int CHAMAMD_dplgsy_cond_Tile(double kappa, int dist, CHAM_desc_t *A, unsigned long long int seed)
{
    const char *fn = "CHAMAMD_dplgsy_cond_Tile";
    if (!ctx()) {
        error(fn, "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (desc_check(A) != CHAMELEON_SUCCESS || cham_get_flttype(A->dtyp) != ChamRealDouble || A->mb != A->nb ||
        A->m != A->n) {
        error(fn, "A must be a square double descriptor with square tiles");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if (!(kappa >= 1.0)) {
        error(fn, "kappa must be >= 1");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if (dist != ChamDistNormal && dist != ChamDistSymmetric) {
        error(fn, "dist must be ChamDistNormal or ChamDistSymmetric");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if (A->m == 0) return CHAMELEON_SUCCESS;
    Context *c = ctx();
    hipStream_t s = c->s_main;
    mirror_prepare_overwrite(A, MIRROR_NORMAL, true);
    Mirror *M = mirror_of(A);
    const int NT = M->mt, ldt = M->ldt, N = A->m, mb = A->mb;
    int rc = latms_q(N, mb, ldt, NT, dist, seed);
    if (rc != CHAMELEON_SUCCESS) return rc;
    const double *Q = cache()->Q;
    /* W = Q diag(sqrt(lambda)) (padded columns scaled by 1), A = W W^T */
    std::vector<double> sc((size_t)NT * ldt, 1.0);
    const double alpha = N > 1 ? std::pow(kappa, -1.0 / (double)(N - 1)) : 1.0;
    for (int i = 0; i < N; ++i) sc[(size_t)(i / mb) * ldt + (i % mb)] = std::sqrt(std::pow(alpha, (double)i));
    const size_t elems = (size_t)ldt * ldt * NT * NT;
    double *W = nullptr, *dsc = nullptr;
    if (hipMalloc(&W, elems * sizeof(double)) != hipSuccess || hipMalloc(&dsc, sc.size() * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        if (W) (void)hipFree(W);
        error(fn, "cannot allocate the generator's workspace");
        return CHAMELEON_ERR_OUT_OF_RESOURCES;
    }
    CHX_HIP(hipMemcpyAsync(dsc, sc.data(), sc.size() * sizeof(double), hipMemcpyHostToDevice, s));
    k::scale_cols(Q, W, NT, ldt, dsc, s);
    double *D = (double *)M->d;
    k::tile_gemm_nt(W, NT, W, NT, D, NT, ldt, 0, NT, 0, NT, 0, NT, 1, 1.0, 0.0, s);
    k::symmetrize(D, NT, ldt, s);
    k::fix_padding(D, NT, NT, ldt, ldt, mb, mb, N, N, 8, true, s);
    CHX_HIP(hipStreamSynchronize(s));
    (void)hipFree(W);
    (void)hipFree(dsc);
    mirror_mark_written(A);
    A->sync = 1;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Cond_Cache_Release(void)
{
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    latms_release();
    return CHAMELEON_SUCCESS;
}

} /* extern "C" */
