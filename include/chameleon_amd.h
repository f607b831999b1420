/*
 * chameleon_amd.h -- entry points of libchameleon_amd.so that have no
 * counterpart in the reference Chameleon API.
 */
#ifndef CHAMELEON_AMD_EXTRAS_H
#define CHAMELEON_AMD_EXTRAS_H

#include <stddef.h>
#include "chameleon.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel classes timed while CHAMELEON_GENERATE_STATS is enabled (the
 * reference's GENERATE_STATS switches on StarPU per-kernel statistics,
 * runtime/starpu/control/runtime_profiling.c:120-150; here each launch of
 * the class is bracketed by HIP events on its own stream). */
enum {
    CHAMAMD_KERNEL_UPDATE = 0,   /* bulk trailing SYRK+GEMM update of the factorisation */
    CHAMAMD_KERNEL_POTRF  = 1,   /* diagonal tile POTRF                             */
    CHAMAMD_KERNEL_TRSM   = 2,   /* panel TRSM                                      */
    CHAMAMD_KERNEL_SOLVE  = 3,   /* triangular-solve GEMMs (forward + backward)     */
    CHAMAMD_KERNEL_PANEL  = 4,   /* fused diagonal POTRF + panel TRSM of one step   */
    CHAMAMD_KERNEL_COLUPD = 5,   /* lookahead column updates (the critical stream)  */
    CHAMAMD_KERNEL_NCLASS = 6
};

/* Accumulated launches, summed device time (ms) and algorithmic flops
 * (Chameleon's flop model, include/chameleon/flops.h) of one kernel class
 * since the last reset.  Results of a sequence are folded in when it is
 * waited on.  Returns CHAMELEON_SUCCESS or CHAMELEON_ERR_ILLEGAL_VALUE. */
int CHAMAMD_Kernel_Stats(int kernel_class, long long *launches, double *total_ms, double *flops);
int CHAMAMD_Kernel_Stats_Reset(void);

/* dposv runs the forward substitution of a tile row right after its panel,
 * beside the rest of the factorisation (as Chameleon's runtime overlaps the
 * pztrsm tasks of zposv with pzpotrf); the last `tiles` tile rows are solved
 * forward after the factorisation instead, by one strip solve.  The first
 * solve of the mixed-precision refinement runs the same way.  tiles = -1
 * (the default, or the CHAMAMD_FWD_TAIL environment variable) picks the
 * count from the tile count and size and, in the mixed-precision solve, the
 * update rate of the chosen products.  Returns CHAMELEON_SUCCESS,
 * CHAMELEON_ERR_ILLEGAL_VALUE for tiles < -1, or
 * CHAMELEON_ERR_NOT_INITIALIZED. */
int CHAMAMD_Set_Forward_Tail(int tiles);

/* Products used by the trailing updates of the mixed-precision solve (the
 * matrix itself stays in FP64 storage, panels are factored in FP64). */
enum {
    CHAMAMD_PREC_FP32   = 0,   /* FP32 operands, v_mfma_f32_16x16x4_f32                      */
    CHAMAMD_PREC_BF16   = 1,   /* BF16 operands, v_mfma_f32_16x16x32_bf16                    */
    CHAMAMD_PREC_BF16X3 = 2,   /* operands split hi+lo in BF16, 3 products (~2^-16 accuracy) */
    CHAMAMD_PREC_FP8    = 3,   /* OCP E4M3 operands, one power-of-two scale per tile,
                                  v_mfma_scale_f32_16x16x128_f8f6f4 (2^-4 accuracy: refinement
                                  converges for condition numbers up to ~10)                  */
    CHAMAMD_PREC_ADAPTIVE = 4  /* per-tile precision by Chameleon's tile-norm rule
                                  (compute/pzhered.c:176-288): each trailing tile is updated
                                  in FP64, FP32 or BF16 as its norm allows for the accuracy
                                  of CHAMAMD_Set_Tile_Accuracy                                */
};

/* Requested accuracy u_high of the per-tile precision choice
 * (CHAMAMD_PREC_ADAPTIVE): tile (m,n) of the working matrix is updated in
 * BF16 when ||A_mn||_F < u_high ||A||_F / (NT 2^-8), else in FP32 when
 * ||A_mn||_F < u_high ||A||_F / (NT 2^-24), else in FP64 -- the rule of
 * the reference's zhered/zgered (default 1e-12, or the CHAMELEON_GERED_ACC
 * environment variable as compute/zhered.c:170-176 reads it).  Returns
 * CHAMELEON_SUCCESS, CHAMELEON_ERR_ILLEGAL_VALUE (u_high <= 0) or
 * CHAMELEON_ERR_NOT_INITIALIZED. */
int CHAMAMD_Set_Tile_Accuracy(double u_high);
/* Lower tiles (m >= n) given FP64, FP32 and BF16 products by the last
 * CHAMAMD_PREC_ADAPTIVE solve, in counts[0..2]. */
int CHAMAMD_Mixed_Tile_Counts(long long *counts);

/* Symmetric diagonal scaling of the working copy before demotion:
 * D^-1/2 A D^-1/2 (D = diag(A)) is factored in reduced precision and the
 * solves are scaled back, so the demoted operands span a range set by the
 * matrix's structure rather than its diagonal's magnitudes; residuals stay
 * on the original A in FP64.  on = 1 (the default) or 0.  Returns
 * CHAMELEON_SUCCESS or CHAMELEON_ERR_NOT_INITIALIZED. */
int CHAMAMD_Set_Mixed_Scaling(int on);

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

/* Symmetric positive definite matrix with eigenvalues lambda_i = alpha^i,
 * alpha = kappa^(-1/(N-1)), i < N, generated on the device:
 * A = Q diag(lambda) Q^T with Q the orthogonal factor of the QR
 * factorisation of a seeded random N x N matrix G -- the construction of
 * the reference's dlatms generator (compute/pzlatms.c:194-239) with the
 * mode-3 spectrum of coreblas/compute/core_zlatm1.c:172-180.
 * dist = ChamDistSymmetric: G = dplrnt(N x N) with the seed advanced 2N
 * states (pzlatms.c:139), the reference's choice; ChamDistNormal: G has
 * N(0,1) entries (Box-Muller on consecutive LCG values), Q is Haar
 * distributed.  Q is formed by shifted CholeskyQR3 on the device (about
 * 7 N^3 flops) and kept for further calls with the same N, tiling, seed
 * and dist; each further kappa costs one N^3 product.  The full matrix
 * (both triangles) is written. */
int CHAMAMD_dplgsy_cond_Tile(double kappa, int dist, CHAM_desc_t *A, unsigned long long int seed);
/* frees the orthogonal factor kept by CHAMAMD_dplgsy_cond_Tile (N^2 doubles) */
int CHAMAMD_Cond_Cache_Release(void);

/* Stopping test of the refinement.  berr = 0 (the default): LAPACK
 * dsposv's test ||R_j||_inf <= ||X_j||_inf ||A||_inf eps sqrt(N) per
 * column.  berr > 0: refine until every column's backward error
 * ||R_j||_inf / (||A||_inf ||X_j||_inf N eps) is at most berr (still at most
 * 30 steps, then the FP64 fallback), e.g. to reach the backward error of an
 * FP64 solve of the same system.  Returns CHAMELEON_SUCCESS,
 * CHAMELEON_ERR_ILLEGAL_VALUE (berr < 0) or CHAMELEON_ERR_NOT_INITIALIZED. */
int CHAMAMD_Set_Refinement_Target(double berr);

/* The same backward error for any X, A symmetric given by its `uplo`
 * triangle, computed on the device in FP64. */
int CHAMAMD_Backward_Error(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *X, CHAM_desc_t *B, double *berr);

/* ------------------------------------------------------------------ */
/* Uneven batches: `count` independent SPD systems A_i X_i = B_i of orders
 * n[i] with nrhs[i] right-hand sides, resident on the device.  No
 * counterpart in the reference (its batch routines take one descriptor of
 * equal tiles and run on the CPU only, compute/zpotrf_batch.c:58-64).
 * Each system is factored and solved by one workgroup, largest systems
 * first; its result is bitwise independent of the rest of the batch.
 * Factors are kept in Lower storage; Set/GetA take either triangle. */
typedef struct CHAMAMD_batch_s CHAMAMD_batch_t;

int CHAMAMD_Batch_Create(CHAMAMD_batch_t **batch, int count, const int *n, const int *nrhs);
int CHAMAMD_Batch_Destroy(CHAMAMD_batch_t **batch);
/* host <-> device copies of system i (column-major, leading dimension ld) */
int CHAMAMD_Batch_SetA(CHAMAMD_batch_t *batch, int i, cham_uplo_t uplo, const double *A, int lda);
int CHAMAMD_Batch_GetA(CHAMAMD_batch_t *batch, int i, cham_uplo_t uplo, double *A, int lda);
int CHAMAMD_Batch_SetB(CHAMAMD_batch_t *batch, int i, const double *B, int ldb);
int CHAMAMD_Batch_GetB(CHAMAMD_batch_t *batch, int i, double *B, int ldb);
/* on the device: A_i = dplgsy(bump = n_i, n_i x n_i, seed + i) and
 * B_i = dplrnt(n_i x nrhs_i, seed + i) (compute/zplgsy.c, zplrnt.c) */
int CHAMAMD_dplgsy_batch(CHAMAMD_batch_t *batch, unsigned long long int seed);
int CHAMAMD_dplrnt_batch(CHAMAMD_batch_t *batch, unsigned long long int seed);
/* factor / solve / both; the blocking forms fill info[i] (may be NULL)
 * with the first non-positive pivot column of system i (0: success) and
 * return the first nonzero one in system order.  Systems that failed to
 * factor are left out of the solve. */
int CHAMAMD_dpotrf_batch(CHAMAMD_batch_t *batch, int *info);
int CHAMAMD_dpotrs_batch(CHAMAMD_batch_t *batch);
int CHAMAMD_dposv_batch(CHAMAMD_batch_t *batch, int *info);
int CHAMAMD_dpotrf_batch_Async(CHAMAMD_batch_t *batch, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);
int CHAMAMD_dpotrs_batch_Async(CHAMAMD_batch_t *batch, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);
int CHAMAMD_dposv_batch_Async(CHAMAMD_batch_t *batch, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);
/* per-system info of the last factorisation (after the sequence completed) */
int CHAMAMD_Batch_Info(CHAMAMD_batch_t *batch, int *info);
/* Mixed-precision solve of every system with FP64 refinement (LAPACK
 * dsposv semantics per system; prec = CHAMAMD_PREC_FP32: the factorisation
 * of a working copy multiplies FP32 operands with FP32 accumulation in its
 * left-looking updates, the 64 x 64 diagonal blocks stay FP64).  Residuals
 * and the stopping test ||r|| <= ||x|| ||A|| eps sqrt(n) run per system on
 * the device, at most 30 corrections; systems whose reduced-precision
 * factorisation failed or did not converge are factored and solved in FP64.
 * B receives the solutions; iter[i] = refinement steps, -3 or -31 (may be
 * NULL); info[i] = the FP64 fallback's info (0 otherwise).  A keeps its
 * values for converged systems and holds the FP64 factor for the others.
 * Returns the first nonzero info in system order, or 0. */
int CHAMAMD_dposv_mixed_batch(CHAMAMD_batch_t *batch, int prec, int *iter, int *info);
/* Size buckets of the batched factorisation, by padded order: <= 64, 128,
 * 256, 512, 1024, larger (6 buckets).  Each bucket is factored with the
 * workgroup shape -- 4 or 8 wavefronts per system -- of a process-wide
 * table.  CHAMAMD_Batch_Tune times both shapes on a sample of every
 * bucket `batch` holds (bumped dplgsy matrices of the same orders, in
 * scratch memory: the batch's matrices are untouched), keeps the faster
 * in the table and copies the table to waves[6] (may be NULL).  It runs by
 * itself on a batch's first factorisation while CHAMELEON_AUTOTUNING is
 * enabled (setting CHAMELEON_TILE_SIZE disables it, as in the reference).
 * Get/Set read and write the table (and the buckets' upper bounds, and
 * whether each was tuned in this process). */
int CHAMAMD_Batch_Tune(CHAMAMD_batch_t *batch, int *waves);
int CHAMAMD_Batch_Get_Tuning(int *waves, int *bucket_top, int *tuned);
int CHAMAMD_Batch_Set_Tuning(const int *waves);
/* on = 1 (the default, or CHAMAMD_BATCH_GRAPHS): a batch's factorisation
 * launches and its solve launch are captured into HIP graphs on first use
 * and replayed; 0: issued directly every time.  Results are the same. */
int CHAMAMD_Set_Batch_Graphs(int on);

/* ------------------------------------------------------------------ */
/* Solve server with continuous batching: requests A X = B arrive over time;
 * factors stay resident in a device pool of a fixed HBM budget (least
 * recently used factors no live request needs are evicted); work is
 * admitted onto free lanes (CHAMAMD_SERVE_LANES streams, default 4) as they
 * and memory free up, while earlier work runs -- new small systems
 * factored together, large ones one per lane, ready small requests solved
 * in one batched launch, large ones by the strip solve.  No counterpart in
 * the reference beyond its sequence/request interface
 * (include/chameleon/runtime_struct.h:78-103), which has no admission
 * control, residency or latency accounting.  Host-thread model as the
 * rest of the API: one thread calls in. */
typedef struct CHAMAMD_server_s CHAMAMD_server_t;
typedef struct {
    long long matrix_id;      /* requests with the same id share one matrix and its factor */
    int n, nrhs;
    const double *A;          /* host matrix (lower triangle read, column-major, lda), or NULL: */
    int lda;                  /*   A = dplgsy(bump = n, n x n, seed) generated on the device    */
    unsigned long long seed;
    double *B;                /* host right-hand sides in, solution out (column-major, ldb)   */
    int ldb;
} CHAMAMD_request_t;
typedef struct {
    long long completed, rejected, factored, reused, resident, pool_used_bytes;
    double p50_ms, p99_ms, max_ms;   /* submit-to-solution latency of completed requests      */
    double flops;                    /* factor + solve flops served (Chameleon's flop model)   */
    long long mixed_solves;          /* large requests solved on a mixed-precision factor      */
    long long ir_steps;              /* refinement corrections over those solves               */
    long long fallbacks;             /* mixed factorisations or refinements redone in FP64     */
    long long max_jobs;              /* most lanes busy at once                                */
    double max_berr;                 /* largest backward error ||r||/(||A|| ||x|| N eps) of a
                                        refined solution                                       */
} CHAMAMD_server_stats_t;

int CHAMAMD_Server_Create(CHAMAMD_server_t **server, size_t hbm_budget_bytes);
int CHAMAMD_Server_Destroy(CHAMAMD_server_t **server);
/* Products of the factorisations of large systems (order > 1024) made
 * after the call: -1 (the default) FP64; CHAMAMD_PREC_FP32, _BF16,
 * _BF16X3 or _FP8 low-precision trailing products with FP64 iterative
 * refinement of every solve on the FP64 matrix kept beside the factor
 * (LAPACK dsposv's stopping test; a matrix whose factorisation or
 * refinement fails is factored again in FP64).  Small systems are
 * factored in FP64.  Returns CHAMELEON_SUCCESS or
 * CHAMELEON_ERR_ILLEGAL_VALUE. */
int CHAMAMD_Server_Set_Precision(CHAMAMD_server_t *server, int prec);
int CHAMAMD_Server_Submit(CHAMAMD_server_t *server, const CHAMAMD_request_t *req, long long *ticket);
/* Non-blocking: collect finished device work, admit pending requests into
 * free memory and lanes, start what can start; *processed = requests
 * finished during the call.  B of a finished request holds its solution. */
int CHAMAMD_Server_Progress(CHAMAMD_server_t *server, int *processed);
/* Progress until every submitted request has finished (waiting on the
 * device in between); requests that can never be placed finish with
 * CHAMELEON_ERR_OUT_OF_RESOURCES. */
int CHAMAMD_Server_Drain(CHAMAMD_server_t *server);
/* *done = 1 once the request finished; *status = 0, the not-positive-
 * definite column, or a negative error (CHAMELEON_ERR_OUT_OF_RESOURCES
 * when its factor cannot fit the budget) */
int CHAMAMD_Server_Test(CHAMAMD_server_t *server, long long ticket, int *done, int *status);
/* tickets of requests finished since the last call, oldest first: up to
 * `cap` of them into tickets[0 .. *count-1] (the rest on the next call) */
int CHAMAMD_Server_Finished(CHAMAMD_server_t *server, long long *tickets, int cap, int *count);
int CHAMAMD_Server_Stats(CHAMAMD_server_t *server, CHAMAMD_server_stats_t *stats);
int CHAMAMD_Server_Reset_Stats(CHAMAMD_server_t *server);

/* dpotrs that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as CHAMELEON_dpotrs_Tile_Async, with the
 * solve's scratch (64 x 64 diagonal-block inverses, 128 x 128 strip
 * inverses of a narrow solve or the strip solve's predecessor products,
 * transposed right-hand sides, publication flags) in a fixed device array
 * of the library (128 MiB + 64 KiB per device, placed when CHAMELEON_Init
 * loads the code object and held until the process ends) instead of
 * allocated workspaces.  A and B must already be resident on the device
 * (left there by an earlier _Tile_Async call), the padded tile size a
 * multiple of 64, and N_padded * (192 + NRHS_padded) <= 16777216 (e.g.
 * N = 32768 with NRHS <= 320); otherwise
 * CHAMELEON_ERR_ILLEGAL_VALUE or
 * CHAMELEON_ERR_OUT_OF_RESOURCES, nothing launched.  Calls that overlap in
 * time on one device share the array and race. */
int CHAMAMD_dpotrs_Tile_Async_NoAlloc(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B,
                                      RUNTIME_sequence_t *sequence, RUNTIME_request_t *request);

#ifdef __cplusplus
}
#endif

#endif
