/*
 * internal.h -- library-private state of libchameleon_amd.
 *
 * Device mirror.  Every descriptor used on the GPU gets a mirror of its
 * (sub)matrix in HBM: mt x nt tiles stored tile after tile in column-major
 * tile order, each tile padded to ldt x wdt elements (ldt = mb rounded up to
 * a multiple of 16, wdt likewise unless the matrix is a single tile column,
 * where wdt = n rounded up to 16).  Padding is zero off the diagonal tiles
 * and the identity on the padded diagonal of diagonal tiles, so the padded
 * matrix has the same Cholesky factor as the real one with identity blocks
 * in the holes; every kernel therefore works on full ldt x ldt tiles.
 *
 * A mirror is either in normal orientation (tile (m,n) holds A(m,n)) or
 * transposed (tile (m,n) holds A(n,m)^T).  The Upper variants of the
 * Cholesky path run the Lower kernels on a transposed mirror: for a
 * symmetric A, the upper factor U satisfies U = L^T.
 */
#ifndef CHAMELEON_AMD_INTERNAL_H
#define CHAMELEON_AMD_INTERNAL_H

#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <vector>
#include "chameleon.h"
#include "chameleon_amd.h"

namespace chx {

void fatal_hip(hipError_t e, const char *what, const char *file, int line);
/* info value left by a kernel whose bounded wait for another workgroup ran
 * out (an internal error, never a property of the matrix) */
#define PANEL_STALLED (-(1 << 30))
/* info value of a low-precision demotion that met a non-finite entry */
#define LP_RANGE (-(1 << 29))
/* maximum that propagates NaN, as LAPACK's norms do (dlange/dlansy check
 * every entry with disnan): a NaN operand makes the result NaN */
__host__ __device__ inline double nan_max(double a, double b)
{
    return (a != a || b != b) ? a + b : (a > b ? a : b);
}
__host__ __device__ inline float nan_maxf(float a, float b)
{
    return (a != a || b != b) ? a + b : (a > b ? a : b);
}
#define CHX_HIP(call) do { hipError_t _e = (call); if (_e != hipSuccess) ::chx::fatal_hip(_e, #call, __FILE__, __LINE__); } while (0)

void latms_release();      /* frees the conditioned generator's kept factor */
/* looks up the no-allocation solve's scratch array on `device` (current)
 * and clears its flags on `s` (noalloc.hip) */
void noalloc_init(int device, hipStream_t s);
void error(const char *func, const char *msg);
void warning(const char *func, const char *msg);

enum MirrorState { MIRROR_NONE = 0, MIRROR_NORMAL = 1, MIRROR_TRANS = 2 };

struct Mirror {
    void   *d = nullptr;       /* device tiles                              */
    int     mt = 0, nt = 0;    /* tiles of the submatrix                    */
    int     ldt = 0, wdt = 0;  /* padded tile rows / columns                */
    size_t  tile_elems = 0;    /* ldt * wdt                                 */
    size_t  bytes = 0;
    int     eltsize = 8;
    int     state = MIRROR_NONE;  /* content valid on device, orientation   */
    bool    host_valid = true;    /* host copy up to date                   */
    bool    diag_identity = false;/* padding of diagonal tiles is identity  */
    bool    lower_only = false;   /* only tiles m >= n hold the matrix (the
                                     upper ones were never uploaded)       */
};

/* Where a factorisation or a strip solve runs: the critical-path stream,
 * an optional bulk-update stream (null: everything on the critical stream,
 * as without lookahead), the events that order the two, the fused panel
 * kernel's ticket counters and the strip solves' publication flags (tagged
 * by launch epoch, so they never need clearing between launches).  The
 * context owns one lane for the library's own calls; the solve server owns
 * one per factorisation or solve it runs at the same time. */
struct Lane {
    hipStream_t sm = nullptr, su = nullptr;
    std::vector<hipEvent_t> events;
    int *flags = nullptr;
    int flags_n = 0;
    int epoch = 0;
    unsigned *panel_ctr = nullptr;
};
Lane *main_lane();
hipEvent_t lane_event(Lane *L, size_t i);
/* at least n flags and the first of `launches` fresh epochs */
int *lane_flags(Lane *L, int n, int launches, int *epoch);
/* four counters: [0..1] the fused panel kernel's tickets, [2] the strip
 * solves' task ticket */
unsigned *lane_counters(Lane *L);
void lane_release(Lane *L);         /* frees flags, counters and events (not the streams) */

struct Context {
    int device = 0;
    int nb = 384;                 /* CHAMELEON_TILE_SIZE default            */
    int ib = 48;
    int translation = ChamInPlace;
    int lookahead = 1;
    int fwd_tail = -1;            /* dposv: tile rows solved forward after potrf (-1: chosen) */
    double ir_target = 0.0;       /* mixed solve: backward-error target (0: LAPACK's test)  */
    int mixed_scaling = 1;        /* mixed solve: symmetric diagonal scaling before demotion */
    double tile_accuracy = 1e-12; /* per-tile precision choice: requested accuracy u_high    */
    long long tile_counts[3] = {0, 0, 0};   /* last per-tile choice: FP64, FP32, BF16 tiles */
    int autotuning = 1;
    int warnings = 1;
    int nworkers = 1;
    hipStream_t s_main = nullptr;     /* critical path / default work     */
    hipStream_t s_update = nullptr;   /* trailing updates (lookahead)      */
    int        *d_info = nullptr;     /* device info flag (one int)        */
    /* workspace reused across calls, grown on demand */
    void  *ws = nullptr;
    size_t ws_bytes = 0;
    void  *ws2 = nullptr;
    size_t ws2_bytes = 0;
    void  *pinned = nullptr;
    size_t pinned_bytes = 0;
    Lane lane;                        /* s_main / s_update and their events, flags, counters */
    CHAM_desc_t *stream_flush = nullptr;  /* a synchronous Lower potrf/posv of this column-major
                                             matrix downloads its factor while it runs */
    /* kernel statistics (CHAMELEON_GENERATE_STATS) */
    bool stats = false;
    struct Pending { int cls; hipEvent_t a, b; double flops; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> timing_pool;
    int panel = 1;                  /* fused panel kernel (CHAMAMD_PANEL=0: POTRF + TRSM)   */
    int pairs = 1;                  /* paired bulk updates (CHAMAMD_PAIRS=0: one per step)  */
    void *latms_cache = nullptr;    /* orthogonal factor kept by the conditioned generator   */
    long long st_launches[CHAMAMD_KERNEL_NCLASS] = {};
    double st_ms[CHAMAMD_KERNEL_NCLASS] = {};
    double st_flops[CHAMAMD_KERNEL_NCLASS] = {};
};

Context *ctx();                       /* nullptr when not initialised */
inline Context *ctx_ptr() { return ctx(); }

/* Bracket the launches issued by `body` on stream s with timing events
 * when statistics are enabled. */
hipEvent_t timing_event();
void stats_fold();
template <typename F>
inline void timed(int cls, double flops, hipStream_t s, F body)
{
    Context *c = ctx_ptr();
    if (!c->stats) { body(); return; }
    hipEvent_t a = timing_event(), b = timing_event();
    (void)hipEventRecord(a, s);
    body();
    (void)hipEventRecord(b, s);
    c->pending.push_back({cls, a, b, flops});
}

void *workspace(size_t bytes);        /* context workspace of >= bytes */
void *workspace2(size_t bytes);       /* second, independent workspace     */
int *strip_flags(int n, int launches, int *epoch);
unsigned *panel_counters();

Mirror *mirror_of(const CHAM_desc_t *A);
/* Make the device mirror of A hold its current values in orientation
 * `state` (MIRROR_NORMAL or MIRROR_TRANS); uploads from host if needed.
 * lower_need: the caller reads and writes only the tiles m >= n (a Lower
 * Cholesky), so a fresh upload of a column-major matrix may move those
 * only; any later full acquire uploads the rest. */
int mirror_acquire(CHAM_desc_t *A, int state, bool diag_identity, hipStream_t s, bool lower_need = false);
/* a mirror holding the lower tiles only gets the others from the host */
void mirror_complete(CHAM_desc_t *A, hipStream_t s);
/* Allocate the mirror (without uploading) for a matrix that the next
 * operation overwrites entirely. */
int mirror_prepare_overwrite(CHAM_desc_t *A, int state, bool diag_identity);
/* Mark device content as the latest (host stale). */
void mirror_mark_written(CHAM_desc_t *A);
/* Copy device content back to the host tiles (normal orientation). */
int mirror_flush(CHAM_desc_t *A, hipStream_t s);
void mirror_release(CHAM_desc_t *A);

int desc_check(const CHAM_desc_t *A);
bool desc_is_cm(const CHAM_desc_t *A);     /* user column-major storage (CM) */

/* transfer.cpp: column-panel transfers of column-major descriptors */
#define XFER_ALL 0
#define XFER_LOWER 1
#define XFER_UPPER 2
bool xfer_eligible(const CHAM_desc_t *A, const Mirror *M);
void xfer_upload(CHAM_desc_t *A, Mirror *M, int part, hipStream_t s);
void xfer_download(CHAM_desc_t *A, Mirror *M, int part, hipStream_t s);
void xfer_release();
/* Streaming flush of a Lower factorisation of a column-major descriptor:
 * begin (after A's upload, lower tiles only), then the factorisation
 * reports that its first ncols tile columns are final in the order of
 * stream s (no-op unless the streamer serves that mirror), end waits for
 * the download of every column (returns false when A is not streamed),
 * drop abandons it (A null: whatever is streamed). */
bool xfer_stream_begin(CHAM_desc_t *A, Mirror *M);
void xfer_stream_columns(const double *mirror, int ncols, hipStream_t s);
bool xfer_stream_end(CHAM_desc_t *A, hipStream_t s);
void xfer_stream_drop(const CHAM_desc_t *A);
int request_fail(RUNTIME_sequence_t *seq, RUNTIME_request_t *req, int status);

/* sequences: device info slot for potrf of this sequence */
struct SeqState {
    int *d_info = nullptr;      /* device int, first failing minor        */
    int *h_info = nullptr;      /* pinned host copy                       */
    bool pending_info = false;
};
SeqState *seq_state(RUNTIME_sequence_t *seq);
int sequence_wait_internal(RUNTIME_sequence_t *seq);

/* ---- scheduler pieces shared by the FP64 and mixed-precision paths ---- */
struct PotrfPlan {
    double *A;
    int NT, ldt, nb, nvalid_last;   /* nb: descriptor tile size (valid columns of a full tile) */
    int *d_info;
    /* mixed precision: -1 = FP64 trailing updates; 0/1/2 = FP32/BF16/BF16x3
     * operands, demoted panel alternating between P[0] and P[1] */
    int lp_mode = -1;
    void *P[4] = {nullptr, nullptr, nullptr, nullptr};   /* panel k in P[k % 2], or P[k % 4] with paired updates */
    /* per-tile precision (lp_mode = CHAMAMD_PREC_ADAPTIVE): FP32 panels in P,
     * BF16 panels in P2, the precision of tile (m,n) at pmap[m + n*NT] */
    void *P2[4] = {nullptr, nullptr, nullptr, nullptr};
    const unsigned char *pmap = nullptr;
    /* called where the main stream is about to wait for a bulk update
     * (with lookahead it usually has time to spare there), when panels
     * 0..k are issued on it: work the callback issues on the main stream
     * runs once they are final (their tiles, and the 64 x 64 inverses of
     * their diagonal blocks in the potrf workspace) */
    void (*after_panels)(void *arg, int k) = nullptr;
    void *after_arg = nullptr;
    /* where it runs (null: the context's lane) and where the 64 x 64
     * diagonal-block inverses go (null: the context workspace; the solves
     * of the same factor read them there) */
    Lane *lane = nullptr;
    double *dinv = nullptr;
    /* potrf_solve_run with at most 8 right-hand sides: room for the 128 x 128
     * strip inverses (NT*ldt*128 doubles), or null for the MFMA strip solve */
    double *d128 = nullptr;
    /* potrf_solve_run, MFMA strip solve: room for the predecessor products
     * (2 x NT*ldt*64 doubles), or null to solve without them */
    double *epred = nullptr;
};
void potrf_run(const PotrfPlan &p);
/* both strip solves of BT (nrp x NT*ldt, transposed right-hand sides) with
 * the Lower factor A and its 64 x 64 diagonal-block inverses, on lane L;
 * ldt a multiple of 64 */
void solve_strips_lane(Lane *L, const double *A, const double *dinv, int NT, int ldt, double *BT, int nrp,
                       const int *d_info, const double *d128 = nullptr, int ncols = 0, const double *ef = nullptr,
                       const double *eb = nullptr);
/* potrf_run with the forward solve of BT (transposed right-hand sides,
 * nrp x NT*ldt, nrhs of them in use) issued beside the factorisation, then
 * the forward strips it left and the backward solve: BT <- A^-1 BT.  The
 * solve kernels return at once when the device info flag is set.  `rate`:
 * the bulk update rate (flop/s) the overlap is sized for.  Requires
 * potrf_solve_eligible(ldt). */
bool potrf_solve_eligible(int ldt);
void potrf_solve_run(PotrfPlan p, double *BT, int nrp, int nrhs, double rate);
struct SolvePlan {
    double *TT, *TN, *BT, *W, *dinv;   /* diagonal-tile inverses, transposed RHS, one block, 64x64 inverses */
    double *d128 = nullptr;            /* 128 x 128 strip inverses (narrow solves), or null */
    double *ef = nullptr, *eb = nullptr;   /* predecessor products (MFMA strip solve), or null */
    int ncols = 0;                     /* right-hand-side columns in use */
    int R = 0;                         /* 64: strip solve; 0: tile-step solve */
};
/* ncols: columns of the right-hand sides in use (at most nrp); with few of
 * them the plan also holds the 128 x 128 strip inverses */
SolvePlan solve_plan(double *A, int NT, int ldt, int nrp, int *d_info, int ncols = 0);
void solve_bt(const SolvePlan &p, double *A, int NT, int ldt, double *BT, int nrp, int *d_info);

/* one system of a batched generation (kernels/batch.hip) */
struct BatchGenSys {
    long long off;                 /* first element of the system                    */
    long long base;                /* origin of the element rule                     */
    unsigned long long seed, abig, cbig;   /* abig, cbig: the LCG map of bigM steps  */
    double bump;
    int ld, rows, cols;            /* leading dimension; rows / columns written      */
    int nvr, nvc;                  /* valid rows / columns (the rest is padding)     */
    int bigM;
    int pad_one;                   /* plgsy: 1 on the padded diagonal (else 0)       */
};

/* ---- kernel launchers (kernels directory) ---- */
namespace k {
/* util.hip */
void fix_padding(void *d, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N, int eltsize,
                 bool diag_identity, hipStream_t s);
void transpose_tiles_inplace(double *d, int nt, int ldt, hipStream_t s);
void plgsy(double *d, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N, double bump, int uplo_skip,
           int bigM, int m0, int n0, unsigned long long seed, int trans, bool diag_identity, hipStream_t s);
void plrnt(double *d, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N, int bigM, int m0, int n0,
           unsigned long long seed, int trans, hipStream_t s);
void set_identity(double *d, int ntiles, int ldt, hipStream_t s);
void norm_partials(const double *d, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N, int uplo,
                   int trans, double *colsum, double *rowsum, double *w, hipStream_t s);
void lacpy_tiles(const double *src, double *dst, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N,
                 int uplo, hipStream_t s);
/* chol.hip */
void potrf_diag(double *Akk, int ldt, double *dinv, int iinfo, int nvalid, int *info, int waves, hipStream_t s);
void panel_factor(double *T, double *P, long long ts, int S, int ldt, double *dinv, int *flags, unsigned *ctr,
                  int epoch, int *info, int iinfo, int nvalid, hipStream_t s);
void trsm_right_lower_trans(double *panel, int ntiles, size_t tile_stride, int ldt, const double *Lkk,
                            const double *dinv, const int *info, hipStream_t s);
void trsm_right_lower_trans_batched(double *P0, int ntiles, long long pstride, int ldt, const double *L0,
                                    long long lstride, const double *dinv0, long long dstride, const int *info,
                                    hipStream_t s);
/* Trailing updates cover the tile pairs n <= m < NT of the columns n0 .. n1-1
 * (n1 clamped to NT).  update_pairs returns how many pairs that is and sets
 * *ncol to the launch mode: the column count, or 0 for the whole triangle
 * n0 .. NT-1 (n1 >= NT and more than one column). */
/* 1: the trailing-update kernels apply C -= product as FP64 atomic adds
 * (same result, bit for bit); 0 (CHAMAMD_ATOMIC_UPDATE=0): load and store */
int atomic_tile_updates();
int update_pairs(int n0, int n1, int NT, int *ncol);
/* pmap (NT x NT bytes, tile (m,n) at m + n*NT, or null): when given, only
 * the tiles marked 0 (FP64) are updated; the low-precision updates take
 * the tiles marked 1 (FP32) and 2 (BF16) */
void trailing_update(double *A, int lmt, int ldt, int k, int n0, int n1, int NT, const int *info, hipStream_t s,
                     int kp = 1, const unsigned char *pmap = nullptr);
void dinv_all(const double *A, int nt, size_t diag_stride, int ldt, double *dinv, hipStream_t s);
/* waves: 4 or 8 wavefronts per workgroup (one workgroup per job); lp: the
 * left-looking updates multiply FP32 operands with FP32 accumulation (the
 * 64 x 64 diagonal blocks, their inverses and the solves below stay FP64) */
void potrf_jobs(double *A, const long long *off, const int *np, const int *ld, const int *nvalid, double *dinv,
                const long long *doff, int count, int *queue, int *info, int grid, hipStream_t s, int waves = 4,
                bool lp = false);
/* occupancy queries of the persistent batch kernels, made before any
 * stream capture (they are host calls, not stream work) */
void potrf_jobs_prepare();
int batch_solve_resident();
/* batch.hip */
void batch_gen(double *A, const BatchGenSys *sys, const long long *unit0, long long units, int count, int sym,
               hipStream_t s);
void batch_solve(const double *A, const long long *aoff, const int *np, const int *ld, const double *dinv,
                 const long long *doff, double *B, const long long *boff, const int *nrhs, const int *info, int count,
                 int *queue, int grid, hipStream_t s);
void zero_ints(int *p, int n, hipStream_t s);
/* refinement pieces of the batched mixed-precision solve (batch.hip) */
/* dst = src on the lower 64-column blocks of every system (the part the
 * batched factorisation reads and writes) */
void batch_copy_lower(const double *src, double *dst, const long long *aoff, const int *np, int count, hipStream_t s);
void batch_residual(const double *A, const long long *aoff, const int *np, const int *n, const double *B,
                    const double *X, double *R, const long long *boff, const int *nrhs, double *anrm, int *state,
                    int *iters, int *unconv, double *berr, int count, int *queue, hipStream_t s);
void batch_update(double *X, const double *R, const long long *boff, const int *np, const int *nrhs,
                  const int *state, int count, int add, hipStream_t s);
void batch_refine_init(const int *info, int *state, int *iters, double *anrm, int count, hipStream_t s);
void tile_pad_diag(double *d, int mt, int nt, int ldt, int wdt, int mb, int nb, int M, int N, int mode,
                   hipStream_t s);
/* dense rows x cols column panel (rows r0.., columns c0.. of tile column n)
 * to (to_tiles = 1) or from (0) the tiles of a mirror */
void panel_tiles(double *P, double *D, int mt, int ldt, size_t tile_elems, int mb, int n, int r0, int rows, int c0,
                 int cols, int to_tiles, hipStream_t s);
/* mixed.hip */
void panel_demote(const double *A, int NT, int ldt, int k, int m0, int m1, void *P, int mode, hipStream_t s);
/* FP8: tiles m0..m1-1 of panel k as E4M3 bytes with a power-of-two scale
 * per tile; a tile with a non-finite entry sets *info = LP_RANGE */
void panel_demote_fp8(const double *A, int NT, int ldt, int k, int m0, int m1, void *P, int *info, hipStream_t s);
/* symmetric diagonal scaling: sc[i] = a_ii^-1/2 over the NT*ldt padded rows,
 * lower tiles of A scaled by sc_i sc_j; scale_bt: rows of transposed
 * right-hand sides (element (c, r) at c + r*nrp) times sc[r] */
void diag_scaling(double *A, int NT, int ldt, double *sc, hipStream_t s);
void diag_rsqrt(const double *A, int NT, int ldt, double *sc, hipStream_t s);   /* sc only */
void scale_bt(double *BT, int nrp, int rows, const double *sc, hipStream_t s);
void trailing_update_lp(double *A, const void *P, const void *P2, int mode, int lmt, int ldt, int n0, int n1, int NT,
                        const int *info, hipStream_t s, const unsigned char *pmap = nullptr);
void tile_sumsq_lower(const double *A, int NT, int ldt, double *ss, hipStream_t s);
void residual(const double *A, int NT, int ldt, const double *XT, const double *BT, double *RT, int nrp, double *part,
              hipStream_t s);
size_t residual_scratch(int NT, int ldt, int nrp);   /* doubles of `part` */
/* column maxima of X into out[0..ncols) and of R into out[nrp..nrp+ncols)
 * (max-combined into out: clear it first) */
void colmax2(const double *X, const double *R, int nrp, int ncols, int nrows, double *out, hipStream_t s);
void axpy(double *Y, const double *X, size_t n, hipStream_t s);
/* lower tiles of S copied to W (with sc: scaled by sc_i sc_j as
 * diag_scaling does, in the same pass) and the symmetric inf-norm's row
 * sums of S into rowsum[N]; returns -1 (nothing launched) unless ldt is a
 * multiple of 128 up to COPY_NORM_MAX_LDT and the tiling of N by mb fills
 * NT tile rows */
#define COPY_NORM_MAX_LDT 2048
int copy_lower_norm(const double *S, double *W, int NT, int ldt, int mb, int N, double *part, double *rowsum,
                    hipStream_t s, const double *sc = nullptr);
size_t copy_lower_norm_scratch(int NT, int ldt);     /* doubles of `part` */
/* latms.hip: square tile mirrors of NT x NT tiles, ldt x ldt, mb valid rows per tile */
void gauss(double *G, int NT, int ldt, int mb, int N, unsigned long long seed, hipStream_t s);
void transpose_full(const double *src, double *dst, int NT, int ldt, hipStream_t s);
void scale_cols(const double *Q, double *W, int NT, int ldt, const double *sc, hipStream_t s);
void symmetrize(double *A, int NT, int ldt, hipStream_t s);
void add_diag(double *A, int NT, int ldt, int mb, int N, double shift, hipStream_t s);
void sumsq_tiles(const double *A, int ntiles, int ldt, double *part, hipStream_t s);
/* C(m,n) = beta C(m,n) + alpha sum_{k0<=k<k1} A(m,k) B(n,k)^T over the tiles
 * r0 <= m < r0+nr, c0 <= n < c0+nc, or (lower) m >= n of that square */
void tile_gemm_nt(const double *A, int amt, const double *B, int bmt, double *C, int cmt, int ldt, int k0, int k1,
                  int r0, int nr, int c0, int nc, int lower, double alpha, double beta, hipStream_t s);
/* solve.hip */
/* ticket: one device counter, cleared by the call before each launch and
 * used by that launch only (launches on one stream may share it) */
/* With D128 (dinv128 of the same factor) and 1 <= ncols <= 8 columns in
 * use with ldt a multiple of 128, the narrow kernel solves the first ncols
 * columns only (128-row strips, vector FMAs); otherwise the MFMA strip
 * kernel solves all nrp columns. */
/* E (optional, MFMA kernel only): the predecessor products of this
 * direction from strip_products (Ef forward, Eb backward). */
int solve_strips(const double *A, const double *D, long long dts, long long doff, int ldd, double *BT, int NT,
                 int ldt, int nrp, int trans, int *flags, int epoch, unsigned *ticket, const int *info, hipStream_t s,
                 int s0 = 0, const double *D128 = nullptr, int ncols = 0, const double *E = nullptr);
/* predecessor products of the 64-row strips (NT*ldt*64 doubles each at Ef
 * and Eb): Ef_s = D_s L(s, s-1), Eb_s = (L(s+1, s) D_s)^T */
void strip_products(const double *A, const double *D, long long dts, long long doff, int ldd, int NT, int ldt,
                    double *Ef, double *Eb, hipStream_t s);
bool narrow_eligible(int ldt, int ncols);
/* 128 x 128 inverses of the diagonal blocks of the 128-row strips (NT*ldt*128
 * doubles at D128) from the 64 x 64 inverses D (addressed as in solve_strips) */
void dinv128(const double *A, const double *D, long long dts, long long doff, int ldd, int NT, int ldt,
             double *D128, hipStream_t s);
void solve_coop(const double *A, const double *TN, const double *TT, double *BT, double *W, int NT, int ldt, int nrhs,
                int trans, const int *info, hipStream_t s);
/* with info != null, nothing is copied once *info is nonzero */
void rhs_transpose(double *Bm, double *BT, int NT, int ldt, int wdt, int bnt, int to_bt, hipStream_t s,
                   const int *info = nullptr);
/* forward substitution of one tile row right after its panel (dposv) */
/* the forward tile solve addresses a tile row of the factor and of the
 * right-hand sides with 32-bit buffer offsets */
inline bool fwd_tile_fits(int ldt, int nrp)
{
    return (size_t)ldt * ldt * 8 < ((size_t)1 << 31) && (size_t)ldt * nrp * 8 < ((size_t)1 << 31);
}
void fwd_tile(const double *A, const double *D, double *BT, int k, int NT, int ldt, int nrp, const int *info,
              hipStream_t s);
void fwd_update(const double *A, double *BT, int k, int NT, int ldt, int nrp, const int *info, hipStream_t s);
void tile_transpose(const double *src, double *dst, int ntiles, int ldt, hipStream_t s);
void gemm_batched(int transA, int M, int N, int K, double alpha, const double *A, int lda, long long strideA,
                  const double *B, int ldb, long long strideB, double beta, double *C, int ldc, long long strideC,
                  int batch, double *copy_dst, const int *info, hipStream_t s);
}

} /* namespace chx */

#endif
