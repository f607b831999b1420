/*
 * batch.cpp -- many independent SPD systems on one device.
 *
 * Two front doors:
 *
 *  - Chameleon's tile batch (compute/zpotrf_batch.c:89-118,
 *    compute/zplghe_batch.c:103-139): every tile of a descriptor is an
 *    independent tempnn x tempnn matrix (tempnn = width of its tile
 *    column).  CHAMELEON_dplgsy_batch_Tile generates tile (m,n) with the
 *    element rule of CORE_dplgsy at origin (n*mt + m)*mb*nb and bigM = 0,
 *    CHAMELEON_dpotrf_batch_Tile factors every tile in place.
 *
 *  - Uneven batches (no counterpart in the reference): systems of any
 *    orders n_i with nrhs_i right-hand sides in one device-resident
 *    CHAMAMD_batch_t.  Systems are stored largest first (the order the
 *    persistent kernels take them from their queues), each padded to a
 *    multiple of 16 with identity, so one kernel shape serves all sizes.
 *
 * Factorisations use the job-queue form of the diagonal-tile POTRF kernel
 * (one workgroup per system), solves the batched strip solve of
 * kernels/batch.hip.  A system is computed by one workgroup in a fixed
 * order of operations, so its result does not depend on the rest of the
 * batch: solved alone or in any mix it is bitwise the same.
 *
 * Size buckets and tuning.  Systems fall in buckets by padded order
 * (<= 64, 128, 256, 512, 1024, larger); each bucket is factored with the
 * workgroup shape (4 or 8 wavefronts per system) of a process-wide table.
 * The table starts from the shapes measured fastest on MI355X and
 * CHAMAMD_Batch_Tune (run by itself on a batch's first factorisation while
 * CHAMELEON_AUTOTUNING is on) times both shapes on a sample of every
 * bucket the batch holds and keeps the faster.  Consecutive buckets of
 * one shape go in one persistent launch; launches of the two shapes
 * alternate between the two streams, so a bucket's launch fills the
 * device while the previous one drains.
 *
 * Graphs.  The factorisation launches and the solve launch of a batch are
 * captured once each (hipGraph; their queue counters are cleared by a
 * kernel, not a memset node) and replayed by every later call.
 */
#include "internal.h"
#include "chameleon_amd.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

using namespace chx;

#define NBUCKET 6
static const int kBucketTop[NBUCKET] = {64, 128, 256, 512, 1024, 1 << 30};
/* device counters of a batch: [0] unused, [1] solve queue, [2 .. 2+NBUCKET)
 * factorisation queues, then the refinement's residual queue and its
 * count of systems not yet converged */
#define QUEUE_INTS (4 + 2 * NBUCKET)
#define Q_RES (2 + 2 * NBUCKET)
#define Q_UNCONV (3 + 2 * NBUCKET)

struct CHAMAMD_batch_s {
    int count = 0;
    int bstart[NBUCKET + 1] = {};              /* slots of bucket b: [bstart[b+1], bstart[b]) -- largest first */
    hipGraphExec_t g_potrf = nullptr, g_potrs = nullptr;
    /* mixed-precision solve: working factor, solution, residual, per system
     * ||A||_inf, refinement state, steps and backward error */
    double *dW = nullptr, *dX = nullptr, *dR = nullptr, *d_anrm = nullptr, *d_berr = nullptr;
    int *d_state = nullptr, *d_iters = nullptr;
    int g_waves[NBUCKET] = {};                 /* shapes the factorisation graph was captured with */
    bool tuned = false;
    std::vector<int> slot;                     /* user index -> storage slot          */
    std::vector<int> user;                     /* slot -> user index                  */
    std::vector<int> n, np, nrhs;              /* per slot                            */
    std::vector<long long> aoff, boff, doff;   /* per slot, in elements               */
    size_t a_elems = 0, b_elems = 0, d_elems = 0;
    double *dA = nullptr, *dB = nullptr, *dD = nullptr;
    int *d_np = nullptr, *d_n = nullptr, *d_nrhs = nullptr, *d_info = nullptr, *d_queue = nullptr;
    long long *d_aoff = nullptr, *d_boff = nullptr, *d_doff = nullptr;
};

namespace {

int rup16(int x) { return (x + 15) & ~15; }

/* the LCG map of `n` steps: x -> A x + C (coreblas/include/coreblas/random.h) */
void lcg_pow(unsigned long long n, unsigned long long *A, unsigned long long *C)
{
    unsigned long long a = 6364136223846793005ULL, c = 1ULL, ra = 1ULL, rc = 0ULL;
    while (n) {
        if (n & 1ULL) { rc = a * rc + c; ra = a * ra; }
        c = c * (a + 1ULL);
        a = a * a;
        n >>= 1;
    }
    *A = ra;
    *C = rc;
}

template <typename T>
T *upload(const std::vector<T> &v)
{
    T *d = nullptr;
    if (v.empty()) return nullptr;
    CHX_HIP(hipMalloc(&d, v.size() * sizeof(T)));
    CHX_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

/* Generate systems described by `sys` (device-side list built here). */
void run_gen(double *base, std::vector<BatchGenSys> &sys, int sym, hipStream_t s)
{
    std::vector<long long> unit0(sys.size() + 1, 0);
    for (size_t i = 0; i < sys.size(); ++i) {
        const long long runs = (sys[i].rows + 31) / 32;
        unit0[i + 1] = unit0[i] + runs * sys[i].cols;
    }
    BatchGenSys *d_sys = upload(sys);
    long long *d_u = upload(unit0);
    k::batch_gen(base, d_sys, d_u, unit0.back(), (int)sys.size(), sym, s);
    CHX_HIP(hipStreamSynchronize(s));
    CHX_HIP(hipFree(d_sys));
    CHX_HIP(hipFree(d_u));
}

BatchGenSys gen_sys(long long off, int ld, int rows, int cols, int nvr, int nvc, long long base, int bigM,
                    unsigned long long seed, double bump)
{
    BatchGenSys g;
    g.off = off;
    g.base = base;
    g.seed = seed;
    lcg_pow((unsigned long long)bigM, &g.abig, &g.cbig);
    g.bump = bump;
    g.ld = ld;
    g.rows = rows;
    g.cols = cols;
    g.nvr = nvr;
    g.nvc = nvc;
    g.bigM = bigM;
    g.pad_one = 1;
    return g;
}

bool check_batch(const char *fn, const CHAMAMD_batch_t *b)
{
    if (!ctx()) {
        error(fn, "CHAMELEON not initialized");
        return false;
    }
    if (!b) {
        error(fn, "NULL batch");
        return false;
    }
    return true;
}

int bucket_of(int np)
{
    int b = 0;
    while (np > kBucketTop[b]) ++b;
    return b;
}

/* Workgroup shape per bucket: wavefronts per system.  Defaults measured on
 * MI355X (CHAMAMD_Batch_Tune on the 20 000-system log-uniform batch). */
int g_bucket_waves[NBUCKET] = {4, 4, 4, 4, 4, 4};
bool g_bucket_tuned[NBUCKET] = {};

/* the factorisation launches of slots [s0, s1) with `waves`, of the
 * matrices at base (the batch's A, or the mixed solve's working copy) */
void potrf_range(CHAMAMD_batch_t *b, int s0, int s1, int waves, int q, hipStream_t s, double *base = nullptr,
                 bool lp = false)
{
    if (s1 <= s0) return;
    k::zero_ints(b->d_queue + q, 1, s);
    k::potrf_jobs(base ? base : b->dA, b->d_aoff + s0, b->d_np + s0, nullptr, b->d_n + s0, b->dD, b->d_doff + s0,
                  s1 - s0, b->d_queue + q, b->d_info + s0, 0, s, waves, lp);
}

/* Every bucket with its shape; runs of buckets with one shape in one
 * launch; launches alternate between the main and the update stream
 * (the update stream first waits for earlier main-stream work, and the
 * main stream for the update stream's launches at the end). */
void issue_potrf(CHAMAMD_batch_t *b, const int *waves, double *base = nullptr, bool lp = false)
{
    Context *c = ctx();
    hipStream_t sm = c->s_main, su = c->s_update;
    Lane *L = main_lane();
    CHX_HIP(hipEventRecord(lane_event(L, 0), sm));
    CHX_HIP(hipStreamWaitEvent(su, lane_event(L, 0), 0));
    int launch = 0, bk = 0;
    while (bk < NBUCKET) {
        /* buckets are stored largest first: bucket NBUCKET-1 at slot 0 */
        const int hb = NBUCKET - 1 - bk;
        if (b->bstart[hb] == b->bstart[hb + 1]) { ++bk; continue; }
        const int w = waves[hb];
        int lo = hb;
        while (lo - 1 >= 0 && (b->bstart[lo - 1] == b->bstart[lo] || waves[lo - 1] == w)) --lo;
        const int s0 = b->bstart[hb + 1], s1 = b->bstart[lo];
        potrf_range(b, s0, s1, w, 2 + launch, (launch & 1) ? su : sm, base, lp);
        ++launch;
        bk = NBUCKET - lo;
    }
    CHX_HIP(hipEventRecord(lane_event(L, 1), su));
    CHX_HIP(hipStreamWaitEvent(sm, lane_event(L, 1), 0));
}

void issue_potrs(CHAMAMD_batch_t *b, hipStream_t s)
{
    k::zero_ints(b->d_queue + 1, 1, s);
    k::batch_solve(b->dA, b->d_aoff, b->d_np, nullptr, b->dD, b->d_doff, b->dB, b->d_boff, b->d_nrhs, b->d_info,
                   b->count, b->d_queue + 1, 0, s);
}

/* graph replay of the batch launches: CHAMAMD_BATCH_GRAPHS (default 1),
 * CHAMAMD_Set_Batch_Graphs */
int g_graphs = -1;

bool graphs_on()
{
    if (g_graphs < 0) {
        const char *e = getenv("CHAMAMD_BATCH_GRAPHS");
        g_graphs = (e && *e) ? (atoi(e) != 0) : 1;
    }
    return g_graphs != 0;
}

void drop_graphs(CHAMAMD_batch_t *b)
{
    if (b->g_potrf) CHX_HIP(hipGraphExecDestroy(b->g_potrf));
    if (b->g_potrs) CHX_HIP(hipGraphExecDestroy(b->g_potrs));
    b->g_potrf = b->g_potrs = nullptr;
}

/* capture `issue` on the main stream into an executable graph */
template <typename F>
hipGraphExec_t capture(F issue)
{
    hipStream_t sm = ctx()->s_main;
    k::potrf_jobs_prepare();
    (void)k::batch_solve_resident();
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    CHX_HIP(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
    issue();
    CHX_HIP(hipStreamEndCapture(sm, &g));
    CHX_HIP(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
    CHX_HIP(hipGraphDestroy(g));
    return e;
}

int tune(CHAMAMD_batch_t *b, int *table);

int batch_potrf(CHAMAMD_batch_t *b, hipStream_t s)
{
    (void)s;
    if (ctx()->autotuning && !b->tuned) tune(b, nullptr);
    if (!graphs_on()) {
        issue_potrf(b, g_bucket_waves);
        return CHAMELEON_SUCCESS;
    }
    if (b->g_potrf && memcmp(b->g_waves, g_bucket_waves, sizeof b->g_waves) != 0) {
        CHX_HIP(hipGraphExecDestroy(b->g_potrf));
        b->g_potrf = nullptr;
    }
    if (!b->g_potrf) {
        memcpy(b->g_waves, g_bucket_waves, sizeof b->g_waves);
        b->g_potrf = capture([&] { issue_potrf(b, b->g_waves); });
    }
    CHX_HIP(hipGraphLaunch(b->g_potrf, ctx()->s_main));
    return CHAMELEON_SUCCESS;
}

int batch_potrs(CHAMAMD_batch_t *b, hipStream_t s)
{
    if (!graphs_on()) {
        issue_potrs(b, s);
        return CHAMELEON_SUCCESS;
    }
    if (!b->g_potrs) b->g_potrs = capture([&] { issue_potrs(b, ctx()->s_main); });
    CHX_HIP(hipGraphLaunch(b->g_potrs, ctx()->s_main));
    return CHAMELEON_SUCCESS;
}

/* Per bucket of this batch: time the factorisation of a sample of its
 * systems (its largest, up to 1024 of them and 4 GB: two persistent grids'
 * worth at the top bucket) with 4 and 8 wavefronts per system, twice each,
 * and keep the faster shape (8 only when 3 % faster).  The sample is laid
 * out like the batch in scratch memory and regenerated (dplgsy, bump = n)
 * before every trial, so the batch's own matrices are not touched. */
int tune(CHAMAMD_batch_t *b, int *table)
{
    Context *c = ctx();
    hipStream_t sm = c->s_main;
    CHX_HIP(hipStreamSynchronize(sm));
    k::potrf_jobs_prepare();
    hipEvent_t e0, e1;
    CHX_HIP(hipEventCreate(&e0));
    CHX_HIP(hipEventCreate(&e1));
    for (int hb = 0; hb < NBUCKET; ++hb) {
        const int s0 = b->bstart[hb + 1], s1 = b->bstart[hb];
        if (s1 <= s0) continue;
        int cnt = std::min(s1 - s0, 1024);
        auto abytes = [&](int m) {
            return (size_t)(b->aoff[s0 + m - 1] - b->aoff[s0] + (long long)b->np[s0 + m - 1] * b->np[s0 + m - 1]) *
                   sizeof(double);
        };
        while (cnt > 1 && abytes(cnt) > ((size_t)4 << 30)) cnt = cnt * 3 / 4;
        std::vector<long long> aoff(cnt), doff(cnt);
        std::vector<int> np(cnt), nv(cnt);
        std::vector<BatchGenSys> sys;
        size_t d_elems = 0;
        for (int i = 0; i < cnt; ++i) {
            aoff[i] = b->aoff[s0 + i] - b->aoff[s0];
            doff[i] = (long long)d_elems;
            np[i] = b->np[s0 + i];
            nv[i] = b->n[s0 + i];
            d_elems += (size_t)((np[i] + 63) / 64) * 4096;
            sys.push_back(gen_sys(aoff[i], np[i], np[i], np[i], nv[i], nv[i], 0, nv[i], 1ULL + i, (double)nv[i]));
        }
        double *tA = nullptr, *tD = nullptr;
        int *tI = nullptr;
        CHX_HIP(hipMalloc(&tA, abytes(cnt)));
        CHX_HIP(hipMalloc(&tD, d_elems * sizeof(double)));
        CHX_HIP(hipMalloc(&tI, (cnt + 1) * sizeof(int)));
        long long *d_aoff = upload(aoff), *d_doff = upload(doff);
        int *d_np = upload(np), *d_nv = upload(nv);
        float t4 = 0.f;
        int best_w = 4;
        for (int w : {4, 8}) {
            float tmin = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                run_gen(tA, sys, 1, sm);
                k::zero_ints(tI, cnt + 1, sm);
                CHX_HIP(hipEventRecord(e0, sm));
                k::potrf_jobs(tA, d_aoff, d_np, nullptr, d_nv, tD, d_doff, cnt, tI + cnt, tI, 0, sm, w);
                CHX_HIP(hipEventRecord(e1, sm));
                CHX_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                CHX_HIP(hipEventElapsedTime(&ms, e0, e1));
                tmin = std::min(tmin, ms);
            }
            if (w == 4) t4 = tmin;
            else if (tmin < 0.97f * t4) best_w = 8;
        }
        void *ptrs[] = {tA, tD, tI, d_aoff, d_doff, d_np, d_nv};
        for (void *p : ptrs) CHX_HIP(hipFree(p));
        g_bucket_waves[hb] = best_w;
        g_bucket_tuned[hb] = true;
    }
    CHX_HIP(hipEventDestroy(e0));
    CHX_HIP(hipEventDestroy(e1));
    b->tuned = true;
    if (table)
        for (int i = 0; i < NBUCKET; ++i) table[i] = g_bucket_waves[i];
    return CHAMELEON_SUCCESS;
}

/* first failing system (in user order) of the last factorisation, info of it */
int batch_status(CHAMAMD_batch_t *b, int *info_user)
{
    std::vector<int> h(b->count);
    CHX_HIP(hipMemcpy(h.data(), b->d_info, b->count * sizeof(int), hipMemcpyDeviceToHost));
    int status = 0;
    for (int i = 0; i < b->count; ++i) {
        const int v = h[b->slot[i]];
        if (info_user) info_user[i] = v;
        if (v != 0 && status == 0) status = v;
    }
    return status;
}

} /* namespace */

extern "C" {

int CHAMAMD_Batch_Create(CHAMAMD_batch_t **batch, int count, const int *n, const int *nrhs)
{
    if (!ctx()) {
        error("CHAMAMD_Batch_Create", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (!batch || count < 0 || (count > 0 && (!n || !nrhs))) {
        error("CHAMAMD_Batch_Create", "illegal arguments");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    for (int i = 0; i < count; ++i)
        if (n[i] < 1 || nrhs[i] < 0) {
            error("CHAMAMD_Batch_Create", "every system needs n >= 1 and nrhs >= 0");
            return CHAMELEON_ERR_ILLEGAL_VALUE;
        }
    CHAMAMD_batch_t *b = new CHAMAMD_batch_t();
    b->count = count;
    b->user.resize(count);
    std::iota(b->user.begin(), b->user.end(), 0);
    /* largest first; ties keep user order */
    std::stable_sort(b->user.begin(), b->user.end(), [&](int x, int y) { return n[x] > n[y]; });
    b->slot.resize(count);
    b->n.resize(count);
    b->np.resize(count);
    b->nrhs.resize(count);
    b->aoff.resize(count);
    b->boff.resize(count);
    b->doff.resize(count);
    for (int sl = 0; sl < count; ++sl) {
        const int u = b->user[sl];
        b->slot[u] = sl;
        b->n[sl] = n[u];
        b->np[sl] = rup16(n[u]);
        b->nrhs[sl] = nrhs[u];
        b->aoff[sl] = (long long)b->a_elems;
        b->boff[sl] = (long long)b->b_elems;
        b->doff[sl] = (long long)b->d_elems;
        b->a_elems += (size_t)b->np[sl] * b->np[sl];
        b->b_elems += (size_t)b->np[sl] * nrhs[u];
        b->d_elems += (size_t)((b->np[sl] + 63) / 64) * 4096;
    }
    /* bucket boundaries over the slots (largest bucket first) */
    for (int hb = 0; hb <= NBUCKET; ++hb) {
        int cnt = 0;
        for (int sl = 0; sl < count; ++sl)
            if (hb < NBUCKET && bucket_of(b->np[sl]) >= hb) ++cnt;
        b->bstart[hb] = cnt;
    }
    if (count > 0) {
        CHX_HIP(hipMalloc(&b->dA, b->a_elems * sizeof(double)));
        CHX_HIP(hipMalloc(&b->dB, (b->b_elems ? b->b_elems : 1) * sizeof(double)));
        CHX_HIP(hipMalloc(&b->dD, b->d_elems * sizeof(double)));
        CHX_HIP(hipMemset(b->dB, 0, (b->b_elems ? b->b_elems : 1) * sizeof(double)));
        b->d_n = upload(b->n);
        b->d_np = upload(b->np);
        b->d_nrhs = upload(b->nrhs);
        b->d_aoff = upload(b->aoff);
        b->d_boff = upload(b->boff);
        b->d_doff = upload(b->doff);
        CHX_HIP(hipMalloc(&b->d_info, count * sizeof(int)));
        CHX_HIP(hipMemset(b->d_info, 0, count * sizeof(int)));
        CHX_HIP(hipMalloc(&b->d_queue, QUEUE_INTS * sizeof(int)));
        /* identity systems until the caller provides matrices */
        std::vector<BatchGenSys> sys;
        for (int sl = 0; sl < count; ++sl)
            sys.push_back(gen_sys(b->aoff[sl], b->np[sl], b->np[sl], b->np[sl], 0, 0, 0, 0, 0, 0.0));
        run_gen(b->dA, sys, 1, ctx()->s_main);
    }
    *batch = b;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_Destroy(CHAMAMD_batch_t **batch)
{
    if (!batch || !*batch) return CHAMELEON_ERR_UNALLOCATED;
    CHAMAMD_batch_t *b = *batch;
    if (ctx()) CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    drop_graphs(b);
    void *ptrs[] = {b->dA, b->dB, b->dD, b->d_np, b->d_n, b->d_nrhs, b->d_info, b->d_queue, b->d_aoff, b->d_boff,
                    b->d_doff, b->dW, b->dX, b->dR, b->d_anrm, b->d_berr, b->d_state, b->d_iters};
    for (void *p : ptrs)
        if (p) CHX_HIP(hipFree(p));
    delete b;
    *batch = nullptr;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_SetA(CHAMAMD_batch_t *b, int i, cham_uplo_t uplo, const double *A, int lda)
{
    if (!check_batch("CHAMAMD_Batch_SetA", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (i < 0 || i >= b->count) return -2;
    if (uplo != ChamUpper && uplo != ChamLower && uplo != ChamUpperLower) return -3;
    const int sl = b->slot[i], n = b->n[sl], np = b->np[sl];
    if (!A) return -4;
    if (lda < n) return -5;
    /* lower storage with identity padding; Upper input is transposed */
    std::vector<double> h((size_t)np * np, 0.0);
    for (int c = 0; c < np; ++c)
        for (int r = c; r < np; ++r) {
            double v;
            if (r >= n || c >= n) v = (r == c) ? 1.0 : 0.0;
            else v = (uplo == ChamUpper) ? A[c + (size_t)r * lda] : A[r + (size_t)c * lda];
            h[r + (size_t)c * np] = v;
        }
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    CHX_HIP(hipMemcpy(b->dA + b->aoff[sl], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_GetA(CHAMAMD_batch_t *b, int i, cham_uplo_t uplo, double *A, int lda)
{
    if (!check_batch("CHAMAMD_Batch_GetA", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (i < 0 || i >= b->count) return -2;
    if (uplo != ChamUpper && uplo != ChamLower) return -3;
    const int sl = b->slot[i], n = b->n[sl], np = b->np[sl];
    if (!A) return -4;
    if (lda < n) return -5;
    std::vector<double> h((size_t)np * np);
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    CHX_HIP(hipMemcpy(h.data(), b->dA + b->aoff[sl], h.size() * sizeof(double), hipMemcpyDeviceToHost));
    /* only the uplo triangle (diagonal included) of A is written */
    for (int c = 0; c < n; ++c)
        for (int r = c; r < n; ++r) {
            if (uplo == ChamLower) A[r + (size_t)c * lda] = h[r + (size_t)c * np];
            else A[c + (size_t)r * lda] = h[r + (size_t)c * np];
        }
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_SetB(CHAMAMD_batch_t *b, int i, const double *B, int ldb)
{
    if (!check_batch("CHAMAMD_Batch_SetB", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (i < 0 || i >= b->count) return -2;
    const int sl = b->slot[i], n = b->n[sl], np = b->np[sl], nr = b->nrhs[sl];
    if (!B && nr > 0) return -3;
    if (ldb < n) return -4;
    if (nr == 0) return CHAMELEON_SUCCESS;
    std::vector<double> h((size_t)np * nr, 0.0);
    for (int c = 0; c < nr; ++c)
        for (int r = 0; r < n; ++r) h[r + (size_t)c * np] = B[r + (size_t)c * ldb];
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    CHX_HIP(hipMemcpy(b->dB + b->boff[sl], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_GetB(CHAMAMD_batch_t *b, int i, double *B, int ldb)
{
    if (!check_batch("CHAMAMD_Batch_GetB", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (i < 0 || i >= b->count) return -2;
    const int sl = b->slot[i], n = b->n[sl], np = b->np[sl], nr = b->nrhs[sl];
    if (!B && nr > 0) return -3;
    if (ldb < n) return -4;
    if (nr == 0) return CHAMELEON_SUCCESS;
    std::vector<double> h((size_t)np * nr);
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    CHX_HIP(hipMemcpy(h.data(), b->dB + b->boff[sl], h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int c = 0; c < nr; ++c)
        for (int r = 0; r < n; ++r) B[r + (size_t)c * ldb] = h[r + (size_t)c * np];
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_dplgsy_batch(CHAMAMD_batch_t *b, unsigned long long seed)
{
    if (!check_batch("CHAMAMD_dplgsy_batch", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    std::vector<BatchGenSys> sys;
    for (int sl = 0; sl < b->count; ++sl) {
        const int n = b->n[sl], np = b->np[sl];
        sys.push_back(gen_sys(b->aoff[sl], np, np, np, n, n, 0, n, seed + (unsigned long long)b->user[sl], (double)n));
    }
    if (!sys.empty()) run_gen(b->dA, sys, 1, ctx()->s_main);
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_dplrnt_batch(CHAMAMD_batch_t *b, unsigned long long seed)
{
    if (!check_batch("CHAMAMD_dplrnt_batch", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    std::vector<BatchGenSys> sys;
    for (int sl = 0; sl < b->count; ++sl) {
        const int n = b->n[sl], np = b->np[sl], nr = b->nrhs[sl];
        if (nr == 0) continue;
        sys.push_back(gen_sys(b->boff[sl], np, np, nr, n, nr, 0, n, seed + (unsigned long long)b->user[sl], 0.0));
    }
    if (!sys.empty()) run_gen(b->dB, sys, 0, ctx()->s_main);
    return CHAMELEON_SUCCESS;
}

static int batch_async(const char *fn, CHAMAMD_batch_t *b, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request,
                       bool factor, bool solve)
{
    if (!check_batch(fn, b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (!sequence || !request) {
        error(fn, "NULL sequence or request");
        return CHAMELEON_ERR_UNALLOCATED;
    }
    if (sequence->status != CHAMELEON_SUCCESS) return request_fail(sequence, request, CHAMELEON_ERR_SEQUENCE_FLUSHED);
    request->status = CHAMELEON_SUCCESS;
    if (b->count == 0) return CHAMELEON_SUCCESS;
    hipStream_t s = ctx()->s_main;
    if (factor) batch_potrf(b, s);
    if (solve) batch_potrs(b, s);
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_dpotrf_batch_Async(CHAMAMD_batch_t *b, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request)
{
    return batch_async("CHAMAMD_dpotrf_batch_Async", b, sequence, request, true, false);
}

int CHAMAMD_dpotrs_batch_Async(CHAMAMD_batch_t *b, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request)
{
    return batch_async("CHAMAMD_dpotrs_batch_Async", b, sequence, request, false, true);
}

int CHAMAMD_dposv_batch_Async(CHAMAMD_batch_t *b, RUNTIME_sequence_t *sequence, RUNTIME_request_t *request)
{
    return batch_async("CHAMAMD_dposv_batch_Async", b, sequence, request, true, true);
}

/* Mixed-precision solve of every system (LAPACK dsposv semantics per
 * system): a working copy of A factored with FP32 products
 * (potrf_jobs lp), X = W^-1 B, then FP64 refinement on the device --
 * residual and dsposv's stopping test per system in one launch, one
 * correction solve for the systems still refining -- at most 30 steps.
 * Systems whose reduced-precision factorisation failed (ITER -3) or that
 * did not converge (-31) are factored in FP64 in place and solved.  B
 * receives the solutions; A is unchanged for the systems that converged
 * and holds the FP64 factor for the others. */
int CHAMAMD_dposv_mixed_batch(CHAMAMD_batch_t *b, int prec, int *iter, int *info)
{
    static const char *fn = "CHAMAMD_dposv_mixed_batch";
    if (!check_batch(fn, b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (prec != CHAMAMD_PREC_FP32) {
        error(fn, "batches refine FP32-product factorisations only (prec = CHAMAMD_PREC_FP32)");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if (b->count == 0) return CHAMELEON_SUCCESS;
    Context *c = ctx();
    hipStream_t sm = c->s_main;
    const int count = b->count;
    if (!b->dW) {
        CHX_HIP(hipMalloc(&b->dW, b->a_elems * sizeof(double)));
        CHX_HIP(hipMalloc(&b->dX, (b->b_elems ? b->b_elems : 1) * sizeof(double)));
        CHX_HIP(hipMalloc(&b->dR, (b->b_elems ? b->b_elems : 1) * sizeof(double)));
        CHX_HIP(hipMalloc(&b->d_anrm, count * sizeof(double)));
        CHX_HIP(hipMalloc(&b->d_berr, count * sizeof(double)));
        CHX_HIP(hipMalloc(&b->d_state, count * sizeof(int)));
        CHX_HIP(hipMalloc(&b->d_iters, count * sizeof(int)));
    }
    /* the working copy: only the part the factorisation touches */
    k::batch_copy_lower(b->dA, b->dW, b->d_aoff, b->d_np, count, sm);
    if (b->b_elems) CHX_HIP(hipMemcpyAsync(b->dX, b->dB, b->b_elems * sizeof(double), hipMemcpyDeviceToDevice, sm));
    issue_potrf(b, g_bucket_waves, b->dW, true);
    k::zero_ints(b->d_queue + 1, 1, sm);
    k::batch_solve(b->dW, b->d_aoff, b->d_np, nullptr, b->dD, b->d_doff, b->dX, b->d_boff, b->d_nrhs, b->d_info,
                   count, b->d_queue + 1, 0, sm);
    k::batch_refine_init(b->d_info, b->d_state, b->d_iters, b->d_anrm, count, sm);
    for (int it = 0;; ++it) {
        k::zero_ints(b->d_queue + Q_RES, 2, sm);
        k::batch_residual(b->dA, b->d_aoff, b->d_np, b->d_n, b->dB, b->dX, b->dR, b->d_boff, b->d_nrhs, b->d_anrm,
                          b->d_state, b->d_iters, b->d_queue + Q_UNCONV, b->d_berr, count, b->d_queue + Q_RES, sm);
        int unconv = 0;
        CHX_HIP(hipMemcpyAsync(&unconv, b->d_queue + Q_UNCONV, sizeof(int), hipMemcpyDeviceToHost, sm));
        CHX_HIP(hipStreamSynchronize(sm));
        if (unconv == 0 || it == 30) break;
        /* correction of the systems still refining (state 0) */
        k::zero_ints(b->d_queue + 1, 1, sm);
        k::batch_solve(b->dW, b->d_aoff, b->d_np, nullptr, b->dD, b->d_doff, b->dR, b->d_boff, b->d_nrhs,
                       b->d_state, count, b->d_queue + 1, 0, sm);
        k::batch_update(b->dX, b->dR, b->d_boff, b->d_np, b->d_nrhs, b->d_state, count, 1, sm);
    }
    std::vector<int> state(count), iters(count);
    CHX_HIP(hipMemcpy(state.data(), b->d_state, count * sizeof(int), hipMemcpyDeviceToHost));
    CHX_HIP(hipMemcpy(iters.data(), b->d_iters, count * sizeof(int), hipMemcpyDeviceToHost));
    /* FP64 fallback of the systems that did not converge */
    std::vector<long long> aoff, doff;
    std::vector<int> np, nv, slot;
    for (int sl = 0; sl < count; ++sl)
        if (state[sl] != 1) {
            aoff.push_back(b->aoff[sl]);
            doff.push_back(b->doff[sl]);
            np.push_back(b->np[sl]);
            nv.push_back(b->n[sl]);
            slot.push_back(sl);
        }
    std::vector<int> finfo(count, 0);
    if (!slot.empty()) {
        const int nf = (int)slot.size();
        long long *d_aoff = upload(aoff), *d_doff = upload(doff);
        int *d_np = upload(np), *d_nv = upload(nv);
        std::vector<long long> boff;
        std::vector<int> nr;
        for (int sl : slot) {
            boff.push_back(b->boff[sl]);
            nr.push_back(b->nrhs[sl]);
        }
        long long *d_boff = upload(boff);
        int *d_nr = upload(nr), *d_fi = nullptr;
        CHX_HIP(hipMalloc(&d_fi, (nf + 2) * sizeof(int)));
        k::zero_ints(d_fi, nf + 2, sm);
        k::potrf_jobs(b->dA, d_aoff, d_np, nullptr, d_nv, b->dD, d_doff, nf, d_fi + nf, d_fi, 0, sm);
        k::batch_update(b->dX, b->dB, b->d_boff, b->d_np, b->d_nrhs, b->d_state, count, 0, sm);
        k::batch_solve(b->dA, d_aoff, d_np, nullptr, b->dD, d_doff, b->dX, d_boff, d_nr, d_fi, nf, d_fi + nf + 1, 0,
                       sm);
        std::vector<int> fi(nf);
        CHX_HIP(hipMemcpyAsync(fi.data(), d_fi, nf * sizeof(int), hipMemcpyDeviceToHost, sm));
        CHX_HIP(hipStreamSynchronize(sm));
        for (int i = 0; i < nf; ++i) finfo[slot[i]] = fi[i];
        void *ptrs[] = {d_aoff, d_doff, d_np, d_nv, d_boff, d_nr, d_fi};
        for (void *p : ptrs) CHX_HIP(hipFree(p));
    }
    if (b->b_elems) CHX_HIP(hipMemcpyAsync(b->dB, b->dX, b->b_elems * sizeof(double), hipMemcpyDeviceToDevice, sm));
    /* the per-system info the batch reports next: the FP64 fallback's */
    CHX_HIP(hipMemcpyAsync(b->d_info, finfo.data(), count * sizeof(int), hipMemcpyHostToDevice, sm));
    CHX_HIP(hipStreamSynchronize(sm));
    int status = 0;
    for (int u = 0; u < count; ++u) {
        const int sl = b->slot[u];
        if (iter) iter[u] = state[sl] == 1 ? iters[sl] : (state[sl] == 2 ? -3 : -31);
        if (info) info[u] = finfo[sl];
        if (finfo[sl] != 0 && status == 0) status = finfo[sl];
    }
    return status;
}

int CHAMAMD_Set_Batch_Graphs(int on)
{
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    g_graphs = on ? 1 : 0;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_Tune(CHAMAMD_batch_t *b, int *waves)
{
    if (!check_batch("CHAMAMD_Batch_Tune", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (b->count == 0) return CHAMELEON_SUCCESS;
    return tune(b, waves);
}

int CHAMAMD_Batch_Get_Tuning(int *waves, int *bucket_top, int *tuned)
{
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    for (int i = 0; i < NBUCKET; ++i) {
        if (waves) waves[i] = g_bucket_waves[i];
        if (bucket_top) bucket_top[i] = kBucketTop[i];
        if (tuned) tuned[i] = g_bucket_tuned[i];
    }
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_Set_Tuning(const int *waves)
{
    if (!ctx()) return CHAMELEON_ERR_NOT_INITIALIZED;
    if (!waves) return CHAMELEON_ERR_ILLEGAL_VALUE;
    for (int i = 0; i < NBUCKET; ++i)
        if (waves[i] != 4 && waves[i] != 8) return CHAMELEON_ERR_ILLEGAL_VALUE;
    for (int i = 0; i < NBUCKET; ++i) g_bucket_waves[i] = waves[i];
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Batch_Info(CHAMAMD_batch_t *b, int *info)
{
    if (!check_batch("CHAMAMD_Batch_Info", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (b->count == 0) return CHAMELEON_SUCCESS;
    CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    return batch_status(b, info);
}

int CHAMAMD_dposv_batch(CHAMAMD_batch_t *b, int *info)
{
    if (!check_batch("CHAMAMD_dposv_batch", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (b->count == 0) return CHAMELEON_SUCCESS;
    hipStream_t s = ctx()->s_main;
    batch_potrf(b, s);
    batch_potrs(b, s);
    CHX_HIP(hipStreamSynchronize(s));
    return batch_status(b, info);
}

int CHAMAMD_dpotrf_batch(CHAMAMD_batch_t *b, int *info)
{
    if (!check_batch("CHAMAMD_dpotrf_batch", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (b->count == 0) return CHAMELEON_SUCCESS;
    hipStream_t s = ctx()->s_main;
    batch_potrf(b, s);
    CHX_HIP(hipStreamSynchronize(s));
    return batch_status(b, info);
}

int CHAMAMD_dpotrs_batch(CHAMAMD_batch_t *b)
{
    if (!check_batch("CHAMAMD_dpotrs_batch", b)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (b->count == 0) return CHAMELEON_SUCCESS;
    hipStream_t s = ctx()->s_main;
    batch_potrs(b, s);
    CHX_HIP(hipStreamSynchronize(s));
    return CHAMELEON_SUCCESS;
}

/* ------------------------------------------------- Chameleon tile batch */

int CHAMELEON_dplgsy_batch_Tile(double bump, CHAM_desc_t *A, unsigned long long int seed)
{
    const char *fn = "CHAMELEON_dplgsy_batch_Tile";
    if (!ctx()) {
        error(fn, "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (desc_check(A) != CHAMELEON_SUCCESS || cham_get_flttype(A->dtyp) != ChamRealDouble) {
        error(fn, "invalid descriptor");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    hipStream_t s = ctx()->s_main;
    mirror_prepare_overwrite(A, MIRROR_NORMAL, false);
    Mirror *M = mirror_of(A);
    std::vector<BatchGenSys> sys;
    for (int n = 0; n < A->nt; ++n)
        for (int m = 0; m < A->mt; ++m) {
            const int vm = std::min(A->mb, A->m - m * A->mb), vn = std::min(A->nb, A->n - n * A->nb);
            const int v = std::min(vm, vn);
            /* zplghe_batch.c:56-57: origin (n*mt + m)*mb*nb, bigM = 0 */
            const long long base = ((long long)n * A->mt + m) * (long long)A->mb * A->nb;
            BatchGenSys g = gen_sys((long long)((size_t)n * M->mt + m) * M->tile_elems, M->ldt, M->ldt, M->wdt, v, v,
                                    base, 0, seed, bump);
            g.pad_one = 0;
            sys.push_back(g);
        }
    run_gen((double *)M->d, sys, 1, s);
    /* the mirror convention: identity on the padded diagonal of diagonal tiles */
    k::tile_pad_diag((double *)M->d, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, 0, s);
    M->diag_identity = true;
    mirror_mark_written(A);
    CHAMELEON_Desc_Flush(A, nullptr);
    CHX_HIP(hipStreamSynchronize(s));
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_dpotrf_batch_Tile(cham_uplo_t uplo, CHAM_desc_t *A)
{
    const char *fn = "CHAMELEON_dpotrf_batch_Tile";
    if (!ctx()) {
        error(fn, "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (desc_check(A) != CHAMELEON_SUCCESS || cham_get_flttype(A->dtyp) != ChamRealDouble) {
        error(fn, "invalid descriptor");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    if (uplo != ChamUpper && uplo != ChamLower) {
        error(fn, "illegal value of uplo");
        return -1;
    }
    hipStream_t s = ctx()->s_main;
    mirror_acquire(A, MIRROR_NORMAL, true, s);
    Mirror *M = mirror_of(A);
    if (uplo == ChamUpper && M->ldt != M->wdt) {
        error(fn, "Upper needs square padded tiles");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    const int ntiles = M->mt * M->nt;
    std::vector<long long> off(ntiles), doff(ntiles);
    std::vector<int> np(ntiles), nvalid(ntiles);
    double *T = (double *)M->d;
    if (uplo == ChamUpper) {
        /* factor the transposed tiles: the upper triangle becomes the lower
         * one, the untouched lower triangle is carried along and restored */
        T = (double *)workspace2((size_t)ntiles * M->tile_elems * sizeof(double));
        k::tile_transpose((const double *)M->d, T, ntiles, M->ldt, s);
    }
    long long dsum = 0;
    for (int n = 0; n < A->nt; ++n)
        for (int m = 0; m < A->mt; ++m) {
            const int t = n * A->mt + m;
            const int vm = std::min(A->mb, A->m - m * A->mb), vn = std::min(A->nb, A->n - n * A->nb);
            const int v = std::min(vm, vn);
            off[t] = (long long)t * (long long)M->tile_elems;
            nvalid[t] = v;
            np[t] = std::min(std::min(M->ldt, M->wdt), rup16(v));
            doff[t] = dsum;
            dsum += (long long)((np[t] + 63) / 64) * 4096;
        }
    long long *d_off = upload(off), *d_doff = upload(doff);
    int *d_np = upload(np), *d_nv = upload(nvalid);
    int *d_info = nullptr, *d_q = nullptr;
    CHX_HIP(hipMalloc(&d_info, ntiles * sizeof(int)));
    CHX_HIP(hipMalloc(&d_q, sizeof(int)));
    double *dinv = (double *)workspace((size_t)dsum * sizeof(double));
    std::vector<int> ldv(ntiles, M->ldt);
    int *d_ld = upload(ldv);
    k::tile_pad_diag(T, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, 1, s);
    k::zero_ints(d_q, 1, s);
    k::potrf_jobs(T, d_off, d_np, d_ld, d_nv, dinv, d_doff, ntiles, d_q, d_info, 0, s);
    k::tile_pad_diag(T, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, 0, s);
    if (uplo == ChamUpper) k::tile_transpose(T, (double *)M->d, ntiles, M->ldt, s);
    std::vector<int> info(ntiles);
    CHX_HIP(hipMemcpyAsync(info.data(), d_info, ntiles * sizeof(int), hipMemcpyDeviceToHost, s));
    CHX_HIP(hipStreamSynchronize(s));
    void *ptrs[] = {d_off, d_doff, d_np, d_nv, d_info, d_q, d_ld};
    for (void *p : ptrs) CHX_HIP(hipFree(p));
    mirror_mark_written(A);
    CHAMELEON_Desc_Flush(A, nullptr);
    CHX_HIP(hipStreamSynchronize(s));
    /* the reference only prints failing tiles (zpotrf_batch.c:46-48); here
     * the first failing tile's column is returned as well */
    for (int t = 0; t < ntiles; ++t)
        if (info[t] > 0) {
            char msg[96];
            snprintf(msg, sizeof msg, "tile %d of the batch is not positive definite (info = %d)", t, info[t]);
            warning(fn, msg);
            return info[t];
        }
    return CHAMELEON_SUCCESS;
}

} /* extern "C" */
