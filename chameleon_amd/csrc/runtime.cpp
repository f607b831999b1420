/*
 * runtime.cpp -- process-global context, device mirrors and sequences.
 *
 * Chameleon keeps one process-global context (control/context.c:145-213)
 * created by CHAMELEON_Init and read by every call; the same holds here,
 * with the HIP device, two streams (critical path and trailing updates), a
 * device info flag and a grow-only workspace.  Sequences
 * (control/async.c:44-221) group asynchronous submissions; waiting on one
 * synchronises both streams and folds the device info flag into its status.
 */
#include "internal.h"
#include "chameleon_amd.h"

#include <cstring>
#include <cstdarg>
#include <mutex>

namespace chx {

static Context *g_ctx = nullptr;

Context *ctx() { return g_ctx; }

void fatal_hip(hipError_t e, const char *what, const char *file, int line)
{
    fprintf(stderr, "CHAMELEON FATAL: HIP error %d (%s) in %s at %s:%d\n", (int)e, hipGetErrorString(e), what, file,
            line);
    abort();
}

void error(const char *func, const char *msg)
{
    if (!g_ctx || g_ctx->warnings >= 0) fprintf(stderr, "CHAMELEON ERROR: %s(): %s\n", func, msg);
}

void warning(const char *func, const char *msg)
{
    if (g_ctx && g_ctx->warnings > 0) fprintf(stderr, "CHAMELEON WARNING: %s(): %s\n", func, msg);
}

void *workspace(size_t bytes)
{
    Context *c = g_ctx;
    if (bytes <= c->ws_bytes) return c->ws;
    if (c->ws) {
        CHX_HIP(hipDeviceSynchronize());
        CHX_HIP(hipFree(c->ws));
    }
    size_t sz = bytes + (bytes >> 3);
    CHX_HIP(hipMalloc(&c->ws, sz));
    c->ws_bytes = sz;
    return c->ws;
}

/* ------------------------------------------------------------ lanes */
Lane *main_lane() { return g_ctx ? &g_ctx->lane : nullptr; }

hipEvent_t lane_event(Lane *L, size_t i)
{
    while (L->events.size() <= i) {
        hipEvent_t e;
        CHX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        L->events.push_back(e);
    }
    return L->events[i];
}

/* Flags of the strip solves and the fused panel kernel: at least `n` of
 * them, and the first of `launches` fresh epochs (flags hold the epoch of
 * the launch that set them, so they never need clearing between launches).
 * Every launch that uses them runs on the lane's critical stream, so one
 * set serves every call on the lane. */
int *lane_flags(Lane *L, int n, int launches, int *epoch)
{
    if (n > L->flags_n || L->epoch > (1 << 30)) {
        CHX_HIP(hipStreamSynchronize(L->sm));   /* no launch still reads the old flags */
        if (L->flags) CHX_HIP(hipFree(L->flags));
        const int cap = n > L->flags_n ? n : L->flags_n;
        CHX_HIP(hipMalloc(&L->flags, (size_t)cap * sizeof(int)));
        k::zero_ints(L->flags, cap, L->sm);
        L->flags_n = cap;
        L->epoch = 0;
    }
    *epoch = L->epoch + 1;
    L->epoch += launches;
    return L->flags;
}

unsigned *lane_counters(Lane *L)
{
    if (!L->panel_ctr) {
        CHX_HIP(hipMalloc(&L->panel_ctr, 4 * sizeof(unsigned)));
        k::zero_ints((int *)L->panel_ctr, 4, L->sm);
    }
    return L->panel_ctr;
}

void lane_release(Lane *L)
{
    for (hipEvent_t e : L->events) CHX_HIP(hipEventDestroy(e));
    L->events.clear();
    if (L->flags) CHX_HIP(hipFree(L->flags));
    if (L->panel_ctr) CHX_HIP(hipFree(L->panel_ctr));
    L->flags = nullptr;
    L->panel_ctr = nullptr;
    L->flags_n = 0;
    L->epoch = 0;
}

int *strip_flags(int n, int launches, int *epoch) { return lane_flags(&g_ctx->lane, n, launches, epoch); }

unsigned *panel_counters() { return lane_counters(&g_ctx->lane); }

void *workspace2(size_t bytes)
{
    Context *c = g_ctx;
    if (bytes <= c->ws2_bytes) return c->ws2;
    if (c->ws2) {
        CHX_HIP(hipDeviceSynchronize());
        CHX_HIP(hipFree(c->ws2));
    }
    CHX_HIP(hipMalloc(&c->ws2, bytes));
    c->ws2_bytes = bytes;
    return c->ws2;
}

/* ------------------------------------------------------------ mirrors */
static int rup16(int x) { return (x + 15) & ~15; }

Mirror *mirror_of(const CHAM_desc_t *A) { return (Mirror *)A->schedopt; }

static Mirror *mirror_get(CHAM_desc_t *A)
{
    Mirror *M = (Mirror *)A->schedopt;
    if (!M) {
        M = new Mirror();
        A->schedopt = M;
    }
    return M;
}

static void mirror_alloc(CHAM_desc_t *A, Mirror *M)
{
    if (M->d) return;
    M->mt = A->mt;
    M->nt = A->nt;
    M->ldt = rup16(A->mt > 1 ? A->mb : (A->m < A->mb ? A->m : A->mb));
    M->wdt = rup16(A->nt > 1 ? A->nb : (A->n < A->nb ? A->n : A->nb));
    if (A->mb == A->nb && A->m == A->n) {
        /* square tile matrices keep square padded tiles */
        int t = M->ldt > M->wdt ? M->ldt : M->wdt;
        M->ldt = M->wdt = t;
    }
    if (M->ldt == 0) M->ldt = 16;
    if (M->wdt == 0) M->wdt = 16;
    M->eltsize = (int)CHAMELEON_Element_Size(A->dtyp);
    M->tile_elems = (size_t)M->ldt * M->wdt;
    M->bytes = M->tile_elems * M->mt * M->nt * M->eltsize;
    if (M->bytes == 0) M->bytes = 256;
    CHX_HIP(hipMalloc(&M->d, M->bytes));
    M->state = MIRROR_NONE;
}

/* host -> device, normal orientation; column-major double matrices by
 * column panels through pinned staging (transfer.cpp), `part` of them */
static void mirror_upload(CHAM_desc_t *A, Mirror *M, hipStream_t s, int part = XFER_ALL)
{
    if (xfer_eligible(A, M)) {
        xfer_upload(A, M, part, s);
        return;
    }
    const int es = M->eltsize;
    for (int n = 0; n < A->nt; ++n) {
        const int tn = A->get_blkdim(A, n, DIM_n, A->n);
        for (int m = 0; m < A->mt; ++m) {
            const int tm = A->get_blkdim(A, m, DIM_m, A->m);
            const void *src = A->get_blkaddr(A, m, n);
            const int ld = A->get_blkldd(A, m);
            char *dst = (char *)M->d + ((size_t)n * M->mt + m) * M->tile_elems * es;
            if (tm > 0 && tn > 0)
                CHX_HIP(hipMemcpy2DAsync(dst, (size_t)M->ldt * es, src, (size_t)ld * es, (size_t)tm * es, tn,
                                         hipMemcpyHostToDevice, s));
        }
    }
}

static void mirror_download(CHAM_desc_t *A, Mirror *M, hipStream_t s)
{
    if (xfer_eligible(A, M)) {
        xfer_download(A, M, M->lower_only ? XFER_LOWER : XFER_ALL, s);
        return;
    }
    const int es = M->eltsize;
    for (int n = 0; n < A->nt; ++n) {
        const int tn = A->get_blkdim(A, n, DIM_n, A->n);
        for (int m = 0; m < A->mt; ++m) {
            const int tm = A->get_blkdim(A, m, DIM_m, A->m);
            void *dst = A->get_blkaddr(A, m, n);
            const int ld = A->get_blkldd(A, m);
            const char *src = (const char *)M->d + ((size_t)n * M->mt + m) * M->tile_elems * es;
            if (tm > 0 && tn > 0)
                CHX_HIP(hipMemcpy2DAsync(dst, (size_t)ld * es, src, (size_t)M->ldt * es, (size_t)tm * es, tn,
                                         hipMemcpyDeviceToHost, s));
        }
    }
}

static void mirror_orient(Mirror *M, int state, hipStream_t s)
{
    if (M->state == state || M->state == MIRROR_NONE) return;
    if (M->eltsize != 8 || M->mt != M->nt || M->ldt != M->wdt) {
        error("mirror_orient", "only square double tile matrices can be transposed");
        return;
    }
    k::transpose_tiles_inplace((double *)M->d, M->nt, M->ldt, s);
    M->state = state;
}

void mirror_complete(CHAM_desc_t *A, hipStream_t s)
{
    Mirror *M = mirror_of(A);
    if (!M || !M->lower_only) return;
    /* the upper tiles were never written on the device: the host holds them */
    xfer_upload(A, M, XFER_UPPER, s);
    k::fix_padding(M->d, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, M->eltsize, M->diag_identity, s);
    M->lower_only = false;
}

int mirror_acquire(CHAM_desc_t *A, int state, bool diag_identity, hipStream_t s, bool lower_need)
{
    Mirror *M = mirror_get(A);
    mirror_alloc(A, M);
    if (M->state == MIRROR_NONE) {
        const bool part = lower_need && state == MIRROR_NORMAL && A->m == A->n && xfer_eligible(A, M);
        mirror_upload(A, M, s, part ? XFER_LOWER : XFER_ALL);
        M->lower_only = part;
        k::fix_padding(M->d, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, M->eltsize, diag_identity, s);
        M->diag_identity = diag_identity;
        M->state = MIRROR_NORMAL;
    } else if (M->lower_only && !(lower_need && state == MIRROR_NORMAL)) {
        mirror_complete(A, s);
    }
    if (M->diag_identity != diag_identity) {
        /* padding does not depend on orientation for square tile matrices */
        k::fix_padding(M->d, M->mt, M->nt, M->ldt, M->wdt, A->mb, A->nb, A->m, A->n, M->eltsize, diag_identity, s);
        M->diag_identity = diag_identity;
    }
    mirror_orient(M, state, s);
    return CHAMELEON_SUCCESS;
}

int mirror_prepare_overwrite(CHAM_desc_t *A, int state, bool diag_identity)
{
    Mirror *M = mirror_get(A);
    mirror_alloc(A, M);
    M->lower_only = false;
    M->state = state;
    M->diag_identity = diag_identity;
    return CHAMELEON_SUCCESS;
}

void mirror_mark_written(CHAM_desc_t *A)
{
    Mirror *M = mirror_of(A);
    if (M) M->host_valid = false;
}

int mirror_flush(CHAM_desc_t *A, hipStream_t s)
{
    Mirror *M = mirror_of(A);
    if (M && xfer_stream_end(A, s)) {
        /* the factor was downloaded column by column while it was formed */
        M->host_valid = true;
        return CHAMELEON_SUCCESS;
    }
    if (!M || M->host_valid || M->state == MIRROR_NONE) return CHAMELEON_SUCCESS;
    if (A->mat == nullptr && A->use_mat) return CHAMELEON_ERR_UNALLOCATED;
    mirror_orient(M, MIRROR_NORMAL, s);
    mirror_download(A, M, s);
    M->host_valid = true;
    return CHAMELEON_SUCCESS;
}

void mirror_release(CHAM_desc_t *A)
{
    xfer_stream_drop(A);
    Mirror *M = mirror_of(A);
    if (!M) return;
    if (M->d) {
        CHX_HIP(hipDeviceSynchronize());
        CHX_HIP(hipFree(M->d));
    }
    delete M;
    A->schedopt = nullptr;
}

/* ---------------------------------------------------------- sequences */
SeqState *seq_state(RUNTIME_sequence_t *seq) { return (SeqState *)seq->schedopt; }

int request_fail(RUNTIME_sequence_t *seq, RUNTIME_request_t *req, int status)
{
    if (req) req->status = status;
    if (seq && seq->status == CHAMELEON_SUCCESS) {
        seq->status = status;
        seq->request = req;
    }
    return status;
}

hipEvent_t timing_event()
{
    Context *c = g_ctx;
    if (c->timing_pool.empty()) {
        hipEvent_t e;
        CHX_HIP(hipEventCreate(&e));
        return e;
    }
    hipEvent_t e = c->timing_pool.back();
    c->timing_pool.pop_back();
    return e;
}

/* fold finished timing pairs into the per-class totals (streams synced) */
void stats_fold()
{
    Context *c = g_ctx;
    for (auto &p : c->pending) {
        float ms = 0.f;
        CHX_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        c->st_launches[p.cls] += 1;
        c->st_ms[p.cls] += ms;
        c->st_flops[p.cls] += p.flops;
        c->timing_pool.push_back(p.a);
        c->timing_pool.push_back(p.b);
    }
    c->pending.clear();
}

int sequence_wait_internal(RUNTIME_sequence_t *seq)
{
    Context *c = g_ctx;
    CHX_HIP(hipStreamSynchronize(c->s_update));
    CHX_HIP(hipStreamSynchronize(c->s_main));
    stats_fold();
    SeqState *st = seq_state(seq);
    if (st && st->pending_info) {
        CHX_HIP(hipMemcpy(st->h_info, st->d_info, sizeof(int), hipMemcpyDeviceToHost));
        if (*st->h_info == PANEL_STALLED) {
            warning("CHAMELEON_Sequence_Wait", "internal error: a panel kernel wait ran out");
            if (seq->status == CHAMELEON_SUCCESS) seq->status = CHAMELEON_ERR_UNEXPECTED;
        } else if (*st->h_info > 0 && seq->status == CHAMELEON_SUCCESS) {
            seq->status = *st->h_info;
        }
        st->pending_info = false;
    }
    return seq->status;
}

} /* namespace chx */

using namespace chx;

/* ============================================================ C API */
extern "C" {

int CHAMELEON_Version(int *ver_major, int *ver_minor, int *ver_micro)
{
    if (ver_major) *ver_major = 1;
    if (ver_minor) *ver_minor = 4;
    if (ver_micro) *ver_micro = 0;
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Initialized(void) { return g_ctx != nullptr; }

static int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return (v && *v) ? atoi(v) : dflt;
}

int __chameleon_initpar(int nworkers, int ncudas, int nthreads_per_worker)
{
    (void)nthreads_per_worker;
    if (g_ctx) {
        error("CHAMELEON_Init", "CHAMELEON already initialized");
        return CHAMELEON_ERR_REINITIALIZED;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        error("CHAMELEON_Init", "no AMD GPU visible: this library computes on a gfx950 device only");
        return CHAMELEON_ERR_NOT_SUPPORTED;
    }
    (void)ncudas;
    Context *c = new Context();
    c->device = 0;
    c->nworkers = nworkers > 0 ? nworkers : 1;
    c->nb = env_int("CHAMELEON_TILE_SIZE", 384);
    c->ib = env_int("CHAMELEON_INNER_BLOCK_SIZE", 48);
    c->lookahead = env_int("CHAMELEON_LOOKAHEAD", 1);
    c->translation = env_int("CHAMELEON_TRANSLATION_MODE", ChamInPlace);
    c->panel = env_int("CHAMAMD_PANEL", 1);
    c->pairs = env_int("CHAMAMD_PAIRS", 1);
    c->fwd_tail = env_int("CHAMAMD_FWD_TAIL", -1);
    {
        /* requested accuracy of the per-tile precision choice, as the
         * reference's zhered reads it (compute/zhered.c:170-176) */
        const char *acc = getenv("CHAMELEON_GERED_ACC");
        if (acc && *acc && strtod(acc, nullptr) > 0.0) c->tile_accuracy = strtod(acc, nullptr);
    }
    CHX_HIP(hipSetDevice(c->device));
    CHX_HIP(hipStreamCreateWithFlags(&c->s_main, hipStreamNonBlocking));
    CHX_HIP(hipStreamCreateWithFlags(&c->s_update, hipStreamNonBlocking));
    c->lane.sm = c->s_main;
    c->lane.su = c->s_update;
    noalloc_init(c->device, c->s_main);
    CHX_HIP(hipMalloc(&c->d_info, sizeof(int)));
    CHX_HIP(hipMemset(c->d_info, 0, sizeof(int)));
    g_ctx = c;
    return CHAMELEON_SUCCESS;
}

int __chameleon_init(int nworkers, int ncudas) { return __chameleon_initpar(nworkers, ncudas, -1); }

int __chameleon_finalize(void)
{
    Context *c = g_ctx;
    if (!c) {
        error("CHAMELEON_Finalize", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    CHX_HIP(hipDeviceSynchronize());
    latms_release();
    xfer_stream_drop(nullptr);
    xfer_release();
    lane_release(&c->lane);
    stats_fold();
    for (hipEvent_t e : c->timing_pool) CHX_HIP(hipEventDestroy(e));
    if (c->ws) CHX_HIP(hipFree(c->ws));
    if (c->ws2) CHX_HIP(hipFree(c->ws2));
    if (c->pinned) CHX_HIP(hipHostFree(c->pinned));
    CHX_HIP(hipFree(c->d_info));
    CHX_HIP(hipStreamDestroy(c->s_main));
    CHX_HIP(hipStreamDestroy(c->s_update));
    delete c;
    g_ctx = nullptr;
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Pause(void) { return CHAMELEON_SUCCESS; }
int CHAMELEON_Resume(void) { return CHAMELEON_SUCCESS; }
int CHAMELEON_GetThreadNbr(void) { return g_ctx ? g_ctx->nworkers : 1; }
int CHAMELEON_Comm_size(void) { return 1; }
int CHAMELEON_Comm_rank(void) { return 0; }

int CHAMELEON_Enable(int option)
{
    if (!g_ctx) { error("CHAMELEON_Enable", "CHAMELEON not initialized"); return CHAMELEON_ERR_NOT_INITIALIZED; }
    switch (option) {
    case CHAMELEON_WARNINGS: g_ctx->warnings = 1; break;
    case CHAMELEON_AUTOTUNING: g_ctx->autotuning = 1; break;
    case CHAMELEON_GENERATE_STATS: g_ctx->stats = true; break;
    case CHAMELEON_ERRORS: case CHAMELEON_DAG: case CHAMELEON_GENERATE_TRACE:
    case CHAMELEON_PARALLEL_KERNEL: case CHAMELEON_BOUND: case CHAMELEON_PROGRESS: case CHAMELEON_GEMM3M:
    case CHAMELEON_GENERIC:
        break;
    default:
        error("CHAMELEON_Enable", "illegal parameter value");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Disable(int option)
{
    if (!g_ctx) { error("CHAMELEON_Disable", "CHAMELEON not initialized"); return CHAMELEON_ERR_NOT_INITIALIZED; }
    switch (option) {
    case CHAMELEON_WARNINGS: g_ctx->warnings = 0; break;
    case CHAMELEON_AUTOTUNING: g_ctx->autotuning = 0; break;
    case CHAMELEON_GENERATE_STATS: g_ctx->stats = false; break;
    case CHAMELEON_ERRORS: case CHAMELEON_DAG: case CHAMELEON_GENERATE_TRACE:
    case CHAMELEON_PARALLEL_KERNEL: case CHAMELEON_BOUND: case CHAMELEON_PROGRESS: case CHAMELEON_GEMM3M:
    case CHAMELEON_GENERIC:
        break;
    default:
        error("CHAMELEON_Disable", "illegal parameter value");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Set(int param, int value)
{
    if (!g_ctx) { error("CHAMELEON_Set", "CHAMELEON not initialized"); return CHAMELEON_ERR_NOT_INITIALIZED; }
    switch (param) {
    case CHAMELEON_TILE_SIZE:
        if (value <= 0) { error("CHAMELEON_Set", "negative tile size"); return CHAMELEON_ERR_ILLEGAL_VALUE; }
        g_ctx->nb = value;
        /* an explicit tile size disables autotuning (control/context.c:410-421) */
        g_ctx->autotuning = 0;
        break;
    case CHAMELEON_INNER_BLOCK_SIZE:
        if (value <= 0) { error("CHAMELEON_Set", "negative inner block size"); return CHAMELEON_ERR_ILLEGAL_VALUE; }
        g_ctx->ib = value;
        break;
    case CHAMELEON_TRANSLATION_MODE:
        if (value != ChamInPlace && value != ChamOutOfPlace) {
            error("CHAMELEON_Set", "illegal value of translation mode");
            return CHAMELEON_ERR_ILLEGAL_VALUE;
        }
        g_ctx->translation = value;
        break;
    case CHAMELEON_LOOKAHEAD:
        if (value < 0) { error("CHAMELEON_Set", "negative lookahead"); return CHAMELEON_ERR_ILLEGAL_VALUE; }
        g_ctx->lookahead = value;
        break;
    case CHAMELEON_HOUSEHOLDER_MODE: case CHAMELEON_HOUSEHOLDER_SIZE: case CHAMELEON_RUNTIME:
    case CHAMELEON_DAG_COLOR_MODE: case CHAMELEON_CONCURRENCY:
        break;
    default:
        error("CHAMELEON_Set", "unknown parameter");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Get(int param, int *value)
{
    if (!g_ctx) { error("CHAMELEON_Get", "CHAMELEON not initialized"); return CHAMELEON_ERR_NOT_INITIALIZED; }
    switch (param) {
    case CHAMELEON_TILE_SIZE: *value = g_ctx->nb; break;
    case CHAMELEON_INNER_BLOCK_SIZE: *value = g_ctx->ib; break;
    case CHAMELEON_TRANSLATION_MODE: *value = g_ctx->translation; break;
    case CHAMELEON_LOOKAHEAD: *value = g_ctx->lookahead; break;
    case CHAMELEON_CONCURRENCY: *value = 1; break;
    default:
        error("CHAMELEON_Get", "unknown parameter");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    return CHAMELEON_SUCCESS;
}

/* ---- sequences ---- */
int CHAMELEON_Sequence_Create(RUNTIME_sequence_t **sequence)
{
    if (!g_ctx) { error("CHAMELEON_Sequence_Create", "CHAMELEON not initialized"); return CHAMELEON_ERR_NOT_INITIALIZED; }
    RUNTIME_sequence_t *s = (RUNTIME_sequence_t *)calloc(1, sizeof(RUNTIME_sequence_t));
    SeqState *st = new SeqState();
    CHX_HIP(hipMalloc(&st->d_info, sizeof(int)));
    CHX_HIP(hipHostMalloc(&st->h_info, sizeof(int), hipHostMallocDefault));
    CHX_HIP(hipMemsetAsync(st->d_info, 0, sizeof(int), g_ctx->s_main));
    s->schedopt = st;
    s->status = CHAMELEON_SUCCESS;
    *sequence = s;
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Sequence_Destroy(RUNTIME_sequence_t *sequence)
{
    if (!g_ctx) { error("CHAMELEON_Sequence_Destroy", "CHAMELEON not initialized"); return CHAMELEON_ERR_NOT_INITIALIZED; }
    if (!sequence) { error("CHAMELEON_Sequence_Destroy", "NULL sequence"); return CHAMELEON_ERR_UNALLOCATED; }
    SeqState *st = seq_state(sequence);
    if (st) {
        CHX_HIP(hipStreamSynchronize(g_ctx->s_main));
        CHX_HIP(hipStreamSynchronize(g_ctx->s_update));
        CHX_HIP(hipFree(st->d_info));
        CHX_HIP(hipHostFree(st->h_info));
        delete st;
    }
    free(sequence);
    return CHAMELEON_SUCCESS;
}

int CHAMELEON_Sequence_Wait(RUNTIME_sequence_t *sequence)
{
    if (!g_ctx) { error("CHAMELEON_Sequence_Wait", "CHAMELEON not initialized"); return CHAMELEON_ERR_NOT_INITIALIZED; }
    if (!sequence) { error("CHAMELEON_Sequence_Wait", "NULL sequence"); return CHAMELEON_ERR_UNALLOCATED; }
    return sequence_wait_internal(sequence);
}

int CHAMELEON_Sequence_Flush(RUNTIME_sequence_t *sequence, RUNTIME_request_t *request)
{
    if (!g_ctx) { error("CHAMELEON_Sequence_Flush", "CHAMELEON not initialized"); return CHAMELEON_ERR_NOT_INITIALIZED; }
    if (!sequence) { error("CHAMELEON_Sequence_Flush", "NULL sequence"); return CHAMELEON_ERR_UNALLOCATED; }
    request_fail(sequence, request, CHAMELEON_ERR_SEQUENCE_FLUSHED);
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Kernel_Stats(int kernel_class, long long *launches, double *total_ms, double *flops)
{
    if (!g_ctx) return CHAMELEON_ERR_NOT_INITIALIZED;
    if (kernel_class < 0 || kernel_class >= CHAMAMD_KERNEL_NCLASS) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (launches) *launches = g_ctx->st_launches[kernel_class];
    if (total_ms) *total_ms = g_ctx->st_ms[kernel_class];
    if (flops) *flops = g_ctx->st_flops[kernel_class];
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Kernel_Stats_Reset(void)
{
    if (!g_ctx) return CHAMELEON_ERR_NOT_INITIALIZED;
    for (int i = 0; i < CHAMAMD_KERNEL_NCLASS; ++i) {
        g_ctx->st_launches[i] = 0;
        g_ctx->st_ms[i] = 0.0;
        g_ctx->st_flops[i] = 0.0;
    }
    return CHAMELEON_SUCCESS;
}

} /* extern "C" */
