/*
 * transfer.cpp -- host <-> device transfers of LAPACK-layout (column-major)
 * matrices: the role of chameleon_zlap2tile / chameleon_ztile2lap
 * (control/compute_z.h:471-541) and of pzlacpy between a user's
 * column-major array and the tile storage, for a device that sits across
 * a bus from the host memory.
 *
 * A column-major descriptor (CHAMELEON_Desc_Create on a user pointer, the
 * layout the LAPACK-layout entry points wrap around their arguments) is
 * moved by column panels of one tile column (split to fit a staging
 * chunk), not tile by tile:
 *   upload:   host worker threads copy the panel's columns out of the
 *             user's array into a pinned staging chunk; a repack kernel
 *             reads the chunk across the bus (coalesced, from every CU at
 *             once: 54 GB/s, where one DMA engine moved 28 GB/s) and
 *             scatters it into the tiles of the device mirror.  While the
 *             kernel works on one chunk, the threads fill the next (three
 *             chunks in turn).  CHAMAMD_XFER_ZC=0: a DMA into device
 *             staging first, then the repack.
 *   download: a gather kernel packs the tiles' columns into device
 *             staging, one DMA moves them into the pinned chunk, worker
 *             threads scatter them into the user's array, the next chunks'
 *             gathers and DMAs already issued.  The download mostly runs
 *             beside the factorisation (streamed, below), where kernels
 *             writing across the bus would hold CUs for the whole transfer;
 *             the DMA engine moves it without them.
 * Only the rows a triangular operation reads need to move: XFER_LOWER
 * takes each tile column from its diagonal tile down (the lower triangle
 * of tiles, all a Lower Cholesky reads or writes), XFER_UPPER the rest.  Pageable user memory cannot be
 * the source or target of an asynchronous DMA, hence the staging; the
 * worker copies run at host memory speed, several at once.
 */
#include "internal.h"

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>

namespace chx {

namespace {

/* fixed pool of worker threads; run(f) calls f(w, nw) on every worker w
 * (the calling thread is worker 0) and returns when all have returned */
class Workers {
public:
    explicit Workers(int n) : n_(std::max(1, n))
    {
        for (int w = 1; w < n_; ++w) th_.emplace_back([this, w] { loop(w); });
    }
    ~Workers()
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    int size() const { return n_; }
    void run(const std::function<void(int, int)> &f)
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            job_ = &f;
            left_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        f(0, n_);
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [this] { return left_ == 0; });
        job_ = nullptr;
    }

private:
    void loop(int w)
    {
        long long seen = 0;
        for (;;) {
            const std::function<void(int, int)> *f;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                f = job_;
            }
            (*f)(w, n_);
            {
                std::lock_guard<std::mutex> g(mu_);
                if (--left_ == 0) done_.notify_one();
            }
        }
    }
    int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int, int)> *job_ = nullptr;
    long long gen_ = 0;
    int left_ = 0;
    bool stop_ = false;
};

#define XFER_CHUNKS 3
#define XFER_CHUNK_BYTES ((size_t)64 << 20)

struct Staging {
    char *host[XFER_CHUNKS] = {};
    char *hdev[XFER_CHUNKS] = {};   /* the same pinned chunk as the device addresses it */
    char *dev[XFER_CHUNKS] = {};
    hipEvent_t ev[XFER_CHUNKS] = {};
    bool used[XFER_CHUNKS] = {};
    bool direct = true;             /* the upload's repack kernels read the pinned chunks */
    Workers *workers = nullptr;
};

Staging *g_st = nullptr;

Staging *staging()
{
    if (g_st) return g_st;
    Staging *st = new Staging();
    const char *zc = getenv("CHAMAMD_XFER_ZC");
    st->direct = !(zc && *zc == '0');
    for (int k = 0; k < XFER_CHUNKS; ++k) {
        CHX_HIP(hipHostMalloc((void **)&st->host[k], XFER_CHUNK_BYTES, hipHostMallocMapped));
        CHX_HIP(hipHostGetDevicePointer((void **)&st->hdev[k], st->host[k], 0));
        CHX_HIP(hipMalloc((void **)&st->dev[k], XFER_CHUNK_BYTES));
        CHX_HIP(hipEventCreateWithFlags(&st->ev[k], hipEventDisableTiming));
    }
    int nw = (int)std::thread::hardware_concurrency();
    const char *e = getenv("CHAMAMD_XFER_THREADS");
    if (e && *e) nw = atoi(e);
    st->workers = new Workers(std::min(16, std::max(1, nw)));
    g_st = st;
    return st;
}

/* one piece: columns c0 .. c0+cols-1 of tile column n, rows r0 .. M-1 */
struct Piece {
    int n, c0, cols, r0, rows;
};

/* part: XFER_ALL every row, XFER_LOWER rows from the diagonal tile down,
 * XFER_UPPER rows above the diagonal tile */
std::vector<Piece> pieces(const CHAM_desc_t *A, int part)
{
    std::vector<Piece> out;
    const size_t cap = XFER_CHUNK_BYTES / sizeof(double);
    for (int n = 0; n < A->nt; ++n) {
        const int tn = A->get_blkdim(A, n, DIM_n, A->n);
        const int diag = std::min(n * A->mb, A->m);
        const int r0 = part == XFER_LOWER ? diag : 0;
        const int rows = (part == XFER_UPPER ? diag : A->m) - r0;
        if (rows <= 0 || tn <= 0) continue;
        const int step = (int)std::max<size_t>(1, std::min<size_t>((size_t)tn, cap / (size_t)rows));
        for (int c0 = 0; c0 < tn; c0 += step) out.push_back({n, c0, std::min(step, tn - c0), r0, rows});
    }
    return out;
}

/* the user's element (row g, column j) of the descriptor's submatrix */
inline double *cm_at(const CHAM_desc_t *A, int g, int j)
{
    double *base = (double *)A->get_blkaddr(A, 0, 0);
    return base + (size_t)g + (size_t)j * A->llm;
}

} /* namespace */

bool xfer_eligible(const CHAM_desc_t *A, const Mirror *M)
{
    return desc_is_cm(A) && M->eltsize == 8 && A->mb == A->nb && A->mt > 0 && A->nt > 0;
}

void xfer_upload(CHAM_desc_t *A, Mirror *M, int part, hipStream_t s)
{
    Staging *st = staging();
    const std::vector<Piece> ps = pieces(A, part);
    for (size_t p = 0; p < ps.size(); ++p) {
        const Piece &pc = ps[p];
        const int k = (int)(p % XFER_CHUNKS);
        if (st->used[k]) CHX_HIP(hipEventSynchronize(st->ev[k]));
        double *h = (double *)st->host[k];
        const int jb = pc.n * A->nb + pc.c0;
        st->workers->run([&](int w, int nw) {
            for (int c = w; c < pc.cols; c += nw)
                memcpy(h + (size_t)c * pc.rows, cm_at(A, pc.r0, jb + c), (size_t)pc.rows * sizeof(double));
        });
        const size_t bytes = (size_t)pc.rows * pc.cols * sizeof(double);
        double *src = (double *)st->hdev[k];
        if (!st->direct) {
            CHX_HIP(hipMemcpyAsync(st->dev[k], h, bytes, hipMemcpyHostToDevice, s));
            src = (double *)st->dev[k];
        }
        k::panel_tiles(src, (double *)M->d, M->mt, M->ldt, M->tile_elems, A->mb, pc.n, pc.r0, pc.rows, pc.c0,
                       pc.cols, 1, s);
        CHX_HIP(hipEventRecord(st->ev[k], s));
        st->used[k] = true;
    }
}

namespace {

/* gather, DMA into pinned staging, scatter into the user's array, the
 * next chunks' gathers and DMAs issued before each scatter */
void download_pieces(Staging *st, CHAM_desc_t *A, Mirror *M, const std::vector<Piece> &ps, hipStream_t s)
{
    auto issue = [&](size_t p) {
        const Piece &pc = ps[p];
        const int k = (int)(p % XFER_CHUNKS);
        k::panel_tiles((double *)st->dev[k], (double *)M->d, M->mt, M->ldt, M->tile_elems, A->mb, pc.n, pc.r0,
                       pc.rows, pc.c0, pc.cols, 0, s);
        CHX_HIP(hipMemcpyAsync(st->host[k], st->dev[k], (size_t)pc.rows * pc.cols * sizeof(double),
                               hipMemcpyDeviceToHost, s));
        CHX_HIP(hipEventRecord(st->ev[k], s));
        st->used[k] = true;
    };
    for (int k = 0; k < XFER_CHUNKS; ++k)
        if (st->used[k]) CHX_HIP(hipEventSynchronize(st->ev[k]));
    for (size_t p = 0; p < ps.size() && p < XFER_CHUNKS; ++p) issue(p);
    for (size_t p = 0; p < ps.size(); ++p) {
        const Piece &pc = ps[p];
        const int k = (int)(p % XFER_CHUNKS);
        CHX_HIP(hipEventSynchronize(st->ev[k]));
        const double *h = (const double *)st->host[k];
        const int jb = pc.n * A->nb + pc.c0;
        st->workers->run([&](int w, int nw) {
            for (int c = w; c < pc.cols; c += nw)
                memcpy(cm_at(A, pc.r0, jb + c), h + (size_t)c * pc.rows, (size_t)pc.rows * sizeof(double));
        });
        if (p + XFER_CHUNKS < ps.size()) issue(p + XFER_CHUNKS);
    }
}

} /* namespace */

void xfer_download(CHAM_desc_t *A, Mirror *M, int part, hipStream_t s)
{
    Staging *st = staging();
    download_pieces(st, A, M, pieces(A, part), s);
}

/* ---------------------------------------------------------- streaming
 * The factor of a Lower Cholesky is final tile column by tile column:
 * column k no longer changes once panel k is done.  While a synchronous
 * call factors a column-major matrix, a background thread downloads the
 * final columns (their lower tiles) on a stream of its own as the
 * factorisation reports them, so most of the flush happens during the
 * factorisation; the flush then only waits for the rest. */
namespace {

struct Streamer {
    CHAM_desc_t *A = nullptr;
    Mirror *M = nullptr;
    hipStream_t xs = nullptr;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::pair<int, hipEvent_t>> q;    /* tile columns final once the event fires */
    bool stop = false;
    int done = 0;                                 /* tile columns downloaded (worker)        */
};

Streamer *g_stream = nullptr;

void stream_worker(Streamer *sr)
{
    Staging *st = staging();
    for (;;) {
        std::pair<int, hipEvent_t> item;
        {
            std::unique_lock<std::mutex> g(sr->mu);
            sr->cv.wait(g, [&] { return sr->stop || !sr->q.empty(); });
            if (sr->q.empty()) return;
            item = sr->q.front();
            sr->q.erase(sr->q.begin());
        }
        CHX_HIP(hipEventSynchronize(item.second));
        CHX_HIP(hipEventDestroy(item.second));
        const int upto = std::min(item.first, sr->A->nt);
        if (upto <= sr->done) continue;
        std::vector<Piece> ps;
        for (const Piece &pc : pieces(sr->A, XFER_LOWER))
            if (pc.n >= sr->done && pc.n < upto) ps.push_back(pc);
        download_pieces(st, sr->A, sr->M, ps, sr->xs);
        sr->done = upto;
    }
}

} /* namespace */

bool xfer_stream_begin(CHAM_desc_t *A, Mirror *M)
{
    if (g_stream || !xfer_eligible(A, M) || !M->lower_only || M->state != MIRROR_NORMAL) return false;
    Staging *st = staging();
    (void)st;
    Streamer *sr = new Streamer();
    sr->A = A;
    sr->M = M;
    CHX_HIP(hipStreamCreateWithFlags(&sr->xs, hipStreamNonBlocking));
    sr->th = std::thread(stream_worker, sr);
    g_stream = sr;
    return true;
}

void xfer_stream_columns(const double *mirror, int ncols, hipStream_t s)
{
    Streamer *sr = g_stream;
    if (!sr || sr->M->d != mirror) return;
    hipEvent_t e;
    CHX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    CHX_HIP(hipEventRecord(e, s));
    {
        std::lock_guard<std::mutex> g(sr->mu);
        sr->q.emplace_back(ncols, e);
    }
    sr->cv.notify_one();
}

bool xfer_stream_end(CHAM_desc_t *A, hipStream_t s)
{
    Streamer *sr = g_stream;
    if (!sr || sr->A != A) return false;
    /* the remaining columns once everything issued so far on s is done */
    xfer_stream_columns((const double *)sr->M->d, A->nt, s);
    {
        std::lock_guard<std::mutex> g(sr->mu);
        sr->stop = true;
    }
    sr->cv.notify_one();
    sr->th.join();
    CHX_HIP(hipStreamSynchronize(sr->xs));
    CHX_HIP(hipStreamDestroy(sr->xs));
    g_stream = nullptr;
    delete sr;
    return true;
}

void xfer_stream_drop(const CHAM_desc_t *A)
{
    Streamer *sr = g_stream;
    if (!sr || (A && sr->A != A)) return;
    {
        std::lock_guard<std::mutex> g(sr->mu);
        sr->stop = true;
    }
    sr->cv.notify_one();
    sr->th.join();
    for (auto &it : sr->q) CHX_HIP(hipEventDestroy(it.second));
    CHX_HIP(hipStreamSynchronize(sr->xs));
    CHX_HIP(hipStreamDestroy(sr->xs));
    g_stream = nullptr;
    delete sr;
}

void xfer_release()
{
    Staging *st = g_st;
    if (!st) return;
    for (int k = 0; k < XFER_CHUNKS; ++k) {
        if (st->used[k]) CHX_HIP(hipEventSynchronize(st->ev[k]));
        CHX_HIP(hipHostFree(st->host[k]));
        CHX_HIP(hipFree(st->dev[k]));
        CHX_HIP(hipEventDestroy(st->ev[k]));
    }
    delete st->workers;
    delete st;
    g_st = nullptr;
}

} /* namespace chx */
