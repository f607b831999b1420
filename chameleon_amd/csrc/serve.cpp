/*
 * serve.cpp -- a solve server on one device: requests A_i X = B_i arrive
 * over time, factors stay resident in HBM for repeated right-hand sides,
 * and pending work is batched round by round.
 *
 * No counterpart in the reference beyond its asynchronous interface
 * (RUNTIME_sequence_t / RUNTIME_request_t, include/chameleon/
 * runtime_struct.h:78-103, several _Tile_Async calls waited once,
 * control/async.c:103-221), which has no admission control, residency or
 * latency accounting.
 *
 * Memory: one device pool of the configured budget holds every resident
 * factor (with the inverses of its 64 x 64 diagonal blocks); a first-fit
 * allocator places new factors and evicts least recently used factors
 * that no pending request needs.  A request whose factor cannot fit even
 * in an empty pool is rejected (CHAMELEON_ERR_OUT_OF_RESOURCES).
 *
 * A round (CHAMAMD_Server_Progress) takes the oldest pending requests (at
 * most CHAMAMD_SERVE_ROUND, default 256):
 *   1. matrices not resident are placed and generated (or uploaded);
 *   2. systems of order <= CHAMAMD_SERVE_SMALL are factored together by
 *      the job-queue POTRF kernel (one workgroup per system, largest
 *      first), larger ones one after the other by the tiled lookahead
 *      factorisation on the whole device;
 *   3. all right-hand sides are solved: the small systems in one batched
 *      solve launch, the large ones by the strip solve;
 *   4. solutions are copied back and each request's latency recorded.
 * Every system is computed by a fixed sequence of kernels that does not
 * depend on what else is in the round, so a request's result is bitwise
 * the same alone or in any mix.
 */
#include "internal.h"
#include "chameleon_amd.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <unordered_map>
#include <vector>

using namespace chx;

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int rup16(int x) { return (x + 15) & ~15; }

/* first-fit allocator over [0, size) in units of 256 bytes */
struct Pool {
    size_t size = 0;
    std::map<size_t, size_t> free_;   /* offset -> length */
    void init(size_t bytes) { size = bytes; free_.clear(); if (bytes) free_[0] = bytes; }
    static size_t rnd(size_t b) { return (b + 255) & ~(size_t)255; }
    bool alloc(size_t bytes, size_t *off)
    {
        bytes = rnd(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes, at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = rnd(bytes);
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Factor {
    long long id = 0;
    int n = 0, np = 0;                /* order, padded order (small) or tile layout (large) */
    bool large = false;
    int NT = 0, ldt = 0, nb = 0;      /* tile layout of a large factor                   */
    size_t off = 0, bytes = 0;        /* pool placement: factor, then inverses           */
    size_t a_elems = 0;
    int info = 0;                     /* 0 factored, > 0 not positive definite           */
    bool ready = false;
    double last_use = 0.0;
    int pins = 0;                     /* pending requests that need it                   */
};

struct Req {
    long long ticket;
    CHAMAMD_request_t r;
    double t_submit;
    Factor *f = nullptr;
};

} /* namespace */

struct CHAMAMD_server_s {
    Pool pool;
    char *base = nullptr;
    int small_max = 1024;
    int large_nb = 512;
    int max_round = 256;          /* requests per round (CHAMAMD_SERVE_ROUND) */
    long long next_ticket = 1;
    std::deque<Req> pending;
    std::unordered_map<long long, Factor *> factors;
    std::list<Factor *> lru;                       /* front: most recent */
    std::unordered_map<long long, int> status;     /* ticket -> status of finished requests */
    std::vector<double> latencies_ms;
    long long completed = 0, rejected = 0, factored = 0, reused = 0;
    double flops = 0.0;
    /* per-round device scratch */
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
};

namespace {

void *scratch(CHAMAMD_server_t *s, size_t bytes)
{
    if (bytes > s->scratch_bytes) {
        if (s->scratch) CHX_HIP(hipFree(s->scratch));
        s->scratch_bytes = bytes + bytes / 4;
        CHX_HIP(hipMalloc(&s->scratch, s->scratch_bytes));
    }
    return s->scratch;
}

template <typename T>
T *put(std::vector<char> &host, const std::vector<T> &v, size_t &cursor)
{
    cursor = (cursor + 255) & ~(size_t)255;
    const size_t at = cursor;
    host.resize(at + v.size() * sizeof(T));
    if (!v.empty()) memcpy(host.data() + at, v.data(), v.size() * sizeof(T));
    cursor = at + v.size() * sizeof(T);
    return reinterpret_cast<T *>(at);   /* offset, rebased after upload */
}

template <typename T>
T *rebase(T *off, void *dev) { return reinterpret_cast<T *>((char *)dev + reinterpret_cast<size_t>(off)); }

void evict(CHAMAMD_server_t *s, Factor *f)
{
    s->pool.release(f->off, f->bytes);
    s->lru.remove(f);
    s->factors.erase(f->id);
    delete f;
}

/* place a new factor, evicting unpinned least recently used ones */
bool place(CHAMAMD_server_t *s, Factor *f)
{
    if (f->bytes > s->pool.size) return false;      /* never fits: evict nothing */
    for (;;) {
        if (s->pool.alloc(f->bytes, &f->off)) return true;
        Factor *victim = nullptr;
        for (auto it = s->lru.rbegin(); it != s->lru.rend(); ++it)
            if ((*it)->pins == 0) {
                victim = *it;
                break;
            }
        if (!victim) return false;
        evict(s, victim);
    }
}

void layout(CHAMAMD_server_t *s, Factor *f)
{
    if (f->n <= s->small_max) {
        f->large = false;
        f->np = rup16(f->n);
        f->a_elems = (size_t)f->np * f->np;
        f->bytes = (f->a_elems + (size_t)((f->np + 63) / 64) * 4096) * sizeof(double);
    } else {
        f->large = true;
        f->nb = s->large_nb;
        f->NT = (f->n + f->nb - 1) / f->nb;
        f->ldt = rup16(f->nb);
        f->a_elems = (size_t)f->NT * f->NT * f->ldt * f->ldt;
        f->bytes = (f->a_elems + (size_t)f->NT * ((f->ldt + 63) / 64) * 4096) * sizeof(double);
    }
    f->bytes = Pool::rnd(f->bytes);
}

double *fptr(CHAMAMD_server_t *s, Factor *f) { return (double *)(s->base + f->off); }
double *dptr(CHAMAMD_server_t *s, Factor *f) { return fptr(s, f) + f->a_elems; }

void finish(CHAMAMD_server_t *s, Req &q, int st)
{
    const double t = now_s();
    s->status[q.ticket] = st;
    if (st == 0) {
        s->latencies_ms.push_back((t - q.t_submit) * 1e3);
        ++s->completed;
    } else if (st < 0) {
        ++s->rejected;
    }
    if (q.f) {
        q.f->pins--;
        q.f->last_use = t;
    }
}

/* one round over everything pending */
int round(CHAMAMD_server_t *s)
{
    Context *c = ctx();
    hipStream_t sm = c->s_main;
    /* continuous batching: the oldest requests, at most max_round of them */
    const size_t take = std::min(s->pending.size(), (size_t)s->max_round);
    std::vector<Req> reqs(s->pending.begin(), s->pending.begin() + take);
    s->pending.erase(s->pending.begin(), s->pending.begin() + take);
    if (reqs.empty()) return 0;

    /* 1. residency */
    std::vector<Factor *> fresh;
    for (Req &q : reqs) {
        auto it = s->factors.find(q.r.matrix_id);
        Factor *f = nullptr;
        if (it != s->factors.end()) {
            f = it->second;
            if (f->n != q.r.n) {
                finish(s, q, CHAMELEON_ERR_ILLEGAL_VALUE);
                q.f = nullptr;
                continue;
            }
            if (f->ready) ++s->reused;
        } else {
            f = new Factor();
            f->id = q.r.matrix_id;
            f->n = q.r.n;
            layout(s, f);
            if (!place(s, f)) {
                delete f;
                q.f = nullptr;
                finish(s, q, CHAMELEON_ERR_OUT_OF_RESOURCES);
                continue;
            }
            s->factors[f->id] = f;
            s->lru.push_front(f);
            fresh.push_back(f);
        }
        f->pins++;
        q.f = f;
        s->lru.remove(f);
        s->lru.push_front(f);
    }
    /* the request that brought a matrix in provides it */
    std::unordered_map<Factor *, const CHAMAMD_request_t *> source;
    for (Req &q : reqs)
        if (q.f && !q.f->ready && !source.count(q.f)) source[q.f] = &q.r;

    /* 2. generate / upload and factor the fresh matrices */
    std::vector<Factor *> small, large;
    for (Factor *f : fresh) (f->large ? large : small).push_back(f);
    std::sort(small.begin(), small.end(), [](Factor *a, Factor *b) { return a->n > b->n; });
    if (!small.empty()) {
        std::vector<BatchGenSys> gen;
        for (Factor *f : small) {
            const CHAMAMD_request_t *r = source[f];
            const size_t off = (size_t)(fptr(s, f) - (double *)s->base);
            if (r->A) {
                std::vector<double> h(f->a_elems, 0.0);
                for (int cc = 0; cc < f->np; ++cc)
                    for (int rr = cc; rr < f->np; ++rr)
                        h[rr + (size_t)cc * f->np] =
                            (rr >= f->n || cc >= f->n) ? (rr == cc ? 1.0 : 0.0) : r->A[rr + (size_t)cc * r->lda];
                CHX_HIP(hipMemcpyAsync(fptr(s, f), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, sm));
                CHX_HIP(hipStreamSynchronize(sm));
            } else {
                BatchGenSys g;
                memset(&g, 0, sizeof g);
                g.off = (long long)off;
                g.seed = r->seed;
                unsigned long long a = 6364136223846793005ULL, cc = 1ULL, ra = 1ULL, rc = 0ULL;
                for (unsigned long long e = (unsigned long long)f->n; e; e >>= 1) {
                    if (e & 1ULL) { rc = a * rc + cc; ra = a * ra; }
                    cc = cc * (a + 1ULL);
                    a = a * a;
                }
                g.abig = ra;
                g.cbig = rc;
                g.bump = (double)f->n;
                g.ld = f->np;
                g.rows = f->np;
                g.cols = f->np;
                g.nvr = g.nvc = f->n;
                g.bigM = f->n;
                g.pad_one = 1;
                gen.push_back(g);
            }
        }
        if (!gen.empty()) {
            std::vector<long long> unit0(gen.size() + 1, 0);
            for (size_t i = 0; i < gen.size(); ++i) unit0[i + 1] = unit0[i] + (long long)((gen[i].rows + 31) / 32) * gen[i].cols;
            std::vector<char> h;
            size_t cur = 0;
            BatchGenSys *o_sys = put(h, gen, cur);
            long long *o_u = put(h, unit0, cur);
            void *d = scratch(s, h.size());
            CHX_HIP(hipMemcpyAsync(d, h.data(), h.size(), hipMemcpyHostToDevice, sm));
            k::batch_gen((double *)s->base, rebase(o_sys, d), rebase(o_u, d), unit0.back(), (int)gen.size(), 1, sm);
            CHX_HIP(hipStreamSynchronize(sm));
        }
        /* factor all small ones in one launch */
        std::vector<long long> off, doff;
        std::vector<int> np, nv, info(small.size() + 1, 0);
        for (Factor *f : small) {
            off.push_back((long long)(fptr(s, f) - (double *)s->base));
            doff.push_back((long long)(dptr(s, f) - (double *)s->base));
            np.push_back(f->np);
            nv.push_back(f->n);
        }
        std::vector<char> h;
        size_t cur = 0;
        long long *o_off = put(h, off, cur), *o_doff = put(h, doff, cur);
        int *o_np = put(h, np, cur), *o_nv = put(h, nv, cur), *o_info = put(h, info, cur);
        void *d = scratch(s, h.size());
        CHX_HIP(hipMemcpyAsync(d, h.data(), h.size(), hipMemcpyHostToDevice, sm));
        int *d_info = rebase(o_info, d);
        k::potrf_jobs((double *)s->base, rebase(o_off, d), rebase(o_np, d), nullptr, rebase(o_nv, d),
                      (double *)s->base, rebase(o_doff, d), (int)small.size(), d_info + small.size(), d_info, 0, sm);
        CHX_HIP(hipMemcpyAsync(info.data(), d_info, small.size() * sizeof(int), hipMemcpyDeviceToHost, sm));
        CHX_HIP(hipStreamSynchronize(sm));
        for (size_t i = 0; i < small.size(); ++i) {
            small[i]->info = info[i];
            small[i]->ready = true;
            const double nn = small[i]->n;
            s->flops += nn * nn * nn / 3.0;
            ++s->factored;
        }
    }
    for (Factor *f : large) {
        const CHAMAMD_request_t *r = source[f];
        double *A = fptr(s, f);
        if (r->A) {
            /* tile layout, lower tiles only, identity padding */
            const size_t ts = (size_t)f->ldt * f->ldt;
            std::vector<double> h(ts);
            for (int tn = 0; tn < f->NT; ++tn)
                for (int tm = tn; tm < f->NT; ++tm) {
                    std::fill(h.begin(), h.end(), 0.0);
                    for (int cc = 0; cc < f->ldt; ++cc)
                        for (int rr = 0; rr < f->ldt; ++rr) {
                            const int gi = tm * f->nb + rr, gj = tn * f->nb + cc;
                            double v = 0.0;
                            if (rr >= f->nb || cc >= f->nb || gi >= f->n || gj >= f->n)
                                v = (tm == tn && rr == cc) ? 1.0 : 0.0;
                            else if (gi >= gj)
                                v = r->A[gi + (size_t)gj * r->lda];
                            h[rr + (size_t)cc * f->ldt] = v;
                        }
                    CHX_HIP(hipMemcpyAsync(A + ((size_t)tn * f->NT + tm) * ts, h.data(), ts * sizeof(double),
                                           hipMemcpyHostToDevice, sm));
                    CHX_HIP(hipStreamSynchronize(sm));
                }
        } else {
            k::plgsy(A, f->NT, f->NT, f->ldt, f->ldt, f->nb, f->nb, f->n, f->n, (double)f->n, ChamLower, f->n, 0, 0,
                     r->seed, 0, true, sm);
        }
        CHX_HIP(hipMemsetAsync(c->d_info, 0, sizeof(int), sm));
        PotrfPlan p;
        p.A = A;
        p.NT = f->NT;
        p.ldt = f->ldt;
        p.nb = f->nb;
        p.nvalid_last = f->n - (f->NT - 1) * f->nb;
        p.d_info = c->d_info;
        potrf_run(p);
        const size_t ts = (size_t)f->ldt * f->ldt;
        k::dinv_all(A, f->NT, (size_t)(f->NT + 1) * ts, f->ldt, dptr(s, f), sm);
        int hinfo = 0;
        CHX_HIP(hipMemcpyAsync(&hinfo, c->d_info, sizeof(int), hipMemcpyDeviceToHost, sm));
        CHX_HIP(hipStreamSynchronize(sm));
        CHX_HIP(hipMemsetAsync(c->d_info, 0, sizeof(int), sm));
        f->info = hinfo;
        f->ready = true;
        const double nn = f->n;
        s->flops += nn * nn * nn / 3.0;
        ++s->factored;
    }

    /* 3. solves */
    std::vector<Req *> sreq, lreq;
    for (Req &q : reqs) {
        if (!q.f) continue;
        if (q.f->info != 0) {
            finish(s, q, q.f->info);
            continue;
        }
        (q.f->large ? lreq : sreq).push_back(&q);
    }
    if (!sreq.empty()) {
        std::vector<long long> aoff, doff, boff;
        std::vector<int> np, nr, zero(sreq.size() + 1, 0);
        size_t belems = 0;
        for (Req *q : sreq) {
            aoff.push_back((long long)(fptr(s, q->f) - (double *)s->base));
            doff.push_back((long long)(dptr(s, q->f) - (double *)s->base));
            boff.push_back((long long)belems);
            np.push_back(q->f->np);
            nr.push_back(q->r.nrhs);
            belems += (size_t)q->f->np * q->r.nrhs;
        }
        std::vector<double> hb(belems, 0.0);
        for (size_t i = 0; i < sreq.size(); ++i) {
            const CHAMAMD_request_t &r = sreq[i]->r;
            for (int cc = 0; cc < r.nrhs; ++cc)
                for (int rr = 0; rr < r.n; ++rr) hb[boff[i] + rr + (size_t)cc * np[i]] = r.B[rr + (size_t)cc * r.ldb];
        }
        std::vector<char> h;
        size_t cur = 0;
        long long *o_aoff = put(h, aoff, cur), *o_doff = put(h, doff, cur), *o_boff = put(h, boff, cur);
        int *o_np = put(h, np, cur), *o_nr = put(h, nr, cur), *o_z = put(h, zero, cur);
        double *o_b = put(h, hb, cur);
        void *d = scratch(s, h.size());
        CHX_HIP(hipMemcpyAsync(d, h.data(), h.size(), hipMemcpyHostToDevice, sm));
        int *d_z = rebase(o_z, d);
        k::batch_solve((double *)s->base, rebase(o_aoff, d), rebase(o_np, d), nullptr, (double *)s->base,
                       rebase(o_doff, d), rebase(o_b, d), rebase(o_boff, d), rebase(o_nr, d), d_z, (int)sreq.size(),
                       d_z + sreq.size(), 0, sm);
        CHX_HIP(hipMemcpyAsync(hb.data(), rebase(o_b, d), belems * sizeof(double), hipMemcpyDeviceToHost, sm));
        CHX_HIP(hipStreamSynchronize(sm));
        for (size_t i = 0; i < sreq.size(); ++i) {
            CHAMAMD_request_t &r = sreq[i]->r;
            for (int cc = 0; cc < r.nrhs; ++cc)
                for (int rr = 0; rr < r.n; ++rr) r.B[rr + (size_t)cc * r.ldb] = hb[boff[i] + rr + (size_t)cc * np[i]];
            s->flops += 2.0 * r.n * (double)r.n * r.nrhs;
            finish(s, *sreq[i], 0);
        }
    }
    for (Req *q : lreq) {
        Factor *f = q->f;
        CHAMAMD_request_t &r = q->r;
        const int nrp = rup16(r.nrhs), rows = f->NT * f->ldt;
        SolvePlan sp = solve_plan(fptr(s, f), f->NT, f->ldt, nrp, c->d_info);
        std::vector<double> hb((size_t)nrp * rows, 0.0);
        for (int cc = 0; cc < r.nrhs; ++cc)
            for (int rr = 0; rr < r.n; ++rr) {
                const int t = rr / f->nb, i = rr % f->nb;
                hb[cc + (size_t)(t * f->ldt + i) * nrp] = r.B[rr + (size_t)cc * r.ldb];
            }
        CHX_HIP(hipMemcpyAsync(sp.BT, hb.data(), hb.size() * sizeof(double), hipMemcpyHostToDevice, sm));
        solve_bt(sp, fptr(s, f), f->NT, f->ldt, sp.BT, nrp, c->d_info);
        CHX_HIP(hipMemcpyAsync(hb.data(), sp.BT, hb.size() * sizeof(double), hipMemcpyDeviceToHost, sm));
        CHX_HIP(hipStreamSynchronize(sm));
        for (int cc = 0; cc < r.nrhs; ++cc)
            for (int rr = 0; rr < r.n; ++rr) {
                const int t = rr / f->nb, i = rr % f->nb;
                r.B[rr + (size_t)cc * r.ldb] = hb[cc + (size_t)(t * f->ldt + i) * nrp];
            }
        s->flops += 2.0 * r.n * (double)r.n * r.nrhs;
        finish(s, *q, 0);
    }
    return (int)reqs.size();
}

} /* namespace */

extern "C" {

int CHAMAMD_Server_Create(CHAMAMD_server_t **server, size_t hbm_budget_bytes)
{
    if (!ctx()) {
        error("CHAMAMD_Server_Create", "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (!server || hbm_budget_bytes < 4096) {
        error("CHAMAMD_Server_Create", "illegal arguments");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    CHAMAMD_server_t *s = new CHAMAMD_server_t();
    if (hipMalloc((void **)&s->base, hbm_budget_bytes) != hipSuccess) {
        delete s;
        error("CHAMAMD_Server_Create", "cannot allocate the factor pool");
        return CHAMELEON_ERR_OUT_OF_RESOURCES;
    }
    s->pool.init(hbm_budget_bytes);
    const char *v = getenv("CHAMAMD_SERVE_SMALL");
    if (v && *v) s->small_max = std::max(16, atoi(v));
    v = getenv("CHAMAMD_SERVE_ROUND");
    if (v && *v) s->max_round = std::max(1, atoi(v));
    s->large_nb = ctx()->nb;
    *server = s;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Destroy(CHAMAMD_server_t **server)
{
    if (!server || !*server) return CHAMELEON_ERR_UNALLOCATED;
    CHAMAMD_server_t *s = *server;
    if (ctx()) CHX_HIP(hipStreamSynchronize(ctx()->s_main));
    for (auto &kv : s->factors) delete kv.second;
    if (s->base) CHX_HIP(hipFree(s->base));
    if (s->scratch) CHX_HIP(hipFree(s->scratch));
    delete s;
    *server = nullptr;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Submit(CHAMAMD_server_t *s, const CHAMAMD_request_t *req, long long *ticket)
{
    if (!s || !req || !ticket) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (req->n < 1 || req->nrhs < 0 || !req->B || req->ldb < req->n || (req->A && req->lda < req->n)) {
        error("CHAMAMD_Server_Submit", "illegal request");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    Req q;
    q.ticket = s->next_ticket++;
    q.r = *req;
    q.t_submit = now_s();
    s->pending.push_back(q);
    *ticket = q.ticket;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Progress(CHAMAMD_server_t *s, int *processed)
{
    if (!s || !ctx()) return CHAMELEON_ERR_ILLEGAL_VALUE;
    const int n = round(s);
    if (processed) *processed = n;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Test(CHAMAMD_server_t *s, long long ticket, int *done, int *status)
{
    if (!s) return CHAMELEON_ERR_ILLEGAL_VALUE;
    auto it = s->status.find(ticket);
    if (done) *done = it != s->status.end();
    if (status) *status = it != s->status.end() ? it->second : 0;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Stats(CHAMAMD_server_t *s, CHAMAMD_server_stats_t *st)
{
    if (!s || !st) return CHAMELEON_ERR_ILLEGAL_VALUE;
    memset(st, 0, sizeof *st);
    std::vector<double> l = s->latencies_ms;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
        if (l.empty()) return 0.0;
        const double x = p / 100.0 * (double)(l.size() - 1);
        const size_t i = (size_t)x;
        const double f = x - (double)i;
        return i + 1 < l.size() ? l[i] * (1.0 - f) + l[i + 1] * f : l[i];
    };
    st->completed = s->completed;
    st->rejected = s->rejected;
    st->factored = s->factored;
    st->reused = s->reused;
    st->resident = (long long)s->factors.size();
    size_t used = s->pool.size;
    for (auto &kv : s->pool.free_) used -= kv.second;
    st->pool_used_bytes = (long long)used;
    st->p50_ms = pct(50.0);
    st->p99_ms = pct(99.0);
    st->max_ms = l.empty() ? 0.0 : l.back();
    st->flops = s->flops;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Reset_Stats(CHAMAMD_server_t *s)
{
    if (!s) return CHAMELEON_ERR_ILLEGAL_VALUE;
    s->latencies_ms.clear();
    s->completed = s->rejected = s->factored = s->reused = 0;
    s->flops = 0.0;
    return CHAMELEON_SUCCESS;
}

} /* extern "C" */
