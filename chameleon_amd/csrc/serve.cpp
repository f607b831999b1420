/*
 * serve.cpp -- a solve server on one device with continuous batching:
 * requests A_i X = B_i arrive over time, factors stay resident in HBM for
 * repeated right-hand sides, and work is admitted whenever a lane and
 * memory are free, while earlier work still runs.
 *
 * No counterpart in the reference beyond its asynchronous interface
 * (RUNTIME_sequence_t / RUNTIME_request_t, include/chameleon/
 * runtime_struct.h:78-103, several _Tile_Async calls waited once,
 * control/async.c:103-221), which has no admission control, residency or
 * latency accounting.
 *
 * Memory: one device pool of the configured budget holds every resident
 * factor (with the inverses of its 64 x 64 diagonal blocks; a
 * mixed-precision factor also keeps A's lower tiles in FP64 for the
 * refinement residuals, and its diagonal scaling); a first-fit allocator
 * places new factors and evicts least recently used factors that no live
 * request needs.  A request whose factor cannot fit even in an empty pool
 * is rejected (CHAMELEON_ERR_OUT_OF_RESOURCES); one that fits only once
 * other requests finish waits for them.
 *
 * Lanes: CHAMAMD_SERVE_LANES (default 8) streams, each with its own panel
 * counters, publication flags, info flag, device scratch and pinned host
 * staging.  A lane runs one job at a time; a job is
 *   - the factorisation of every new small matrix (order <= 1024) in one
 *     job-queue POTRF launch (generation or upload first),
 *   - the tiled factorisation of one large matrix (single-stream form of
 *     the lookahead scheduler, FP64 or with low-precision trailing products),
 *   - the solve of every ready small request in one batched launch,
 *   - the strip solve of one large request (FP64), or
 *   - the refinement of one large request on a mixed-precision factor:
 *     solve, FP64 residual, convergence test on the host between steps
 *     (LAPACK dsposv's test, at most 30 corrections, else the matrix is
 *     factored again in FP64 and the request solved with that).
 * Host data moves through the lane's pinned staging with asynchronous
 * copies, completion is polled with events: CHAMAMD_Server_Progress never
 * waits for the device.  Large factorisations take at most lanes - 1 lanes
 * so that solves keep moving: lane 0 never takes one.  Every kernel of the
 * server waits on nothing (job queues) or only on work taken by
 * workgroups that started before it (the panel kernel's and the strip
 * solve's tickets), so any mix of them may share the device.
 *
 * Every system is computed by a fixed sequence of kernels that does not
 * depend on what else runs, so a request's result is bitwise the same
 * alone or in any mix, whatever overlaps it in time.
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
size_t rup256(size_t b) { return (b + 255) & ~(size_t)255; }

/* first-fit allocator over [0, size) in units of 256 bytes */
struct Pool {
    size_t size = 0;
    std::map<size_t, size_t> free_;   /* offset -> length */
    void init(size_t bytes) { size = bytes; free_.clear(); if (bytes) free_[0] = bytes; }
    bool alloc(size_t bytes, size_t *off)
    {
        bytes = rup256(bytes);
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
        bytes = rup256(bytes);
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

enum { F_NEW = 0, F_BUSY = 1, F_READY = 2 };

struct Factor {
    long long id = 0;
    int n = 0, np = 0;                /* order, padded order (small)                        */
    bool large = false;
    int NT = 0, ldt = 0, nb = 0;      /* tile layout of a large factor                      */
    int prec = -1;                    /* -1 FP64; else the low-precision products' mode     */
    size_t off = 0, bytes = 0;        /* pool placement                                     */
    size_t a_elems = 0;               /* elements of the factor                             */
    size_t dinv_off = 0, orig_off = 0, sc_off = 0;   /* element offsets after the factor     */
    int info = 0;                     /* 0 factored, > 0 not positive definite              */
    int state = F_NEW;
    bool redo = false;                /* factor again in FP64 from the kept lower tiles     */
    double anrm = 0.0;                /* ||A||_inf (mixed: the refinement's stopping test)  */
    double last_use = 0.0;
    int pins = 0;                     /* live requests attached to it                       */
    int users = 0;                    /* solve jobs in flight that read it                  */
    const CHAMAMD_request_t *source = nullptr;   /* request that supplies A (host matrix)  */
};

enum { R_WAIT = 0, R_SOLVE = 1 };

struct Req {
    long long ticket = 0;
    CHAMAMD_request_t r;
    double t_submit = 0.0;
    Factor *f = nullptr;
    int state = R_WAIT;
};

enum { J_SMALL_FACTOR, J_LARGE_FACTOR, J_SMALL_SOLVE, J_LARGE_SOLVE, J_MIXED_SOLVE };

struct Job {
    int kind = 0;
    std::vector<Factor *> fs;
    std::vector<Req *> rs;
    std::vector<size_t> boff;         /* small solve: element offset of each request's B     */
    size_t host_off = 0;              /* staging offset of what comes back                   */
    int stage = 0, iter = 0;
    int nrp = 0, rows = 0;
    double *BT = nullptr, *XT = nullptr, *RT = nullptr, *part = nullptr, *nw = nullptr, *rowsum = nullptr;
    double *d128 = nullptr;           /* 128 x 128 strip inverses of a narrow mixed solve    */
    double *ef = nullptr, *eb = nullptr;   /* predecessor products of a wide mixed solve     */
};

struct Slot {
    Lane lane;
    hipEvent_t done = nullptr;
    Job *job = nullptr;
    char *dev = nullptr;
    size_t dev_bytes = 0;
    char *host = nullptr;
    size_t host_bytes = 0;
    int *d_info = nullptr;
    int *h_info = nullptr;
};

} /* namespace */

struct CHAMAMD_server_s {
    Pool pool;
    char *base = nullptr;
    int small_max = 1024;
    int large_nb = 512;
    int prec = -1;
    long long next_ticket = 1;
    std::deque<Req *> live;                         /* unfinished requests, submission order */
    std::unordered_map<long long, Factor *> factors;
    std::list<Factor *> lru;                        /* front: most recent */
    std::unordered_map<long long, int> status;      /* ticket -> status of finished requests */
    std::vector<long long> finished;                /* tickets finished since the last query */
    std::vector<double> latencies_ms;
    long long completed = 0, rejected = 0, factored = 0, reused = 0;
    long long mixed_solves = 0, ir_steps = 0, fallbacks = 0, max_jobs = 0;
    double flops = 0.0, max_berr = 0.0;
    std::vector<Slot> slots;
};

namespace {

typedef CHAMAMD_server_t Server;

double *fptr(Server *s, Factor *f) { return (double *)(s->base + f->off); }
double *dptr(Server *s, Factor *f) { return fptr(s, f) + f->dinv_off; }
double *optr(Server *s, Factor *f) { return fptr(s, f) + f->orig_off; }
double *scptr(Server *s, Factor *f) { return fptr(s, f) + f->sc_off; }

/* Device scratch and pinned staging of an idle lane, grown on demand (the
 * lane has nothing in flight when a job is laid out on it). */
char *slot_dev(Slot &sl, size_t bytes)
{
    if (bytes > sl.dev_bytes) {
        if (sl.dev) CHX_HIP(hipFree(sl.dev));
        sl.dev_bytes = rup256(bytes + bytes / 4);
        CHX_HIP(hipMalloc(&sl.dev, sl.dev_bytes));
    }
    return sl.dev;
}

char *slot_host(Slot &sl, size_t bytes)
{
    if (bytes > sl.host_bytes) {
        if (sl.host) CHX_HIP(hipHostFree(sl.host));
        sl.host_bytes = rup256(bytes + bytes / 4);
        CHX_HIP(hipHostMalloc((void **)&sl.host, sl.host_bytes, hipHostMallocDefault));
    }
    return sl.host;
}

/* a packing plan: byte offsets of consecutive 256-aligned pieces */
struct Layout {
    size_t used = 0;
    size_t add(size_t bytes)
    {
        const size_t at = used;
        used = rup256(used + bytes);
        return at;
    }
};

void layout(Server *s, Factor *f)
{
    if (f->n <= s->small_max) {
        f->large = false;
        f->prec = -1;
        f->np = rup16(f->n);
        f->a_elems = (size_t)f->np * f->np;
        f->dinv_off = f->a_elems;
        f->bytes = (f->a_elems + (size_t)((f->np + 63) / 64) * 4096) * sizeof(double);
    } else {
        f->large = true;
        f->nb = s->large_nb;
        f->NT = (f->n + f->nb - 1) / f->nb;
        f->ldt = rup16(f->nb);
        const size_t ts = (size_t)f->ldt * f->ldt;
        f->a_elems = (size_t)f->NT * f->NT * ts;
        f->dinv_off = f->a_elems;
        const size_t dinv = (size_t)f->NT * (f->ldt / 64) * 4096;
        /* a mixed factor keeps A's tiles for the residuals, and the scaling */
        f->prec = (s->prec >= 0 && f->ldt % 128 == 0 && f->ldt <= COPY_NORM_MAX_LDT) ? s->prec : -1;
        f->orig_off = f->dinv_off + dinv;
        f->sc_off = f->orig_off + (f->prec >= 0 ? f->a_elems : 0);
        f->bytes = (f->sc_off + (f->prec >= 0 ? (size_t)f->NT * f->ldt : 0)) * sizeof(double);
    }
    f->bytes = rup256(f->bytes);
}

void evict(Server *s, Factor *f)
{
    s->pool.release(f->off, f->bytes);
    s->lru.remove(f);
    s->factors.erase(f->id);
    delete f;
}

/* place a new factor, evicting unpinned least recently used ones */
bool place(Server *s, Factor *f)
{
    for (;;) {
        if (s->pool.alloc(f->bytes, &f->off)) return true;
        Factor *victim = nullptr;
        for (auto it = s->lru.rbegin(); it != s->lru.rend(); ++it)
            if ((*it)->pins == 0 && (*it)->state != F_BUSY) {
                victim = *it;
                break;
            }
        if (!victim) return false;
        evict(s, victim);
    }
}

void touch(Server *s, Factor *f)
{
    s->lru.remove(f);
    s->lru.push_front(f);
}

/* a request leaves the live list with `st` */
void finish(Server *s, Req *q, int st)
{
    const double t = now_s();
    s->status[q->ticket] = st;
    s->finished.push_back(q->ticket);
    if (st == 0) {
        s->latencies_ms.push_back((t - q->t_submit) * 1e3);
        ++s->completed;
    } else if (st < 0) {
        ++s->rejected;
    }
    if (q->f) {
        q->f->pins--;
        q->f->last_use = t;
        if (q->f->source == &q->r) q->f->source = nullptr;
    }
    q->f = nullptr;
    q->state = -1;     /* removed from the live list by the caller's sweep */
}

/* ---------------------------------------------------------------- jobs */

void launch_small_factor(Server *s, Slot &sl, Job *job)
{
    hipStream_t st = sl.lane.sm;
    const int count = (int)job->fs.size();
    std::vector<BatchGenSys> gen;
    std::vector<long long> off, doff;
    std::vector<int> np, nv;
    std::vector<Factor *> host_src;
    for (Factor *f : job->fs) {
        off.push_back((long long)(fptr(s, f) - (double *)s->base));
        doff.push_back((long long)(dptr(s, f) - (double *)s->base));
        np.push_back(f->np);
        nv.push_back(f->n);
        if (f->source && f->source->A) {
            host_src.push_back(f);
            continue;
        }
        BatchGenSys g;
        memset(&g, 0, sizeof g);
        g.off = off.back();
        g.seed = f->source ? f->source->seed : 0ULL;
        /* the LCG map of bigM = n steps (coreblas/include/coreblas/random.h:32-73) */
        unsigned long long a = 6364136223846793005ULL, cc = 1ULL, ra = 1ULL, rc = 0ULL;
        for (unsigned long long e = (unsigned long long)f->n; e; e >>= 1) {
            if (e & 1ULL) { rc = a * rc + cc; ra = a * ra; }
            cc = cc * (a + 1ULL);
            a = a * a;
        }
        g.abig = ra;
        g.cbig = rc;
        g.bump = (double)f->n;
        g.ld = g.rows = g.cols = f->np;
        g.nvr = g.nvc = f->n;
        g.bigM = f->n;
        g.pad_one = 1;
        gen.push_back(g);
    }
    std::vector<long long> unit0(gen.size() + 1, 0);
    for (size_t i = 0; i < gen.size(); ++i)
        unit0[i + 1] = unit0[i] + (long long)((gen[i].rows + 31) / 32) * gen[i].cols;
    /* staging: [descriptors ... info] uploaded as one block, then each host
     * matrix packed with identity padding */
    Layout d;
    const size_t o_gen = d.add(gen.size() * sizeof(BatchGenSys)), o_u = d.add(unit0.size() * 8),
                 o_off = d.add(off.size() * 8), o_doff = d.add(doff.size() * 8), o_np = d.add(np.size() * 4),
                 o_nv = d.add(nv.size() * 4), o_info = d.add((count + 1) * 4);
    const size_t desc_bytes = d.used;
    std::vector<size_t> o_a;
    for (Factor *f : host_src) o_a.push_back(d.add(f->a_elems * sizeof(double)));
    char *h = slot_host(sl, d.used);
    char *dv = slot_dev(sl, desc_bytes);
    if (!gen.empty()) memcpy(h + o_gen, gen.data(), gen.size() * sizeof(BatchGenSys));
    memcpy(h + o_u, unit0.data(), unit0.size() * 8);
    memcpy(h + o_off, off.data(), off.size() * 8);
    memcpy(h + o_doff, doff.data(), doff.size() * 8);
    memcpy(h + o_np, np.data(), np.size() * 4);
    memcpy(h + o_nv, nv.data(), nv.size() * 4);
    memset(h + o_info, 0, (count + 1) * 4);
    CHX_HIP(hipMemcpyAsync(dv, h, desc_bytes, hipMemcpyHostToDevice, st));
    for (size_t i = 0; i < host_src.size(); ++i) {
        Factor *f = host_src[i];
        const CHAMAMD_request_t *r = f->source;
        double *p = (double *)(h + o_a[i]);
        for (int cc = 0; cc < f->np; ++cc)
            for (int rr = 0; rr < f->np; ++rr)
                p[rr + (size_t)cc * f->np] = (rr >= f->n || cc >= f->n || rr < cc)
                                                 ? (rr == cc ? 1.0 : 0.0)
                                                 : r->A[rr + (size_t)cc * r->lda];
        CHX_HIP(hipMemcpyAsync(fptr(s, f), p, f->a_elems * sizeof(double), hipMemcpyHostToDevice, st));
    }
    if (!gen.empty())
        k::batch_gen((double *)s->base, (const BatchGenSys *)(dv + o_gen), (const long long *)(dv + o_u),
                     unit0.back(), (int)gen.size(), 1, st);
    int *d_info = (int *)(dv + o_info);
    k::potrf_jobs((double *)s->base, (const long long *)(dv + o_off), (const int *)(dv + o_np), nullptr,
                  (const int *)(dv + o_nv), (double *)s->base, (const long long *)(dv + o_doff), count, d_info + count,
                  d_info, 0, st);
    CHX_HIP(hipMemcpyAsync(h + o_info, d_info, count * sizeof(int), hipMemcpyDeviceToHost, st));
    job->host_off = o_info;
}

void launch_large_factor(Server *s, Slot &sl, Job *job)
{
    hipStream_t st = sl.lane.sm;
    Factor *f = job->fs[0];
    const int NT = f->NT, ldt = f->ldt;
    const size_t ts = (size_t)ldt * ldt;
    const bool mixed = f->prec >= 0 && !f->redo;
    const bool keep = f->prec >= 0;            /* A's tiles kept in the orig area */
    double *W = fptr(s, f), *target = keep ? optr(s, f) : W;
    Layout d;
    const size_t o_part = d.add(k::copy_lower_norm_scratch(NT, ldt) * sizeof(double));
    const size_t o_rows = d.add((size_t)NT * ldt * sizeof(double));
    const size_t pbytes = (size_t)NT * ts * 4;
    const size_t o_p0 = d.add(mixed ? pbytes : 0), o_p1 = d.add(mixed ? pbytes : 0);
    char *dv = slot_dev(sl, d.used);
    Layout hl;
    const size_t o_hinfo = hl.add(64);
    const size_t o_hrows = hl.add((size_t)f->n * sizeof(double));
    const bool upload = !f->redo && f->source && f->source->A;
    const size_t o_ha = hl.add(upload ? f->a_elems * sizeof(double) : 0);
    char *h = slot_host(sl, hl.used);
    if (!f->redo) {
        if (upload) {
            /* lower tiles in the tile layout, identity padding, upper tiles 0 */
            const CHAMAMD_request_t *r = f->source;
            double *p = (double *)(h + o_ha);
            memset(p, 0, f->a_elems * sizeof(double));
            for (int tn = 0; tn < NT; ++tn)
                for (int tm = tn; tm < NT; ++tm) {
                    double *T = p + ((size_t)tn * NT + tm) * ts;
                    for (int cc = 0; cc < ldt; ++cc)
                        for (int rr = 0; rr < ldt; ++rr) {
                            const int gi = tm * f->nb + rr, gj = tn * f->nb + cc;
                            double v = 0.0;
                            if (rr >= f->nb || cc >= f->nb || gi >= f->n || gj >= f->n)
                                v = (tm == tn && rr == cc) ? 1.0 : 0.0;
                            else if (gi >= gj)
                                v = r->A[gi + (size_t)gj * r->lda];
                            T[rr + (size_t)cc * ldt] = v;
                        }
                }
            CHX_HIP(hipMemcpyAsync(target, p, f->a_elems * sizeof(double), hipMemcpyHostToDevice, st));
        } else {
            k::plgsy(target, NT, NT, ldt, ldt, f->nb, f->nb, f->n, f->n, (double)f->n, ChamLower, f->n, 0, 0,
                     f->source ? f->source->seed : 0ULL, 0, true, st);
        }
    }
    if (keep) {
        /* working copy of the lower tiles (mixed: scaled by diag(A)^-1/2 on
         * both sides in the same pass) and ||A||_inf in one pass */
        double *rowsum = (double *)(dv + o_rows);
        if (mixed) k::diag_rsqrt(optr(s, f), NT, ldt, scptr(s, f), st);
        if (k::copy_lower_norm(optr(s, f), W, NT, ldt, f->nb, f->n, (double *)(dv + o_part), rowsum, st,
                               mixed ? scptr(s, f) : nullptr) != 0)
            fatal_hip(hipErrorInvalidValue, "copy_lower_norm", __FILE__, __LINE__);
        CHX_HIP(hipMemcpyAsync(h + o_hrows, rowsum, f->n * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    k::zero_ints(sl.d_info, 1, st);
    PotrfPlan p;
    p.A = W;
    p.NT = NT;
    p.ldt = ldt;
    p.nb = f->nb;
    p.nvalid_last = f->n - (NT - 1) * f->nb;
    p.d_info = sl.d_info;
    p.lane = &sl.lane;
    p.dinv = dptr(s, f);
    if (mixed) {
        p.lp_mode = f->prec;
        p.P[0] = dv + o_p0;
        p.P[1] = dv + o_p1;
    }
    potrf_run(p);
    CHX_HIP(hipMemcpyAsync(h + o_hinfo, sl.d_info, sizeof(int), hipMemcpyDeviceToHost, st));
    job->host_off = o_hrows;
    job->stage = keep ? 1 : 0;
}

void launch_small_solve(Server *s, Slot &sl, Job *job)
{
    hipStream_t st = sl.lane.sm;
    const int count = (int)job->rs.size();
    std::vector<long long> aoff, doff, boff;
    std::vector<int> np, nr;
    size_t belems = 0;
    for (Req *q : job->rs) {
        aoff.push_back((long long)(fptr(s, q->f) - (double *)s->base));
        doff.push_back((long long)(dptr(s, q->f) - (double *)s->base));
        boff.push_back((long long)belems);
        np.push_back(q->f->np);
        nr.push_back(q->r.nrhs);
        belems += (size_t)q->f->np * q->r.nrhs;
    }
    Layout d;
    const size_t o_aoff = d.add(count * 8), o_doff = d.add(count * 8), o_boff = d.add(count * 8),
                 o_np = d.add(count * 4), o_nr = d.add(count * 4), o_z = d.add((count + 1) * 4),
                 o_b = d.add(belems * sizeof(double));
    char *h = slot_host(sl, d.used);
    char *dv = slot_dev(sl, d.used);
    memcpy(h + o_aoff, aoff.data(), count * 8);
    memcpy(h + o_doff, doff.data(), count * 8);
    memcpy(h + o_boff, boff.data(), count * 8);
    memcpy(h + o_np, np.data(), count * 4);
    memcpy(h + o_nr, nr.data(), count * 4);
    memset(h + o_z, 0, (count + 1) * 4);
    double *hb = (double *)(h + o_b);
    for (int i = 0; i < count; ++i) {
        const CHAMAMD_request_t &r = job->rs[i]->r;
        double *p = hb + boff[i];
        for (int cc = 0; cc < r.nrhs; ++cc) {
            for (int rr = 0; rr < r.n; ++rr) p[rr + (size_t)cc * np[i]] = r.B[rr + (size_t)cc * r.ldb];
            for (int rr = r.n; rr < np[i]; ++rr) p[rr + (size_t)cc * np[i]] = 0.0;
        }
    }
    CHX_HIP(hipMemcpyAsync(dv, h, d.used, hipMemcpyHostToDevice, st));
    int *d_z = (int *)(dv + o_z);
    k::batch_solve((double *)s->base, (const long long *)(dv + o_aoff), (const int *)(dv + o_np), nullptr,
                   (double *)s->base, (const long long *)(dv + o_doff), (double *)(dv + o_b),
                   (const long long *)(dv + o_boff), (const int *)(dv + o_nr), d_z, count, d_z + count, 0, st);
    CHX_HIP(hipMemcpyAsync(hb, dv + o_b, belems * sizeof(double), hipMemcpyDeviceToHost, st));
    job->host_off = o_b;
    job->boff.assign(boff.begin(), boff.end());
}

/* transposed right-hand sides of a large request: element (c, r) at
 * c + r nrp, row r = tile t, row i at t ldt + i */
void pack_bt(const Factor *f, const CHAMAMD_request_t &r, int nrp, double *hb)
{
    memset(hb, 0, (size_t)nrp * f->NT * f->ldt * sizeof(double));
    for (int cc = 0; cc < r.nrhs; ++cc)
        for (int rr = 0; rr < r.n; ++rr) {
            const int t = rr / f->nb, i = rr % f->nb;
            hb[cc + (size_t)(t * f->ldt + i) * nrp] = r.B[rr + (size_t)cc * r.ldb];
        }
}

void unpack_bt(const Factor *f, CHAMAMD_request_t &r, int nrp, const double *hb)
{
    for (int cc = 0; cc < r.nrhs; ++cc)
        for (int rr = 0; rr < r.n; ++rr) {
            const int t = rr / f->nb, i = rr % f->nb;
            r.B[rr + (size_t)cc * r.ldb] = hb[cc + (size_t)(t * f->ldt + i) * nrp];
        }
}

void launch_large_solve(Server *s, Slot &sl, Job *job)
{
    hipStream_t st = sl.lane.sm;
    Req *q = job->rs[0];
    Factor *f = q->f;
    const int nrp = rup16(std::max(q->r.nrhs, 1)), rows = f->NT * f->ldt;
    const size_t vb = (size_t)nrp * rows;
    /* few right-hand sides: the narrow strip solve, with the 128 x 128
     * strip inverses formed here from the factor's 64 x 64 ones; else the
     * MFMA strip solve with its predecessor products, formed here too */
    const bool narrow = k::narrow_eligible(f->ldt, q->r.nrhs);
    Layout d;
    const size_t o_bt = d.add(vb * 8), o_aux = d.add((size_t)rows * 128 * 8);
    char *dv = slot_dev(sl, d.used);
    char *h = slot_host(sl, vb * sizeof(double));
    pack_bt(f, q->r, nrp, (double *)h);
    double *BT = (double *)(dv + o_bt), *aux = (double *)(dv + o_aux);
    CHX_HIP(hipMemcpyAsync(BT, h, vb * sizeof(double), hipMemcpyHostToDevice, st));
    k::zero_ints(sl.d_info, 1, st);
    const long long dper = (long long)(f->ldt / 64) * 4096;
    if (narrow) {
        k::dinv128(fptr(s, f), dptr(s, f), dper, 4096, 64, f->NT, f->ldt, aux, st);
        solve_strips_lane(&sl.lane, fptr(s, f), dptr(s, f), f->NT, f->ldt, BT, nrp, sl.d_info, aux, q->r.nrhs);
    } else {
        double *eb = aux + (size_t)rows * 64;
        k::strip_products(fptr(s, f), dptr(s, f), dper, 4096, 64, f->NT, f->ldt, aux, eb, st);
        solve_strips_lane(&sl.lane, fptr(s, f), dptr(s, f), f->NT, f->ldt, BT, nrp, sl.d_info, nullptr, 0, aux, eb);
    }
    CHX_HIP(hipMemcpyAsync(h, BT, vb * sizeof(double), hipMemcpyDeviceToHost, st));
    job->nrp = nrp;
    job->rows = rows;
}

/* mixed-precision refinement of one large request: launch, or continue
 * after the host's convergence test */
void mixed_step(Server *s, Slot &sl, Job *job, bool first)
{
    hipStream_t st = sl.lane.sm;
    Req *q = job->rs[0];
    Factor *f = q->f;
    const size_t vb = (size_t)job->nrp * job->rows;
    double *sc = scptr(s, f);
    if (first) {
        CHX_HIP(hipMemcpyAsync(job->BT, sl.host, vb * sizeof(double), hipMemcpyHostToDevice, st));
        CHX_HIP(hipMemcpyAsync(job->XT, job->BT, vb * sizeof(double), hipMemcpyDeviceToDevice, st));
        k::zero_ints(sl.d_info, 1, st);
        k::scale_bt(job->XT, job->nrp, job->rows, sc, st);
        const long long dper = (long long)(f->ldt / 64) * 4096;
        if (job->d128)
            k::dinv128(fptr(s, f), dptr(s, f), dper, 4096, 64, f->NT, f->ldt, job->d128, st);
        else
            k::strip_products(fptr(s, f), dptr(s, f), dper, 4096, 64, f->NT, f->ldt, job->ef, job->eb, st);
        solve_strips_lane(&sl.lane, fptr(s, f), dptr(s, f), f->NT, f->ldt, job->XT, job->nrp, sl.d_info, job->d128,
                          q->r.nrhs, job->ef, job->eb);
        k::scale_bt(job->XT, job->nrp, job->rows, sc, st);
    } else {
        /* correction: X += A_lp^-1 R */
        k::scale_bt(job->RT, job->nrp, job->rows, sc, st);
        solve_strips_lane(&sl.lane, fptr(s, f), dptr(s, f), f->NT, f->ldt, job->RT, job->nrp, sl.d_info, job->d128,
                          q->r.nrhs, job->ef, job->eb);
        k::scale_bt(job->RT, job->nrp, job->rows, sc, st);
        k::axpy(job->XT, job->RT, vb, st);
    }
    /* FP64 residual on the kept lower tiles of A, column max-norms of X and R */
    k::residual(optr(s, f), f->NT, f->ldt, job->XT, job->BT, job->RT, job->nrp, job->part, st);
    k::zero_ints((int *)job->nw, 4 * job->nrp, st);
    k::colmax2(job->XT, job->RT, job->nrp, job->nrp, job->rows, job->nw, st);
    CHX_HIP(hipMemcpyAsync(sl.host + job->host_off, job->nw, 2 * job->nrp * sizeof(double), hipMemcpyDeviceToHost,
                           st));
    job->stage = 1;
}

void launch_mixed_solve(Server *s, Slot &sl, Job *job)
{
    Req *q = job->rs[0];
    Factor *f = q->f;
    job->nrp = rup16(std::max(q->r.nrhs, 1));
    job->rows = f->NT * f->ldt;
    const size_t vb = (size_t)job->nrp * job->rows;
    Layout d;
    const bool narrow = k::narrow_eligible(f->ldt, q->r.nrhs);
    const size_t o_bt = d.add(vb * 8), o_xt = d.add(vb * 8), o_rt = d.add(vb * 8),
                 o_part = d.add(k::residual_scratch(f->NT, f->ldt, job->nrp) * 8), o_nw = d.add(2 * job->nrp * 8),
                 o_aux = d.add((size_t)job->rows * 128 * 8);
    char *dv = slot_dev(sl, d.used);
    /* the narrow solve's strip inverses, or the MFMA solve's predecessor
     * products (formed in the first step, read by every correction) */
    job->d128 = narrow ? (double *)(dv + o_aux) : nullptr;
    job->ef = narrow ? nullptr : (double *)(dv + o_aux);
    job->eb = narrow ? nullptr : job->ef + (size_t)job->rows * 64;
    job->BT = (double *)(dv + o_bt);
    job->XT = (double *)(dv + o_xt);
    job->RT = (double *)(dv + o_rt);
    job->part = (double *)(dv + o_part);
    job->nw = (double *)(dv + o_nw);
    Layout hl;
    hl.add(vb * 8);
    job->host_off = hl.add(2 * job->nrp * 8);
    char *h = slot_host(sl, hl.used);
    pack_bt(f, q->r, job->nrp, (double *)h);
    job->iter = 0;
    mixed_step(s, sl, job, true);
    ++s->mixed_solves;
}

/* ---------------------------------------------------------- completion */

/* Job on `sl` completed (its event fired).  Returns the number of requests
 * finished; a job that issues a further step stays on the lane. */
int complete(Server *s, Slot &sl)
{
    Job *job = sl.job;
    int done = 0;
    bool keep = false;
    switch (job->kind) {
    case J_SMALL_FACTOR: {
        const int *info = (const int *)(sl.host + job->host_off);
        for (size_t i = 0; i < job->fs.size(); ++i) {
            Factor *f = job->fs[i];
            f->info = info[i];
            f->state = F_READY;
            const double nn = f->n;
            s->flops += nn * nn * nn / 3.0;
            ++s->factored;
        }
        break;
    }
    case J_LARGE_FACTOR: {
        Factor *f = job->fs[0];
        f->info = *(const int *)sl.host;
        if (job->stage == 1) {
            const double *rs = (const double *)(sl.host + job->host_off);
            double a = 0.0;
            for (int i = 0; i < f->n; ++i) a = nan_max(a, rs[i]);
            f->anrm = a;
        }
        const double nn = f->n;
        s->flops += nn * nn * nn / 3.0;
        ++s->factored;
        if (f->prec >= 0 && !f->redo && f->info != 0) {
            /* the reduced-precision factorisation failed: FP64 from the kept tiles */
            f->redo = true;
            f->state = F_NEW;
            ++s->fallbacks;
        } else {
            if (f->redo) f->prec = -1;
            f->state = F_READY;
        }
        break;
    }
    case J_SMALL_SOLVE: {
        const double *hb = (const double *)(sl.host + job->host_off);
        for (size_t i = 0; i < job->rs.size(); ++i) {
            Req *q = job->rs[i];
            CHAMAMD_request_t &r = q->r;
            const int np = q->f->np;
            for (int cc = 0; cc < r.nrhs; ++cc)
                for (int rr = 0; rr < r.n; ++rr) r.B[rr + (size_t)cc * r.ldb] = hb[job->boff[i] + rr + (size_t)cc * np];
            s->flops += 2.0 * r.n * (double)r.n * r.nrhs;
            finish(s, q, 0);
            ++done;
        }
        break;
    }
    case J_LARGE_SOLVE: {
        Req *q = job->rs[0];
        q->f->users--;
        unpack_bt(q->f, q->r, job->nrp, (const double *)sl.host);
        s->flops += 2.0 * q->r.n * (double)q->r.n * q->r.nrhs;
        finish(s, q, 0);
        ++done;
        break;
    }
    case J_MIXED_SOLVE: {
        Req *q = job->rs[0];
        Factor *f = q->f;
        if (job->stage == 2) {
            f->users--;
            unpack_bt(f, q->r, job->nrp, (const double *)sl.host);
            s->flops += 2.0 * q->r.n * (double)q->r.n * q->r.nrhs * (1 + job->iter);
            finish(s, q, 0);
            ++done;
            break;
        }
        /* LAPACK dsposv's test ||r_j|| <= ||x_j|| ||A|| eps sqrt(N) per column */
        const double *xn = (const double *)(sl.host + job->host_off), *rn = xn + job->nrp;
        const double eps = std::ldexp(1.0, -53), N = f->n;
        const double cte = f->anrm * eps * std::sqrt(N);
        bool ok = true;
        double worst = 0.0;
        for (int j = 0; j < q->r.nrhs; ++j) {
            if (!(rn[j] <= xn[j] * cte)) ok = false;
            const double dd = f->anrm * xn[j] * N * eps;
            worst = nan_max(worst, dd > 0 ? rn[j] / dd : (rn[j] > 0 ? INFINITY : 0.0));
        }
        if (ok && *sl.h_info == 0) {
            s->max_berr = nan_max(s->max_berr, worst);
            const size_t vb = (size_t)job->nrp * job->rows;
            CHX_HIP(hipMemcpyAsync(sl.host, job->XT, vb * sizeof(double), hipMemcpyDeviceToHost, sl.lane.sm));
            CHX_HIP(hipEventRecord(sl.done, sl.lane.sm));
            job->stage = 2;
            keep = true;
        } else if (job->iter < 30 && *sl.h_info == 0) {
            ++job->iter;
            ++s->ir_steps;
            mixed_step(s, sl, job, false);
            CHX_HIP(hipMemcpyAsync(sl.h_info, sl.d_info, sizeof(int), hipMemcpyDeviceToHost, sl.lane.sm));
            CHX_HIP(hipEventRecord(sl.done, sl.lane.sm));
            keep = true;
        } else {
            /* no convergence: factor again in FP64 and solve with that */
            f->users--;
            ++s->fallbacks;
            if (f->prec >= 0 && !f->redo && f->state == F_READY) {
                f->redo = true;
                f->state = F_NEW;
            }
            q->state = R_WAIT;
        }
        break;
    }
    }
    if (!keep) {
        delete job;
        sl.job = nullptr;
    }
    return done;
}

/* ------------------------------------------------------------- rounds */

int poll(Server *s)
{
    int done = 0;
    for (Slot &sl : s->slots) {
        if (!sl.job) continue;
        const hipError_t e = hipEventQuery(sl.done);
        if (e == hipErrorNotReady) continue;
        CHX_HIP(e);
        done += complete(s, sl);
    }
    return done;
}

void admit(Server *s)
{
    /* first every live request whose matrix is resident pins it, so that
     * no placement below evicts a factor a live request needs */
    for (Req *q : s->live) {
        if (q->f || q->state < 0) continue;
        auto it = s->factors.find(q->r.matrix_id);
        if (it == s->factors.end()) continue;
        Factor *f = it->second;
        if (f->n != q->r.n) {
            finish(s, q, CHAMELEON_ERR_ILLEGAL_VALUE);
            continue;
        }
        ++s->reused;                           /* served by a factor another request brought */
        f->pins++;
        q->f = f;
        touch(s, f);
    }
    /* new factors are placed in arrival order: after the first that does
     * not fit, later requests only attach to factors already placed, so an
     * overloaded server (thousands of requests waiting for memory) spends
     * one failed placement per pass, not one per waiting request, and a
     * large factor is not starved by smaller ones arriving after it */
    bool blocked = false;
    for (Req *q : s->live) {
        if (q->f || q->state < 0) continue;
        auto it = s->factors.find(q->r.matrix_id);
        if (it == s->factors.end() && blocked) continue;
        if (it != s->factors.end()) {          /* placed for an earlier request just now */
            Factor *f = it->second;
            if (f->n != q->r.n) {
                finish(s, q, CHAMELEON_ERR_ILLEGAL_VALUE);
                continue;
            }
            ++s->reused;
            f->pins++;
            q->f = f;
            continue;
        }
        Factor *f = new Factor();
        f->id = q->r.matrix_id;
        f->n = q->r.n;
        layout(s, f);
        if (f->bytes > s->pool.size) {
            delete f;
            finish(s, q, CHAMELEON_ERR_OUT_OF_RESOURCES);
            continue;
        }
        if (!place(s, f)) {            /* fits once live requests release memory */
            delete f;
            blocked = true;
            continue;
        }
        f->source = &q->r;
        s->factors[f->id] = f;
        s->lru.push_front(f);
        f->pins++;
        q->f = f;
    }
    /* failed factorisations answer their requests */
    for (Req *q : s->live)
        if (q->f && q->state == R_WAIT && q->f->state == F_READY && q->f->info != 0) finish(s, q, q->f->info);
    while (!s->live.empty() && s->live.front()->state < 0) {
        delete s->live.front();
        s->live.pop_front();
    }
    for (auto it = s->live.begin(); it != s->live.end();) {
        if ((*it)->state < 0) {
            delete *it;
            it = s->live.erase(it);
        } else {
            ++it;
        }
    }
}

void start(Server *s, Slot &sl, Job *job)
{
    sl.job = job;
    switch (job->kind) {
    case J_SMALL_FACTOR: launch_small_factor(s, sl, job); break;
    case J_LARGE_FACTOR: launch_large_factor(s, sl, job); break;
    case J_SMALL_SOLVE: launch_small_solve(s, sl, job); break;
    case J_LARGE_SOLVE: launch_large_solve(s, sl, job); break;
    case J_MIXED_SOLVE: launch_mixed_solve(s, sl, job); break;
    }
    if (job->kind == J_MIXED_SOLVE)
        CHX_HIP(hipMemcpyAsync(sl.h_info, sl.d_info, sizeof(int), hipMemcpyDeviceToHost, sl.lane.sm));
    CHX_HIP(hipEventRecord(sl.done, sl.lane.sm));
}

void launch(Server *s)
{
    /* lane 0 never takes a large factorisation, so that small work and
     * solves always find a lane; the others take any job */
    std::vector<Slot *> idle;
    int busy_large = 0;
    for (size_t i = 1; i <= s->slots.size(); ++i) {
        Slot &sl = s->slots[i % s->slots.size()];
        if (!sl.job) idle.push_back(&sl);
        else if (sl.job->kind == J_LARGE_FACTOR) ++busy_large;
    }
    size_t next = 0;
    auto take = [&]() -> Slot * { return next < idle.size() ? idle[next++] : nullptr; };
    /* small solves: every ready small request in one launch */
    std::vector<Req *> small_ready;
    for (Req *q : s->live)
        if (q->f && q->state == R_WAIT && !q->f->large && q->f->state == F_READY && q->f->info == 0)
            small_ready.push_back(q);
    if (!small_ready.empty()) {
        if (Slot *sl = take()) {
            Job *job = new Job();
            job->kind = J_SMALL_SOLVE;
            job->rs = small_ready;
            for (Req *q : small_ready) q->state = R_SOLVE;
            start(s, *sl, job);
        }
    }
    /* new small matrices: one factorisation launch */
    std::vector<Factor *> small_new;
    for (Req *q : s->live)
        if (q->f && !q->f->large && q->f->state == F_NEW) {
            q->f->state = F_BUSY;          /* listed once */
            small_new.push_back(q->f);
        }
    if (!small_new.empty()) {
        if (Slot *sl = take()) {
            std::sort(small_new.begin(), small_new.end(), [](Factor *a, Factor *b) { return a->n > b->n; });
            Job *job = new Job();
            job->kind = J_SMALL_FACTOR;
            job->fs = small_new;
            start(s, *sl, job);
        } else {
            for (Factor *f : small_new) f->state = F_NEW;
        }
    }
    /* large solves, oldest first, one per lane (the strip solve's tasks
     * are ordered by a ticket, so several run side by side) */
    for (Req *q : s->live) {
        if (!(q->f && q->state == R_WAIT && q->f->large && q->f->state == F_READY && q->f->info == 0)) continue;
        Slot *sl = take();
        if (!sl) break;
        Job *job = new Job();
        job->kind = q->f->prec >= 0 ? J_MIXED_SOLVE : J_LARGE_SOLVE;
        job->rs.push_back(q);
        q->state = R_SOLVE;
        q->f->users++;
        start(s, *sl, job);
    }
    /* large factorisations on lanes 1.., a factor that is factored again in
     * FP64 waiting for the solves still reading it */
    const int max_large = std::max(1, (int)s->slots.size() - 1);
    for (Req *q : s->live) {
        if (busy_large >= max_large) break;
        if (!(q->f && q->f->large && q->f->state == F_NEW && q->f->users == 0)) continue;
        Slot *sl = nullptr;
        for (size_t i = next; i < idle.size(); ++i)
            if (idle[i] != &s->slots[0] || s->slots.size() == 1) {
                sl = idle[i];
                idle.erase(idle.begin() + i);
                break;
            }
        if (!sl) break;
        Job *job = new Job();
        job->kind = J_LARGE_FACTOR;
        job->fs.push_back(q->f);
        q->f->state = F_BUSY;
        ++busy_large;
        start(s, *sl, job);
    }
    long long running = 0;
    for (Slot &sl : s->slots) running += sl.job != nullptr;
    s->max_jobs = std::max(s->max_jobs, running);
}

int progress(Server *s)
{
    const int done = poll(s);
    admit(s);
    launch(s);
    return done;
}

bool idle(Server *s)
{
    for (Slot &sl : s->slots)
        if (sl.job) return false;
    return s->live.empty();
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
    int lanes = 8;
    v = getenv("CHAMAMD_SERVE_LANES");
    if (v && *v) lanes = std::min(16, std::max(1, atoi(v)));
    /* the strip solve needs tiles a multiple of 64 */
    const int nb = ctx()->nb;
    s->large_nb = (nb % 64 == 0) ? nb : 512;
    s->slots.resize(lanes);
    for (Slot &sl : s->slots) {
        CHX_HIP(hipStreamCreateWithFlags(&sl.lane.sm, hipStreamNonBlocking));
        sl.lane.su = nullptr;
        CHX_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        CHX_HIP(hipMalloc(&sl.d_info, 256));
        CHX_HIP(hipHostMalloc((void **)&sl.h_info, 256, hipHostMallocDefault));
        *sl.h_info = 0;
        k::zero_ints(sl.d_info, 64, sl.lane.sm);
    }
    *server = s;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Destroy(CHAMAMD_server_t **server)
{
    if (!server || !*server) return CHAMELEON_ERR_UNALLOCATED;
    CHAMAMD_server_t *s = *server;
    for (Slot &sl : s->slots) {
        CHX_HIP(hipStreamSynchronize(sl.lane.sm));
        delete sl.job;
        lane_release(&sl.lane);
        CHX_HIP(hipStreamDestroy(sl.lane.sm));
        CHX_HIP(hipEventDestroy(sl.done));
        CHX_HIP(hipFree(sl.d_info));
        CHX_HIP(hipHostFree(sl.h_info));
        if (sl.dev) CHX_HIP(hipFree(sl.dev));
        if (sl.host) CHX_HIP(hipHostFree(sl.host));
    }
    for (auto &kv : s->factors) delete kv.second;
    for (Req *q : s->live) delete q;
    if (s->base) CHX_HIP(hipFree(s->base));
    delete s;
    *server = nullptr;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Set_Precision(CHAMAMD_server_t *s, int prec)
{
    if (!s) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (prec != -1 && (prec < CHAMAMD_PREC_FP32 || prec > CHAMAMD_PREC_FP8)) {
        error("CHAMAMD_Server_Set_Precision", "illegal value of prec");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    s->prec = prec;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Submit(CHAMAMD_server_t *s, const CHAMAMD_request_t *req, long long *ticket)
{
    if (!s || !req || !ticket) return CHAMELEON_ERR_ILLEGAL_VALUE;
    if (req->n < 1 || req->nrhs < 0 || !req->B || req->ldb < req->n || (req->A && req->lda < req->n)) {
        error("CHAMAMD_Server_Submit", "illegal request");
        return CHAMELEON_ERR_ILLEGAL_VALUE;
    }
    Req *q = new Req();
    q->ticket = s->next_ticket++;
    q->r = *req;
    q->t_submit = now_s();
    s->live.push_back(q);
    *ticket = q->ticket;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Progress(CHAMAMD_server_t *s, int *processed)
{
    if (!s || !ctx()) return CHAMELEON_ERR_ILLEGAL_VALUE;
    const int n = progress(s);
    if (processed) *processed = n;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Drain(CHAMAMD_server_t *s)
{
    if (!s || !ctx()) return CHAMELEON_ERR_ILLEGAL_VALUE;
    progress(s);
    while (!idle(s)) {
        /* wait for the oldest job in flight, then move on */
        Slot *wait = nullptr;
        for (Slot &sl : s->slots)
            if (sl.job) {
                wait = &sl;
                break;
            }
        if (!wait) {
            /* nothing in flight and nothing could start: requests that wait
             * for memory no live request will release */
            for (Req *q : s->live)
                if (q->state >= 0) finish(s, q, CHAMELEON_ERR_OUT_OF_RESOURCES);
            for (Req *q : s->live) delete q;
            s->live.clear();
            break;
        }
        CHX_HIP(hipEventSynchronize(wait->done));
        progress(s);
    }
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

int CHAMAMD_Server_Finished(CHAMAMD_server_t *s, long long *tickets, int cap, int *count)
{
    if (!s || !count || cap < 0 || (cap > 0 && !tickets)) return CHAMELEON_ERR_ILLEGAL_VALUE;
    const int n = (int)std::min<size_t>((size_t)cap, s->finished.size());
    if (n) memcpy(tickets, s->finished.data(), (size_t)n * sizeof(long long));
    s->finished.erase(s->finished.begin(), s->finished.begin() + n);
    *count = n;
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
    st->mixed_solves = s->mixed_solves;
    st->ir_steps = s->ir_steps;
    st->fallbacks = s->fallbacks;
    st->max_jobs = s->max_jobs;
    st->max_berr = s->max_berr;
    return CHAMELEON_SUCCESS;
}

int CHAMAMD_Server_Reset_Stats(CHAMAMD_server_t *s)
{
    if (!s) return CHAMELEON_ERR_ILLEGAL_VALUE;
    s->latencies_ms.clear();
    s->completed = s->rejected = s->factored = s->reused = 0;
    s->mixed_solves = s->ir_steps = s->fallbacks = s->max_jobs = 0;
    s->flops = s->max_berr = 0.0;
    return CHAMELEON_SUCCESS;
}

} /* extern "C" */
