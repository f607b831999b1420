/*
 * noalloc.hip -- dpotrs that allocates nothing.
 *
 * CHAMAMD_dpotrs_Tile_Async_NoAlloc runs the same kernels, in the same
 * order, as CHAMELEON_dpotrs_Tile_Async with tiles a multiple of 64 (the
 * 64 x 64 diagonal-block inverses, then with at most 8 right-hand sides and
 * tiles a multiple of 128 the 128 x 128 strip inverses, else the strip
 * solve's predecessor products, the transposition of B, one strip-solve
 * launch per direction, the transposition back), but takes every buffer
 * beyond A and B from `scratch`, a device array of fixed size that the
 * code object carries, instead of the library's growable workspaces.  It
 * is for callers that cannot call hipMalloc: code under stream capture,
 * tight serving loops, harnesses that forbid allocation.
 *
 * Layout of `scratch` on each device:
 *   [0, NOALLOC_FLAG_BYTES)       publication flags of the strip solves,
 *                                 tagged by launch epoch (never cleared
 *                                 between calls; cleared once at set-up,
 *                                 CHAMELEON_Init), and in its last int the
 *                                 launches' task ticket (cleared before
 *                                 each launch)
 *   [NOALLOC_FLAG_BYTES, ...)     doubles: the 64 x 64 inverses
 *                                 (Npad * 64), the 128 x 128 strip
 *                                 inverses of a narrow solve or the two
 *                                 directions' predecessor products
 *                                 (Npad * 128), then the transposed
 *                                 right-hand sides (nrp * Npad)
 * with Npad = NT * ldt (padded rows) and nrp = B's padded column count.
 * Calls are rejected (CHAMELEON_ERR_OUT_OF_RESOURCES) when that exceeds
 * NOALLOC_MAX_DOUBLES (128 MiB of doubles), e.g. N = 32768 is served with
 * up to 320 right-hand sides.
 *
 * Every call on a device shares that device's array, which keeps the last
 * call's contents: calls that overlap in time on one device race (issue
 * them on one host thread; the library's main stream orders them).
 */
#include "internal.h"
#include "chameleon_amd.h"

#include <mutex>

using namespace chx;

#define NOALLOC_FLAG_BYTES (64u << 10)
#define NOALLOC_MAX_DOUBLES ((size_t)1 << 24)      /* 128 MiB of doubles */
#define NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[NOALLOC_FLAG_BYTES + NOALLOC_MAX_DOUBLES * sizeof(double)];

namespace {

struct DeviceScratch {
    unsigned char *ptr = nullptr;   /* this device's address of `scratch` */
    int epoch = 0;                  /* last strip-solve epoch used        */
};

std::mutex g_lock;
DeviceScratch g_dev[NOALLOC_MAX_DEVICES];

/* Address of `scratch` on `device` (the current device): looked up once,
 * and the flags cleared by a kernel on `s`.  The lookup makes the runtime
 * load the code object on the device, which places its global arrays, so
 * it runs at the library's init, not inside the entry point. */
DeviceScratch *device_scratch(int device, hipStream_t s)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    DeviceScratch *d = &g_dev[device];
    if (d->ptr) return d;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        error("CHAMAMD_dpotrs_Tile_Async_NoAlloc", hipGetErrorString(e));
        return nullptr;
    }
    d->ptr = (unsigned char *)p;
    k::zero_ints((int *)d->ptr, (int)(NOALLOC_FLAG_BYTES / sizeof(int)), s);
    d->epoch = 0;
    return d;
}

} /* namespace */

namespace chx {
/* set-up at CHAMELEON_Init, with `device` current */
void noalloc_init(int device, hipStream_t s)
{
    std::lock_guard<std::mutex> guard(g_lock);
    (void)device_scratch(device, s);
}
} /* namespace chx */

extern "C" int CHAMAMD_dpotrs_Tile_Async_NoAlloc(cham_uplo_t uplo, CHAM_desc_t *A, CHAM_desc_t *B,
                                                RUNTIME_sequence_t *sequence, RUNTIME_request_t *request)
{
    static const char *fn = "CHAMAMD_dpotrs_Tile_Async_NoAlloc";
    Context *c = ctx();
    if (!c) {
        error(fn, "CHAMELEON not initialized");
        return CHAMELEON_ERR_NOT_INITIALIZED;
    }
    if (!sequence || !request) {
        error(fn, "NULL sequence or request");
        return CHAMELEON_ERR_UNALLOCATED;
    }
    if (sequence->status != CHAMELEON_SUCCESS) return request_fail(sequence, request, CHAMELEON_ERR_SEQUENCE_FLUSHED);
    request->status = CHAMELEON_SUCCESS;
    if (uplo != ChamUpper && uplo != ChamLower) {
        error(fn, "illegal value of uplo");
        return request_fail(sequence, request, -1);
    }
    if (desc_check(A) != CHAMELEON_SUCCESS || desc_check(B) != CHAMELEON_SUCCESS) {
        error(fn, "invalid descriptor");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (A->m != A->n || A->mb != A->nb || B->mb != B->nb || B->m != A->n || B->mb != A->mb ||
        cham_get_flttype(A->dtyp) != ChamRealDouble || cham_get_flttype(B->dtyp) != ChamRealDouble) {
        error(fn, "A must be square, B must have A's rows and tiling, both real double with square tiles");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (A->m == 0 || B->n == 0) return CHAMELEON_SUCCESS;
    Mirror *MA = mirror_of(A), *MB = mirror_of(B);
    const int want = uplo == ChamUpper ? MIRROR_TRANS : MIRROR_NORMAL;
    /* both operands must already be resident: making a device copy would
     * allocate; an orientation change is an in-place transposition */
    if (!MA || !MB || !MA->d || !MB->d || MA->state == MIRROR_NONE || MB->state == MIRROR_NONE) {
        error(fn, "A and B must be resident on the device (left there by an earlier _Tile_Async call)");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (MA->lower_only && want != MIRROR_NORMAL) {
        error(fn, "A's device copy holds its lower tiles only (an Upper solve would need an upload)");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    if (MA->ldt % 64 != 0 || MB->ldt != MA->ldt) {
        error(fn, "the padded tile size must be a multiple of 64");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    const int NT = MA->mt, ldt = MA->ldt;
    const size_t npad = (size_t)NT * ldt;
    const int nrp = MB->nt * MB->wdt;
    const bool narrow = k::narrow_eligible(ldt, B->n);
    /* the narrow solve's strip inverses or the MFMA solve's predecessor
     * products: Npad * 128 doubles either way */
    const size_t ndinv = npad * 64, naux = npad * 128, nbt = (size_t)nrp * npad;
    const int S = (int)(npad / 64);
    if (ndinv + naux + nbt > NOALLOC_MAX_DOUBLES || (size_t)(4 * S + 1) * sizeof(int) > NOALLOC_FLAG_BYTES) {
        error(fn, "problem larger than the fixed scratch array (N_padded * (192 + NRHS_padded) <= 16777216)");
        return request_fail(sequence, request, CHAMELEON_ERR_OUT_OF_RESOURCES);
    }
    hipStream_t s = c->s_main;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) device = c->device;
    std::lock_guard<std::mutex> guard(g_lock);
    DeviceScratch *d = (device >= 0 && device < NOALLOC_MAX_DEVICES) ? &g_dev[device] : nullptr;
    if (!d || !d->ptr) {
        error(fn, "the scratch array was not set up on this device at CHAMELEON_Init");
        return request_fail(sequence, request, CHAMELEON_ERR_UNEXPECTED);
    }
    /* two epochs per direction-and-column-group; flags hold the epoch of
     * the launch that set them, so a fresh range needs no clearing until
     * the counter wraps */
    const int chunks = (nrp + 63) / 64;
    if (d->epoch > (1 << 30)) {
        k::zero_ints((int *)d->ptr, (int)(NOALLOC_FLAG_BYTES / sizeof(int)), s);
        d->epoch = 0;
    }
    const int epoch = d->epoch + 1;
    d->epoch += 2 * chunks;
    int *flags = (int *)d->ptr;
    /* the strip solves' task ticket: the last int of the flag region, past
     * every flag (at most 4 * S <= 4096 of them) */
    unsigned *ticket = (unsigned *)(d->ptr + NOALLOC_FLAG_BYTES) - 1;
    double *dinv = (double *)(d->ptr + NOALLOC_FLAG_BYTES);
    double *aux = dinv + ndinv;
    double *d128 = narrow ? aux : nullptr;
    double *ef = narrow ? nullptr : aux, *eb = narrow ? nullptr : aux + npad * 64;
    double *BT = dinv + ndinv + naux;
    SeqState *st = seq_state(sequence);

    if (MB->state != MIRROR_NORMAL) {
        error(fn, "B's device copy is not in normal orientation");
        return request_fail(sequence, request, CHAMELEON_ERR_ILLEGAL_VALUE);
    }
    /* the same launches as the allocating path: the padding and orientation
     * fix-ups of mirror_acquire (in place), then solve_plan and solve_run */
    if (!MA->diag_identity) {
        k::fix_padding(MA->d, MA->mt, MA->nt, MA->ldt, MA->wdt, A->mb, A->nb, A->m, A->n, MA->eltsize, true, s);
        MA->diag_identity = true;
    }
    if (MA->state != want) {
        k::transpose_tiles_inplace((double *)MA->d, NT, ldt, s);
        MA->state = want;
    }
    if (MB->diag_identity) {
        k::fix_padding(MB->d, MB->mt, MB->nt, MB->ldt, MB->wdt, B->mb, B->nb, B->m, B->n, MB->eltsize, false, s);
        MB->diag_identity = false;
    }
    const double *Ad = (const double *)MA->d;
    const size_t ts = (size_t)ldt * ldt, dper = (size_t)(ldt / 64) * 64 * 64;
    k::dinv_all(Ad, NT, (size_t)(NT + 1) * ts, ldt, dinv, s);
    if (narrow)
        k::dinv128(Ad, dinv, (long long)dper, 64LL * 64, 64, NT, ldt, d128, s);
    else
        k::strip_products(Ad, dinv, (long long)dper, 64LL * 64, 64, NT, ldt, ef, eb, s);
    k::rhs_transpose((double *)MB->d, BT, NT, ldt, MB->wdt, MB->nt, 1, s);
    for (int dir = 0; dir < 2; ++dir)
        k::solve_strips(Ad, dinv, (long long)dper, 64LL * 64, 64, BT, NT, ldt, nrp, dir, flags, epoch + dir * chunks,
                        ticket, st->d_info, s, 0, d128, B->n, dir ? eb : ef);
    k::rhs_transpose((double *)MB->d, BT, NT, ldt, MB->wdt, MB->nt, 0, s);
    st->pending_info = true;
    mirror_mark_written(B);
    B->sync = 1;
    return CHAMELEON_SUCCESS;
}
