"""Parity at the benchmark configurations' own sizes.

* CPU configuration (Chameleon's CPU dposv case): N = 4096, NB = 320,
  NRHS = 16, LAPACK-layout CHAMELEON_dposv against the oracle (OpenBLAS
  tile kernels) and Chameleon's two checks (threshold 60).
* FP64 GPU configuration: N = 8192 with NB 256 / 384 / 512, NRHS = 64,
  against the oracle; N = 16384 and 32768 with NB = 512: the factor against
  the oracle's (relative Frobenius of the lower triangles, column block by
  column block so that only two N^2 host copies exist) and the solution's
  scaled residual computed on the device.
* Mixed precision at N = 32768 (kappa 1e4, A = Q diag(lambda) Q^T with Q
  from the QR factorisation of a Gaussian matrix): split-BF16 products +
  FP64 refinement reach the FP64 solve's accuracy.
* Batched: the full 20 000-system batch (n log-uniform in [16, 1024],
  nrhs 1..8): every system's scaled residual <= 60, and 50 systems spread
  over the size range against the oracle.
* Serving: 10 s of Poisson arrivals at 2000 requests/s (n log-uniform in
  [64, 8192], nrhs in {1, 4, 16}, 30 % reusing an earlier matrix, 200 GB
  factor budget): every request served with status 0 and every solution's
  scaled residual <= 60, computed on the device with A regenerated.

Tolerances as in test_gpu_cholesky.py: factor 1e-12, solution 1e-11.
"""
import ctypes
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

L_, UL_ = 122, 123
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def ch():
    import chameleon_amd as ch
    ch.init()
    yield ch


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def rel_frob_lower(F, R, blk=1024):
    """||tril(F) - tril(R)||_F / ||tril(R)||_F without full-size temporaries."""
    n = R.shape[1]
    num = den = 0.0
    for j0 in range(0, n, blk):
        j1 = min(n, j0 + blk)
        f, r = np.tril(F[j0:j1, j0:j1]), np.tril(R[j0:j1, j0:j1])
        num += float(np.sum((f - r) ** 2))
        den += float(np.sum(r ** 2))
        if j1 < n:
            d = F[j1:, j0:j1] - R[j1:, j0:j1]
            num += float(np.sum(d * d))
            den += float(np.sum(R[j1:, j0:j1] ** 2))
    return (num / den) ** 0.5


def oracle_posv(n, nrhs, nb):
    A = oracle.dplgsy(float(n), L_, n, 51, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    oracle.use_blas(True)
    try:
        assert oracle.dposv(L_, A, n, B, nrhs, nb, nthreads=THREADS) == 0
    finally:
        oracle.use_blas(False)
    return A, B


def gpu_posv(ch, n, nrhs, nb):
    ch.set_param(ch.TILE_SIZE, nb)
    A = oracle.dplgsy(float(n), L_, n, 51, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    assert ch.lib.CHAMELEON_dposv(L_, n, nrhs, _p(A), n, _p(B), n) == 0
    return A, B


def test_cpu_configuration_n4096(ch):
    n, nb, nrhs = 4096, 320, 16
    say("config cpu", n, nb, nrhs)
    F, X = gpu_posv(ch, n, nrhs, nb)
    Fr, Xr = oracle_posv(n, nrhs, nb)
    assert rel_frob_lower(F, Fr) <= 1e-12
    assert oracle.rel_frobenius(X, Xr) <= 1e-11
    A0 = oracle.dplgsy(float(n), UL_, n, 51)
    assert oracle.check_factorization(A0, F, L_)[0]
    assert oracle.check_solve(A0, X, oracle.dplrnt(n, nrhs, 5673))[0]


@pytest.mark.parametrize("nb", [256, 384, 512])
def test_fp64_configuration_n8192(ch, nb):
    n, nrhs = 8192, 64
    say("config fp64", n, nb, nrhs)
    F, X = gpu_posv(ch, n, nrhs, nb)
    Fr, Xr = oracle_posv(n, nrhs, nb)
    assert rel_frob_lower(F, Fr) <= 1e-12
    assert oracle.rel_frobenius(X, Xr) <= 1e-11
    del F, Fr
    A0 = oracle.dplgsy(float(n), UL_, n, 51)
    assert oracle.check_solve(A0, X, oracle.dplrnt(n, nrhs, 5673))[0]


def device_backward_error(ch, n, nb, X):
    """Scaled residual max_j ||B_j - A X_j||_inf / (||A||_inf ||X_j||_inf N eps)
    with A and B regenerated on the device (no further host copies)."""
    L = ch.lib
    with ch.Desc.alloc(n, n, nb) as dA, ch.Desc.alloc(n, X.shape[1], nb) as dB, ch.Desc.from_array(X, nb) as dX:
        seq = ch.Sequence()
        ch.check(L.CHAMELEON_dplgsy_Tile_Async(float(n), L_, dA.ptr, 51, seq.ptr, seq.req))
        ch.check(L.CHAMELEON_dplrnt_Tile_Async(dB.ptr, 5673, seq.ptr, seq.req))
        ch.check(seq.wait())
        seq.destroy()
        be = ctypes.c_double(-1.0)
        ch.check(L.CHAMAMD_Backward_Error(L_, dA.ptr, dX.ptr, dB.ptr, ctypes.byref(be)))
    return be.value


@pytest.mark.parametrize("n", [16384, 32768])
def test_fp64_configuration_large(ch, n):
    nb, nrhs = 512, 64
    say("config fp64 large", n, nb, nrhs)
    F, X = gpu_posv(ch, n, nrhs, nb)
    be = device_backward_error(ch, n, nb, X)
    assert 0.0 < be <= 60.0, be
    Fr, Xr = oracle_posv(n, nrhs, nb)
    assert oracle.rel_frobenius(X, Xr) <= 1e-11
    del Xr
    assert rel_frob_lower(F, Fr) <= 1e-12


@pytest.mark.parametrize("nrhs", [1, 8])
def test_narrow_solve_n32768(ch, nrhs):
    """dpotrs at N = 32768, NB = 512 with 1 and 8 right-hand sides (the
    narrow strip solve) on device-resident descriptors: the same columns
    solved among 16 (the MFMA strip solve) agree to 1e-12, and the scaled
    residual computed on the device is <= 60."""
    n, nb = 32768, 512
    say("config narrow dpotrs", n, nb, nrhs)
    L = ch.lib
    ch.set_param(ch.TILE_SIZE, nb)
    with ch.Desc.alloc(n, n, nb) as dA, ch.Desc.alloc(n, nrhs, nb) as dB, ch.Desc.alloc(n, 16, nb) as dW:
        seq = ch.Sequence()
        ch.check(L.CHAMELEON_dplgsy_Tile_Async(float(n), L_, dA.ptr, 51, seq.ptr, seq.req))
        ch.check(L.CHAMELEON_dpotrf_Tile_Async(L_, dA.ptr, seq.ptr, seq.req))
        ch.check(L.CHAMELEON_dplrnt_Tile_Async(dB.ptr, 5673, seq.ptr, seq.req))
        ch.check(L.CHAMELEON_dplrnt_Tile_Async(dW.ptr, 5673, seq.ptr, seq.req))
        ch.check(L.CHAMELEON_dpotrs_Tile_Async(L_, dA.ptr, dB.ptr, seq.ptr, seq.req))
        ch.check(L.CHAMELEON_dpotrs_Tile_Async(L_, dA.ptr, dW.ptr, seq.ptr, seq.req))
        assert seq.wait() == 0
        seq.destroy()
        X = dB.to_array(n, nrhs)
        W = dW.to_array(n, 16)
    assert oracle.rel_frobenius(X, W[:, :nrhs]) <= 1e-12
    be = device_backward_error(ch, n, nb, X)
    assert 0.0 < be <= 60.0, be


def test_mixed_configuration_n32768(ch):
    """kappa 1e4, NRHS 1: the split-BF16 solve refined in FP64 matches the
    FP64 dposv of the same A and b to the conditioning of the problem, and
    its backward error is at FP64 level."""
    n, nb, kappa = 32768, 512, 1e4
    L = ch.lib
    say("config mixed", n, nb, kappa)
    ch.set_param(ch.TILE_SIZE, nb)

    def run(fn):
        seq = ch.Sequence()
        rc = fn(seq)
        info = seq.wait()
        seq.destroy()
        ch.check(rc)
        return info

    try:
        with ch.Desc.alloc(n, n, nb) as dA, ch.Desc.alloc(n, n, nb) as dA0, ch.Desc.alloc(n, 1, nb) as dB, \
                ch.Desc.alloc(n, 1, nb) as dX, ch.Desc.alloc(n, 1, nb) as dX64:
            ch.check(L.CHAMAMD_dplgsy_cond_Tile(kappa, 203, dA0.ptr, 1234))
            run(lambda q: L.CHAMELEON_dplrnt_Tile_Async(dB.ptr, 5673, q.ptr, q.req))
            run(lambda q: L.CHAMELEON_dlacpy_Tile_Async(UL_, dA0.ptr, dA.ptr, q.ptr, q.req))
            run(lambda q: L.CHAMELEON_dlacpy_Tile_Async(UL_, dB.ptr, dX64.ptr, q.ptr, q.req))
            assert run(lambda q: L.CHAMELEON_dposv_Tile_Async(L_, dA.ptr, dX64.ptr, q.ptr, q.req)) == 0
            run(lambda q: L.CHAMELEON_dlacpy_Tile_Async(UL_, dA0.ptr, dA.ptr, q.ptr, q.req))
            it, be = ctypes.c_int(-99), ctypes.c_double(-1.0)
            assert run(lambda q: L.CHAMAMD_dposv_mixed_Tile_Async(L_, dA.ptr, dB.ptr, dX.ptr, 2, ctypes.byref(it),
                                                                  ctypes.byref(be), q.ptr, q.req)) == 0
            assert it.value >= 0, it.value
            be64 = ctypes.c_double(-1.0)
            ch.check(L.CHAMAMD_Backward_Error(L_, dA0.ptr, dX64.ptr, dB.ptr, ctypes.byref(be64)))
            ch.check(L.CHAMAMD_Backward_Error(L_, dA0.ptr, dX.ptr, dB.ptr, ctypes.byref(be)))
            assert be.value <= 1.0 and be64.value <= 1.0, (be.value, be64.value)
            x, x64 = dX.to_array(n, 1), dX64.to_array(n, 1)
            assert np.abs(x - x64).max() / np.abs(x64).max() <= kappa * 1e-12
    finally:
        L.CHAMAMD_Cond_Cache_Release()


def test_batch_configuration_20000(ch):
    count, seed = 20000, 2024
    rng = np.random.default_rng(11)
    n = np.floor(np.exp(rng.uniform(np.log(16), np.log(1024), count))).astype(np.int32)
    nrhs = rng.integers(1, 9, count).astype(np.int32)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    say("config batch", count)
    b = ctypes.c_void_p()
    ch.check(ch.lib.CHAMAMD_Batch_Create(ctypes.byref(b), count, ip(n), ip(nrhs)))
    try:
        ch.check(ch.lib.CHAMAMD_dplgsy_batch(b, seed))
        ch.check(ch.lib.CHAMAMD_dplrnt_batch(b, seed))
        info = np.zeros(count, dtype=np.int32)
        assert ch.lib.CHAMAMD_dposv_batch(b, ip(info)) == 0 and not info.any()
        X = []
        for i in range(count):
            ni, ri = int(n[i]), int(nrhs[i])
            x = np.zeros((ni, ri), order="F")
            ch.check(ch.lib.CHAMAMD_Batch_GetB(b, i, _p(x), ni))
            X.append(x)

        def residual(i):
            ni, ri = int(n[i]), int(nrhs[i])
            A0 = oracle.dplgsy(float(ni), UL_, ni, seed + i)
            B0 = oracle.dplrnt(ni, ri, seed + i)
            return oracle.check_solve(A0, X[i], B0)[1]

        with ThreadPoolExecutor(THREADS) as ex:
            res = np.array(list(ex.map(residual, range(count))))
        assert np.all(np.isfinite(res)) and res.max() <= 60.0, (int(res.argmax()), res.max())
        # 50 systems spread over the size range, against the oracle's dposv
        order = np.argsort(n, kind="stable")
        sample = order[np.linspace(0, count - 1, 50).astype(int)]
        oracle.use_blas(True)
        try:
            for i in sample:
                ni, ri = int(n[i]), int(nrhs[i])
                Ar = oracle.dplgsy(float(ni), UL_, ni, seed + i)
                Br = oracle.dplrnt(ni, ri, seed + i)
                assert oracle.dposv(L_, Ar, ni, Br, ri, 64) == 0
                F = np.zeros((ni, ni), order="F")
                ch.check(ch.lib.CHAMAMD_Batch_GetA(b, int(i), L_, _p(F), ni))
                assert oracle.rel_frobenius(F, np.tril(Ar)) <= 1e-12, i
                assert oracle.rel_frobenius(X[i], Br) <= 1e-11, i
        finally:
            oracle.use_blas(False)
    finally:
        ch.lib.CHAMAMD_Batch_Destroy(ctypes.byref(b))


def test_serve_configuration_10s(ch):
    """The serving configuration at its own size (FP64 factors).  Requests
    arrive open-loop at 2000/s for 10 s; after the drain every request's
    scaled residual ||B - A X||_inf / (||A||_inf ||X||_inf N eps) is formed
    on the device (CHAMAMD_Backward_Error, A = dplgsy(bump = n, seed)
    regenerated once per matrix id) and must be <= 60."""
    L = ch.lib
    rng = np.random.default_rng(11)
    rate, seconds = 2000.0, 10.0
    arrivals = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.5) + 16))
    arrivals = arrivals[arrivals < seconds]
    reqs, seen = [], []
    for i in range(len(arrivals)):
        if seen and rng.random() < 0.3:
            mid, n = seen[rng.integers(len(seen))]
        else:
            n = int(np.floor(np.exp(rng.uniform(np.log(64), np.log(8192)))))
            mid = i
            seen.append((mid, n))
        reqs.append((float(arrivals[i]), mid, n, int(rng.choice([1, 4, 16]))))
    base = np.asfortranarray(rng.uniform(-0.5, 0.5, (8192, 16)))
    say("config serve", len(reqs), "requests,", len(seen), "matrices")
    ch.set_param(ch.TILE_SIZE, 256)
    srv = ch.Server(200 << 30)
    out = {}
    try:
        t0 = time.perf_counter()
        nxt = 0
        while nxt < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt][0] <= now:
                _, mid, n, r = reqs[nxt]
                B = base[:n, :r].copy(order="F")
                out[srv.submit(mid, n, B, seed=1000 + mid)] = (mid, n, r, B)
                nxt += 1
            srv.progress()
        srv.drain()
        for t in out:
            assert srv.test(t) == (True, 0), t
        st = srv.stats()
        wall = time.perf_counter() - t0
    finally:
        srv.destroy()
    say("served", st["completed"], "in", round(wall, 1), "s; factored", st["factored"], "reused", st["reused"])
    assert st["completed"] == len(reqs)
    by_mid = {}
    for mid, n, r, X in out.values():
        by_mid.setdefault((mid, n), []).append((r, X))
    worst = 0.0
    for (mid, n), items in by_mid.items():
        with ch.Desc.alloc(n, n, 256) as dA:
            seq = ch.Sequence()
            ch.check(L.CHAMELEON_dplgsy_Tile_Async(float(n), UL_, dA.ptr, 1000 + mid, seq.ptr, seq.req))
            ch.check(seq.wait())
            seq.destroy()
            for r, X in items:
                assert np.all(np.isfinite(X)), (mid, n)
                B0 = base[:n, :r].copy(order="F")
                with ch.Desc.from_array(X, 256) as dX, ch.Desc.from_array(B0, 256) as dB:
                    be = ctypes.c_double(-1.0)
                    ch.check(L.CHAMAMD_Backward_Error(L_, dA.ptr, dX.ptr, dB.ptr, ctypes.byref(be)))
                assert 0.0 <= be.value <= 60.0, (mid, n, r, be.value)
                worst = max(worst, be.value)
    say("worst scaled residual", worst)
