"""GPU tests of the solve server (CHAMAMD_Server_*): continuous batching,
resident factors, HBM admission and per-request results.

Parity bar: every served solution within 1e-11 of the oracle's solve of
the same system and passing Chameleon's residual check; a request served
among others that overlap it in time on the device is bitwise the same as
served alone; a matrix id seen again reuses its factor; a factor larger
than the budget is rejected with CHAMELEON_ERR_OUT_OF_RESOURCES and
placements never evict a factor a live request needs.  With
mixed-precision factors the refined solutions meet the same bar (FP64
backward error, or the FP64 fallback)."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

L_, UL_ = 122, 123
ERR_OUT_OF_RESOURCES = -106
BF16X3, FP8 = 2, 3


@pytest.fixture(scope="module")
def ch():
    import chameleon_amd as ch
    ch.init()
    ch.set_param(ch.TILE_SIZE, 256)
    yield ch


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rhs(n, nrhs, seed):
    return np.asfortranarray(oracle.dplrnt(n, nrhs, seed))


def reference(n, seed, B):
    A0 = oracle.dplgsy(float(n), UL_, n, seed)
    return A0, np.linalg.solve(A0, B)


def test_server_mixed_sizes_vs_oracle(ch):
    sizes = [(16, 1), (100, 4), (333, 16), (1024, 1), (1100, 4), (2000, 16), (64, 4), (700, 1)]
    say("server mixed sizes", sizes)
    with ch.Server(8 << 30) as srv:
        reqs = []
        for i, (n, r) in enumerate(sizes):
            B = rhs(n, r, 500 + i)
            reqs.append((srv.submit(i, n, B, seed=900 + i), n, B.copy(order="F"), B, 900 + i))
        srv.drain()
        for t, n, B0, X, seed in reqs:
            assert srv.test(t) == (True, 0)
            A0, Xr = reference(n, seed, B0)
            assert oracle.rel_frobenius(X, Xr) <= 1e-11, n
            assert oracle.check_solve(A0, X, B0)[0], n
        st = srv.stats()
        assert st["completed"] == len(sizes) and st["factored"] == len(sizes) and st["p99_ms"] > 0


def test_server_progress_does_not_block(ch):
    """Progress returns before the device work it started has finished:
    right after the first call a large factorisation is still in flight,
    and repeated calls finish every request."""
    say("server non-blocking")
    with ch.Server(4 << 30) as srv:
        B = rhs(6000, 2, 3)
        B0 = B.copy(order="F")
        t = srv.submit(1, 6000, B, seed=31)
        assert srv.progress() == 0
        assert srv.test(t) == (False, 0)
        for _ in range(200000):
            srv.progress()
            if srv.test(t)[0]:
                break
        assert srv.test(t) == (True, 0)
        assert oracle.rel_frobenius(B, reference(6000, 31, B0)[1]) <= 1e-11


def test_server_reuses_resident_factor(ch):
    say("server reuse")
    with ch.Server(4 << 30) as srv:
        B1 = rhs(1500, 3, 1)
        srv.submit(42, 1500, B1, seed=7)
        srv.drain()
        B2 = rhs(1500, 5, 2)
        B2c = B2.copy(order="F")
        t = srv.submit(42, 1500, B2, seed=7)
        srv.drain()
        assert srv.test(t) == (True, 0)
        st = srv.stats()
        assert st["factored"] == 1 and st["reused"] == 1
        assert oracle.rel_frobenius(B2, reference(1500, 7, B2c)[1]) <= 1e-11


PROBE = [(300, 2, 11), (1300, 3, 12), (3000, 4, 13)]
MIX = [(50, 1, 1), (800, 8, 2), (300, 2, 11), (2500, 1, 3), (1300, 3, 12), (4000, 2, 4), (3000, 4, 13),
       (900, 16, 5), (1800, 1, 6)]


def serve_overlapped(ch, prec=-1):
    """The MIX requests submitted in three waves, the later ones while the
    earlier ones are on the device; returns the probes' solutions and the
    most lanes busy at once."""
    out = {}
    with ch.Server(8 << 30) as srv:
        srv.set_precision(prec)
        waves = [MIX[:3], MIX[3:6], MIX[6:]]
        for w, wave in enumerate(waves):
            for j, (n, r, seed) in enumerate(wave):
                B = rhs(n, r, seed)
                srv.submit(100 + 10 * w + j, n, B, seed=seed)
                if (n, r, seed) in PROBE:
                    out[(n, r, seed)] = B
            srv.progress()
            srv.progress()
        srv.drain()
        st = srv.stats()
        assert st["completed"] == len(MIX)
    return out, st


def serve_alone(ch, key, prec=-1):
    n, r, seed = key
    with ch.Server(8 << 30) as srv:
        srv.set_precision(prec)
        B = rhs(n, r, seed)
        srv.submit(1, n, B, seed=seed)
        srv.drain()
    return B


@pytest.mark.parametrize("prec", [-1, BF16X3])
def test_server_alone_equals_mix(ch, prec):
    """Requests admitted while others run (several lanes busy at once) give
    bitwise the results they give alone."""
    say("server alone == mix", prec)
    mixed, st = serve_overlapped(ch, prec)
    assert st["max_jobs"] >= 2, st
    for key in PROBE:
        assert np.array_equal(serve_alone(ch, key, prec), mixed[key]), key
    for key, X in mixed.items():
        n, r, seed = key
        B0 = rhs(n, r, seed)
        A0, Xr = reference(n, seed, B0)
        assert oracle.rel_frobenius(X, Xr) <= 1e-11, key
        assert oracle.check_solve(A0, X, B0)[0], key


def test_server_budget_admission_and_eviction(ch):
    say("server budget")
    n = 1000                                   # small path: 1008^2 doubles + inverses ~ 8.2 MB
    with ch.Server(20 << 20) as srv:           # room for two such factors
        for i in range(3):
            B = rhs(n, 1, i)
            t = srv.submit(i, n, B, seed=i)
            srv.drain()
            assert srv.test(t) == (True, 0)
        st = srv.stats()
        assert st["resident"] == 2 and st["factored"] == 3
        B = rhs(n, 1, 9)
        t = srv.submit(0, n, B, seed=0)         # evicted earlier: factored again
        srv.drain()
        assert srv.test(t) == (True, 0) and srv.stats()["factored"] == 4
        Bbig = rhs(3000, 1, 1)
        t = srv.submit(99, 3000, Bbig, seed=1)   # 72 MB factor: never fits
        srv.drain()
        assert srv.test(t) == (True, ERR_OUT_OF_RESOURCES)
        # two live requests that need both resident slots at once
        ta = srv.submit(0, n, rhs(n, 1, 3), seed=0)
        tb = srv.submit(5, n, rhs(n, 1, 4), seed=5)
        srv.drain()
        assert srv.test(ta) == (True, 0) and srv.test(tb) == (True, 0)


def test_server_placement_spares_factors_live_requests_need(ch):
    """Budget of two factors, ids 0 and 1 resident (0 least recently used).
    Submitted together: a new id 2, then id 0 again.  Placing id 2 must
    evict id 1 (no live request needs it), not id 0, which the second
    request needs: only id 2 is factored, id 0 is reused."""
    say("server pins before placement")
    n = 1000
    with ch.Server(20 << 20) as srv:
        for i in (0, 1):
            srv.submit(i, n, rhs(n, 1, i), seed=i)
            srv.drain()
        st0 = srv.stats()
        B2, B0 = rhs(n, 1, 20), rhs(n, 1, 21)
        B0c = B0.copy(order="F")
        t2 = srv.submit(2, n, B2, seed=2)
        t0 = srv.submit(0, n, B0, seed=0)
        srv.drain()
        st = srv.stats()
        assert srv.test(t2) == (True, 0) and srv.test(t0) == (True, 0)
        assert st["factored"] == st0["factored"] + 1, (st0, st)
        assert st["reused"] == st0["reused"] + 1
        assert oracle.rel_frobenius(B0, reference(n, 0, B0c)[1]) <= 1e-11


def test_server_host_matrix_and_failure(ch):
    say("server host matrix")
    rng = np.random.default_rng(3)
    with ch.Server(2 << 30) as srv:
        out = []
        for i, n in enumerate((37, 1200)):
            G = rng.standard_normal((n, n))
            A = np.asfortranarray(G @ G.T + n * np.eye(n))
            B = np.asfortranarray(rng.standard_normal((n, 2)))
            out.append((srv.submit(10 + i, n, B, A=A), A, B.copy(order="F"), B))
        bad = np.asfortranarray(np.eye(200))
        bad[150, 150] = -1.0
        tbad = srv.submit(77, 200, np.ones((200, 1), order="F"), A=bad)
        badL = np.asfortranarray(np.eye(1500))
        badL[1200, 1200] = -1.0
        tbadL = srv.submit(78, 1500, np.ones((1500, 1), order="F"), A=badL)
        srv.drain()
        for t, A, B0, X in out:
            assert srv.test(t) == (True, 0)
            assert oracle.rel_frobenius(X, np.linalg.solve(A, B0)) <= 1e-11
        assert srv.test(tbad) == (True, 151)
        assert srv.test(tbadL) == (True, 1201)


@pytest.mark.parametrize("prec", [BF16X3, FP8])
def test_server_mixed_precision(ch, prec):
    """Large systems factored with low-precision trailing products and
    refined in FP64 per request: the FP64 answer (oracle), FP64 backward
    error, repeated right-hand sides on the resident mixed factor."""
    say("server mixed precision", prec)
    sizes = [(1100, 1, 41), (2048, 4, 42), (3333, 16, 43), (500, 2, 44)]
    with ch.Server(8 << 30) as srv:
        srv.set_precision(prec)
        reqs = []
        for i, (n, r, seed) in enumerate(sizes):
            B = rhs(n, r, seed)
            reqs.append((srv.submit(i, n, B, seed=seed), n, B.copy(order="F"), B, seed))
        B = rhs(2048, 2, 99)
        reqs.append((srv.submit(1, 2048, B, seed=42), 2048, B.copy(order="F"), B, 42))   # reuse
        srv.drain()
        for t, n, B0, X, seed in reqs:
            assert srv.test(t) == (True, 0), n
            A0, Xr = reference(n, seed, B0)
            assert oracle.rel_frobenius(X, Xr) <= 1e-11, n
            assert oracle.check_solve(A0, X, B0)[0], n
        st = srv.stats()
        assert st["mixed_solves"] == 4 and st["factored"] == 4 and st["reused"] == 1, st
        assert 0.0 < st["max_berr"] <= 1.0, st


def test_server_mixed_fallback(ch):
    """A host matrix whose BF16-product factorisation cannot be refined to
    FP64 accuracy (condition number 1e10) is factored again in FP64: the
    answer still meets Chameleon's residual check."""
    say("server mixed fallback")
    n = 1536
    rng = np.random.default_rng(8)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = 1e10 ** (-np.arange(n) / (n - 1))
    A = np.asfortranarray(((Q * lam) @ Q.T + ((Q * lam) @ Q.T).T) / 2)
    B = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, 2)))
    B0 = B.copy(order="F")
    with ch.Server(2 << 30) as srv:
        srv.set_precision(1)       # plain BF16 products
        t = srv.submit(5, n, B, A=A)
        srv.drain()
        assert srv.test(t) == (True, 0)
        st = srv.stats()
        assert st["fallbacks"] >= 1, st
    assert oracle.check_solve(A, B, B0)[0]
