"""GPU tests of the solve server (CHAMAMD_Server_*): continuous batching,
resident factors, HBM admission and per-request results.

Parity bar: every served solution within 1e-11 of the oracle's solve of
the same system and passing Chameleon's residual check; a request served
in a mixed round is bitwise the same as served alone; a matrix id seen
again reuses its factor; a factor larger than the budget is rejected
with CHAMELEON_ERR_OUT_OF_RESOURCES and evictions never touch a factor a
pending request needs."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

L_, UL_ = 122, 123
ERR_OUT_OF_RESOURCES = -106


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


def test_server_mixed_round_vs_oracle(ch):
    sizes = [(16, 1), (100, 4), (333, 16), (1024, 1), (1100, 4), (2000, 16), (64, 4), (700, 1)]
    say("server mixed round", sizes)
    with ch.Server(8 << 30) as srv:
        reqs = []
        for i, (n, r) in enumerate(sizes):
            B = rhs(n, r, 500 + i)
            reqs.append((srv.submit(i, n, B, seed=900 + i), n, B.copy(order="F"), B, 900 + i))
        assert srv.progress() == len(sizes)
        for t, n, B0, X, seed in reqs:
            assert srv.test(t) == (True, 0)
            A0, Xr = reference(n, seed, B0)
            assert oracle.rel_frobenius(X, Xr) <= 1e-11, n
            assert oracle.check_solve(A0, X, B0)[0], n
        st = srv.stats()
        assert st["completed"] == len(sizes) and st["factored"] == len(sizes) and st["p99_ms"] > 0


def test_server_reuses_resident_factor(ch):
    say("server reuse")
    with ch.Server(4 << 30) as srv:
        B1 = rhs(1500, 3, 1)
        srv.submit(42, 1500, B1, seed=7)
        srv.progress()
        B2 = rhs(1500, 5, 2)
        B2c = B2.copy(order="F")
        t = srv.submit(42, 1500, B2, seed=7)
        srv.progress()
        assert srv.test(t) == (True, 0)
        st = srv.stats()
        assert st["factored"] == 1 and st["reused"] == 1
        assert oracle.rel_frobenius(B2, reference(1500, 7, B2c)[1]) <= 1e-11


def test_server_alone_equals_mix(ch):
    say("server alone == mix")
    probe = [(300, 2, 11), (1300, 3, 12)]
    mixed = {}
    with ch.Server(8 << 30) as srv:
        for j, (n, r, seed) in enumerate([(50, 1, 1), (800, 8, 2), (300, 2, 11), (2500, 1, 3), (1300, 3, 12)]):
            B = rhs(n, r, seed)
            srv.submit(100 + j, n, B, seed=seed)
            if (n, r, seed) in probe:
                mixed[(n, r, seed)] = B
        srv.progress()
    for key in probe:
        n, r, seed = key
        with ch.Server(8 << 30) as srv:
            B = rhs(n, r, seed)
            srv.submit(1, n, B, seed=seed)
            srv.progress()
        assert np.array_equal(B, mixed[key]), key


def test_server_budget_admission_and_eviction(ch):
    say("server budget")
    n = 1000                                   # small path: 1008^2 doubles + inverses ~ 8.6 MB
    with ch.Server(20 << 20) as srv:           # room for two such factors
        for i in range(3):
            B = rhs(n, 1, i)
            t = srv.submit(i, n, B, seed=i)
            srv.progress()
            assert srv.test(t) == (True, 0)
        st = srv.stats()
        assert st["resident"] == 2 and st["factored"] == 3
        B = rhs(n, 1, 9)
        t = srv.submit(0, n, B, seed=0)         # evicted earlier: factored again
        srv.progress()
        assert srv.test(t) == (True, 0) and srv.stats()["factored"] == 4
        Bbig = rhs(3000, 1, 1)
        t = srv.submit(99, 3000, Bbig, seed=1)   # 72 MB factor: never fits
        srv.progress()
        assert srv.test(t) == (True, ERR_OUT_OF_RESOURCES)
        # two pending requests that need both resident slots at once
        ta = srv.submit(0, n, rhs(n, 1, 3), seed=0)
        tb = srv.submit(5, n, rhs(n, 1, 4), seed=5)
        srv.progress()
        assert srv.test(ta) == (True, 0) and srv.test(tb) == (True, 0)


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
        srv.progress()
        for t, A, B0, X in out:
            assert srv.test(t) == (True, 0)
            assert oracle.rel_frobenius(X, np.linalg.solve(A, B0)) <= 1e-11
        assert srv.test(tbad) == (True, 151)
