"""GPU tests of the batched Cholesky paths: Chameleon's tile batch
(CHAMELEON_dplgsy_batch_Tile / CHAMELEON_dpotrf_batch_Tile) and the
uneven device-resident batches (CHAMAMD_Batch_*).

Parity bar: generators bit-exact with the oracle; every factor within
1e-12 relative Frobenius of the oracle's factor of the same system alone;
every solution within 1e-11 of the oracle's and passing Chameleon's
residual check; a system solved in a batch is bitwise the same as solved
alone; failures reported per system without disturbing the others."""
import ctypes
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

L_, U_, UL_ = 122, 121, 123


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


def ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Batch:
    def __init__(self, ch, n, nrhs):
        self.ch = ch
        self.n = np.asarray(n, dtype=np.int32)
        self.nrhs = np.asarray(nrhs, dtype=np.int32)
        self.ptr = ctypes.c_void_p()
        ch.check(ch.lib.CHAMAMD_Batch_Create(ctypes.byref(self.ptr), len(self.n), ip(self.n), ip(self.nrhs)))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.ch.lib.CHAMAMD_Batch_Destroy(ctypes.byref(self.ptr))

    def get_a(self, i, uplo=L_):
        n = int(self.n[i])
        A = np.zeros((n, n), order="F")
        self.ch.check(self.ch.lib.CHAMAMD_Batch_GetA(self.ptr, i, uplo, _p(A), n))
        return A

    def get_b(self, i):
        n, r = int(self.n[i]), int(self.nrhs[i])
        B = np.zeros((n, max(r, 1)), order="F")
        self.ch.check(self.ch.lib.CHAMAMD_Batch_GetB(self.ptr, i, _p(B), n))
        return B[:, :r]

    def set_a(self, i, A, uplo=L_):
        self.ch.check(self.ch.lib.CHAMAMD_Batch_SetA(self.ptr, i, uplo, _p(A), A.shape[0]))

    def set_b(self, i, B):
        self.ch.check(self.ch.lib.CHAMAMD_Batch_SetB(self.ptr, i, _p(B), B.shape[0]))

    def posv(self):
        info = np.zeros(len(self.n), dtype=np.int32)
        rc = self.ch.lib.CHAMAMD_dposv_batch(self.ptr, ip(info))
        return rc, info


def log_uniform_sizes(count, lo, hi, seed):
    rng = np.random.default_rng(seed)
    n = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int32)
    nrhs = rng.integers(1, 9, count).astype(np.int32)
    return n, nrhs


# ------------------------------------------------------------ tile batch
@pytest.mark.parametrize("M,N,nb", [(96, 64, 32), (300, 200, 100), (512, 256, 128), (48, 48, 48)])
def test_tile_batch_generator_bit_exact(ch, M, N, nb):
    say("tile batch gen", M, N, nb)
    seed, bump = 4242, float(nb)
    with ch.Desc.alloc(M, N, nb) as dA:
        ch.check(ch.lib.CHAMELEON_dplgsy_batch_Tile(bump, dA.ptr, seed))
        got = dA.to_array(M, N)
    ref = oracle.dplgsy_batch(bump, M, N, nb, nb, seed)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("uplo", [L_, U_])
@pytest.mark.parametrize("M,N,nb", [(96, 64, 32), (300, 200, 100), (512, 256, 128), (200, 200, 40)])
def test_tile_batch_potrf_vs_oracle(ch, uplo, M, N, nb):
    say("tile batch potrf", uplo, M, N, nb)
    seed = 99
    A = oracle.dplgsy_batch(float(nb), M, N, nb, nb, seed)
    Ar = A.copy(order="F")
    with ch.Desc.from_array(A, nb) as dA:
        assert ch.lib.CHAMELEON_dpotrf_batch_Tile(uplo, dA.ptr) == 0
    oracle.use_blas(True)
    try:
        info = oracle.dpotrf_batch(uplo, Ar, M, N, nb, nb)
    finally:
        oracle.use_blas(False)
    assert not info.any()
    tri = np.tril if uplo == L_ else np.triu
    other = (lambda X: np.triu(X, 1)) if uplo == L_ else (lambda X: np.tril(X, -1))
    A0 = oracle.dplgsy_batch(float(nb), M, N, nb, nb, seed)
    for n in range(N // nb):
        for m in range(M // nb):
            sl = (slice(m * nb, (m + 1) * nb), slice(n * nb, (n + 1) * nb))
            assert oracle.rel_frobenius(tri(A[sl]), tri(Ar[sl])) <= 1e-12
            assert np.array_equal(other(A[sl]), other(A0[sl]))      # other triangle untouched
            assert oracle.check_factorization(A0[sl], A[sl], uplo)[0]


def test_tile_batch_potrf_reports_failing_tile(ch):
    M, N, nb = 128, 64, 32
    say("tile batch failure")
    A = oracle.dplgsy_batch(float(nb), M, N, nb, nb, 7)
    A[2 * nb + 5, 5] = -1.0                                # tile (2, 0): fails at column 6
    Ar = A.copy(order="F")
    with ch.Desc.from_array(A, nb) as dA:
        rc = ch.lib.CHAMELEON_dpotrf_batch_Tile(L_, dA.ptr)
    info = oracle.dpotrf_batch(L_, Ar, M, N, nb, nb)
    assert rc == info[2] == 6
    assert info.sum() == 6                                  # the other tiles factor
    for t in (0, 1, 3):
        sl = (slice(t * nb, (t + 1) * nb), slice(0, nb))
        assert oracle.rel_frobenius(np.tril(A[sl]), np.tril(Ar[sl])) <= 1e-12


# ---------------------------------------------------------- uneven batch
def test_uneven_batch_vs_oracle(ch):
    n, nrhs = log_uniform_sizes(64, 16, 700, seed=1)
    say("uneven batch", len(n), int(n.sum()))
    seed = 1000
    with Batch(ch, n, nrhs) as b:
        ch.check(ch.lib.CHAMAMD_dplgsy_batch(b.ptr, seed))
        ch.check(ch.lib.CHAMAMD_dplrnt_batch(b.ptr, seed))
        for i in (0, 5, 17):                                  # generators bit-exact
            ni = int(n[i])
            assert np.array_equal(np.tril(b.get_a(i)), np.tril(oracle.dplgsy(float(ni), UL_, ni, seed + i)))
            assert np.array_equal(b.get_b(i), oracle.dplrnt(ni, int(nrhs[i]), seed + i))
        rc, info = b.posv()
        assert rc == 0 and not info.any()
        oracle.use_blas(True)
        try:
            for i in range(len(n)):
                ni, ri = int(n[i]), int(nrhs[i])
                A0 = oracle.dplgsy(float(ni), UL_, ni, seed + i)
                B0 = oracle.dplrnt(ni, ri, seed + i)
                Ar, Br = A0.copy(order="F"), B0.copy(order="F")
                assert oracle.dposv(L_, Ar, ni, Br, ri, 64) == 0
                assert oracle.rel_frobenius(b.get_a(i), np.tril(Ar)) <= 1e-12, i
                X = b.get_b(i)
                assert oracle.rel_frobenius(X, Br) <= 1e-11, i
                assert oracle.check_solve(A0, X, B0)[0], i
        finally:
            oracle.use_blas(False)


def test_uneven_batch_alone_equals_in_mix(ch):
    n, nrhs = log_uniform_sizes(40, 16, 1024, seed=2)
    say("alone == mix")
    seed = 77
    with Batch(ch, n, nrhs) as b:
        ch.check(ch.lib.CHAMAMD_dplgsy_batch(b.ptr, seed))
        ch.check(ch.lib.CHAMAMD_dplrnt_batch(b.ptr, seed))
        assert b.posv()[0] == 0
        mix = {i: (b.get_a(i), b.get_b(i)) for i in (0, 3, 11, 39)}
    for i, (Fm, Xm) in mix.items():
        with Batch(ch, n[i:i + 1], nrhs[i:i + 1]) as one:
            ch.check(ch.lib.CHAMAMD_dplgsy_batch(one.ptr, seed + i))
            ch.check(ch.lib.CHAMAMD_dplrnt_batch(one.ptr, seed + i))
            assert one.posv()[0] == 0
            assert np.array_equal(one.get_a(0), Fm)
            assert np.array_equal(one.get_b(0), Xm)


def test_uneven_batch_host_matrices_upper(ch):
    rng = np.random.default_rng(5)
    n = np.array([17, 64, 100, 33], dtype=np.int32)
    nrhs = np.array([1, 8, 3, 16], dtype=np.int32)
    say("host upper")
    mats, rhs = [], []
    with Batch(ch, n, nrhs) as b:
        for i, (ni, ri) in enumerate(zip(n, nrhs)):
            G = rng.standard_normal((ni, ni))
            A = np.asfortranarray(G @ G.T + ni * np.eye(ni))
            B = np.asfortranarray(rng.standard_normal((ni, ri)))
            Au = np.asfortranarray(np.triu(A) + np.tril(np.full((ni, ni), np.nan), -1))   # lower part never read
            b.set_a(i, Au, U_)
            b.set_b(i, B)
            mats.append(A)
            rhs.append(B)
        rc, info = b.posv()
        assert rc == 0 and not info.any()
        for i, (A, B) in enumerate(zip(mats, rhs)):
            U = b.get_a(i, U_)
            assert oracle.rel_frobenius(np.triu(U), np.linalg.cholesky(A).T) <= 1e-12
            assert oracle.rel_frobenius(b.get_b(i), np.linalg.solve(A, B)) <= 1e-11


def test_uneven_batch_failures_are_per_system(ch):
    n = np.array([40, 130, 64, 200], dtype=np.int32)
    nrhs = np.array([2, 2, 2, 2], dtype=np.int32)
    say("per-system failures")
    seed = 31
    with Batch(ch, n, nrhs) as b:
        ch.check(ch.lib.CHAMAMD_dplgsy_batch(b.ptr, seed))
        ch.check(ch.lib.CHAMAMD_dplrnt_batch(b.ptr, seed))
        bad = oracle.dplgsy(float(n[1]), UL_, int(n[1]), seed + 1)
        bad[70, 70] = -5.0
        b.set_a(1, bad, L_)
        rc, info = b.posv()
        ref_info = oracle.dpotrf(L_, bad.copy(order="F"), int(n[1]), 32)
        assert info[1] == ref_info == 71 and rc == 71
        assert info[0] == info[2] == info[3] == 0
        for i in (0, 2, 3):
            ni = int(n[i])
            A0 = oracle.dplgsy(float(ni), UL_, ni, seed + i)
            B0 = oracle.dplrnt(ni, 2, seed + i)
            assert oracle.rel_frobenius(b.get_b(i), np.linalg.solve(A0, B0)) <= 1e-11


def test_uneven_batch_async_and_empty(ch):
    n, nrhs = log_uniform_sizes(12, 16, 300, seed=9)
    say("async")
    with Batch(ch, n, nrhs) as b:
        ch.check(ch.lib.CHAMAMD_dplgsy_batch(b.ptr, 3))
        ch.check(ch.lib.CHAMAMD_dplrnt_batch(b.ptr, 3))
        seq = ch.Sequence()
        ch.check(ch.lib.CHAMAMD_dpotrf_batch_Async(b.ptr, seq.ptr, seq.req))
        ch.check(ch.lib.CHAMAMD_dpotrs_batch_Async(b.ptr, seq.ptr, seq.req))
        ch.check(seq.wait())
        seq.destroy()
        info = np.ones(len(n), dtype=np.int32)
        assert ch.lib.CHAMAMD_Batch_Info(b.ptr, ip(info)) == 0 and not info.any()
        i = 4
        ni = int(n[i])
        X = np.linalg.solve(oracle.dplgsy(float(ni), UL_, ni, 3 + i), oracle.dplrnt(ni, int(nrhs[i]), 3 + i))
        assert oracle.rel_frobenius(b.get_b(i), X) <= 1e-11
    with Batch(ch, [], []) as e:
        assert e.posv()[0] == 0


def _batch_sizes(count, seed):
    rng = np.random.default_rng(seed)
    n = np.floor(np.exp(rng.uniform(np.log(16), np.log(1300), count))).astype(np.int32)
    n[:3] = [1100, 1300, 17]            # the bucket above 1024 and the smallest one
    nrhs = rng.integers(1, 9, count).astype(np.int32)
    return n, nrhs


def _solve_all(ch, b, seed):
    ch.check(ch.lib.CHAMAMD_dplgsy_batch(b.ptr, seed))
    ch.check(ch.lib.CHAMAMD_dplrnt_batch(b.ptr, seed))
    rc, info = b.posv()
    assert rc == 0 and not info.any()
    return [b.get_b(i).copy() for i in range(len(b.n))], [b.get_a(i) for i in (0, 1, 5)]


def test_batch_graph_replays_equal_eager(ch):
    """The factorisation and solve launches replayed from their captured
    graphs, five times with eager work (regeneration, reads) in between,
    give bitwise the results of launching them directly."""
    n, nrhs = _batch_sizes(400, 21)
    say("batch graphs", len(n))
    with Batch(ch, n, nrhs) as b:
        try:
            assert ch.lib.CHAMAMD_Set_Batch_Graphs(0) == 0
            X0, F0 = _solve_all(ch, b, 77)
            assert ch.lib.CHAMAMD_Set_Batch_Graphs(1) == 0
            for rep in range(5):
                X, F = _solve_all(ch, b, 77)
                for i in range(len(n)):
                    assert np.array_equal(X[i], X0[i]), (rep, i)
                for f, f0 in zip(F, F0):
                    assert np.array_equal(f, f0), rep
        finally:
            ch.lib.CHAMAMD_Set_Batch_Graphs(1)
    for i in (0, 1, 2, 9):
        ni, ri = int(n[i]), int(nrhs[i])
        A0 = oracle.dplgsy(float(ni), UL_, ni, 77 + i)
        assert oracle.check_solve(A0, X0[i], oracle.dplrnt(ni, ri, 77 + i))[0], i


def test_batch_tuner_and_shapes(ch):
    """CHAMAMD_Batch_Tune picks 4 or 8 wavefronts for every bucket the batch
    holds; both shapes compute every system in the same order of
    operations, so a batch factored with either gives bitwise the same
    factors and solutions, equal to the oracle's."""
    n, nrhs = _batch_sizes(300, 22)
    say("batch tuner", len(n))
    tops = (ctypes.c_int * 6)()
    before = (ctypes.c_int * 6)()
    ch.check(ch.lib.CHAMAMD_Batch_Get_Tuning(before, tops, None))
    assert list(tops)[:5] == [64, 128, 256, 512, 1024]
    try:
        with Batch(ch, n, nrhs) as b:
            table = (ctypes.c_int * 6)()
            ch.check(ch.lib.CHAMAMD_Batch_Tune(b.ptr, table))
            tuned = (ctypes.c_int * 6)()
            ch.check(ch.lib.CHAMAMD_Batch_Get_Tuning(None, None, tuned))
            assert all(w in (4, 8) for w in table) and all(tuned), (list(table), list(tuned))
            res = {}
            for w in (4, 8):
                ch.check(ch.lib.CHAMAMD_Batch_Set_Tuning((ctypes.c_int * 6)(*([w] * 6))))
                res[w] = _solve_all(ch, b, 55)
            ch.check(ch.lib.CHAMAMD_Batch_Set_Tuning((ctypes.c_int * 6)(4, 8, 4, 8, 4, 8)))
            res["mix"] = _solve_all(ch, b, 55)
            for key in (8, "mix"):
                for i in range(len(n)):
                    assert np.array_equal(res[key][0][i], res[4][0][i]), (key, i)
                for f, f0 in zip(res[key][1], res[4][1]):
                    assert np.array_equal(f, f0), key
            assert ch.lib.CHAMAMD_Batch_Set_Tuning((ctypes.c_int * 6)(*([5] * 6))) == ch.ERR_ILLEGAL_VALUE
    finally:
        ch.lib.CHAMAMD_Batch_Set_Tuning(before)
    oracle.use_blas(True)
    try:
        for i in (0, 1, 2, 40):
            ni, ri = int(n[i]), int(nrhs[i])
            Ar = oracle.dplgsy(float(ni), UL_, ni, 55 + i)
            Br = oracle.dplrnt(ni, ri, 55 + i)
            assert oracle.dposv(L_, Ar, ni, Br, ri, 64) == 0
            assert oracle.rel_frobenius(res[4][0][i], Br) <= 1e-11, i
    finally:
        oracle.use_blas(False)


FP32 = 0


def _mixed(ch, b):
    count = len(b.n)
    it = np.zeros(count, dtype=np.int32)
    info = np.zeros(count, dtype=np.int32)
    rc = ch.lib.CHAMAMD_dposv_mixed_batch(b.ptr, FP32, ip(it), ip(info))
    return rc, it, info


def test_batch_mixed_fp32_refined(ch):
    """FP32-product factorisations refined in FP64 per system: every system
    converges (bumped generator, condition number ~1), its solution is the
    FP64 one (oracle) and passes Chameleon's check, A is left unchanged,
    and a system refined alone is bitwise the same as in the batch."""
    n, nrhs = _batch_sizes(300, 23)
    say("batch mixed fp32", len(n))
    with Batch(ch, n, nrhs) as b:
        ch.check(ch.lib.CHAMAMD_dplgsy_batch(b.ptr, 91))
        ch.check(ch.lib.CHAMAMD_dplrnt_batch(b.ptr, 91))
        A_before = [b.get_a(i) for i in (0, 1, 7)]
        rc, it, info = _mixed(ch, b)
        assert rc == 0 and not info.any()
        assert it.min() >= 0 and it.max() <= 30, (it.min(), it.max())
        X = [b.get_b(i).copy() for i in range(len(n))]
        for k, i in enumerate((0, 1, 7)):
            assert np.array_equal(b.get_a(i), A_before[k]), i
    oracle.use_blas(True)
    try:
        for i in (0, 1, 2, 7, 50, 299):
            ni, ri = int(n[i]), int(nrhs[i])
            Ar = oracle.dplgsy(float(ni), UL_, ni, 91 + i)
            Br = oracle.dplrnt(ni, ri, 91 + i)
            B0 = Br.copy(order="F")
            assert oracle.dposv(L_, Ar.copy(order="F"), ni, Br, ri, 64) == 0
            assert oracle.rel_frobenius(X[i], Br) <= 1e-11, i
            assert oracle.check_solve(Ar, X[i], B0)[0], i
    finally:
        oracle.use_blas(False)
    # alone == in the batch
    for i in (1, 7):
        ni, ri = int(n[i]), int(nrhs[i])
        with Batch(ch, [ni], [ri]) as one:
            one.set_a(0, oracle.dplgsy(float(ni), L_, ni, 91 + i))
            one.set_b(0, np.asfortranarray(oracle.dplrnt(ni, ri, 91 + i)))
            rc, it1, info1 = _mixed(ch, one)
            assert rc == 0 and it1[0] == it[i]
            assert np.array_equal(one.get_b(0), X[i]), i


def test_batch_mixed_fallbacks(ch):
    """A non-SPD system stops its reduced-precision factorisation (ITER -3)
    and reports LAPACK's info from the FP64 fallback; a system of condition
    number 1e12 cannot be solved from FP32 products (ITER -3 or -31) and is
    solved in FP64; the systems beside them are not disturbed."""
    say("batch mixed fallbacks")
    sizes = [300, 200, 150, 90]
    with Batch(ch, sizes, [2, 1, 3, 1]) as b:
        rng = np.random.default_rng(4)
        mats = []
        for i, ni in enumerate(sizes):
            if i == 0:
                Q, _ = np.linalg.qr(rng.standard_normal((ni, ni)))
                lam = 1e12 ** (-np.arange(ni) / (ni - 1))
                A = (Q * lam) @ Q.T
                A = np.asfortranarray((A + A.T) / 2)
            elif i == 2:
                A = np.asfortranarray(np.eye(ni))
                A[100, 100] = -2.0
            else:
                A = oracle.dplgsy(float(ni), UL_, ni, 40 + i)
            mats.append(A)
            b.set_a(i, A, UL_)
            b.set_b(i, np.asfortranarray(rng.uniform(-0.5, 0.5, (ni, int(b.nrhs[i])))))
        B0 = [b.get_b(i).copy() for i in range(len(sizes))]
        rc, it, info = _mixed(ch, b)
        assert rc == 101 and list(info) == [0, 0, 101, 0], (rc, list(info))
        # condition 1e12: the FP32-product factorisation either meets a
        # non-positive pivot (-3) or cannot be refined (-31)
        assert it[0] in (-3, -31) and it[2] == -3 and it[1] >= 0 and it[3] >= 0, list(it)
        for i in (0, 1, 3):
            assert oracle.check_solve(mats[i], b.get_b(i), B0[i])[0], i
