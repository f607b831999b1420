"""The oracle pinned against the golden fixtures (LAPACK/BLAS outputs and a
pure-Python LCG), against invariants the reference's own tests check
(generator tile invariance, testing/testing_zcheck_aux.c:831-894), and
against Chameleon's acceptance checks (threshold 60)."""
import numpy as np
import pytest

import oracle

L, U = oracle.LOWER, oracle.UPPER


def test_lcg_jump_matches_sequential(golden):
    for key in golden.files:
        if key.endswith("_pos") and key.startswith("lcg_seed"):
            seed = int(key[len("lcg_seed"):-len("_pos")])
            pos = golden[key]
            states = golden[key.replace("_pos", "_state")]
            for p, s in zip(pos, states):
                assert oracle.rnd64_jump(int(p), seed) == int(s)


def test_lcg_jump_composes():
    # jump(a+b) == jump(b) started from jump(a)
    for a, b in [(3, 5), (1000, 12345), (2**40, 2**33 + 7)]:
        assert oracle.rnd64_jump(a + b, 77) == oracle.rnd64_jump(b, oracle.rnd64_jump(a, 77))


@pytest.mark.parametrize("n,bump,seed", [(15, 15.0, 51), (37, 37.0, 51), (21, 0.0, 3)])
@pytest.mark.parametrize("nb", [1, 4, 16, 17, 40])
def test_plgsy_element_rule_and_tile_invariance(golden, n, bump, seed, nb):
    ref = golden[f"plgsy_n{n}_b{int(bump)}_s{seed}"]
    A = oracle.dplgsy(bump, oracle.UPPERLOWER, n, seed, nb=nb)
    assert np.array_equal(A, ref)


@pytest.mark.parametrize("m,n,seed", [(37, 13, 5673), (16, 22, 7)])
@pytest.mark.parametrize("nb", [1, 5, 16, 17, 64])
def test_plrnt_element_rule_and_tile_invariance(golden, m, n, seed, nb):
    ref = golden[f"plrnt_m{m}_n{n}_s{seed}"]
    assert np.array_equal(oracle.dplrnt(m, n, seed, nb=nb), ref)


def test_plgsy_uplo_touches_only_its_tiles():
    n, nb = 37, 16
    A = np.full((n, n), 7.0, order="F")
    oracle.dplgsy(n, L, n, 51, nb=nb, out=A)
    full = oracle.dplgsy(n, oracle.UPPERLOWER, n, 51, nb=nb)
    for m in range(3):
        for k in range(3):
            blk = (slice(m * nb, (m + 1) * nb), slice(k * nb, (k + 1) * nb))
            if k > m:
                assert np.all(A[blk] == 7.0)
            else:
                assert np.array_equal(A[blk], full[blk])
    vals = full - np.diag(np.diag(full))
    assert vals.max() <= 0.5 and vals.min() > -0.5


def _tile_kernels_through(golden):
    lib = oracle.lib()
    p = oracle._p
    T = np.asfortranarray(golden["trsm_T"])
    B = np.array(golden["trsm_B"], order="F")
    lib.orc_dtrsm_tile(142, L, 112, 131, 13, 17, 1.0, p(T), 17, p(B), 13)
    assert np.allclose(B, golden["trsm_RLTN"], rtol=1e-13, atol=1e-14)
    for trans, key in ((111, "trsm_LLNN"), (112, "trsm_LLTN")):
        Bl = np.array(golden["trsm_Bl"], order="F")
        lib.orc_dtrsm_tile(141, L, trans, 131, 17, 5, 1.0, p(T), 17, p(Bl), 17)
        assert np.allclose(Bl, golden[key], rtol=1e-13, atol=1e-14)
    A = np.asfortranarray(golden["syrk_A"])
    C = np.array(golden["syrk_C"], order="F")
    lib.orc_dsyrk_tile(L, 111, 17, 11, -1.0, p(A), 17, 1.0, p(C), 17)
    ref = golden["syrk_LN"]
    assert np.allclose(np.tril(C), np.tril(ref), rtol=1e-13, atol=1e-14)
    assert np.array_equal(np.triu(C, 1), np.triu(golden["syrk_C"], 1))
    Ag, Bg = np.asfortranarray(golden["gemm_A"]), np.asfortranarray(golden["gemm_B"])
    Cg = np.array(golden["gemm_C"], order="F")
    lib.orc_dgemm_tile(111, 112, 13, 9, 11, -1.0, p(Ag), 13, p(Bg), 9, 1.0, p(Cg), 13)
    assert np.allclose(Cg, golden["gemm_NT"], rtol=1e-13, atol=1e-14)


def test_tile_kernels_loops_vs_blas_golden(golden):
    oracle.use_blas(False)
    _tile_kernels_through(golden)


def test_tile_kernels_openblas_vs_blas_golden(golden):
    if oracle.use_blas(True) is None:
        pytest.skip("no host OpenBLAS")
    try:
        _tile_kernels_through(golden)
    finally:
        oracle.use_blas(False)


CASES = [(15, 1), (19, 13), (21, 22), (35, 33), (37, 13), (64, 16), (100, 7)]


@pytest.mark.parametrize("n,nrhs", CASES)
@pytest.mark.parametrize("uplo", [L, U])
@pytest.mark.parametrize("nb", [16, 17])
def test_tiled_posv_vs_lapack(golden, n, nrhs, uplo, nb):
    ch = "L" if uplo == L else "U"
    lda, ldb = n + 4, n + 6     # lda > n as in testing/input/posv.in
    A = oracle.dplgsy(float(n), uplo, n, 51, lda=lda, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, lda=ldb, nb=nb)
    A0 = oracle.dplgsy(float(n), oracle.UPPERLOWER, n, 51)
    B0 = B[:n].copy()
    other = (np.triu(A[:n, :n], 1) if uplo == L else np.tril(A[:n, :n], -1)).copy()
    info = oracle.dposv(uplo, A, n, B, nrhs, nb)
    assert info == 0
    F = A[:n, :n]
    ref_F = golden[f"chol_n{n}_{ch}"]
    tri = np.tril if uplo == L else np.triu
    assert oracle.rel_frobenius(tri(F), ref_F) < 1e-13
    # the other triangle is not referenced and stays untouched
    assert np.array_equal(np.triu(F, 1) if uplo == L else np.tril(F, -1), other)
    assert oracle.rel_frobenius(B[:n], golden[f"sol_n{n}_r{nrhs}_{ch}"]) < 1e-12
    assert oracle.check_factorization(A0, F, uplo)[0]
    assert oracle.check_solve(A0, B[:n], B0)[0]


def test_potrf_then_potrs_equals_posv():
    n, nrhs, nb = 70, 5, 16
    A1 = oracle.dplgsy(float(n), L, n, 51, nb=nb)
    A2 = A1.copy(order="F")
    B1 = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    B2 = B1.copy(order="F")
    assert oracle.dposv(L, A1, n, B1, nrhs, nb) == 0
    assert oracle.dpotrf(L, A2, n, nb) == 0
    assert oracle.dpotrs(L, A2, n, B2, nrhs, nb) == 0
    assert np.array_equal(A1, A2) and np.array_equal(B1, B2)


@pytest.mark.parametrize("key,n,bump,seed", [("nonspd_n21_info", 21, 0.0, 3), ("nonspd_n40_b2_info", 40, 2.0, 9)])
@pytest.mark.parametrize("nb", [4, 16])
def test_non_spd_reports_leading_minor(golden, key, n, bump, seed, nb):
    A = oracle.dplgsy(bump, oracle.UPPERLOWER, n, seed)
    info = oracle.dpotrf(L, A, n, nb)
    assert info == int(golden[key][0]) > 0


def test_argument_errors():
    A = np.zeros((4, 4), order="F")
    assert oracle.lib().orc_dpotrf(99, 4, oracle._p(A), 4, 2, 0) == -1
    assert oracle.lib().orc_dpotrf(L, -1, oracle._p(A), 4, 2, 0) == -2
    assert oracle.lib().orc_dpotrf(L, 4, oracle._p(A), 3, 2, 0) == -4
    assert oracle.lib().orc_dposv(L, 4, -1, oracle._p(A), 4, oracle._p(A), 4, 2, 0) == -3
    assert oracle.lib().orc_dposv(L, 4, 1, oracle._p(A), 4, oracle._p(A), 3, 2, 0) == -7


def test_checks_catch_planted_errors():
    n = 30
    A0 = oracle.dplgsy(float(n), oracle.UPPERLOWER, n, 51)
    F = np.linalg.cholesky(A0)
    assert oracle.check_factorization(A0, F, L)[0]
    bad = F.copy()
    bad[17, 3] += 1e-9
    assert not oracle.check_factorization(A0, bad, L)[0]
    B = oracle.dplrnt(n, 3, 5)
    X = np.linalg.solve(A0, B)
    assert oracle.check_solve(A0, X, B)[0]
    Xb = X.copy()
    Xb[5, 1] *= 1 + 1e-10
    assert not oracle.check_solve(A0, Xb, B)[0]


# ------------------------------------------------------------------ batches
@pytest.mark.parametrize("M,N,mb,nb", [(96, 64, 32, 32), (40, 40, 20, 20), (48, 16, 16, 16)])
def test_batch_generator_closed_form(M, N, mb, nb):
    """zplghe_batch.c:56-57 puts tile (m,n) at origin m0 = n0 = (n*mt+m)*mb*nb
    with bigM = 0, so element (i,j) of a tile is the LCG value at position
    m0 + max(i,j) (diagonal + bump): checked against the jump-ahead alone."""
    seed, bump = 77, 3.0
    A = oracle.dplgsy_batch(bump, M, N, mb, nb, seed)
    mt = M // mb
    two64 = 2.0 ** -64
    for n in range(N // nb):
        for m in range(mt):
            m0 = (n * mt + m) * mb * nb
            T = A[m * mb:(m + 1) * mb, n * nb:(n + 1) * nb]
            for i in range(nb):
                for j in range(nb):
                    v = 0.5 - oracle.rnd64_jump(m0 + max(i, j), seed) * two64
                    assert T[i, j] == v + (bump if i == j else 0.0)


def test_batch_potrf_tiles_are_independent():
    M, N, mb, nb = 90, 60, 30, 30
    A = oracle.dplgsy_batch(float(nb), M, N, mb, nb, 5)
    A0 = A.copy(order="F")
    info = oracle.dpotrf_batch(L, A, M, N, mb, nb)
    assert not info.any()
    for n in range(N // nb):
        for m in range(M // mb):
            T0 = A0[m * mb:(m + 1) * mb, n * nb:(n + 1) * nb]
            T = A[m * mb:(m + 1) * mb, n * nb:(n + 1) * nb]
            Lr = np.linalg.cholesky(T0)          # each tile alone, textbook
            assert np.allclose(np.tril(T), Lr, rtol=0, atol=1e-13)
            assert np.array_equal(np.triu(T, 1), np.triu(T0, 1))   # upper triangle untouched
    bad = oracle.dplgsy_batch(1.0, M, N, mb, nb, 5)
    info = oracle.dpotrf_batch(L, bad, M, N, mb, nb)
    assert (info > 0).all()
