"""GPU tests of the mixed-precision solve with FP64 iterative refinement
(CHAMELEON_dsposv / CHAMAMD_dposv_mixed_Tile) and of the FP64 backward
error report.

Parity bar: whatever the product precision, a converged refinement must
give the FP64 answer -- the solution within cond(A) * 1e-13 of an FP64
reference solve and a scaled residual under Chameleon's threshold (60);
a refinement that cannot converge must fall back to FP64 (ITER < 0) with
the same guarantees.  A is unchanged when refinement succeeded.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

L_, U_, UL_ = 122, 121, 123
FP32, BF16, BF16X3, FP8 = 0, 1, 2, 3


@pytest.fixture(scope="module")
def ch():
    import chameleon_amd as ch
    ch.init()
    L = ch.load()
    L.CHAMELEON_dsposv.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int)]
    L.CHAMAMD_dposv_mixed_Tile.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    L.CHAMAMD_Backward_Error.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.POINTER(ctypes.c_double)]
    yield ch


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def conditioned_spd(n, kappa, seed):
    """A = Q diag(lambda) Q^T, Q from QR of a seeded Gaussian matrix,
    lambda geometric from 1 to 1/kappa (the construction of
    compute/pzlatms.c mode 3, core_zlatm1.c:172-180)."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = kappa ** (-np.arange(n) / (n - 1))
    A = (Q * lam) @ Q.T
    return np.asfortranarray((A + A.T) / 2)


def mixed(ch, uplo, A, B, prec, nb):
    n, nrhs = B.shape
    X = np.zeros_like(B, order="F")
    it, be = ctypes.c_int(-99), ctypes.c_double(-2.0)
    with ch.Desc.from_array(A, nb) as dA, ch.Desc.from_array(B, nb) as dB, ch.Desc.from_array(X, nb) as dX:
        rc = ch.lib.CHAMAMD_dposv_mixed_Tile(uplo, dA.ptr, dB.ptr, dX.ptr, prec, ctypes.byref(it), ctypes.byref(be))
    return rc, X, it.value, be.value


@pytest.mark.parametrize("prec", [FP32, BF16X3, BF16, FP8])
@pytest.mark.parametrize("uplo", [L_, U_])
@pytest.mark.parametrize("n,nb,nrhs", [(1024, 256, 1), (1536, 256, 16), (2048, 512, 64), (1000, 128, 3),
                                       (1150, 128, 7), (2500, 256, 32), (1536, 128, 128)])
def test_mixed_well_conditioned(ch, prec, uplo, n, nb, nrhs):
    say("mixed bumped", prec, uplo, n, nb, nrhs)
    A = oracle.dplgsy(float(n), uplo, n, 51, nb=nb)
    A_in = A.copy(order="F")
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    rc, X, it, be = mixed(ch, uplo, A, B, prec, nb)
    assert rc == 0
    assert 0 <= it <= 30, it
    assert np.array_equal(A, A_in)                     # A unchanged after refinement
    A0 = oracle.dplgsy(float(n), UL_, n, 51)
    Xr = np.linalg.solve(A0, B)
    assert oracle.rel_frobenius(X, Xr) <= 1e-13
    assert oracle.check_solve(A0, X, B)[0]
    assert 0.0 <= be <= 1.0


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
@pytest.mark.parametrize("prec", [FP32, BF16X3, BF16, FP8])
def test_mixed_conditioned(ch, kappa, prec):
    n, nb, nrhs = 1024, 256, 4
    say("mixed conditioned", kappa, prec)
    A = conditioned_spd(n, kappa, 7)
    A0 = A.copy(order="F")
    rng = np.random.default_rng(3)
    B = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, nrhs)))
    rc, X, it, be = mixed(ch, L_, A, B, prec, nb)
    assert rc == 0
    Xr = np.linalg.solve(A0, B)
    # forward error bounded by cond(A) times a modest multiple of eps
    assert oracle.rel_frobenius(X, Xr) <= kappa * 1e-13
    assert oracle.check_solve(A0, X, B)[0]
    if it >= 0:
        assert np.array_equal(A, A0)
    else:
        assert it in (-3, -31)
    if prec == FP32:
        assert it >= 0      # kappa * 2^-24 << 1: FP32 products always converge here


def test_dsposv_lapack_layout(ch):
    n, nrhs, nb = 1200, 5, 256
    say("dsposv lapack")
    ch.set_param(ch.TILE_SIZE, nb)
    A = oracle.dplgsy(float(n), L_, n, 51, lda=n + 3, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, lda=n + 5, nb=nb)
    X = np.zeros((n + 7, nrhs), order="F")
    it = ctypes.c_int(-99)
    rc = ch.lib.CHAMELEON_dsposv(L_, n, nrhs, _p(A), n + 3, _p(B), n + 5, _p(X), n + 7, ctypes.byref(it))
    assert rc == 0 and it.value >= 0
    A0 = oracle.dplgsy(float(n), UL_, n, 51)
    assert oracle.rel_frobenius(X[:n], np.linalg.solve(A0, B[:n])) <= 1e-13
    assert ch.lib.CHAMELEON_dsposv(L_, n, nrhs, _p(A), n + 3, _p(B), n + 5, _p(X), n - 1, ctypes.byref(it)) == -9


def test_mixed_untileable_uses_fp64(ch):
    n, nb, nrhs = 500, 100, 2        # padded tile 112: not a multiple of 128
    say("mixed fallback -1")
    A = oracle.dplgsy(float(n), L_, n, 51, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    rc, X, it, be = mixed(ch, L_, A, B, FP32, nb)
    assert rc == 0 and it == -1
    A0 = oracle.dplgsy(float(n), UL_, n, 51)
    assert oracle.rel_frobenius(X, np.linalg.solve(A0, B)) <= 1e-12
    F = oracle.dplgsy(float(n), L_, n, 51, nb=nb)
    assert oracle.dpotrf(L_, F, n, nb) == 0
    assert oracle.rel_frobenius(np.tril(A), np.tril(F)) <= 1e-12   # A holds the FP64 factor


def test_mixed_non_spd_reports_info(ch, golden):
    n, nb = 256, 128
    say("mixed non-spd")
    A = oracle.dplgsy(0.0, UL_, n, 3)
    _, info_ref = __import__("scipy.linalg.lapack", fromlist=["dpotrf"]).dpotrf(A.copy(), lower=True)
    B = np.ones((n, 1), order="F")
    rc, X, it, be = mixed(ch, L_, A, B, FP32, nb)
    assert it == -3
    assert rc == info_ref > 0


@pytest.mark.parametrize("uplo,n,nb,nrhs,prec", [(L_, 1150, 128, 3, BF16X3), (U_, 1150, 128, 3, FP32),
                                                  (L_, 2048, 512, 8, BF16X3), (U_, 1000, 256, 2, BF16X3),
                                                  (L_, 1000, 120, 2, BF16X3), (U_, 1000, 120, 2, FP32)])
def test_mixed_norm_matches_backward_error(ch, uplo, n, nb, nrhs, prec):
    """The refinement forms ||A||_inf in the same pass as its working copy
    of A's lower tiles.  The backward error it reports for the returned X
    must equal CHAMAMD_Backward_Error's, which recomputes the same residual
    (same kernel, same operands) with the separate norm kernel: the ratio of
    the two is the ratio of the two norms.  A dropped transposed half, a
    counted padding row or a doubled diagonal moves it by percents here
    (dense A, last tile ragged for 1150/128 and 1000/256; nb 120 pads every
    tile to 128 rows)."""
    say("mixed norm", uplo, n, nb, nrhs, prec)
    A = conditioned_spd(n, 1e3, 11)
    if uplo == L_:
        A[np.triu_indices(n, 1)] = 7.0       # outside the stored triangle: must not count
    else:
        A[np.tril_indices(n, -1)] = 7.0
    A0 = A.copy(order="F")
    rng = np.random.default_rng(5)
    B = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, nrhs)))
    rc, X, it, be = mixed(ch, uplo, A, B, prec, nb)
    assert rc == 0 and it >= 0, it
    assert np.array_equal(A, A0)
    ref = ctypes.c_double(-1)
    with ch.Desc.from_array(A0, nb) as dA, ch.Desc.from_array(X, nb) as dX, ch.Desc.from_array(B, nb) as dB:
        assert ch.lib.CHAMAMD_Backward_Error(uplo, dA.ptr, dX.ptr, dB.ptr, ctypes.byref(ref)) == 0
    assert 0.0 < be < 1.0
    assert abs(be / ref.value - 1.0) <= 1e-12, (be, ref.value)
    # and the norm itself: the backward error with numpy's ||A||_inf of the symmetric A
    S = np.tril(A0) + np.tril(A0, -1).T if uplo == L_ else np.triu(A0) + np.triu(A0, 1).T
    R = B - S @ X
    ref_np = max(np.abs(R[:, j]).max() / (np.abs(S).sum(1).max() * np.abs(X[:, j]).max() * n * 2.0 ** -53)
                 for j in range(nrhs))
    assert 0.25 <= be / ref_np <= 4.0


@pytest.mark.parametrize("n,nb,nrhs", [(1536, 128, 48), (2000, 256, 5), (1150, 128, 7)])
@pytest.mark.parametrize("uplo", [L_, U_])
def test_mixed_forward_beside_factorisation(ch, n, nb, nrhs, uplo):
    """The first solve of the refinement runs forward beside the
    factorisation (tile rows up to the tail) and by strip solve after it.
    A wrong first solve would still be corrected by refinement, so besides
    the answer the step count must match the run that solves every row
    after the factorisation (tail 1000 tile rows: none beside it)."""
    A0 = oracle.dplgsy(float(n), uplo, n, 51, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    Xr = np.linalg.solve(oracle.dplgsy(float(n), UL_, n, 51), B)
    try:
        ref = None
        for tail in (1000, 0, 2, -1):
            say("mixed forward beside potrf", n, nb, nrhs, uplo, tail)
            assert ch.lib.CHAMAMD_Set_Forward_Tail(tail) == 0
            A = A0.copy(order="F")
            rc, X, it, be = mixed(ch, uplo, A, B, BF16X3, nb)
            assert rc == 0 and it >= 0, it
            assert np.array_equal(A, A0)
            assert oracle.rel_frobenius(X, Xr) <= 1e-13
            if ref is None:
                ref = (it, X)
            else:
                assert it == ref[0], (tail, it, ref[0])
                assert oracle.rel_frobenius(X, ref[1]) <= 1e-13
    finally:
        ch.lib.CHAMAMD_Set_Forward_Tail(-1)


def test_backward_error_of_fp64_solution(ch):
    n, nb, nrhs = 2048, 256, 8
    say("backward error")
    A = oracle.dplgsy(float(n), L_, n, 51, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    X = B.copy(order="F")
    ch.set_param(ch.TILE_SIZE, nb)
    assert ch.lib.CHAMELEON_dposv(L_, n, nrhs, _p(A), n, _p(X), n) == 0
    A0 = oracle.dplgsy(float(n), L_, n, 51, nb=nb)
    be = ctypes.c_double(-1)
    with ch.Desc.from_array(A0, nb) as dA, ch.Desc.from_array(X, nb) as dX, ch.Desc.from_array(B, nb) as dB:
        assert ch.lib.CHAMAMD_Backward_Error(L_, dA.ptr, dX.ptr, dB.ptr, ctypes.byref(be)) == 0
    Af = oracle.dplgsy(float(n), UL_, n, 51)
    R = B - Af @ X
    ref = max(np.abs(R[:, j]).max() / (np.abs(Af).sum(1).max() * np.abs(X[:, j]).max() * n * 2.0 ** -53)
              for j in range(nrhs))
    # both are rounding-level quantities summed in different orders
    assert 0.0 < be.value < 1.0 and 0.25 <= be.value / ref <= 4.0


@pytest.mark.parametrize("prec", [FP32, BF16, FP8])
def test_mixed_diagonal_scaling(ch, prec):
    """A = D^1/2 A0 D^1/2, A0 the bumped dplgsy matrix (condition number
    ~1) and D spanning 1e-6 .. 1e6: the symmetric diagonal scaling hands the
    low-precision factorisation A0 itself, so refinement converges in every
    precision, FP8 included; without the scaling the answer is still the
    FP64 one (converged or by fallback).

    The scaled matrix has a normwise condition number near 1e12 and
    refinement stops at LAPACK's normwise test ||r|| <= ||x|| ||A|| eps
    sqrt(N), which bounds the forward error only through that number.  With
    the scaling the answer is held to 1e-10 of numpy's LU solution (6.4e-12
    measured for FP32 products on MI355X); without it, refinement from an
    FP8 factor of the unscaled matrix stops at that test further from it
    (1.8e-9 measured), so 1e-7 there.  Both meet Chameleon's residual
    check."""
    n, nb, nrhs = 1536, 256, 4
    A0 = oracle.dplgsy(float(n), UL_, n, 51)
    d = np.logspace(-6, 6, n)[np.random.default_rng(2).permutation(n)]
    As = np.asfortranarray(np.sqrt(d)[:, None] * A0 * np.sqrt(d)[None, :])
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    Xr = np.linalg.solve(As, B)
    try:
        for on in (1, 0):
            say("mixed scaling", prec, on)
            assert ch.lib.CHAMAMD_Set_Mixed_Scaling(on) == 0
            A = As.copy(order="F")
            rc, X, it, be = mixed(ch, L_, A, B, prec, nb)
            assert rc == 0
            if on:
                assert 0 <= it <= 30, it
            assert oracle.rel_frobenius(X, Xr) <= (1e-10 if on else 1e-7)
            assert oracle.check_solve(As, X, B)[0]
    finally:
        ch.lib.CHAMAMD_Set_Mixed_Scaling(1)


def test_mixed_fp8_nonfinite_falls_back(ch):
    """A NaN below the diagonal of the first panel reaches the FP8 demotion,
    whose range check stops the low-precision factorisation (ITER = -3);
    the FP64 fallback then reports the info of reference LAPACK for the same
    matrix: dpotrf2/dpotf2 stop at the first pivot that is not positive or
    is NaN (DISNAN), here column 301, which the oracle's textbook tile
    kernels reproduce.  (OpenBLAS's own dpotrf, behind scipy, has no NaN
    test and returns 0 on this matrix.)"""
    n, nb = 512, 128
    say("mixed fp8 nan")
    A = oracle.dplgsy(float(n), UL_, n, 51)
    A[300, 5] = A[5, 300] = np.nan
    B = np.ones((n, 1), order="F")
    info_ref = oracle.dposv(L_, A.copy(order="F"), n, B.copy(order="F"), 1, nb)
    assert info_ref == 301
    rc, X, it, be = mixed(ch, L_, A.copy(order="F"), B, FP8, nb)
    assert it == -3
    assert rc == info_ref > 0


ADAPTIVE = 4


@pytest.mark.parametrize("u_high", [1e-12, 1e-6, 1e-3])
@pytest.mark.parametrize("n,nb", [(2048, 256), (1150, 128)])
def test_mixed_per_tile_precision(ch, u_high, n, nb):
    """Per-tile precision (CHAMAMD_PREC_ADAPTIVE): the tiles given FP64,
    FP32 and BF16 products are those of the oracle's restatement of
    Chameleon's zhered rule on the scaled working matrix S A S, and the
    refined answer is the FP64 one whatever the mix."""
    say("mixed per-tile", u_high, n, nb)
    A0 = oracle.dplgsy(float(n), UL_, n, 51)
    B = oracle.dplrnt(n, 3, 5673, nb=nb)
    d = 1.0 / np.sqrt(np.diag(A0))
    want = oracle.tile_precision_map(d[:, None] * A0 * d[None, :], nb, u_high)
    counts = [sum(1 for v in want.values() if v == p) for p in (0, 1, 2)]
    try:
        assert ch.lib.CHAMAMD_Set_Tile_Accuracy(u_high) == 0
        A = oracle.dplgsy(float(n), L_, n, 51, nb=nb)
        rc, X, it, be = mixed(ch, L_, A, B, ADAPTIVE, nb)
        got = (ctypes.c_longlong * 3)()
        assert ch.lib.CHAMAMD_Mixed_Tile_Counts(got) == 0
    finally:
        ch.lib.CHAMAMD_Set_Tile_Accuracy(1e-12)
    assert rc == 0 and 0 <= it <= 30, it
    assert list(got) == counts, (list(got), counts)
    assert oracle.rel_frobenius(X, np.linalg.solve(A0, B)) <= 1e-13
    assert oracle.check_solve(A0, X, B)[0]


@pytest.mark.parametrize("kappa", [1e2, 1e6])
def test_mixed_per_tile_precision_conditioned(ch, kappa):
    n, nb = 1024, 256
    say("mixed per-tile conditioned", kappa)
    A = conditioned_spd(n, kappa, 7)
    A0 = A.copy(order="F")
    B = np.asfortranarray(np.random.default_rng(3).uniform(-0.5, 0.5, (n, 2)))
    try:
        assert ch.lib.CHAMAMD_Set_Tile_Accuracy(1e-5) == 0
        rc, X, it, be = mixed(ch, L_, A, B, ADAPTIVE, nb)
    finally:
        ch.lib.CHAMAMD_Set_Tile_Accuracy(1e-12)
    assert rc == 0
    assert oracle.rel_frobenius(X, np.linalg.solve(A0, B)) <= kappa * 1e-13
    assert oracle.check_solve(A0, X, B)[0]
