"""GPU parity of the FP64 Cholesky path, through the C ABI.

Every case compares libchameleon_amd on the MI355X with the oracle (the C
restatement of Chameleon's CPU path, see oracle/cham_oracle.c) on the same
seeded dplgsy/dplrnt inputs, and applies Chameleon's own acceptance checks
(testing/testing_zcheck_facto.c, threshold 60).  Tolerances:

* generators: bit-exact;
* factor: relative Frobenius distance to the oracle factor <= 1e-12;
* solution: relative Frobenius distance to the oracle solution <= 1e-11
  (it is the factor's error amplified by cond(A), which is O(1) for the
  diagonally bumped generator);
* non-SPD inputs: the same `info` as LAPACK dpotrf.

The parameter grids are those of testing/input/{potrf,potrs,posv}.in
(nb 16/17, odd n, lda > n, nrhs > nb) plus the large tile sizes of the
MI355X path.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

L_, U_, UL_ = 122, 121, 123
TOL_FACTOR = 1e-12
TOL_SOLVE = 1e-11


@pytest.fixture(scope="module")
def ch():
    import chameleon_amd as ch
    ch.init()
    yield ch


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def reference_solution(uplo, n, nrhs, nb, lda=None, ldb=None):
    """Oracle run of CHAMELEON_dposv on the generator inputs."""
    lda = lda or n
    ldb = ldb or n
    A = oracle.dplgsy(float(n), uplo, n, 51, lda=lda, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, lda=ldb, nb=nb)
    info = oracle.dposv(uplo, A, n, B, nrhs, nb)
    return A, B, info


# ----------------------------------------------------------------- generators
@pytest.mark.parametrize("n,nb,lda", [(15, 16, 15), (37, 17, 41), (100, 32, 100), (1000, 128, 1003), (700, 320, 700)])
@pytest.mark.parametrize("uplo", [L_, U_, UL_])
def test_dplgsy_bitexact(ch, n, nb, lda, uplo):
    say("dplgsy", n, nb, lda, uplo)
    ch.set_param(ch.TILE_SIZE, nb)
    A = np.full((lda, n), 3.25, order="F")
    ch.check(ch.lib.CHAMELEON_dplgsy(float(n), uplo, n, _p(A), lda, 51))
    R = np.full((lda, n), 3.25, order="F")
    oracle.dplgsy(float(n), uplo, n, 51, lda=lda, nb=nb, out=R)
    assert np.array_equal(A, R)


@pytest.mark.parametrize("m,n,nb", [(37, 13, 17), (16, 22, 16), (1000, 64, 320), (333, 777, 128)])
def test_dplrnt_bitexact(ch, m, n, nb):
    say("dplrnt", m, n, nb)
    ch.set_param(ch.TILE_SIZE, nb)
    A = np.zeros((m, n), order="F")
    ch.check(ch.lib.CHAMELEON_dplrnt(m, n, _p(A), m, 5673))
    assert np.array_equal(A, oracle.dplrnt(m, n, 5673, nb=nb))


# ------------------------------------------------------- posv.in grid (tile API)
POSV_GRID = [(nb, n, nrhs) for nb in (16, 17) for n in (15, 21, 35) for nrhs in (1, 13, 22, 33)]


@pytest.mark.parametrize("nb,n,nrhs", POSV_GRID)
@pytest.mark.parametrize("uplo", [L_, U_])
def test_posv_tile_grid(ch, nb, n, nrhs, uplo):
    lda, ldb = 37, 39
    say("posv_tile", nb, n, nrhs, uplo)
    ch.set_param(ch.TILE_SIZE, nb)
    A = oracle.dplgsy(float(n), uplo, n, 51, lda=lda, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, lda=ldb, nb=nb)
    A_in = A.copy(order="F")
    dA = ch.Desc.from_array(A, nb, n, n)
    dB = ch.Desc.from_array(B, nb, n, nrhs)
    try:
        assert ch.lib.CHAMELEON_dposv_Tile(uplo, dA.ptr, dB.ptr) == 0
    finally:
        dA.destroy()
        dB.destroy()
    Ar, Br, info = reference_solution(uplo, n, nrhs, nb, lda, ldb)
    assert info == 0
    tri = np.tril if uplo == L_ else np.triu
    assert oracle.rel_frobenius(tri(A[:n, :n]), tri(Ar[:n, :n])) <= TOL_FACTOR
    assert oracle.rel_frobenius(B[:n], Br[:n]) <= TOL_SOLVE
    # the other triangle and the rows beyond n are untouched
    other = np.triu if uplo == L_ else np.tril
    k = 1 if uplo == L_ else -1
    assert np.array_equal(other(A[:n, :n], k), other(A_in[:n, :n], k))
    assert np.array_equal(A[n:], A_in[n:])
    A0 = oracle.dplgsy(float(n), UL_, n, 51)
    assert oracle.check_factorization(A0, A[:n, :n], uplo)[0]
    B0 = oracle.dplrnt(n, nrhs, 5673)
    assert oracle.check_solve(A0, B[:n], B0)[0]


@pytest.mark.parametrize("nb,n", [(nb, n) for nb in (16, 17) for n in (15, 19, 37)])
@pytest.mark.parametrize("uplo", [L_, U_])
def test_potrf_then_potrs_ccrb(ch, nb, n, uplo):
    """Library-allocated (CCRB) descriptors filled through Lap2Desc, as the
    reference testing drivers do, potrf.in/potrs.in sizes (lda 41, ldb 43)."""
    say("potrf_potrs_ccrb", nb, n, uplo)
    ch.set_param(ch.TILE_SIZE, nb)
    for nrhs in (1, 13, 21, 35):
        A = oracle.dplgsy(float(n), UL_, n, 51, lda=41, nb=nb)
        B = oracle.dplrnt(n, nrhs, 5673, lda=43, nb=nb)
        with ch.Desc.alloc(n, n, nb) as dA, ch.Desc.alloc(n, nrhs, nb) as dB:
            ch.check(ch.lib.CHAMELEON_Lap2Desc(UL_, _p(A), 41, dA.ptr))
            ch.check(ch.lib.CHAMELEON_Lap2Desc(UL_, _p(B), 43, dB.ptr))
            assert ch.lib.CHAMELEON_dpotrf_Tile(uplo, dA.ptr) == 0
            assert ch.lib.CHAMELEON_dpotrs_Tile(uplo, dA.ptr, dB.ptr) == 0
            F = dA.to_array(n, n)
            X = dB.to_array(n, nrhs)
        Ar, Br, info = reference_solution(uplo, n, nrhs, nb, 41, 43)
        tri = np.tril if uplo == L_ else np.triu
        assert oracle.rel_frobenius(tri(F), tri(Ar[:n, :n])) <= TOL_FACTOR
        assert oracle.rel_frobenius(X, Br[:n]) <= TOL_SOLVE


# ------------------------------------------------------------- LAPACK layout
@pytest.mark.parametrize("uplo", [L_, U_])
def test_lapack_layout_and_lapacke(ch, uplo):
    n, nrhs, nb = 150, 7, 64
    say("lapack", uplo)
    ch.set_param(ch.TILE_SIZE, nb)
    Ar, Br, _ = reference_solution(uplo, n, nrhs, nb, n + 3, n + 5)
    tri = np.tril if uplo == L_ else np.triu
    ch_u = b"L" if uplo == L_ else b"U"
    # CHAMELEON_dposv
    A = oracle.dplgsy(float(n), uplo, n, 51, lda=n + 3, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, lda=n + 5, nb=nb)
    assert ch.lib.CHAMELEON_dposv(uplo, n, nrhs, _p(A), n + 3, _p(B), n + 5) == 0
    assert oracle.rel_frobenius(tri(A[:n]), tri(Ar[:n])) <= TOL_FACTOR
    assert oracle.rel_frobenius(B[:n], Br[:n]) <= TOL_SOLVE
    # CHAMELEON_dpotrf + CHAMELEON_dpotrs
    A = oracle.dplgsy(float(n), uplo, n, 51, lda=n + 3, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, lda=n + 5, nb=nb)
    assert ch.lib.CHAMELEON_dpotrf(uplo, n, _p(A), n + 3) == 0
    assert ch.lib.CHAMELEON_dpotrs(uplo, n, nrhs, _p(A), n + 3, _p(B), n + 5) == 0
    assert oracle.rel_frobenius(B[:n], Br[:n]) <= TOL_SOLVE
    # LAPACKE
    A = oracle.dplgsy(float(n), uplo, n, 51, lda=n + 3, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, lda=n + 5, nb=nb)
    assert ch.lib.CHAMELEON_lapacke_dposv(ch.COL_MAJOR, ch_u, n, nrhs, _p(A), n + 3, _p(B), n + 5) == 0
    assert oracle.rel_frobenius(B[:n], Br[:n]) <= TOL_SOLVE
    assert ch.lib.CHAMELEON_lapacke_dposv(101, ch_u, n, nrhs, _p(A), n + 3, _p(B), n + 5) == -1
    # Fortran
    A = oracle.dplgsy(float(n), uplo, n, 51, lda=n + 3, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, lda=n + 5, nb=nb)
    c_u = ctypes.create_string_buffer(ch_u)
    iv = lambda v: ctypes.byref(ctypes.c_int(v))
    info = ctypes.c_int(-99)
    ch.lib.chameleon_lapack_dposv_(c_u, iv(n), iv(nrhs), _p(A), iv(n + 3), _p(B), iv(n + 5), ctypes.byref(info))
    assert info.value == 0
    assert oracle.rel_frobenius(B[:n], Br[:n]) <= TOL_SOLVE


def test_argument_errors(ch):
    say("argument errors")
    A = np.zeros((4, 4), order="F")
    L = ch.lib
    assert L.CHAMELEON_dposv(99, 4, 1, _p(A), 4, _p(A), 4) == -1
    assert L.CHAMELEON_dposv(L_, -1, 1, _p(A), 4, _p(A), 4) == -2
    assert L.CHAMELEON_dposv(L_, 4, -1, _p(A), 4, _p(A), 4) == -3
    assert L.CHAMELEON_dposv(L_, 4, 1, _p(A), 3, _p(A), 4) == -5
    assert L.CHAMELEON_dposv(L_, 4, 1, _p(A), 4, _p(A), 3) == -7
    assert L.CHAMELEON_dpotrf(L_, 4, _p(A), 3) == -4
    assert L.CHAMELEON_dposv(L_, 0, 1, _p(A), 1, _p(A), 1) == 0
    # non-square tiles are refused by the tile interface
    p = ctypes.c_void_p()
    assert L.CHAMELEON_Desc_Create(ctypes.byref(p), _p(A), ch.REAL_DOUBLE, 2, 4, 8, 4, 4, 0, 0, 4, 4, 1, 1) == 0
    assert L.CHAMELEON_dpotrf_Tile(L_, p) == ch.ERR_ILLEGAL_VALUE
    L.CHAMELEON_Desc_Destroy(ctypes.byref(p))


# ------------------------------------------------------------------ non-SPD
@pytest.mark.parametrize("key,n,bump,seed", [("nonspd_n21_info", 21, 0.0, 3), ("nonspd_n40_b2_info", 40, 2.0, 9)])
@pytest.mark.parametrize("nb", [4, 16, 17])
@pytest.mark.parametrize("uplo", [L_, U_])
def test_non_spd_info(ch, golden, key, n, bump, seed, nb, uplo):
    say("non-spd", key, nb, uplo)
    ch.set_param(ch.TILE_SIZE, nb)
    A = oracle.dplgsy(bump, UL_, n, seed)
    info = ch.lib.CHAMELEON_dpotrf(uplo, n, _p(A), n)
    assert info == int(golden[key][0])
    A = oracle.dplgsy(bump, UL_, n, seed)
    B = np.ones((n, 2), order="F")
    assert ch.lib.CHAMELEON_dposv(uplo, n, 2, _p(A), n, _p(B), n) == int(golden[key][0])


@pytest.mark.parametrize("nb", [64, 128, 320])
@pytest.mark.parametrize("uplo", [L_, U_])
def test_non_spd_info_fused_panel(ch, nb, uplo):
    """A diagonally dominant SPD matrix with one diagonal entry set to -1 has
    its first non-positive leading minor exactly there: info = j + 1.  The
    positions fall in the first, a middle and the last tile, at different
    64-column blocks of the multi-workgroup panel factorisation."""
    n = 700
    for j in (5, 200, 650):
        say("non-spd fused", nb, uplo, j)
        ch.set_param(ch.TILE_SIZE, nb)
        A = oracle.dplgsy(float(n), UL_, n, 51)
        A[j, j] = -1.0
        assert ch.lib.CHAMELEON_dpotrf(uplo, n, _p(A), n) == j + 1
        A = oracle.dplgsy(float(n), UL_, n, 51)
        A[j, j] = -1.0
        B = np.ones((n, 3), order="F")
        assert ch.lib.CHAMELEON_dposv(uplo, n, 3, _p(A), n, _p(B), n) == j + 1
        # the solve is not applied (LAPACK dposv calls dpotrs only when info
        # is 0), also when the forward solve runs beside the factorisation
        assert np.array_equal(B, np.ones((n, 3), order="F"))
    # the context is usable afterwards
    A = oracle.dplgsy(float(n), UL_, n, 51)
    assert ch.lib.CHAMELEON_dpotrf(uplo, n, _p(A), n) == 0


# ---------------------------------------------------------- large tile sizes
@pytest.mark.parametrize("n,nb,nrhs", [(1000, 128, 16), (1000, 320, 64), (2048, 256, 64), (2100, 384, 64),
                                       (2048, 512, 1), (1536, 192, 3), (1100, 128, 5), (2250, 256, 64),
                                       (2600, 512, 17)])
@pytest.mark.parametrize("uplo", [L_, U_])
def test_large_tiles_vs_oracle(ch, n, nb, nrhs, uplo):
    say("large", n, nb, nrhs, uplo)
    ch.set_param(ch.TILE_SIZE, nb)
    A = oracle.dplgsy(float(n), uplo, n, 51, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    assert ch.lib.CHAMELEON_dposv(uplo, n, nrhs, _p(A), n, _p(B), n) == 0
    oracle.use_blas(True)
    try:
        Ar, Br, info = reference_solution(uplo, n, nrhs, nb)
    finally:
        oracle.use_blas(False)
    tri = np.tril if uplo == L_ else np.triu
    assert oracle.rel_frobenius(tri(A), tri(Ar)) <= TOL_FACTOR
    assert oracle.rel_frobenius(B, Br) <= TOL_SOLVE
    A0 = oracle.dplgsy(float(n), UL_, n, 51)
    assert oracle.check_factorization(A0, A, uplo)[0]
    assert oracle.check_solve(A0, B, oracle.dplrnt(n, nrhs, 5673))[0]


@pytest.mark.parametrize("n,nb,nrhs", [(1536, 256, 65), (1024, 512, 130), (960, 320, 200), (1344, 192, 48)])
def test_solve_rhs_chunks(ch, n, nb, nrhs):
    """The strip solve handles the right-hand sides in chunks of 64 columns
    (and a narrower last chunk); every chunk must match the oracle."""
    say("rhs chunks", n, nb, nrhs)
    ch.set_param(ch.TILE_SIZE, nb)
    A = oracle.dplgsy(float(n), L_, n, 51, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    assert ch.lib.CHAMELEON_dposv(L_, n, nrhs, _p(A), n, _p(B), n) == 0
    oracle.use_blas(True)
    try:
        Ar, Br, _ = reference_solution(L_, n, nrhs, nb)
    finally:
        oracle.use_blas(False)
    assert oracle.rel_frobenius(B, Br) <= TOL_SOLVE
    for j in range(0, nrhs, 16):
        assert oracle.rel_frobenius(B[:, j:j + 16], Br[:, j:j + 16]) <= TOL_SOLVE


@pytest.mark.parametrize("nb", [40, 80, 100, 144, 200])
def test_partial_outer_blocks(ch, nb):
    """Tile sizes whose padded width is not a multiple of 64 exercise the
    narrower last column block of the POTRF/TRSM kernels."""
    n, nrhs = 611, 5
    say("partial outer block", nb)
    ch.set_param(ch.TILE_SIZE, nb)
    for uplo in (L_, U_):
        A = oracle.dplgsy(float(n), uplo, n, 51, nb=nb)
        B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
        assert ch.lib.CHAMELEON_dposv(uplo, n, nrhs, _p(A), n, _p(B), n) == 0
        oracle.use_blas(True)
        try:
            Ar, Br, _ = reference_solution(uplo, n, nrhs, nb)
        finally:
            oracle.use_blas(False)
        tri = np.tril if uplo == L_ else np.triu
        assert oracle.rel_frobenius(tri(A), tri(Ar)) <= TOL_FACTOR
        assert oracle.rel_frobenius(B, Br) <= TOL_SOLVE


# -------------------------------------------------- device-resident workflow
def test_async_device_resident_matches_sync(ch):
    n, nb, nrhs = 1024, 256, 64
    say("async", n, nb)
    ch.set_param(ch.TILE_SIZE, nb)
    L = ch.lib
    with ch.Desc.alloc(n, n, nb) as dA, ch.Desc.alloc(n, nrhs, nb) as dB:
        seq = ch.Sequence()
        assert L.CHAMELEON_dplgsy_Tile_Async(float(n), UL_, dA.ptr, 51, seq.ptr, seq.req) == 0
        assert L.CHAMELEON_dplrnt_Tile_Async(dB.ptr, 5673, seq.ptr, seq.req) == 0
        assert L.CHAMELEON_dposv_Tile_Async(L_, dA.ptr, dB.ptr, seq.ptr, seq.req) == 0
        assert L.CHAMELEON_Desc_Flush(dA.ptr, seq.ptr) == 0
        assert L.CHAMELEON_Desc_Flush(dB.ptr, seq.ptr) == 0
        assert seq.wait() == 0
        seq.destroy()
        F = dA.to_array(n, n)
        X = dB.to_array(n, nrhs)
    A = oracle.dplgsy(float(n), UL_, n, 51, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    assert ch.lib.CHAMELEON_dposv(L_, n, nrhs, _p(A), n, _p(B), n) == 0
    assert np.array_equal(np.tril(F), np.tril(A))
    assert np.array_equal(X, B)


def test_lookahead_schedules_agree(ch):
    """With lookahead the bulk update takes two panels per pass (one read and
    write of each trailing tile for two steps), which changes the order in
    which the two rank-nb terms are added: the factors of the two schedules
    agree to rounding and both pass the factorisation check."""
    n, nb = 1500, 128
    say("lookahead", n, nb)
    ch.set_param(ch.TILE_SIZE, nb)
    out = []
    for la in (1, 0):
        ch.set_param(ch.LOOKAHEAD, la)
        A = oracle.dplgsy(float(n), L_, n, 51, nb=nb)
        assert ch.lib.CHAMELEON_dpotrf(L_, n, _p(A), n) == 0
        out.append(np.tril(A))
    ch.set_param(ch.LOOKAHEAD, 1)
    assert oracle.rel_frobenius(out[0], out[1]) <= 1e-14
    A0 = oracle.dplgsy(float(n), UL_, n, 51)
    for F in out:
        assert oracle.check_factorization(A0, F, L_)[0]


@pytest.mark.parametrize("n,nb,nrhs", [(3000, 128, 1), (2500, 256, 2), (1000, 128, 3), (4100, 384, 4),
                                       (2048, 256, 6), (3333, 512, 8)])
@pytest.mark.parametrize("uplo", [L_, U_])
def test_narrow_solves_vs_oracle(ch, n, nb, nrhs, uplo):
    """One to eight right-hand sides with tiles a multiple of 128 take the
    narrow strip solve (128-row strips, vector FMAs, 128 x 128 strip
    inverses built from the 64 x 64 ones): dpotrf then dpotrs against the
    oracle's dposv, and Chameleon's residual check."""
    say("narrow solves", n, nb, nrhs, uplo)
    ch.set_param(ch.TILE_SIZE, nb)
    A = oracle.dplgsy(float(n), uplo, n, 51, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    assert ch.lib.CHAMELEON_dpotrf(uplo, n, _p(A), n) == 0
    assert ch.lib.CHAMELEON_dpotrs(uplo, n, nrhs, _p(A), n, _p(B), n) == 0
    oracle.use_blas(True)
    try:
        Ar, Br, info = reference_solution(uplo, n, nrhs, nb)
    finally:
        oracle.use_blas(False)
    assert info == 0
    assert oracle.rel_frobenius(B, Br) <= TOL_SOLVE
    for j in range(nrhs):
        assert oracle.rel_frobenius(B[:, j:j + 1], Br[:, j:j + 1]) <= TOL_SOLVE, j
    assert oracle.check_solve(oracle.dplgsy(float(n), UL_, n, 51), B, oracle.dplrnt(n, nrhs, 5673))[0]


@pytest.mark.parametrize("n,nb,nrhs", [(100, 64, 20), (129, 64, 33), (190, 64, 48), (200, 128, 64), (300, 128, 1),
                                       (256, 128, 9)])
@pytest.mark.parametrize("uplo", [L_, U_])
def test_strip_solve_few_strips(ch, n, nb, nrhs, uplo):
    """Two or three strips only (the chain's first and last strips are its
    only ones; the predecessor products of 16- and 32-column chunks, the
    48-column chunk without them, the narrow kernel with one 128-row strip
    per tile and a partial last tile), dpotrf then dpotrs against the
    oracle's dposv."""
    say("few strips", n, nb, nrhs, uplo)
    ch.set_param(ch.TILE_SIZE, nb)
    A = oracle.dplgsy(float(n), uplo, n, 51, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    assert ch.lib.CHAMELEON_dpotrf(uplo, n, _p(A), n) == 0
    assert ch.lib.CHAMELEON_dpotrs(uplo, n, nrhs, _p(A), n, _p(B), n) == 0
    Ar, Br, info = reference_solution(uplo, n, nrhs, nb)
    assert info == 0
    assert oracle.rel_frobenius(B, Br) <= TOL_SOLVE
    assert oracle.check_solve(oracle.dplgsy(float(n), UL_, n, 51), B, oracle.dplrnt(n, nrhs, 5673))[0]


@pytest.mark.parametrize("n,nb,nrhs", [(500, 512, 7), (1700, 256, 64), (1345, 192, 130), (2000, 320, 1),
                                       (1536, 128, 48), (3000, 256, 1), (2600, 512, 4), (3000, 384, 2)])
@pytest.mark.parametrize("uplo", [L_, U_])
def test_posv_forward_beside_factorisation(ch, n, nb, nrhs, uplo):
    """dposv solves tile rows forward right after their panels (beside the
    rest of the factorisation, with the panel's 64 x 64 block inverses) and
    the last `tail` tile rows by one strip solve after it; dpotrf then dpotrs
    solves everything after the factorisation.  Same factor, solutions equal
    to rounding, for every split (tail 0: every tile row the schedule
    reaches; -1: the default choice; a single tile has no rows below)."""
    ch.set_param(ch.TILE_SIZE, nb)
    A0 = oracle.dplgsy(float(n), uplo, n, 51, nb=nb)
    B0 = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    A2, B2 = A0.copy(order="F"), B0.copy(order="F")
    assert ch.lib.CHAMELEON_dpotrf(uplo, n, _p(A2), n) == 0
    assert ch.lib.CHAMELEON_dpotrs(uplo, n, nrhs, _p(A2), n, _p(B2), n) == 0
    tri = np.tril if uplo == L_ else np.triu
    Afull = oracle.dplgsy(float(n), UL_, n, 51)
    try:
        for tail in (0, 2, -1):
            say("posv forward beside potrf", n, nb, nrhs, uplo, tail)
            assert ch.lib.CHAMAMD_Set_Forward_Tail(tail) == 0
            A, B = A0.copy(order="F"), B0.copy(order="F")
            assert ch.lib.CHAMELEON_dposv(uplo, n, nrhs, _p(A), n, _p(B), n) == 0
            assert np.array_equal(tri(A), tri(A2))
            assert oracle.rel_frobenius(B, B2) <= 1e-13
            assert oracle.check_solve(Afull, B, oracle.dplrnt(n, nrhs, 5673))[0]
    finally:
        ch.lib.CHAMAMD_Set_Forward_Tail(-1)
    assert ch.lib.CHAMAMD_Set_Forward_Tail(-2) == ch.ERR_ILLEGAL_VALUE


def test_upper_is_transpose_of_lower(ch):
    n, nb = 600, 64
    say("upper vs lower", n, nb)
    ch.set_param(ch.TILE_SIZE, nb)
    AL = oracle.dplgsy(float(n), UL_, n, 51)
    AU = AL.copy(order="F")
    assert ch.lib.CHAMELEON_dpotrf(L_, n, _p(AL), n) == 0
    assert ch.lib.CHAMELEON_dpotrf(U_, n, _p(AU), n) == 0
    assert np.array_equal(np.tril(AL), np.triu(AU).T)


def test_norms(ch):
    n, nb = 300, 64
    say("norms")
    ch.set_param(ch.TILE_SIZE, nb)
    A = oracle.dplgsy(float(n), UL_, n, 51)
    G = oracle.dplrnt(n, 70, 3)
    L = ch.lib
    assert np.isclose(L.CHAMELEON_dlange(ch.ONE_NORM, n, 70, _p(G), n), np.abs(G).sum(0).max(), rtol=1e-13)
    assert np.isclose(L.CHAMELEON_dlange(ch.INF_NORM, n, 70, _p(G), n), np.abs(G).sum(1).max(), rtol=1e-13)
    assert L.CHAMELEON_dlange(ch.MAX_NORM, n, 70, _p(G), n) == np.abs(G).max()
    assert np.isclose(L.CHAMELEON_dlange(ch.FROBENIUS_NORM, n, 70, _p(G), n), np.linalg.norm(G), rtol=1e-13)
    for uplo in (L_, U_):
        assert np.isclose(L.CHAMELEON_dlansy(ch.ONE_NORM, uplo, n, _p(A), n), oracle.dlansy_one(A, uplo), rtol=1e-13)


def test_norms_propagate_nan(ch):
    """A NaN entry makes every norm NaN, as LAPACK's dlange/dlansy do (they
    test each entry with disnan); a max that drops NaN would hide it."""
    n, nb = 200, 64
    say("norms nan")
    ch.set_param(ch.TILE_SIZE, nb)
    G = oracle.dplrnt(n, 70, 3)
    G[137, 41] = np.nan
    L = ch.lib
    for norm in (ch.ONE_NORM, ch.INF_NORM, ch.MAX_NORM, ch.FROBENIUS_NORM):
        assert np.isnan(L.CHAMELEON_dlange(norm, n, 70, _p(G), n)), norm
    A = oracle.dplgsy(float(n), UL_, n, 51)
    A[150, 20] = A[20, 150] = np.nan
    for uplo in (L_, U_):
        for norm in (ch.ONE_NORM, ch.MAX_NORM):
            assert np.isnan(L.CHAMELEON_dlansy(norm, uplo, n, _p(A), n)), (uplo, norm)
