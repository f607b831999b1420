"""Oracle for the tiled Cholesky path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package.  It wraps ``libcham_oracle.so`` (the C
restatement of Chameleon's CPU path, see cham_oracle.c) and restates
Chameleon's numerical acceptance checks in numpy:

* :func:`check_factorization` -- testing/testing_zcheck_facto.c:198-300,
  ``||A0 - L L^T||_1 / (||A0||_1 N eps) <= 60`` with the symmetric one-norm
  of the ``uplo`` triangle (dlansy).
* :func:`check_solve` -- testing/testing_zcheck_facto.c:392-450,
  ``||B - A X||_1 / (||A||_1 ||X||_1 max(M,N) eps) <= 60``.
* ``EPS = 2**-53`` -- control/lamch.c:49-53 (LAPACK dlamch('e')).
"""
import ctypes
import glob
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcham_oracle.so")

UPPER, LOWER, UPPERLOWER = 121, 122, 123
EPS = 2.0 ** -53
THRESHOLD = 60.0

_lib = None


def build():
    """Compile the C restatement (gcc, OpenMP)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "libcham_oracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64, i32, dbl, ptr = ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        L.orc_rnd64_jump.restype = u64
        L.orc_rnd64_jump.argtypes = [u64, u64]
        L.orc_dlaran_at.restype = dbl
        L.orc_dlaran_at.argtypes = [u64, u64]
        L.orc_dplgsy.argtypes = [dbl, i32, i32, ptr, i32, i32, u64]
        L.orc_dplrnt.argtypes = [i32, i32, ptr, i32, i32, u64]
        L.orc_dplgsy_tile.argtypes = [dbl, i32, i32, ptr, i32, i32, i32, i32, u64]
        L.orc_dplrnt_tile.argtypes = [i32, i32, ptr, i32, i32, i32, i32, u64]
        L.orc_dpotrf_tile.argtypes = [i32, i32, ptr, i32]
        L.orc_dtrsm_tile.argtypes = [i32, i32, i32, i32, i32, i32, dbl, ptr, i32, ptr, i32]
        L.orc_dsyrk_tile.argtypes = [i32, i32, i32, i32, dbl, ptr, i32, dbl, ptr, i32]
        L.orc_dgemm_tile.argtypes = [i32, i32, i32, i32, i32, dbl, ptr, i32, ptr, i32, dbl, ptr, i32]
        L.orc_dpotrf.argtypes = [i32, i32, ptr, i32, i32, i32]
        L.orc_dpotrs.argtypes = [i32, i32, i32, ptr, i32, ptr, i32, i32, i32]
        L.orc_dposv.argtypes = [i32, i32, i32, ptr, i32, ptr, i32, i32, i32]
        L.orc_set_blas.argtypes = [ptr, ptr, ptr, ptr, ptr]
        L.orc_dplgsy_batch.argtypes = [dbl, i32, i32, i32, i32, ptr, i32, u64]
        L.orc_dpotrf_batch.argtypes = [i32, i32, i32, i32, i32, ptr, i32, ptr]
        L.orc_wtime.restype = dbl
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _fortran(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


def rnd64_jump(n, seed):
    return lib().orc_rnd64_jump(n, seed)


def dplgsy(bump, uplo, n, seed, lda=None, nb=None, out=None):
    """Column-major symmetric random matrix of CHAMELEON_dplgsy."""
    lda = lda or max(1, n)
    A = out if out is not None else _fortran((lda, n))
    lib().orc_dplgsy(float(bump), uplo, n, _p(A), lda, nb or max(1, n), seed)
    return A


def dplrnt(m, n, seed, lda=None, nb=None, out=None):
    """Column-major uniform random matrix of CHAMELEON_dplrnt."""
    lda = lda or max(1, m)
    A = out if out is not None else _fortran((lda, n))
    lib().orc_dplrnt(m, n, _p(A), lda, nb or max(1, m), seed)
    return A


def dplgsy_batch(bump, M, N, mb, nb, seed, lda=None):
    """Tile batch of CHAMELEON_dplgsy_batch_Tile: every mb x nb tile of an
    M x N column-major matrix is an independent symmetric matrix."""
    lda = lda or max(1, M)
    A = _fortran((lda, N))
    lib().orc_dplgsy_batch(float(bump), M, N, mb, nb, _p(A), lda, seed)
    return A


def dpotrf_batch(uplo, A, M, N, mb, nb):
    """CHAMELEON_dpotrf_batch_Tile on a column-major matrix: every tile
    factored alone; returns the per-tile info array (tile n*mt + m)."""
    assert A.flags.f_contiguous and A.dtype == np.float64
    mt, nt = -(-M // mb), -(-N // nb)
    info = np.zeros(mt * nt, dtype=np.int32)
    lib().orc_dpotrf_batch(uplo, M, N, mb, nb, _p(A), A.shape[0], _p(info))
    return info


def dpotrf(uplo, A, n, nb, nthreads=0):
    """In-place tiled Cholesky of the leading n x n block of Fortran-ordered A."""
    assert A.flags.f_contiguous and A.dtype == np.float64
    return lib().orc_dpotrf(uplo, n, _p(A), A.shape[0], nb, nthreads)


def dpotrs(uplo, A, n, B, nrhs, nb, nthreads=0):
    assert A.flags.f_contiguous and B.flags.f_contiguous
    return lib().orc_dpotrs(uplo, n, nrhs, _p(A), A.shape[0], _p(B), B.shape[0], nb, nthreads)


def dposv(uplo, A, n, B, nrhs, nb, nthreads=0):
    assert A.flags.f_contiguous and B.flags.f_contiguous
    return lib().orc_dposv(uplo, n, nrhs, _p(A), A.shape[0], _p(B), B.shape[0], nb, nthreads)


def find_openblas():
    """Path of the LP64 OpenBLAS shipped with scipy (Fortran interface with a
    ``scipy_`` prefix), or None."""
    try:
        import scipy
    except ImportError:
        return None
    root = os.path.dirname(os.path.dirname(scipy.__file__))
    hits = sorted(glob.glob(os.path.join(root, "scipy.libs", "libscipy_openblas-*.so")))
    return hits[0] if hits else None


_blas_handle = None


def blas_config():
    """Build string of the OpenBLAS that find_openblas() picks (its
    ``scipy_openblas_get_config``), or None."""
    path = find_openblas()
    if path is None:
        return None
    h = ctypes.CDLL(path)
    fn = getattr(h, "scipy_openblas_get_config", None)
    if fn is None:
        return None
    fn.restype = ctypes.c_char_p
    return fn().decode(errors="replace")


def use_blas(enable=True):
    """Route the oracle's tile kernels through host OpenBLAS (single-threaded
    inside each task, as coreblas assumes) or back to the textbook loops."""
    global _blas_handle
    L = lib()
    if not enable:
        L.orc_set_blas(None, None, None, None, None)
        return None
    path = find_openblas()
    if path is None:
        return None
    _blas_handle = ctypes.CDLL(path)
    h = _blas_handle
    L.orc_set_blas(ctypes.cast(h.scipy_dpotrf_, ctypes.c_void_p), ctypes.cast(h.scipy_dtrsm_, ctypes.c_void_p),
                   ctypes.cast(h.scipy_dsyrk_, ctypes.c_void_p), ctypes.cast(h.scipy_dgemm_, ctypes.c_void_p),
                   ctypes.cast(h.scipy_openblas_set_num_threads, ctypes.c_void_p))
    return path


# ------------------------------------------------------------------ checks
def dlansy_one(A, uplo):
    """One-norm of the symmetric matrix whose `uplo` triangle is stored in A."""
    T = np.tril(A) if uplo == LOWER else np.triu(A)
    S = T + T.T - np.diag(np.diag(A))
    return np.abs(S).sum(axis=0).max() if S.size else 0.0


def check_factorization(A0, F, uplo):
    """Chameleon's check_dxxtrf for a symmetric matrix: returns
    (ok, ||A0 - L L^T||_1 / (||A0||_1 N eps)).  A0 is the full symmetric
    input, F holds the factor in its `uplo` triangle."""
    n = A0.shape[0]
    if uplo == LOWER:
        L = np.tril(F)
        R = A0 - L @ L.T
    else:
        U = np.triu(F)
        R = A0 - U.T @ U
    anorm = dlansy_one(A0, uplo)
    rnorm = dlansy_one(R, uplo)
    res = rnorm / (anorm * n * EPS)
    return (np.isfinite(res) and res <= THRESHOLD), res


def check_solve(A0, X, B):
    """Chameleon's check_dsolve (ChamSymmetric, NoTrans): returns
    (ok, ||B - A X||_1 / (||A||_1 ||X||_1 N eps))."""
    n = A0.shape[0]
    anorm = np.abs(A0).sum(axis=0).max()
    xnorm = np.abs(X).sum(axis=0).max()
    rnorm = np.abs(B - A0 @ X).sum(axis=0).max()
    res = rnorm / (anorm * xnorm * n * EPS)
    return (np.isfinite(xnorm) and np.isfinite(res) and res <= THRESHOLD), res


def rel_frobenius(X, Y):
    d = np.linalg.norm(Y)
    return np.linalg.norm(X - Y) / (d if d > 0 else 1.0)


# ------------------------------------------------------- conditioned SPD
# Restatement of the conditioned generator CHAMAMD_dplgsy_cond_Tile: the
# construction of the reference's chameleon_pzlatms (compute/pzlatms.c:
# 194-239: U <- random matrix, U <- qr(U), A <- U D U^T) with the mode-3
# spectrum of coreblas/compute/core_zlatm1.c:172-180.  Q comes from numpy's
# Householder QR here (LAPACK dgeqrf/dorgqr), an algorithm independent of
# the device's CholeskyQR3; both give the same Q up to column signs, which
# cancel in Q D Q^T.
_RND_A = np.uint64(6364136223846793005)
_RND_C = np.uint64(1)


def lcg_states(positions, seed):
    """LCG state after `positions[k]` steps from `seed` (coreblas/include/
    coreblas/random.h:32-52, binary powering of the affine map), vectorised."""
    n = np.asarray(positions, dtype=np.uint64).copy()
    x = np.full(n.shape, np.uint64(seed), dtype=np.uint64)
    a, c = _RND_A, _RND_C
    with np.errstate(over="ignore"):
        for _ in range(64):
            bit = (n & np.uint64(1)).astype(bool)
            x = np.where(bit, a * x + c, x)
            c = c * (a + np.uint64(1))
            a = a * a
            n = n >> np.uint64(1)
    return x


def gauss_matrix(n, seed):
    """N(0,1) matrix of the generator's ChamDistNormal rule: element (i, j)
    takes the LCG states x1, x2 after 2p and 2p + 1 steps, p = i + j n,
    u = ((x >> 11) + 1/2) 2^-53, G = sqrt(-2 ln u1) cos(2 pi u2)."""
    i, j = np.meshgrid(np.arange(n, dtype=np.uint64), np.arange(n, dtype=np.uint64), indexing="ij")
    p = i + j * np.uint64(n)
    x1 = lcg_states(2 * p, seed)
    with np.errstate(over="ignore"):
        x2 = _RND_A * x1 + _RND_C
    u1 = ((x1 >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = ((x2 >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    return np.asfortranarray(np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2))


def latm1_mode3(n, kappa):
    """D of core_zlatm1.c:172-180 (mode 3): D[0] = 1, D[i] = alpha^i,
    alpha = kappa^(-1/(n-1))."""
    if n <= 1:
        return np.ones(n)
    alpha = kappa ** (-1.0 / (n - 1))
    return np.array([alpha ** i for i in range(n)])


def latms(kappa, n, seed, dist=203):
    """A = Q diag(D) Q^T, Q = qr(G); G Gaussian (dist 203, ChamDistNormal) or
    Chameleon's dplrnt matrix with the seed advanced 2n states (dist 202,
    pzlatms.c:139,197)."""
    if dist == 203:
        G = gauss_matrix(n, seed)
    else:
        G = dplrnt(n, n, rnd64_jump(2 * n, seed))
    Q, _ = np.linalg.qr(G)
    D = latm1_mode3(n, kappa)
    A = (Q * D) @ Q.T
    A = 0.5 * (A + A.T)
    return np.asfortranarray(A), D


# ------------------------------------------------- per-tile precision rule
def tile_precision_map(A, nb, u_high=1e-12):
    """Precision of every lower tile (m >= n) of the symmetric matrix whose
    lower triangle is A, by the rule of compute/pzhered.c:176-288 and
    runtime/starpu/codelets/codelet_zgered.c:62-163: gnorm = ||A||_F,
    threshold = u_high gnorm / NT, threshold_diag = threshold while
    u_high < 2^-24 else 2^-24 gnorm / NT (pzhered.c:249); a tile goes to the
    half-width type if ||A_mn||_F < t / u16 (here BF16, u16 = 2^-8), else to
    FP32 if ||A_mn||_F < t / 2^-24, else stays FP64.  Returns a dict
    (m, n) -> 0 / 1 / 2 (FP64 / FP32 / BF16)."""
    n = A.shape[0]
    S = np.tril(A) + np.tril(A, -1).T
    nt = -(-n // nb)
    ss = {}
    for j in range(nt):
        for i in range(j, nt):
            ss[(i, j)] = float(np.sum(S[i * nb:(i + 1) * nb, j * nb:(j + 1) * nb] ** 2))
    gnorm = np.sqrt(sum((1.0 if i == j else 2.0) * v for (i, j), v in ss.items()))
    u32, u16 = 2.0 ** -24, 2.0 ** -8
    thr = u_high * gnorm / nt
    thr_diag = thr if u_high < u32 else u32 * gnorm / nt
    out = {}
    for (i, j), v in ss.items():
        t = thr_diag if i == j else thr
        ln = np.sqrt(v)
        out[(i, j)] = 2 if ln < t / u16 else (1 if ln < t / u32 else 0)
    return out
