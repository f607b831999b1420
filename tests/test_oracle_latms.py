"""Pinning of the oracle's conditioned generator (oracle.latms and its
Gaussian element rule) without a GPU.

* lcg_states (vectorised jump-ahead) equals plain sequential stepping of
  the LCG for small positions;
* gauss_matrix equals a scalar restatement of its element rule on sampled
  entries, and its entries have the moments of N(0,1);
* latms: A is symmetric, its eigenvalues are the mode-3 spectrum (closed
  form alpha^i), and the dplrnt-based variant uses Chameleon's generator
  with the seed advanced 2n states.
"""
import math

import numpy as np

import oracle

A_LCG = 6364136223846793005
MASK = (1 << 64) - 1


def step(x):
    return (A_LCG * x + 1) & MASK


def test_lcg_states_match_sequential_steps():
    seed = 987654321
    xs, x = [], seed
    for _ in range(300):
        xs.append(x)
        x = step(x)
    got = oracle.lcg_states(np.arange(300), seed)
    assert [int(v) for v in got] == xs


def test_gauss_element_rule_scalar_restatement():
    n, seed = 37, 4242
    G = oracle.gauss_matrix(n, seed)
    for (i, j) in [(0, 0), (5, 0), (0, 7), (36, 36), (12, 29)]:
        p = i + j * n
        x1 = int(oracle.rnd64_jump(2 * p, seed))
        x2 = step(x1)
        u1 = ((x1 >> 11) + 0.5) * 2.0 ** -53
        u2 = ((x2 >> 11) + 0.5) * 2.0 ** -53
        v = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
        assert G[i, j] == v


def test_gauss_moments():
    G = oracle.gauss_matrix(300, 7)
    assert abs(G.mean()) < 0.02
    assert abs(G.var() - 1.0) < 0.02
    assert abs(np.mean(G ** 4) - 3.0) < 0.1


def test_latm1_mode3_closed_form():
    D = oracle.latm1_mode3(11, 1e4)
    assert D[0] == 1.0 and abs(D[-1] - 1e-4) < 1e-16
    assert np.allclose(D[1:] / D[:-1], 1e4 ** (-0.1), rtol=1e-14, atol=0)


def test_latms_spectrum_and_symmetry():
    for dist in (203, 202):
        for kappa in (1e2, 1e6):
            A, D = oracle.latms(kappa, 160, 99, dist)
            assert np.array_equal(A, A.T)
            ev = np.sort(np.linalg.eigvalsh(A))
            assert np.max(np.abs(ev - np.sort(D))) < 1e-13
            assert abs(np.linalg.cond(A) / kappa - 1.0) < 1e-6


def test_latms_symmetric_dist_uses_shifted_dplrnt():
    n, seed = 24, 5
    G = oracle.dplrnt(n, n, oracle.rnd64_jump(2 * n, seed))
    Q, _ = np.linalg.qr(G)
    A, D = oracle.latms(10.0, n, seed, 202)
    assert np.allclose(Q.T @ A @ Q, np.diag(D), atol=1e-14)


def test_tile_precision_rule_closed_form():
    """Block matrix with known tile norms: diagonal tiles 4 I (norm 8 for
    nb = 4), off-diagonal tiles c J (all entries c, norm 4 c).  gnorm^2 =
    NT * 64 + NT (NT - 1) * 16 c^2; the choice follows from the thresholds."""
    nb, nt = 4, 3
    for c, u_high in [(1e-3, 1e-12), (1e-3, 1e-6), (1e-8, 1e-6), (1e-9, 1e-3)]:
        A = np.full((nb * nt, nb * nt), c)
        for t in range(nt):
            A[t * nb:(t + 1) * nb, t * nb:(t + 1) * nb] = 4.0 * np.eye(nb)
        got = oracle.tile_precision_map(A, nb, u_high)
        g = np.sqrt(nt * 64.0 + nt * (nt - 1) * 16.0 * c * c)
        thr = u_high * g / nt
        thr_d = thr if u_high < 2.0 ** -24 else 2.0 ** -24 * g / nt
        want_off = 2 if 4 * c < thr / 2.0 ** -8 else (1 if 4 * c < thr / 2.0 ** -24 else 0)
        want_diag = 2 if 8 < thr_d / 2.0 ** -8 else (1 if 8 < thr_d / 2.0 ** -24 else 0)
        for (i, j), v in got.items():
            assert v == (want_diag if i == j else want_off), (c, u_high, i, j, v)
    # the cases above reach all three precisions
    def block(c):
        A = np.full((12, 12), c)
        for t in range(3):
            A[t * 4:(t + 1) * 4, t * 4:(t + 1) * 4] = 4.0 * np.eye(4)
        return A
    assert set(oracle.tile_precision_map(block(1e-9), 4, 1e-3).values()) == {0, 2}
    assert set(oracle.tile_precision_map(block(1e-3), 4, 1e-6).values()) == {0, 1}
    assert set(oracle.tile_precision_map(block(1e-3), 4, 1e-12).values()) == {0}
