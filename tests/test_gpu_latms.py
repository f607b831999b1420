"""GPU tests of the conditioned SPD generator CHAMAMD_dplgsy_cond_Tile
(A = Q diag(lambda) Q^T, Q = qr(G), the construction of the reference's
compute/pzlatms.c:194-239 with the mode-3 spectrum of core_zlatm1.c).

Checked against what fixes A independently of the device algorithm:
* A is exactly symmetric;
* its eigenvalues are lambda_i = alpha^i (numpy eigvalsh against the
  closed form), so cond(A) = kappa;
* A equals the oracle's A built from numpy's Householder QR of the same G
  (the Gaussian element rule, or Chameleon's dplrnt with the seed advanced
  2n states), to the conditioning of Q as a function of G;
* a second kappa from the same seed reuses the kept orthogonal factor and
  still matches the oracle.
"""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

NORMAL, SYMMETRIC = 203, 202


@pytest.fixture(scope="module")
def ch():
    import chameleon_amd as ch
    ch.init()
    yield ch
    ch.lib.CHAMAMD_Cond_Cache_Release()


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def device_latms(ch, kappa, n, nb, seed, dist):
    with ch.Desc.alloc(n, n, nb) as dA:
        ch.check(ch.lib.CHAMAMD_dplgsy_cond_Tile(kappa, dist, dA.ptr, seed))
        return dA.to_array(n, n)


@pytest.mark.parametrize("dist", [NORMAL, SYMMETRIC])
@pytest.mark.parametrize("n,nb", [(512, 128), (1000, 100), (777, 256), (1024, 512)])
def test_cond_generator_matches_qr_construction(ch, dist, n, nb):
    seed = 1234
    for kappa in (1e2, 1e6):
        say("latms", dist, n, nb, kappa)
        A = device_latms(ch, kappa, n, nb, seed, dist)
        assert np.array_equal(A, A.T)
        lam = oracle.latm1_mode3(n, kappa)
        ev = np.sort(np.linalg.eigvalsh(A))
        assert np.max(np.abs(ev - np.sort(lam))) <= 1e-12
        Ar, _ = oracle.latms(kappa, n, seed, dist)
        err = oracle.rel_frobenius(A, Ar)
        assert err <= 1e-10, err


def test_cond_generator_seed_and_cache(ch):
    n, nb = 640, 128
    say("latms seeds")
    A1 = device_latms(ch, 1e4, n, nb, 7, NORMAL)
    A2 = device_latms(ch, 1e4, n, nb, 8, NORMAL)
    A1b = device_latms(ch, 1e4, n, nb, 7, NORMAL)
    assert not np.allclose(A1, A2)
    assert np.array_equal(A1, A1b)          # deterministic, cache or not
    ch.lib.CHAMAMD_Cond_Cache_Release()
    A1c = device_latms(ch, 1e4, n, nb, 7, NORMAL)
    assert np.array_equal(A1, A1c)
