"""CHAMAMD_dpotrs_Tile_Async_NoAlloc against CHAMELEON_dpotrs_Tile_Async.

The no-allocation entry point runs the same kernels in the same order on a
fixed device array instead of the library's workspaces, so its solution
must equal the allocating call's bit for bit.  Two calls in a row, the
second on a smaller system (the array keeps the first call's contents),
then the rejections: operands not resident on the device, and a problem
larger than the array.  Device free memory is the same before and after a
call (nothing allocated).
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

L_ = 122


@pytest.fixture(scope="module")
def ch():
    import chameleon_amd as ch
    ch.init()
    yield ch


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def hip_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def run(ch, fn):
    seq = ch.Sequence()
    rc = fn(seq)
    st = seq.wait()
    seq.destroy()
    return rc, st


def solve_both(ch, n, nrhs, nb):
    L = ch.lib
    ch.set_param(ch.TILE_SIZE, nb)
    with ch.Desc.alloc(n, n, nb) as dA, ch.Desc.alloc(n, nrhs, nb) as dB1, ch.Desc.alloc(n, nrhs, nb) as dB2:
        rc, st = run(ch, lambda q: (L.CHAMELEON_dplgsy_Tile_Async(float(n), L_, dA.ptr, 51, q.ptr, q.req) or
                                    L.CHAMELEON_dpotrf_Tile_Async(L_, dA.ptr, q.ptr, q.req) or
                                    L.CHAMELEON_dplrnt_Tile_Async(dB1.ptr, 5673, q.ptr, q.req) or
                                    L.CHAMELEON_dplrnt_Tile_Async(dB2.ptr, 5673, q.ptr, q.req)))
        assert rc == 0 and st == 0
        rc, st = run(ch, lambda q: L.CHAMELEON_dpotrs_Tile_Async(L_, dA.ptr, dB1.ptr, q.ptr, q.req))
        assert rc == 0 and st == 0
        free0 = hip_free_bytes()
        rc, st = run(ch, lambda q: L.CHAMAMD_dpotrs_Tile_Async_NoAlloc(L_, dA.ptr, dB2.ptr, q.ptr, q.req))
        assert rc == 0 and st == 0
        assert hip_free_bytes() == free0
        return dB1.to_array(n, nrhs), dB2.to_array(n, nrhs)


@pytest.mark.parametrize("n,nrhs,nb", [(4096, 64, 512), (1000, 7, 128), (3000, 1, 256)])
def test_noalloc_dpotrs_equals_dpotrs(ch, n, nrhs, nb):
    """Two calls in a row (this parametrisation runs the larger first)."""
    say("noalloc dpotrs", n, nrhs, nb)
    X1, X2 = solve_both(ch, n, nrhs, nb)
    assert np.array_equal(X1, X2)
    A = oracle.dplgsy(float(n), L_, n, 51, nb=nb)
    B = oracle.dplrnt(n, nrhs, 5673, nb=nb)
    assert oracle.dposv(L_, A, n, B, nrhs, nb) == 0
    assert oracle.rel_frobenius(X2, B) <= 1e-11


def test_noalloc_dpotrs_rejections(ch):
    L = ch.lib
    say("noalloc rejections")
    nb = 512
    ch.set_param(ch.TILE_SIZE, nb)
    # not resident: a fresh descriptor has no device copy
    with ch.Desc.alloc(1024, 1024, nb) as dA, ch.Desc.alloc(1024, 4, nb) as dB:
        rc, st = run(ch, lambda q: L.CHAMAMD_dpotrs_Tile_Async_NoAlloc(L_, dA.ptr, dB.ptr, q.ptr, q.req))
        assert rc == ch.ERR_ILLEGAL_VALUE and st == rc
    # larger than the array: 32768 * (192 + 512) doubles > 2^24
    n, nrhs = 32768, 500
    with ch.Desc.alloc(n, n, nb) as dA, ch.Desc.alloc(n, nrhs, nb) as dB:
        rc, st = run(ch, lambda q: (L.CHAMELEON_dplgsy_Tile_Async(float(n), L_, dA.ptr, 51, q.ptr, q.req) or
                                    L.CHAMELEON_dplrnt_Tile_Async(dB.ptr, 5673, q.ptr, q.req)))
        assert rc == 0 and st == 0
        rc, st = run(ch, lambda q: L.CHAMAMD_dpotrs_Tile_Async_NoAlloc(L_, dA.ptr, dB.ptr, q.ptr, q.req))
        assert rc == ch.ERR_OUT_OF_RESOURCES and st == rc
