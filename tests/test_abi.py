"""The C ABI without a GPU: the library loads and exports every function
that include/*.h declares; no compute call is made here."""
import ctypes
import os

import pytest


def test_library_exports_every_declared_function():
    import chameleon_amd as ch
    if not os.path.exists(ch.LIB_PATH):
        ch.build_library()
    lib = ctypes.CDLL(ch.LIB_PATH)
    names = ch.header_functions()
    assert len(names) > 50
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_calls_before_init_fail_cleanly():
    import chameleon_amd as ch
    L = ch.load()
    if L.CHAMELEON_Initialized():
        pytest.skip("already initialised in this process")
    assert L.CHAMELEON_Set(ch.TILE_SIZE, 128) == ch.ERR_NOT_INITIALIZED
    assert L.CHAMELEON_dposv(ch.LOWER, 4, 1, None, 4, None, 4) == ch.ERR_NOT_INITIALIZED
    assert L.CHAMELEON_dpotrf_Tile(ch.LOWER, None) == ch.ERR_NOT_INITIALIZED


def test_enum_values_match_reference_contract():
    import chameleon_amd as ch
    assert (ch.UPPER, ch.LOWER, ch.UPPERLOWER) == (121, 122, 123)
    assert ch.REAL_DOUBLE == 7 and ch.REAL_FLOAT == 6
    assert ch.TILE_SIZE == 2
