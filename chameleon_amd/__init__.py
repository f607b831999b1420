"""Python binding of libchameleon_amd.so, the MI355X-native Chameleon
Cholesky library.

The shared library exports the Chameleon C API (include/chameleon.h); this
module declares those entry points with ctypes and adds a thin helper layer
for numpy arrays:

    import chameleon_amd as ch
    ch.init()
    A = ch.Desc.from_array(a, nb=384)          # CM descriptor over a Fortran array
    ch.check(ch.lib.CHAMELEON_dposv_Tile(ch.LOWER, A.ptr, B.ptr))

There is no fallback: if the library is not built, or no GPU is visible,
every call raises.
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libchameleon_amd.so")
INCLUDE = os.path.join(ROOT, "include")

# enum values (include/chameleon/constants.h)
UPPER, LOWER, UPPERLOWER = 121, 122, 123
NOTRANS, TRANS = 111, 112
ONE_NORM, FROBENIUS_NORM, INF_NORM, MAX_NORM = 171, 174, 175, 177
REAL_DOUBLE = (1 << 2) | 3
REAL_FLOAT = (1 << 2) | 2
MAT_ALLOC_GLOBAL = None
MAT_ALLOC_TILE = ctypes.c_void_p(-1 & ((1 << 64) - 1))
TILE_SIZE, LOOKAHEAD = 2, 8
SUCCESS = 0
ERR_NOT_INITIALIZED = -101
ERR_ILLEGAL_VALUE = -104
ERR_OUT_OF_RESOURCES = -106
COL_MAJOR = 102
DIST_UNIFORM, DIST_SYMMETRIC, DIST_NORMAL = 201, 202, 203


class ChameleonError(RuntimeError):
    pass


def build_library():
    """Compile libchameleon_amd.so (see build.py)."""
    import importlib
    return importlib.import_module(__name__ + ".build").build()


_lib = None


class ServerRequest(ctypes.Structure):
    """CHAMAMD_request_t (include/chameleon_amd.h)."""
    _fields_ = [("matrix_id", ctypes.c_longlong), ("n", ctypes.c_int), ("nrhs", ctypes.c_int),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_int), ("seed", ctypes.c_ulonglong),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    """CHAMAMD_server_stats_t (include/chameleon_amd.h)."""
    _fields_ = [("completed", ctypes.c_longlong), ("rejected", ctypes.c_longlong), ("factored", ctypes.c_longlong),
                ("reused", ctypes.c_longlong), ("resident", ctypes.c_longlong),
                ("pool_used_bytes", ctypes.c_longlong), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("flops", ctypes.c_double), ("mixed_solves", ctypes.c_longlong),
                ("ir_steps", ctypes.c_longlong), ("fallbacks", ctypes.c_longlong), ("max_jobs", ctypes.c_longlong),
                ("max_berr", ctypes.c_double)]


def _declare(L):
    i32, dbl, vp, u64 = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_ulonglong
    pp = ctypes.POINTER(ctypes.c_void_p)
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    sig = {
        "__chameleon_init": (i32, [i32, i32]),
        "__chameleon_finalize": (i32, []),
        "CHAMELEON_Initialized": (i32, []),
        "CHAMELEON_Set": (i32, [i32, i32]),
        "CHAMELEON_Get": (i32, [i32, ctypes.POINTER(i32)]),
        "CHAMELEON_Desc_Create": (i32, [pp, vp, i32, i32, i32, i32, i32, i32, i32, i32, i32, i32, i32, i32]),
        "CHAMELEON_Desc_Destroy": (i32, [pp]),
        "CHAMELEON_Desc_Flush": (i32, [vp, vp]),
        "CHAMELEON_Desc_Acquire": (i32, [vp]),
        "CHAMELEON_Desc_Release": (i32, [vp]),
        "CHAMELEON_Lap2Desc": (i32, [i32, vp, i32, vp]),
        "CHAMELEON_Desc2Lap": (i32, [i32, vp, vp, i32]),
        "CHAMELEON_Sequence_Create": (i32, [pp]),
        "CHAMELEON_Sequence_Destroy": (i32, [vp]),
        "CHAMELEON_Sequence_Wait": (i32, [vp]),
        "CHAMELEON_dposv": (i32, [i32, i32, i32, vp, i32, vp, i32]),
        "CHAMELEON_dpotrf": (i32, [i32, i32, vp, i32]),
        "CHAMELEON_dpotrs": (i32, [i32, i32, i32, vp, i32, vp, i32]),
        "CHAMELEON_dplgsy": (i32, [dbl, i32, i32, vp, i32, u64]),
        "CHAMELEON_dplrnt": (i32, [i32, i32, vp, i32, u64]),
        "CHAMELEON_dlange": (dbl, [i32, i32, i32, vp, i32]),
        "CHAMELEON_dlansy": (dbl, [i32, i32, i32, vp, i32]),
        "CHAMELEON_dposv_Tile": (i32, [i32, vp, vp]),
        "CHAMELEON_dpotrf_Tile": (i32, [i32, vp]),
        "CHAMELEON_dpotrs_Tile": (i32, [i32, vp, vp]),
        "CHAMELEON_dplgsy_Tile": (i32, [dbl, i32, vp, u64]),
        "CHAMELEON_dplrnt_Tile": (i32, [vp, u64]),
        "CHAMELEON_dlange_Tile": (dbl, [i32, vp]),
        "CHAMELEON_dlansy_Tile": (dbl, [i32, i32, vp]),
        "CHAMELEON_dlacpy_Tile": (i32, [i32, vp, vp]),
        "CHAMELEON_dlacpy_Tile_Async": (i32, [i32, vp, vp, vp, vp]),
        "CHAMELEON_dposv_Tile_Async": (i32, [i32, vp, vp, vp, vp]),
        "CHAMELEON_dpotrf_Tile_Async": (i32, [i32, vp, vp, vp]),
        "CHAMELEON_dpotrs_Tile_Async": (i32, [i32, vp, vp, vp, vp]),
        "CHAMAMD_dpotrs_Tile_Async_NoAlloc": (i32, [i32, vp, vp, vp, vp]),
        "CHAMELEON_dplgsy_Tile_Async": (i32, [dbl, i32, vp, u64, vp, vp]),
        "CHAMELEON_dplrnt_Tile_Async": (i32, [vp, u64, vp, vp]),
        "CHAMELEON_lapacke_dposv": (i32, [i32, ctypes.c_char, i32, i32, vp, i32, vp, i32]),
        "CHAMELEON_lapacke_dpotrf": (i32, [i32, ctypes.c_char, i32, vp, i32]),
        "CHAMELEON_lapacke_dpotrs": (i32, [i32, ctypes.c_char, i32, i32, vp, i32, vp, i32]),
        "chameleon_lapack_dposv_": (None, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "chameleon_lapack_dpotrf_": (None, [vp, vp, vp, vp, vp]),
        "chameleon_lapack_dpotrs_": (None, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "CHAMELEON_dsposv": (i32, [i32, i32, i32, vp, i32, vp, i32, vp, i32, ip]),
        "CHAMELEON_dsposv_Tile": (i32, [i32, vp, vp, vp, ip]),
        "CHAMAMD_dposv_mixed_Tile": (i32, [i32, vp, vp, vp, i32, ip, dp]),
        "CHAMAMD_dposv_mixed_Tile_Async": (i32, [i32, vp, vp, vp, i32, ip, dp, vp, vp]),
        "CHAMAMD_dplgsy_cond_Tile": (i32, [dbl, i32, vp, u64]),
        "CHAMAMD_Cond_Cache_Release": (i32, []),
        "CHAMAMD_Set_Refinement_Target": (i32, [dbl]),
        "CHAMAMD_Set_Mixed_Scaling": (i32, [i32]),
        "CHAMAMD_Set_Tile_Accuracy": (i32, [dbl]),
        "CHAMAMD_Mixed_Tile_Counts": (i32, [ctypes.POINTER(ctypes.c_longlong)]),
        "CHAMAMD_Backward_Error": (i32, [i32, vp, vp, vp, dp]),
        "CHAMAMD_Kernel_Stats": (i32, [i32, ctypes.POINTER(ctypes.c_longlong), dp, dp]),
        "CHAMAMD_Kernel_Stats_Reset": (i32, []),
        "CHAMELEON_dplgsy_batch_Tile": (i32, [dbl, vp, u64]),
        "CHAMELEON_dpotrf_batch_Tile": (i32, [i32, vp]),
        "CHAMAMD_Batch_Create": (i32, [pp, i32, ip, ip]),
        "CHAMAMD_Batch_Destroy": (i32, [pp]),
        "CHAMAMD_Batch_SetA": (i32, [vp, i32, i32, vp, i32]),
        "CHAMAMD_Batch_GetA": (i32, [vp, i32, i32, vp, i32]),
        "CHAMAMD_Batch_SetB": (i32, [vp, i32, vp, i32]),
        "CHAMAMD_Batch_GetB": (i32, [vp, i32, vp, i32]),
        "CHAMAMD_dplgsy_batch": (i32, [vp, u64]),
        "CHAMAMD_dplrnt_batch": (i32, [vp, u64]),
        "CHAMAMD_dpotrf_batch": (i32, [vp, ip]),
        "CHAMAMD_dpotrs_batch": (i32, [vp]),
        "CHAMAMD_dposv_batch": (i32, [vp, ip]),
        "CHAMAMD_dpotrf_batch_Async": (i32, [vp, vp, vp]),
        "CHAMAMD_dpotrs_batch_Async": (i32, [vp, vp, vp]),
        "CHAMAMD_dposv_batch_Async": (i32, [vp, vp, vp]),
        "CHAMAMD_Batch_Info": (i32, [vp, ip]),
        "CHAMAMD_Batch_Tune": (i32, [vp, ip]),
        "CHAMAMD_dposv_mixed_batch": (i32, [vp, i32, ip, ip]),
        "CHAMAMD_Batch_Get_Tuning": (i32, [ip, ip, ip]),
        "CHAMAMD_Batch_Set_Tuning": (i32, [ip]),
        "CHAMAMD_Set_Batch_Graphs": (i32, [i32]),
        "CHAMAMD_Server_Create": (i32, [pp, ctypes.c_size_t]),
        "CHAMAMD_Server_Destroy": (i32, [pp]),
        "CHAMAMD_Server_Submit": (i32, [vp, ctypes.POINTER(ServerRequest), ctypes.POINTER(ctypes.c_longlong)]),
        "CHAMAMD_Server_Progress": (i32, [vp, ip]),
        "CHAMAMD_Server_Drain": (i32, [vp]),
        "CHAMAMD_Server_Finished": (i32, [vp, ctypes.POINTER(ctypes.c_longlong), i32, ip]),
        "CHAMAMD_Server_Set_Precision": (i32, [vp, i32]),
        "CHAMAMD_Server_Test": (i32, [vp, ctypes.c_longlong, ip, ip]),
        "CHAMAMD_Server_Stats": (i32, [vp, ctypes.POINTER(ServerStats)]),
        "CHAMAMD_Server_Reset_Stats": (i32, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def load():
    """Load libchameleon_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ChameleonError(f"{LIB_PATH} is missing: run chameleon_amd/build.py")
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        _declare(L)
        _lib = L
    return _lib


class _LibProxy:
    def __getattr__(self, name):
        return getattr(load(), name)


lib = _LibProxy()


def check(rc, what="call"):
    if rc != SUCCESS:
        raise ChameleonError(f"{what} returned {rc}")
    return rc


def init(ncpus=1, ngpus=1):
    L = load()
    if L.CHAMELEON_Initialized():
        return
    check(L.__chameleon_init(ncpus, ngpus), "CHAMELEON_Init")


def finalize():
    L = load()
    if L.CHAMELEON_Initialized():
        check(L.__chameleon_finalize(), "CHAMELEON_Finalize")


def set_param(param, value):
    check(load().CHAMELEON_Set(param, value), "CHAMELEON_Set")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Desc:
    """Owning wrapper of a CHAM_desc_t*.  `array` keeps the user's host
    matrix alive for CM descriptors."""

    def __init__(self, ptr, array=None):
        self.ptr = ptr
        self.array = array

    @classmethod
    def from_array(cls, a, nb, m=None, n=None):
        """CM descriptor over the Fortran-ordered float64 array `a`
        (lm = a.shape[0] is the leading dimension)."""
        assert a.dtype == np.float64 and a.flags.f_contiguous
        lm, ln = a.shape
        m = lm if m is None else m
        n = ln if n is None else n
        p = ctypes.c_void_p()
        check(load().CHAMELEON_Desc_Create(ctypes.byref(p), _ptr(a), REAL_DOUBLE, nb, nb, nb * nb, lm, ln, 0, 0, m, n,
                                           1, 1), "CHAMELEON_Desc_Create")
        return cls(p, a)

    @classmethod
    def alloc(cls, m, n, nb, dtype=REAL_DOUBLE):
        """Library-allocated (CCRB) descriptor."""
        p = ctypes.c_void_p()
        check(load().CHAMELEON_Desc_Create(ctypes.byref(p), MAT_ALLOC_TILE, dtype, nb, nb, nb * nb, m, n, 0, 0, m, n,
                                           1, 1), "CHAMELEON_Desc_Create")
        return cls(p, None)

    def to_array(self, m, n, uplo=UPPERLOWER, out=None):
        a = out if out is not None else np.zeros((m, n), order="F")
        check(load().CHAMELEON_Desc2Lap(uplo, self.ptr, _ptr(a), a.shape[0]), "CHAMELEON_Desc2Lap")
        return a

    def destroy(self):
        if self.ptr:
            load().CHAMELEON_Desc_Destroy(ctypes.byref(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


class Sequence:
    def __init__(self):
        self.ptr = ctypes.c_void_p()
        check(load().CHAMELEON_Sequence_Create(ctypes.byref(self.ptr)), "CHAMELEON_Sequence_Create")
        # RUNTIME_request_t { int status, flush, priority; void *parent }
        self.request = (ctypes.c_char * 24)()
        ctypes.memset(self.request, 0, 24)
        ctypes.c_int.from_buffer(self.request, 4).value = 1

    @property
    def req(self):
        return ctypes.cast(self.request, ctypes.c_void_p)

    def wait(self):
        return load().CHAMELEON_Sequence_Wait(self.ptr)

    def destroy(self):
        if self.ptr:
            load().CHAMELEON_Sequence_Destroy(self.ptr)
            self.ptr = None


def header_functions():
    """Names of every function declared in include/ (the exported C ABI)."""
    names = set()
    for dirpath, _, files in os.walk(INCLUDE):
        for f in files:
            if not f.endswith(".h"):
                continue
            text = open(os.path.join(dirpath, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            text = re.sub(r"static inline[^{]*\{[^}]*\}", "", text, flags=re.S)
            for m in re.finditer(r"^\s*[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
                name = m.group(1)
                if name in ("if", "while", "for", "return", "sizeof", "defined", "typedef"):
                    continue
                line = text[m.start():text.find("\n", m.start())]
                if line.lstrip().startswith(("#", "typedef")):
                    continue
                names.add(name)
    return names


class Server:
    """Owning wrapper of a CHAMAMD_server_t*.  Keeps the right-hand-side
    arrays of submitted requests alive until they are served."""

    def __init__(self, budget_bytes):
        self.ptr = ctypes.c_void_p()
        check(load().CHAMAMD_Server_Create(ctypes.byref(self.ptr), int(budget_bytes)), "CHAMAMD_Server_Create")
        self.keep = {}

    def submit(self, matrix_id, n, B, seed=0, A=None):
        """B: float64 column-major array or column block of one (n x nrhs,
        leading dimension from its strides), overwritten by the solution."""
        assert B.dtype == np.float64 and B.strides[0] == 8 and B.shape[0] >= n
        ldb = B.strides[1] // 8 if B.shape[1] > 1 else max(B.shape[0], 1)
        r = ServerRequest(matrix_id, n, B.shape[1], A.ctypes.data if A is not None else None,
                          A.shape[0] if A is not None else 0, seed, B.ctypes.data, ldb)
        t = ctypes.c_longlong()
        check(load().CHAMAMD_Server_Submit(self.ptr, ctypes.byref(r), ctypes.byref(t)), "CHAMAMD_Server_Submit")
        self.keep[t.value] = (A, B)
        return t.value

    def progress(self):
        """Non-blocking step; returns the number of requests finished."""
        n = ctypes.c_int()
        check(load().CHAMAMD_Server_Progress(self.ptr, ctypes.byref(n)), "CHAMAMD_Server_Progress")
        return n.value

    def finished(self, cap=4096):
        """Tickets of the requests finished since the last call."""
        buf = (ctypes.c_longlong * cap)()
        out = []
        while True:
            n = ctypes.c_int()
            check(load().CHAMAMD_Server_Finished(self.ptr, buf, cap, ctypes.byref(n)), "CHAMAMD_Server_Finished")
            out.extend(buf[:n.value])
            if n.value < cap:
                for t in out:
                    self.keep.pop(t, None)
                return out

    def drain(self):
        """Serve until every submitted request has finished."""
        check(load().CHAMAMD_Server_Drain(self.ptr), "CHAMAMD_Server_Drain")

    def set_precision(self, prec):
        check(load().CHAMAMD_Server_Set_Precision(self.ptr, prec), "CHAMAMD_Server_Set_Precision")

    def test(self, ticket):
        d, st = ctypes.c_int(), ctypes.c_int()
        check(load().CHAMAMD_Server_Test(self.ptr, ticket, ctypes.byref(d), ctypes.byref(st)))
        if d.value:
            self.keep.pop(ticket, None)
        return bool(d.value), st.value

    def stats(self):
        st = ServerStats()
        check(load().CHAMAMD_Server_Stats(self.ptr, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def reset_stats(self):
        check(load().CHAMAMD_Server_Reset_Stats(self.ptr))

    def destroy(self):
        if self.ptr:
            load().CHAMAMD_Server_Destroy(ctypes.byref(self.ptr))
            self.ptr = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()
