"""ctypes binding of the CPU oracle (oracle/fmm_oracle.c -> oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")


class OrcAlg(ctypes.Structure):
    _fields_ = [("mt", ctypes.c_int), ("kt", ctypes.c_int), ("nt", ctypes.c_int), ("R", ctypes.c_int),
                ("U", ctypes.POINTER(ctypes.c_int)), ("V", ctypes.POINTER(ctypes.c_int)),
                ("W", ctypes.POINTER(ctypes.c_int))]

    def tables(self):
        R = self.R
        def grab(p, parts):
            return np.array([[p[s * R + r] for r in range(R)] for s in range(parts)], dtype=np.int64)
        return (grab(self.U, self.mt * self.kt), grab(self.V, self.kt * self.nt),
                grab(self.W, self.mt * self.nt))


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_L = ctypes.c_long


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_parse_alg.argtypes = [ctypes.c_char_p, ctypes.POINTER(OrcAlg)]
        L.orc_free_alg.argtypes = [ctypes.POINTER(OrcAlg)]
        L.orc_compose.argtypes = [ctypes.POINTER(OrcAlg)] * 3
        L.orc_brent_violations.argtypes = [ctypes.POINTER(OrcAlg)]
        L.orc_brent_violations.restype = ctypes.c_long
        L.orc_part_geometry.argtypes = [_L, ctypes.c_int, ctypes.POINTER(_L), ctypes.POINTER(_L)]
        L.orc_dgemm.argtypes = [_L, _L, _L, ctypes.c_double, _dp, _L, _dp, _L, ctypes.c_double, _dp, _L]
        L.orc_dgemm_naive.argtypes = [_L, _L, _L, _dp, _L, _dp, _L, _dp, _L]
        L.orc_fmm_dgemm.argtypes = [ctypes.POINTER(OrcAlg), _L, _L, _L, ctypes.c_double, _dp, _L, _dp, _L,
                                    ctypes.c_double, _dp, _L]
        L.orc_fill_uniform.argtypes = [_dp, _L, _L, _L, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_fill_int.argtypes = [_dp, _L, _L, _L, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
        L.orc_fill_uniform_f32.argtypes = [ctypes.POINTER(ctypes.c_float), _L, _L, _L, ctypes.c_uint64,
                                           ctypes.c_uint64]
        L.orc_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    assert a.flags["F_CONTIGUOUS"] and a.dtype == np.float64
    return a.ctypes.data_as(_dp)


def parse(path):
    a = OrcAlg()
    rc = lib().orc_parse_alg(os.fsencode(path), ctypes.byref(a))
    if rc != 0:
        raise ValueError(f"oracle parse of {path} failed ({rc})")
    return a


def compose(a, b):
    out = OrcAlg()
    rc = lib().orc_compose(ctypes.byref(a), ctypes.byref(b), ctypes.byref(out))
    if rc != 0:
        raise MemoryError("compose")
    return out


def brent_violations(a):
    return lib().orc_brent_violations(ctypes.byref(a))


def part_geometry(whole, tilde):
    sz = (_L * tilde)()
    off = (_L * tilde)()
    lib().orc_part_geometry(whole, tilde, sz, off)
    return list(sz), list(off)


def uniform(m, n, seed, stream):
    X = np.empty((m, n), dtype=np.float64, order="F")
    if m and n:
        lib().orc_fill_uniform(_p(X), m, n, m, seed, stream)
    return X


def uniform_f32(m, n, seed, stream):
    X = np.empty((m, n), dtype=np.float32, order="F")
    if m and n:
        lib().orc_fill_uniform_f32(X.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), m, n, m, seed, stream)
    return X


def integers(m, n, seed, stream, amp):
    X = np.empty((m, n), dtype=np.float64, order="F")
    if m and n:
        lib().orc_fill_int(_p(X), m, n, m, seed, stream, amp)
    return X


def dgemm(alpha, A, B, beta, C):
    """C := beta*C + alpha*A*B in place (Fortran-ordered float64 arrays)."""
    m, k = A.shape
    n = B.shape[1]
    lib().orc_dgemm(m, n, k, alpha, _p(A), max(m, 1), _p(B), max(k, 1), beta, _p(C), max(m, 1))
    return C


def dgemm_naive(A, B, C):
    m, k = A.shape
    n = B.shape[1]
    lib().orc_dgemm_naive(m, n, k, _p(A), max(m, 1), _p(B), max(k, 1), _p(C), max(m, 1))
    return C


def fmm_dgemm(alg, alpha, A, B, beta, C):
    m, k = A.shape
    n = B.shape[1]
    rc = lib().orc_fmm_dgemm(ctypes.byref(alg), m, n, k, alpha, _p(A), max(m, 1), _p(B), max(k, 1), beta,
                             _p(C), max(m, 1))
    if rc != 0:
        raise MemoryError("orc_fmm_dgemm")
    return C


def threads():
    return lib().orc_threads()


def normalized_error(C, C_ref, A, B):
    """||C - C_ref||_F / (||A||_F * ||B||_F)."""
    return float(np.linalg.norm(C - C_ref) / (np.linalg.norm(A) * np.linalg.norm(B)))
