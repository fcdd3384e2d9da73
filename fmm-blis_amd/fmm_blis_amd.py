"""ctypes binding of libfmm_blis_amd.so (include/fmm_blis_amd.h).

This is the Python face of the C ABI: the fmm_blis drop-in (fmm_t, obj_t,
new_fmm, bli_strassen_ab_ex, ...) and the native fmmhip_* layer (device
buffers, streams, events, seeded fills, the FMM and classical GEMMs).  Tests
and bench.py call the library only through here.  Loading fails loudly when
the shared library has not been built; there is no fallback path.

Reference interface mirrored (dnparikh/fmm-blis): fmm_t
plugin/bli_plugin_fmm_blis.h:58-68, new_fmm/free_fmm
frame/base/bli_read_fmm.c:24-86, bli_strassen_ab_ex/bli_strassen_ab
frame/base/bli_strassen_ab.c:123-300, bli_plugin_register_fmm_blis
plugin/bli_plugin_register.c:43-64.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfmm_blis_amd.so")

# Datatype codes of obj_t.info bits 0-2 (blis/frame/include/bli_type_defs.h:378-381)
BLIS_FLOAT = 0
BLIS_DOUBLE = 2
BLIS_CONSTANT = 5
BLIS_SUCCESS = -1
BLIS_FAILURE = -2
TRANS_BIT = 1 << 3

LEAF_NATIVE, LEAF_F32, LEAF_BF16X3, LEAF_BF16X3_6, LEAF_BF16X3_8, LEAF_INT8X4 = 0, 1, 2, 3, 4, 5

STATUS = {0: "success", 1: "invalid argument", 2: "out of memory", 3: "HIP runtime error",
          4: "unsupported", 5: "no gfx950 device", 6: "coefficient file unreadable or malformed",
          7: "problem exceeds the no-allocation scratch"}


class FmmT(ctypes.Structure):
    """fmm_t: <m~, k~, n~; R> with int tables U, V, W (row-major [part][r])."""
    _fields_ = [("m_tilde", ctypes.c_int), ("n_tilde", ctypes.c_int), ("k_tilde", ctypes.c_int),
                ("R", ctypes.c_int), ("U", ctypes.POINTER(ctypes.c_int)),
                ("V", ctypes.POINTER(ctypes.c_int)), ("W", ctypes.POINTER(ctypes.c_int))]

    def tables(self):
        """(U, V, W) as nested Python lists [part][r]."""
        mt, kt, nt, R = self.m_tilde, self.k_tilde, self.n_tilde, self.R
        def grab(p, parts):
            return [[p[s * R + r] for r in range(R)] for s in range(parts)]
        return grab(self.U, mt * kt), grab(self.V, kt * nt), grab(self.W, mt * nt)


class _Atom(ctypes.Structure):
    _fields_ = [("real", ctypes.c_double), ("imag", ctypes.c_double)]


class ObjT(ctypes.Structure):
    """Layout-identical mirror of BLIS obj_t (bli_type_defs.h:1215-1245)."""
    pass


ObjT._fields_ = [("root", ctypes.POINTER(ObjT)), ("off", ctypes.c_int64 * 2), ("dim", ctypes.c_int64 * 2),
                 ("diag_off", ctypes.c_int64), ("info", ctypes.c_uint32), ("info2", ctypes.c_uint32),
                 ("elem_size", ctypes.c_uint64), ("buffer", ctypes.c_void_p), ("rs", ctypes.c_int64),
                 ("cs", ctypes.c_int64), ("is_", ctypes.c_int64), ("scalar", _Atom),
                 ("m_padded", ctypes.c_int64), ("n_padded", ctypes.c_int64), ("ps", ctypes.c_int64),
                 ("pd", ctypes.c_int64), ("m_panel", ctypes.c_int64), ("n_panel", ctypes.c_int64)]


class ConstData(ctypes.Structure):
    """Mirror of BLIS constdata_t (blis/frame/include/bli_type_defs.h:1200-1208),
    the buffer of a BLIS_CONSTANT object such as BLIS_ONE: one value per
    datatype, read by type (bli_obj_macro_defs.h:1286-1297)."""
    _fields_ = [("s", ctypes.c_float), ("d", ctypes.c_double), ("c", ctypes.c_float * 2),
                ("z", ctypes.c_double * 2), ("i", ctypes.c_int64)]


def obj_constant(value_d, value_s=None):
    """A BLIS_CONSTANT scalar object; returns (obj, storage) -- keep storage alive."""
    cd = ConstData()
    cd.s = value_d if value_s is None else value_s
    cd.d = value_d
    cd.c[0] = cd.s
    cd.z[0] = value_d
    cd.i = int(value_d)
    o = ObjT()
    lib().fmm_obj_attach(BLIS_CONSTANT, 1, 1, ctypes.cast(ctypes.pointer(cd), _P), 1, 1, ctypes.byref(o))
    return o, cd


class Request(ctypes.Structure):
    """fmmhip_request_t."""
    _fields_ = [("is_double", ctypes.c_int), ("m", ctypes.c_long), ("n", ctypes.c_long), ("k", ctypes.c_long),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("A", ctypes.c_void_p), ("lda", ctypes.c_long),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_long), ("C", ctypes.c_void_p), ("ldc", ctypes.c_long),
                ("cand", ctypes.c_int), ("keep", ctypes.c_int), ("arrival_s", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    """fmmhip_server_stats_t."""
    _fields_ = [("submitted", ctypes.c_long), ("completed", ctypes.c_long), ("failed", ctypes.c_long),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("span_s", ctypes.c_double), ("arena_bytes", ctypes.c_size_t),
                ("arena_peak", ctypes.c_size_t), ("max_inflight", ctypes.c_int)]


class FmmError(RuntimeError):
    pass


_lib = None

_P = ctypes.c_void_p
_L = ctypes.c_long
_D = ctypes.c_double
_F = ctypes.c_float
_I = ctypes.c_int
_U64 = ctypes.c_uint64

_SIGS = {
    # drop-in
    "bli_plugin_register_fmm_blis": (_I, []),
    "new_fmm": (FmmT, [ctypes.c_char_p]),
    "free_fmm": (None, [FmmT]),
    "bli_strassen_ab_ex": (None, [ctypes.POINTER(ObjT)] * 5 + [FmmT]),
    "bli_strassen_ab": (None, [ctypes.POINTER(ObjT)] * 5),
    "bli_fmm_ab_ex": (None, [ctypes.POINTER(ObjT)] * 5 + [ctypes.POINTER(FmmT), _I, _I]),
    "bli_gemm_hip": (None, [ctypes.POINTER(ObjT)] * 5),
    "fmm_blis_amd_set_abort_on_error": (None, [_I]),
    "fmm_blis_amd_last_error": (_I, []),
    "fmm_obj_attach": (None, [ctypes.c_uint32, ctypes.c_int64, ctypes.c_int64, _P, ctypes.c_int64,
                              ctypes.c_int64, ctypes.POINTER(ObjT)]),
    "fmm_obj_scalar": (None, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ObjT)]),
    "fmm_blis_amd_alg_cache_stats": (None, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "bli_plugin_init_fmm_blis_gfx950_ref": (None, []),
    "fmm_blis_get_ukr_dt": (_P, [ctypes.c_uint32, ctypes.c_uint64]),
    # native
    "fmmhip_status_string": (ctypes.c_char_p, [_I]),
    "fmmhip_init": (_I, []),
    "fmmhip_device_arch": (_I, [ctypes.c_char_p, ctypes.c_size_t]),
    "fmmhip_device_cu_count": (_I, [ctypes.POINTER(_I)]),
    "fmmhip_alg_create": (_I, [ctypes.POINTER(_P), ctypes.POINTER(FmmT), _I]),
    "fmmhip_alg_create_classical": (_I, [ctypes.POINTER(_P)]),
    "fmmhip_alg_destroy": (None, [_P]),
    "fmmhip_alg_info": (_I, [_P] + [ctypes.POINTER(_I)] * 4 + [ctypes.POINTER(_L)] * 3),
    "fmmhip_alg_brent_violations": (_L, [_P]),
    "fmmhip_alg_set_shape": (_I, [_P, _I]),
    "fmmhip_alg_set_pack": (_I, [_P, _I]),
    "fmmhip_alg_pack_for": (_I, [_P, _I, _L, _L, _L, ctypes.POINTER(_I)]),
    "fmmhip_alg_last_pack": (_I, [_P, ctypes.POINTER(_I)]),
    "fmmhip_alg_last_shape": (_I, [_P, ctypes.POINTER(_I)]),
    "fmmhip_alg_set_split": (_I, [_P, _I]),
    "fmmhip_workspace_release": (_I, []),
    "fmmhip_timing_enable": (_I, [_I]),
    "fmmhip_timing_read": (_I, [ctypes.POINTER(_L)] + [ctypes.POINTER(_D)] * 3),
    "fmmhip_memcpy_d2d": (_I, [_P, _P, ctypes.c_size_t, _P]),
    "fmmhip_copy_kernel": (_I, [_P, _P, ctypes.c_size_t, _P]),
    "fmmhip_measure_peaks": (_I, [ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "fmmhip_measure_hbm": (_I, [ctypes.POINTER(_D), ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "fmmhip_dgemm_fmm": (_I, [_P, _I, _L, _L, _L, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _P]),
    "fmmhip_sgemm_fmm": (_I, [_P, _L, _L, _L, _F, _P, _L, _L, _P, _L, _L, _F, _P, _L, _L, _P]),
    "fmmhip_dgemm": (_I, [_L, _L, _L, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _P]),
    "fmmhip_dgemm_fmm_noalloc": (_I, [_P, _L, _L, _L, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _P]),
    "fmmhip_sgemm": (_I, [_L, _L, _L, _F, _P, _L, _L, _P, _L, _L, _F, _P, _L, _L, _P]),
    "fmmhip_fmm_workspace_bytes": (_I, [_P, _I, _L, _L, _L, _L, _L, _L, _L, _L, _L,
                                        ctypes.POINTER(ctypes.c_size_t)]),
    "fmmhip_dpack_fmm": (_I, [_P, _L, _L, _L, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _P, ctypes.c_size_t, _P]),
    "fmmhip_dgemm_fmm_packed": (_I, [_P, _L, _L, _L, _D, _P, _L, _L, _P, _L, _L, _D, _P, _L, _L, _P,
                                     ctypes.c_size_t, _P]),
    "fmmhip_spack_fmm": (_I, [_P, _L, _L, _L, _F, _P, _L, _L, _P, _L, _L, _F, _P, _L, _L, _P, ctypes.c_size_t, _P]),
    "fmmhip_sgemm_fmm_packed": (_I, [_P, _L, _L, _L, _F, _P, _L, _L, _P, _L, _L, _F, _P, _L, _L, _P,
                                     ctypes.c_size_t, _P]),
    "fmmhip_batch_create": (_I, [ctypes.POINTER(_P)]),
    "fmmhip_batch_destroy": (None, [_P]),
    "fmmhip_batch_add_d": (_I, [_P, _L, _L, _L, _D, _P, _L, _P, _L, _D, _P, _L, _I]),
    "fmmhip_batch_add_s": (_I, [_P, _L, _L, _L, _F, _P, _L, _P, _L, _F, _P, _L, _I]),
    "fmmhip_batch_plan": (_I, [_P]),
    "fmmhip_batch_run": (_I, [_P, _P]),
    "fmmhip_batch_info": (_I, [_P, ctypes.POINTER(_L), ctypes.POINTER(_L), ctypes.POINTER(_L),
                               ctypes.POINTER(_D), ctypes.POINTER(ctypes.c_size_t)]),
    "fmmhip_batch_choice": (_I, [_P, _L, ctypes.POINTER(_I)]),
    "fmmhip_batch_set_tile": (_I, [_P, _L, _I]),
    "fmmhip_batch_tile": (_I, [_P, _L, ctypes.POINTER(_I)]),
    "fmmhip_batch_fast_items": (_I, [_P, ctypes.POINTER(_L), ctypes.POINTER(_L)]),
    "fmmhip_tuner_choose_tile": (_I, [_I, _L, _L, _L, _I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                      ctypes.POINTER(_D)]),
    "fmmhip_tuner_choose": (_I, [_I, _L, _L, _L, _I, ctypes.POINTER(_I), ctypes.POINTER(_D)]),
    "fmmhip_tuner_estimate": (_I, [_I, _I, _L, _L, _L, _I, ctypes.POINTER(_D)]),
    "fmmhip_tuner_calibrate": (_I, [_P]),
    "fmmhip_tuner_save": (_I, [ctypes.c_char_p]),
    "fmmhip_tuner_load": (_I, [ctypes.c_char_p]),
    "fmmhip_server_create": (_I, [ctypes.POINTER(_P), ctypes.c_size_t, _I]),
    "fmmhip_server_now": (_D, [_P]),
    "fmmhip_server_submit": (_I, [_P, ctypes.POINTER(Request), ctypes.POINTER(_L)]),
    "fmmhip_server_wait": (_I, [_P, _L, ctypes.POINTER(_I)]),
    "fmmhip_server_result": (_I, [_P, _L, ctypes.POINTER(_P), ctypes.POINTER(_L), ctypes.POINTER(_D),
                                  ctypes.POINTER(_I)]),
    "fmmhip_server_release": (_I, [_P, _L]),
    "fmmhip_server_drain": (_I, [_P]),
    "fmmhip_server_stats": (_I, [_P, ctypes.POINTER(ServerStats)]),
    "fmmhip_server_destroy": (None, [_P]),
    "fmmhip_malloc": (_I, [ctypes.POINTER(_P), ctypes.c_size_t]),
    "fmmhip_free": (_I, [_P]),
    "fmmhip_memcpy_h2d": (_I, [_P, _P, ctypes.c_size_t]),
    "fmmhip_memcpy_d2h": (_I, [_P, _P, ctypes.c_size_t]),
    "fmmhip_memset": (_I, [_P, _I, ctypes.c_size_t]),
    "fmmhip_stream_create": (_I, [ctypes.POINTER(_P)]),
    "fmmhip_stream_destroy": (_I, [_P]),
    "fmmhip_stream_sync": (_I, [_P]),
    "fmmhip_device_sync": (_I, []),
    "fmmhip_event_create": (_I, [ctypes.POINTER(_P)]),
    "fmmhip_event_destroy": (_I, [_P]),
    "fmmhip_event_record": (_I, [_P, _P]),
    "fmmhip_event_elapsed_ms": (_I, [ctypes.POINTER(_F), _P, _P]),
    "fmmhip_mem_info": (_I, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "fmmhip_fill_uniform": (_I, [_P, _L, _L, _L, _U64, _U64, _P]),
    "fmmhip_fill_uniform_f32": (_I, [_P, _L, _L, _L, _U64, _U64, _P]),
    "fmmhip_fill_int": (_I, [_P, _L, _L, _L, _U64, _U64, _I, _P]),
    "fmmhip_scale_rows_cols": (_I, [_P, _L, _L, _L, _P, _P, _P]),
    "fmmhip_fro_diff": (_I, [_P, _P, _L, _L, _L, _L, ctypes.POINTER(_D)]),
    "fmmhip_fro": (_I, [_P, _L, _L, _L, ctypes.POINTER(_D)]),
}

EXPORTED = sorted(_SIGS) + ["STRASSEN_FMM", "FMM_BLIS_PACK_UKR", "FMM_BLIS_GEMM_UKR"]


def lib():
    """Loads the shared library (once).  Raises FmmError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FmmError(f"{LIB_PATH} is not built: run __graft_entry__.build() (make -C fmm-blis_amd)")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = lib().fmmhip_status_string(rc).decode()
        raise FmmError(f"{what}: {msg} ({rc})")


def strassen():
    """The library's built-in STRASSEN_FMM (frame/base/bli_strassen_ab.c:12-24)."""
    return FmmT.in_dll(lib(), "STRASSEN_FMM")


def new_fmm(path):
    return lib().new_fmm(os.fsencode(path))


def free_fmm(f):
    lib().free_fmm(f)


class Alg:
    """Composed L-level algorithm (fmmhip_alg_t)."""

    def __init__(self, levels=None, classical=False):
        self.h = _P()
        if classical:
            check(lib().fmmhip_alg_create_classical(ctypes.byref(self.h)), "alg_create_classical")
        else:
            arr = (FmmT * len(levels))(*levels)
            check(lib().fmmhip_alg_create(ctypes.byref(self.h), arr, len(levels)), "alg_create")

    def info(self):
        mt, kt, nt, R = _I(), _I(), _I(), _I()
        nu, nv, nw = _L(), _L(), _L()
        check(lib().fmmhip_alg_info(self.h, *(ctypes.byref(x) for x in (mt, kt, nt, R, nu, nv, nw))))
        return dict(mt=mt.value, kt=kt.value, nt=nt.value, R=R.value, nnz_u=nu.value, nnz_v=nv.value,
                    nnz_w=nw.value)

    def brent_violations(self):
        return lib().fmmhip_alg_brent_violations(self.h)

    def set_shape(self, shape):
        check(lib().fmmhip_alg_set_shape(self.h, shape), "alg_set_shape")

    def set_pack(self, pack):
        check(lib().fmmhip_alg_set_pack(self.h, pack), "alg_set_pack")

    def pack_for(self, m, n, k, leaf=0):
        """Pack mode (0 fused, 1 pre-summed, 2 products materialised) a
        column-major FP64 call of m x n x k resolves to."""
        p = ctypes.c_int(-1)
        check(lib().fmmhip_alg_pack_for(self.h, leaf, m, n, k, ctypes.byref(p)), "alg_pack_for")
        return p.value

    def last_pack(self):
        """Pack mode the last FP64/FP32-leaf call took (-1 before the first)."""
        p = ctypes.c_int(-1)
        check(lib().fmmhip_alg_last_pack(self.h, ctypes.byref(p)), "alg_last_pack")
        return p.value

    def last_shape(self):
        """Kernel shape the last FP64/FP32 call launched (-1 before the first)."""
        p = ctypes.c_int(-1)
        check(lib().fmmhip_alg_last_shape(self.h, ctypes.byref(p)), "alg_last_shape")
        return p.value

    def set_split(self, groups):
        """Product groups per C tile: 1, 2, 4, 8, or -1 (chosen per call)."""
        check(lib().fmmhip_alg_set_split(self.h, groups), "alg_set_split")

    def close(self):
        if self.h:
            lib().fmmhip_alg_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Raw device allocation through fmmhip_malloc."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = _P()
        check(lib().fmmhip_malloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), f"malloc {nbytes}")

    @property
    def addr(self):
        return self.ptr.value

    def upload(self, arr):
        import numpy as np
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(lib().fmmhip_memcpy_h2d(self.ptr, a.ctypes.data_as(_P), a.nbytes), "h2d")

    def download(self, dtype, count):
        import numpy as np
        out = np.empty(count, dtype=dtype)
        check(lib().fmmhip_memcpy_d2h(out.ctypes.data_as(_P), self.ptr, out.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            lib().fmmhip_free(self.ptr)
            self.ptr = _P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceMatrix:
    """Column-major m x n matrix in device memory (leading dimension ld)."""

    def __init__(self, m, n, dtype="f8", ld=None):
        import numpy as np
        self.m, self.n = int(m), int(n)
        self.dtype = np.dtype(dtype)
        self.ld = int(ld) if ld is not None else max(self.m, 1)
        self.buf = DeviceBuffer(self.dtype.itemsize * self.ld * max(self.n, 1))

    @property
    def ptr(self):
        return self.buf.ptr

    def fill_uniform(self, seed, stream_id, stream=None):
        f = lib().fmmhip_fill_uniform if self.dtype.itemsize == 8 else lib().fmmhip_fill_uniform_f32
        check(f(self.ptr, self.m, self.n, self.ld, seed, stream_id, stream), "fill_uniform")
        return self

    def fill_int(self, seed, stream_id, amp, stream=None):
        assert self.dtype.itemsize == 8
        check(lib().fmmhip_fill_int(self.ptr, self.m, self.n, self.ld, seed, stream_id, amp, stream), "fill_int")
        return self

    def upload(self, arr):
        import numpy as np
        a = np.asarray(arr, dtype=self.dtype)
        full = np.zeros((self.n, self.ld), dtype=self.dtype)
        full[:, : self.m] = a.T
        self.buf.upload(full)
        return self

    def to_numpy(self):
        check(lib().fmmhip_device_sync(), "sync")
        flat = self.buf.download(self.dtype, self.ld * max(self.n, 1))
        return flat.reshape(max(self.n, 1), self.ld)[: self.n, : self.m].T.copy()

    def free(self):
        self.buf.free()


def dgemm_fmm(alg, m, n, k, alpha, A, B, beta, C, leaf=LEAF_NATIVE, stream=None):
    """C := beta*C + alpha*A*B on column-major DeviceMatrix operands."""
    check(lib().fmmhip_dgemm_fmm(alg.h, leaf, m, n, k, alpha, A.ptr, 1, A.ld, B.ptr, 1, B.ld, beta,
                                 C.ptr, 1, C.ld, stream), "dgemm_fmm")


def sgemm_fmm(alg, m, n, k, alpha, A, B, beta, C, stream=None):
    check(lib().fmmhip_sgemm_fmm(alg.h, m, n, k, alpha, A.ptr, 1, A.ld, B.ptr, 1, B.ld, beta,
                                 C.ptr, 1, C.ld, stream), "sgemm_fmm")


def dgemm(m, n, k, alpha, A, B, beta, C, stream=None):
    check(lib().fmmhip_dgemm(m, n, k, alpha, A.ptr, 1, A.ld, B.ptr, 1, B.ld, beta, C.ptr, 1, C.ld, stream),
          "dgemm")


def sgemm(m, n, k, alpha, A, B, beta, C, stream=None):
    check(lib().fmmhip_sgemm(m, n, k, alpha, A.ptr, 1, A.ld, B.ptr, 1, B.ld, beta, C.ptr, 1, C.ld, stream),
          "sgemm")


CAND_CLASSICAL, CAND_STRASSEN1, CAND_STRASSEN2 = 0, 1, 2


class Batch:
    """Many independent column-major GEMMs captured into one HIP graph
    (fmmhip_batch_*).  add() takes DeviceMatrix operands of dtype f8
    (FP64) or f4 (FP32); cand -1 lets the tuner choose the algorithm."""

    def __init__(self):
        self.h = _P()
        check(lib().fmmhip_batch_create(ctypes.byref(self.h)), "batch_create")
        self._keep = []

    def add(self, m, n, k, alpha, A, B, beta, C, cand=-1):
        if not (A.dtype == B.dtype == C.dtype):
            raise FmmError("batch_add: operands of one problem share one datatype")
        if C.dtype.itemsize == 8:
            rc = lib().fmmhip_batch_add_d(self.h, m, n, k, alpha, A.ptr, A.ld, B.ptr, B.ld, beta, C.ptr, C.ld, cand)
        else:
            rc = lib().fmmhip_batch_add_s(self.h, m, n, k, alpha, A.ptr, A.ld, B.ptr, B.ld, beta, C.ptr, C.ld, cand)
        check(rc, "batch_add")
        self._keep.append((A, B, C))

    def plan(self):
        check(lib().fmmhip_batch_plan(self.h), "batch_plan")

    def run(self, stream=None):
        check(lib().fmmhip_batch_run(self.h, stream), "batch_run")

    def info(self):
        n, f64, f32, est, ws = _L(), _L(), _L(), _D(), ctypes.c_size_t()
        check(lib().fmmhip_batch_info(self.h, ctypes.byref(n), ctypes.byref(f64), ctypes.byref(f32),
                                      ctypes.byref(est), ctypes.byref(ws)), "batch_info")
        return {"problems": n.value, "items_f64": f64.value, "items_f32": f32.value, "est_ms": est.value,
                "workspace_bytes": ws.value}

    def choice(self, i):
        c = _I()
        check(lib().fmmhip_batch_choice(self.h, i, ctypes.byref(c)), "batch_choice")
        return c.value

    def set_tile(self, i, tile):
        """Force problem i's tile edge (64 or 128; 0 = the tuner's choice)."""
        check(lib().fmmhip_batch_set_tile(self.h, i, tile), "batch_set_tile")

    def fast_items(self):
        """(FP64, FP32) work items of the plan that run on the fast kernel."""
        f64, f32 = _L(), _L()
        check(lib().fmmhip_batch_fast_items(self.h, ctypes.byref(f64), ctypes.byref(f32)), "batch_fast_items")
        return f64.value, f32.value

    def tile(self, i):
        """The planned tile edge of problem i."""
        t = _I()
        check(lib().fmmhip_batch_tile(self.h, i, ctypes.byref(t)), "batch_tile")
        return t.value

    def close(self):
        if self.h:
            lib().fmmhip_batch_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving form (fmmhip_server_*): requests over device-resident
    DeviceMatrix operands, admitted by HBM budget onto `slots` streams."""

    def __init__(self, hbm_budget, slots=4):
        self.h = _P()
        check(lib().fmmhip_server_create(ctypes.byref(self.h), hbm_budget, slots), "server_create")

    def now(self):
        return lib().fmmhip_server_now(self.h)

    def submit(self, m, n, k, alpha, A, B, beta=0.0, C=None, cand=-1, keep=1, arrival_s=0.0, lda=None, ldb=None):
        """A and B are DeviceMatrix objects or (pointer, ld) pairs; C a
        DeviceMatrix or None (server-allocated, beta 0).  Returns the id."""
        def ptr_ld(X, ld):
            if isinstance(X, tuple):
                return X
            return X.ptr.value if hasattr(X.ptr, "value") else X.ptr, X.ld
        pa, la = ptr_ld(A, lda)
        pb, lb = ptr_ld(B, ldb)
        dt = A.dtype if not isinstance(A, tuple) else None
        r = Request()
        r.is_double = 1 if (dt is None or dt.itemsize == 8) else 0
        r.m, r.n, r.k, r.alpha, r.beta = m, n, k, alpha, beta
        r.A, r.lda, r.B, r.ldb = pa, la, pb, lb
        if C is not None:
            pc = C.ptr.value if hasattr(C.ptr, "value") else C.ptr
            r.C, r.ldc = pc, C.ld
        r.cand, r.keep, r.arrival_s = cand, keep, arrival_s
        rid = _L()
        check(lib().fmmhip_server_submit(self.h, ctypes.byref(r), ctypes.byref(rid)), "server_submit")
        return rid.value

    def submit_request(self, req):
        rid = _L()
        check(lib().fmmhip_server_submit(self.h, ctypes.byref(req), ctypes.byref(rid)), "server_submit")
        return rid.value

    def wait(self, rid):
        st = _I()
        check(lib().fmmhip_server_wait(self.h, rid, ctypes.byref(st)), "server_wait")
        return st.value

    def result(self, rid):
        """(status, C pointer, ldc, latency ms, candidate)."""
        c, ld, lat, cand = _P(), _L(), _D(), _I()
        st = lib().fmmhip_server_result(self.h, rid, ctypes.byref(c), ctypes.byref(ld), ctypes.byref(lat),
                                        ctypes.byref(cand))
        return st, c.value, ld.value, lat.value, cand.value

    def release(self, rid):
        check(lib().fmmhip_server_release(self.h, rid), "server_release")

    def drain(self):
        check(lib().fmmhip_server_drain(self.h), "server_drain")

    def stats(self):
        s = ServerStats()
        check(lib().fmmhip_server_stats(self.h, ctypes.byref(s)), "server_stats")
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().fmmhip_server_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tuner_choose(is_double, m, n, k, lone=True):
    """(candidate, estimated ms) the tuner picks for one problem."""
    c, est = _I(), _D()
    check(lib().fmmhip_tuner_choose(1 if is_double else 0, m, n, k, 1 if lone else 0, ctypes.byref(c),
                                    ctypes.byref(est)), "tuner_choose")
    return c.value, est.value


def tuner_choose_tile(is_double, m, n, k, lone=True):
    """(candidate, tile edge, estimated ms) the tuner picks for one problem."""
    c, t, est = _I(), _I(), _D()
    check(lib().fmmhip_tuner_choose_tile(1 if is_double else 0, m, n, k, 1 if lone else 0, ctypes.byref(c),
                                         ctypes.byref(t), ctypes.byref(est)), "tuner_choose_tile")
    return c.value, t.value, est.value


def tuner_estimate(is_double, cand, m, n, k, lone=True):
    est = _D()
    check(lib().fmmhip_tuner_estimate(1 if is_double else 0, cand, m, n, k, 1 if lone else 0, ctypes.byref(est)),
          "tuner_estimate")
    return est.value


def tuner_calibrate(stream=None):
    check(lib().fmmhip_tuner_calibrate(stream), "tuner_calibrate")


def tuner_save(path):
    check(lib().fmmhip_tuner_save(path.encode()), "tuner_save")


def tuner_load(path):
    check(lib().fmmhip_tuner_load(path.encode()), "tuner_load")


def fro_diff(X, Y):
    out = _D()
    check(lib().fmmhip_fro_diff(X.ptr, Y.ptr, X.m, X.n, X.ld, Y.ld, ctypes.byref(out)), "fro_diff")
    return out.value


def fro(X):
    out = _D()
    check(lib().fmmhip_fro(X.ptr, X.m, X.n, X.ld, ctypes.byref(out)), "fro")
    return out.value


def sync():
    check(lib().fmmhip_device_sync(), "device_sync")


def measure_peaks():
    """(FP64 MFMA TFLOP/s, copy GB/s) measured on the device."""
    t, g = _D(), _D()
    check(lib().fmmhip_measure_peaks(ctypes.byref(t), ctypes.byref(g)), "measure_peaks")
    return t.value, g.value


def measure_hbm():
    """(copy, read-only, write-only) GB/s: the device's streaming ceilings."""
    c, r, w = _D(), _D(), _D()
    check(lib().fmmhip_measure_hbm(ctypes.byref(c), ctypes.byref(r), ctypes.byref(w)), "measure_hbm")
    return c.value, r.value, w.value


def timing_enable(on=True):
    check(lib().fmmhip_timing_enable(1 if on else 0), "timing_enable")


def timing_read():
    """(launches, kernel_ms_total, presum_ms_total, kernel_ms_max) since the last read."""
    n, k, p, mx = _L(), _D(), _D(), _D()
    check(lib().fmmhip_timing_read(ctypes.byref(n), ctypes.byref(k), ctypes.byref(p), ctypes.byref(mx)),
          "timing_read")
    return n.value, k.value, p.value, mx.value


class Event:
    def __init__(self):
        self.h = _P()
        check(lib().fmmhip_event_create(ctypes.byref(self.h)), "event_create")

    def record(self, stream=None):
        check(lib().fmmhip_event_record(self.h, stream), "event_record")

    def elapsed_ms(self, later):
        ms = _F()
        check(lib().fmmhip_event_elapsed_ms(ctypes.byref(ms), self.h, later.h), "event_elapsed")
        return ms.value


def obj_attach(arr_or_ptr, m, n, rs, cs, dt=BLIS_DOUBLE):
    """obj_t view of a numpy array (host) or an integer device address."""
    o = ObjT()
    if hasattr(arr_or_ptr, "ctypes"):
        p = arr_or_ptr.ctypes.data_as(_P)
    else:
        p = _P(arr_or_ptr)
    lib().fmm_obj_attach(dt, m, n, p, rs, cs, ctypes.byref(o))
    return o


def obj_scalar(value):
    """1x1 double scalar object; returns (obj, storage) -- keep storage alive."""
    st = ctypes.c_double(value)
    o = ObjT()
    lib().fmm_obj_scalar(ctypes.byref(st), ctypes.byref(o))
    return o, st
