"""GPU tests of the drop-in boundary beyond the reference's valid domain:
general strides on device and host objects, the kernel registry (slots
registered by bli_plugin_register_fmm_blis, filled by the per-configuration
init, read back and called), and the algorithm cache under many distinct
tables from several host threads.

Bar: integer-valued inputs, so every result must equal the oracle's FMM (or
classical product) bit for bit.
"""
import ctypes
import sys
import threading

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O
from conftest import alg_path

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    rc = F.lib().fmmhip_init()
    if rc != 0:
        pytest.fail(f"no usable gfx950 device: {F.STATUS.get(rc)}")
    F.lib().fmm_blis_amd_set_abort_on_error(0)
    yield
    F.lib().fmm_blis_amd_set_abort_on_error(1)


def strided_host(X, rs, cs):
    """A flat buffer holding X with element (i,j) at i*rs + j*cs (other
    elements NaN, so a read outside the view shows)."""
    m, n = X.shape
    buf = np.full((m - 1) * rs + (n - 1) * cs + 1, np.nan)
    for j in range(n):
        buf[np.arange(m) * rs + j * cs] = X[:, j]
    return buf


def strided_read(buf, m, n, rs, cs):
    out = np.empty((m, n))
    for j in range(n):
        out[:, j] = buf[np.arange(m) * rs + j * cs]
    return out


@pytest.mark.parametrize("name,levels", [("222", 1), ("222", 2), ("333", 1), ("classical", 1)])
@pytest.mark.parametrize("where", ["device", "host"])
def test_general_strides(name, levels, where):
    """A with rs = 2, cs = 3m (every other row of a wider view), B with
    rs = 3, cs = 2k+1, C with rs = 2, cs = 2m+3, alpha = 2, beta = -1, through
    bli_fmm_ab_ex (or bli_gemm_hip for the classical product), device and
    host buffers: bit for bit the oracle's FMM on the same values, and the
    elements of the buffers outside the views untouched."""
    m, n, k = 157, 133, 149
    say(f"general strides {name} L{levels} {where}")
    A, B, C = O.integers(m, k, 30, 0, 3), O.integers(k, n, 30, 1, 3), O.integers(m, n, 30, 2, 3)
    if name == "classical":
        want = O.dgemm(2.0, A, B, -1.0, C.copy(order="F"))
    else:
        og = O.parse(alg_path(name))
        og = og if levels == 1 else O.compose(og, og)
        want = O.fmm_dgemm(og, 2.0, A, B, -1.0, C.copy(order="F"))
    sA, sB, sC = (2, 3 * m), (3, 2 * k + 1), (2, 2 * m + 3)
    hA, hB, hC = strided_host(A, *sA), strided_host(B, *sB), strided_host(C, *sC)
    keep = hC.copy()
    al, s1 = F.obj_scalar(2.0)
    be, s2 = F.obj_scalar(-1.0)
    bufs = []
    if where == "device":
        for h in (hA, hB, hC):
            d = F.DeviceBuffer(h.nbytes)
            d.upload(h)
            bufs.append(d)
        pa, pb, pc = (d.ptr.value for d in bufs)
    else:
        pa, pb, pc = hA, hB, hC
    oa, ob, oc = F.obj_attach(pa, m, k, *sA), F.obj_attach(pb, k, n, *sB), F.obj_attach(pc, m, n, *sC)
    L = F.lib()
    if name == "classical":
        L.bli_gemm_hip(al, oa, ob, be, oc)
    else:
        f = F.strassen() if name == "222" else F.new_fmm(alg_path(name))
        lv = (F.FmmT * levels)(*([f] * levels))
        L.bli_fmm_ab_ex(al, oa, ob, be, oc, lv, levels, F.LEAF_NATIVE)
    assert L.fmm_blis_amd_last_error() == 0
    outC = bufs[2].download(np.float64, hC.size) if where == "device" else hC
    got = strided_read(outC, m, n, *sC)
    assert np.array_equal(got, want)
    inside = np.zeros(hC.size, dtype=bool)
    for j in range(n):
        inside[np.arange(m) * sC[0] + j * sC[1]] = True
    assert np.array_equal(np.isnan(outC[~inside]), np.isnan(keep[~inside]))


def test_general_strides_native_entry():
    """fmmhip_dgemm_fmm with a general-strided device A and B (gathered on
    the device), FP64 and FP32, equals the oracle's FMM bit for bit."""
    m, n, k = 300, 260, 280
    A, B, C = O.integers(m, k, 31, 0, 3), O.integers(k, n, 31, 1, 3), O.integers(m, n, 31, 2, 3)
    want = O.fmm_dgemm(O.parse(alg_path("222")), 1.0, A, B, 1.0, C.copy(order="F"))
    sA, sB = (3, 3 * m + 1), (2, 2 * k)
    alg = F.Alg([F.strassen()])
    for dt in (np.float64, np.float32):
        hA = strided_host(A, *sA).astype(dt)
        hB = strided_host(B, *sB).astype(dt)
        dA, dB = F.DeviceBuffer(hA.nbytes), F.DeviceBuffer(hB.nbytes)
        dA.upload(hA)
        dB.upload(hB)
        dC = F.DeviceMatrix(m, n, dtype=dt)
        dC.upload(C.astype(dt))
        if dt == np.float64:
            rc = F.lib().fmmhip_dgemm_fmm(alg.h, 0, m, n, k, 1.0, dA.ptr, sA[0], sA[1], dB.ptr, sB[0], sB[1], 1.0,
                                          dC.ptr, 1, m, None)
        else:
            rc = F.lib().fmmhip_sgemm_fmm(alg.h, m, n, k, 1.0, dA.ptr, sA[0], sA[1], dB.ptr, sB[0], sB[1], 1.0,
                                          dC.ptr, 1, m, None)
        F.check(rc, "strided native call")
        F.sync()
        assert np.array_equal(dC.to_numpy().astype(np.float64), want), dt


def test_kernel_registry_slots_and_calls():
    """bli_plugin_register_fmm_blis registers the two slots (ids 1 and 2)
    and the per-configuration init fills them for float and double; the
    kernels read back through fmm_blis_get_ukr_dt, called in order (pack,
    then gemm) on one stream, equal the one-call pre-summed product bit for
    bit (uniform inputs, alpha 0.5, beta 2) and the oracle's FMM on
    integers."""
    L = F.lib()
    assert L.bli_plugin_register_fmm_blis() == F.BLIS_SUCCESS
    pack_id = ctypes.c_uint64.in_dll(L, "FMM_BLIS_PACK_UKR").value
    gemm_id = ctypes.c_uint64.in_dll(L, "FMM_BLIS_GEMM_UKR").value
    assert (pack_id, gemm_id) == (1, 2)
    assert not L.fmm_blis_get_ukr_dt(F.BLIS_DOUBLE, 7)
    assert not L.fmm_blis_get_ukr_dt(3, pack_id)           # complex: no kernel
    D = ctypes.c_double
    Fl = ctypes.c_float
    P, Lg = ctypes.c_void_p, ctypes.c_long

    def ukr(dt, slot, scalar):
        addr = L.fmm_blis_get_ukr_dt(dt, slot)
        assert addr
        proto = ctypes.CFUNCTYPE(ctypes.c_int, P, Lg, Lg, Lg, scalar, P, Lg, Lg, P, Lg, Lg, scalar, P, Lg, Lg, P,
                                 ctypes.c_size_t, P)
        return proto(addr)

    m, n, k = 390, 410, 370
    alg = F.Alg([F.strassen()] * 2)
    for dt, npdt, scalar in ((F.BLIS_DOUBLE, np.float64, D), (F.BLIS_FLOAT, np.float32, Fl)):
        pack, gemm = ukr(dt, pack_id, scalar), ukr(dt, gemm_id, scalar)
        need = ctypes.c_size_t()
        F.check(L.fmmhip_fmm_workspace_bytes(alg.h, int(npdt == np.float64), m, n, k, 1, m, 1, k, 1, m,
                                             ctypes.byref(need)), "workspace_bytes")
        ws = F.DeviceBuffer(need.value)
        for integer in (False, True):
            if integer:
                A, B, C = O.integers(m, k, 32, 0, 2), O.integers(k, n, 32, 1, 2), O.integers(m, n, 32, 2, 2)
                alpha, beta = 1.0, 1.0
            else:
                A, B, C = O.uniform(m, k, 32, 0), O.uniform(k, n, 32, 1), O.uniform(m, n, 32, 2)
                alpha, beta = 0.5, 2.0
            dA, dB = F.DeviceMatrix(m, k, dtype=npdt), F.DeviceMatrix(k, n, dtype=npdt)
            dA.upload(A.astype(npdt))
            dB.upload(B.astype(npdt))
            dC1, dC2 = F.DeviceMatrix(m, n, dtype=npdt), F.DeviceMatrix(m, n, dtype=npdt)
            dC1.upload(C.astype(npdt))
            dC2.upload(C.astype(npdt))
            args = (alg.h, m, n, k, alpha, dA.ptr, 1, m, dB.ptr, 1, k, beta, dC1.ptr, 1, m, ws.ptr, need.value, None)
            F.check(pack(*args), "pack ukr")
            F.check(gemm(*args), "gemm ukr")
            alg.set_pack(1)
            alg.set_split(1)
            if npdt == np.float64:
                F.dgemm_fmm(alg, m, n, k, alpha, dA, dB, beta, dC2)
            else:
                F.sgemm_fmm(alg, m, n, k, alpha, dA, dB, beta, dC2)
            alg.set_pack(-1)
            alg.set_split(-1)
            F.sync()
            got = dC1.to_numpy()
            assert np.array_equal(got, dC2.to_numpy()), (dt, integer)
            if integer:
                og = O.parse(alg_path("222"))
                want = O.fmm_dgemm(O.compose(og, og), 1.0, A, B, 1.0, C.copy(order="F"))
                assert np.array_equal(got.astype(np.float64), want), dt
        too_small = L.fmmhip_dpack_fmm(alg.h, m, n, k, 1.0, dA.ptr, 1, m, dB.ptr, 1, k, 1.0, dC1.ptr, 1, m, ws.ptr,
                                       16, None)
        assert too_small == 1                              # FMMHIP_ERR_INVALID: workspace too small


def permuted_strassen(perm):
    """Strassen with its products renumbered by perm: distinct tables, same
    algorithm (the Brent equations hold for any product order)."""
    base = F.strassen()
    U, V, W = base.tables()
    U, V, W = (np.ascontiguousarray(np.asarray(X)[:, perm]).astype(np.int32) for X in (U, V, W))
    f = F.FmmT()
    f.m_tilde, f.n_tilde, f.k_tilde, f.R = 2, 2, 2, 7
    keep = [U, V, W]
    f.U, f.V, f.W = (X.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) for X in keep)
    return f, keep


def test_algorithm_cache_many_tables_from_threads():
    """Four host threads each run bli_strassen_ab_ex with 12 distinct tables
    (48 in all, three times the cache's 16 slots), so calls evict each
    other's algorithms while they run: pinned algorithms are never evicted,
    every result equals the oracle's FMM bit for bit, and nothing stays
    pinned afterwards."""
    import itertools
    perms = list(itertools.islice(itertools.permutations(range(7)), 1, 49))
    tables = [permuted_strassen(list(p)) for p in perms]
    m, n, k = 96, 88, 80
    og = O.parse(alg_path("222"))
    A, B = O.integers(m, k, 33, 0, 3), O.integers(k, n, 33, 1, 3)
    want = O.fmm_dgemm(og, 1.0, A, B, 0.0, np.zeros((m, n), order="F"))
    dA, dB = F.DeviceMatrix(m, k), F.DeviceMatrix(k, n)
    dA.upload(A)
    dB.upload(B)
    outs = [F.DeviceMatrix(m, n) for _ in tables]
    errors = []
    one, s1 = F.obj_scalar(1.0)
    zero, s0 = F.obj_scalar(0.0)

    def worker(t):
        try:
            for rep in range(2):
                for i in range(t, len(tables), 4):
                    oa = F.obj_attach(dA.ptr.value, m, k, 1, m)
                    ob = F.obj_attach(dB.ptr.value, k, n, 1, k)
                    oc = F.obj_attach(outs[i].ptr.value, m, n, 1, m)
                    F.lib().bli_strassen_ab_ex(one, oa, ob, zero, oc, tables[i][0])
                    if F.lib().fmm_blis_amd_last_error():
                        errors.append((t, i, F.lib().fmm_blis_amd_last_error()))
        except Exception as e:          # reported on the main thread
            errors.append(e)

    ths = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    assert not errors, errors
    for i, o in enumerate(outs):
        assert np.array_equal(o.to_numpy(), want), i
    cached, pinned = ctypes.c_int(), ctypes.c_int()
    F.lib().fmm_blis_amd_alg_cache_stats(ctypes.byref(cached), ctypes.byref(pinned))
    say(f"alg cache: {cached.value} cached, {pinned.value} pinned")
    assert cached.value == 16 and pinned.value == 0
