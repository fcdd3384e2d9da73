"""GPU parity: the HIP path through the C ABI against the pinned CPU oracle
on the same seeded inputs.

Bars (stated per test):
  * integer-valued inputs make every algorithm exact in FP64, so the device
    result must equal the oracle's bit for bit (any indexing, sign, clipping
    or beta mistake shows);
  * uniform[-1,1] inputs: normalised error ||C - C_ref||_F/(||A||_F ||B||_F)
    against the oracle's classical product, below the per-level bounds;
  * the device fill reproduces the oracle's seeded bits exactly.
"""
import ctypes
import sys

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O
from conftest import ALG_NAMES, alg_path

pytestmark = pytest.mark.gpu

# normalised-error bounds for uniform[-1,1] inputs at the sizes below
TOL_FP64 = {1: 1e-15, 2: 5e-15}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    rc = F.lib().fmmhip_init()
    if rc != 0:
        pytest.fail(f"no usable gfx950 device: {F.STATUS.get(rc)}")
    buf = ctypes.create_string_buffer(64)
    F.lib().fmmhip_device_arch(buf, 64)
    say("device", buf.value.decode())
    yield


_algs = {}


def fmm(name):
    if name not in _algs:
        _algs[name] = F.new_fmm(alg_path(name))
    return _algs[name]


def dev(X):
    D = F.DeviceMatrix(X.shape[0], X.shape[1], dtype=X.dtype)
    D.upload(X)
    return D


def run_fmm(alg, alpha, A, B, beta, C, leaf=F.LEAF_NATIVE):
    m, k = A.shape
    n = B.shape[1]
    dA, dB, dC = dev(A), dev(B), dev(C)
    F.dgemm_fmm(alg, m, n, k, alpha, dA, dB, beta, dC, leaf=leaf)
    F.sync()
    return dC.to_numpy()


@pytest.mark.parametrize("m,n,seed,sid", [(1, 1, 0, 0), (37, 29, 0, 3), (1000, 17, 5, 1)])
def test_device_fill_bit_identical(m, n, seed, sid):
    say("fill", m, n, seed, sid)
    D = F.DeviceMatrix(m, n).fill_uniform(seed, sid)
    assert np.array_equal(D.to_numpy(), O.uniform(m, n, seed, sid))
    I = F.DeviceMatrix(m, n).fill_int(seed, sid, 4)
    assert np.array_equal(I.to_numpy(), O.integers(m, n, seed, sid, 4))
    S = F.DeviceMatrix(m, n, dtype="f4").fill_uniform(seed, sid)
    assert np.array_equal(S.to_numpy(), O.uniform_f32(m, n, seed, sid))


CLASSICAL_SHAPES = [(1, 1, 1), (3, 13, 7), (128, 128, 16), (129, 130, 17), (200, 64, 333), (512, 384, 256),
                    (1, 300, 5), (300, 1, 77), (33, 65, 1)]


@pytest.mark.parametrize("m,n,k", CLASSICAL_SHAPES)
def test_classical_dgemm_exact_integers(m, n, k):
    say("classical int", m, n, k)
    A, B, C = O.integers(m, k, 1, 0, 4), O.integers(k, n, 1, 1, 4), O.integers(m, n, 1, 2, 4)
    alg = F.Alg(classical=True)
    for alpha, beta in [(1.0, 1.0), (-2.0, 0.5), (1.0, 0.0)]:
        want = O.dgemm(alpha, A, B, beta, C.copy(order="F"))
        got = run_fmm(alg, alpha, A, B, beta, C)
        assert np.array_equal(got, want), (alpha, beta)


@pytest.mark.parametrize("m,n,k", [(257, 255, 300), (512, 512, 512)])
def test_classical_dgemm_random(m, n, k):
    say("classical rand", m, n, k)
    A, B, C = O.uniform(m, k, 0, 0), O.uniform(k, n, 0, 1), O.uniform(m, n, 0, 2)
    want = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    dA, dB, dC = dev(A), dev(B), dev(C)
    F.dgemm(m, n, k, 1.0, dA, dB, 1.0, dC)
    got = dC.to_numpy()
    assert O.normalized_error(got, want, A, B) < 1e-15


FMM_SHAPES = [(3, 13, 7), (31, 33, 35), (64, 64, 64), (97, 101, 103), (258, 260, 262), (1, 5, 3), (17, 1, 19),
              (300, 280, 36), (96, 96, 1)]


@pytest.mark.parametrize("name", ALG_NAMES)
@pytest.mark.parametrize("levels", [1, 2])
@pytest.mark.parametrize("pack", [0, 1, 2])
def test_fmm_exact_integers(name, levels, pack):
    """Bit-exact agreement with the oracle's FMM on integer inputs (uneven
    splits included), in the fused, the pre-summed and the
    products-materialised pack modes."""
    f = fmm(name)
    alg = F.Alg([f] * levels)
    alg.set_pack(pack)
    og = O.parse(alg_path(name))
    og = og if levels == 1 else O.compose(og, og)
    shapes = FMM_SHAPES if not (name == "444" and levels == 2) else [(20, 18, 17), (130, 140, 150)]
    for (m, n, k) in shapes:
        say("fmm int", name, levels, m, n, k)
        A, B, C = O.integers(m, k, 3, 0, 4), O.integers(k, n, 3, 1, 4), O.integers(m, n, 3, 2, 4)
        for alpha, beta in [(1.0, 1.0), (2.0, -3.0), (0.5, 0.0)]:
            want = O.fmm_dgemm(og, alpha, A, B, beta, C.copy(order="F"))
            got = run_fmm(alg, alpha, A, B, beta, C)
            assert np.array_equal(got, want), (m, n, k, alpha, beta)


@pytest.mark.parametrize("name", ALG_NAMES)
@pytest.mark.parametrize("levels", [1, 2])
@pytest.mark.parametrize("pack", [0, 1, 2])
def test_fmm_random_error(name, levels, pack):
    f = fmm(name)
    alg = F.Alg([f] * levels)
    alg.set_pack(pack)
    m = n = k = 432 if name != "444" else 512
    say("fmm rand", name, levels, m)
    A, B, C = O.uniform(m, k, 0, 0), O.uniform(k, n, 0, 1), O.uniform(m, n, 0, 2)
    ref = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    got = run_fmm(alg, 1.0, A, B, 1.0, C)
    err = O.normalized_error(got, ref, A, B)
    say(f"  normalised error {err:.3e}")
    assert err < TOL_FP64[levels]


def test_beta_zero_ignores_nan():
    m, n, k = 70, 90, 50
    A, B = O.uniform(m, k, 2, 0), O.uniform(k, n, 2, 1)
    C = np.full((m, n), np.nan, order="F")
    got = run_fmm(F.Alg([F.strassen()]), 1.0, A, B, 0.0, C)
    assert np.all(np.isfinite(got))
    assert O.normalized_error(got, A @ B, A, B) < 1e-15


def test_alpha_zero_and_k_zero_scale_c():
    A, B, C = O.uniform(20, 0, 1, 0), O.uniform(0, 30, 1, 1), O.uniform(20, 30, 1, 2)
    alg = F.Alg([F.strassen()])
    dA, dB, dC = F.DeviceMatrix(20, 0), F.DeviceMatrix(0, 30), dev(C)
    F.dgemm_fmm(alg, 20, 30, 0, 1.0, dA, dB, 3.0, dC)
    assert np.array_equal(dC.to_numpy(), 3.0 * C)
    A = O.uniform(20, 10, 1, 0)
    B = O.uniform(10, 30, 1, 1)
    got = run_fmm(alg, 0.0, A, B, -1.0, C)
    assert np.array_equal(got, -C)


@pytest.mark.parametrize("ta,tb,tc", [(True, False, False), (False, True, False), (True, True, True),
                                      (False, False, True)])
def test_operand_layouts(ta, tb, tc):
    """Row-major (transposed) operands through the native strides."""
    m, n, k = 150, 170, 130
    say("layouts", ta, tb, tc)
    A, B, C = O.integers(m, k, 4, 0, 3), O.integers(k, n, 4, 1, 3), O.integers(m, n, 4, 2, 3)
    want = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    def put(X, t):
        # row-major storage = column-major storage of X^T
        return dev(np.asfortranarray(X.T)) if t else dev(X)
    rsA, csA = (k, 1) if ta else (1, m)
    rsB, csB = (n, 1) if tb else (1, k)
    rsC, csC = (n, 1) if tc else (1, m)
    alg = F.Alg([fmm("232")])
    for pack in (0, 1, 2):
        dA, dB, dC = put(A, ta), put(B, tb), put(C, tc)
        alg.set_pack(pack)
        F.check(F.lib().fmmhip_dgemm_fmm(alg.h, 0, m, n, k, 1.0, dA.ptr, rsA, csA, dB.ptr, rsB, csB, 1.0,
                                         dC.ptr, rsC, csC, None), "dgemm_fmm layouts")
        got = dC.to_numpy()
        if tc:
            got = got.T
        assert np.array_equal(got, want), pack


def test_dropin_reference_case_host_objects():
    """test/test_strassen_oapi.c:44-224 restated: 3 x 13 x 7, alpha = beta = 1,
    column-stored objects in host memory, ||C - C_ref||_F printed."""
    m, n, k = 3, 13, 7
    A, B, C = O.uniform(m, k, 0, 0), O.uniform(k, n, 0, 1), O.uniform(m, n, 0, 2)
    C_ref = O.dgemm_naive(A, B, C.copy(order="F"))
    Cw = C.copy(order="F")
    L = F.lib()
    assert L.bli_plugin_register_fmm_blis() == F.BLIS_SUCCESS
    one, s1 = F.obj_scalar(1.0)
    L.bli_strassen_ab(one, F.obj_attach(A, m, k, 1, m), F.obj_attach(B, k, n, 1, k), one,
                      F.obj_attach(Cw, m, n, 1, m))
    resid = np.linalg.norm(Cw - C_ref)
    say(f"{m:5d}\t {n:5d}\t {k:5d}\t resid {resid:.2g}")
    assert resid < 1e-14


@pytest.mark.parametrize("name", ["222", "232", "333"])
def test_dropin_ex_device_objects_and_trans(name):
    m, n, k = 201, 190, 180
    A, B, C = O.integers(m, k, 6, 0, 3), O.integers(k, n, 6, 1, 3), O.integers(m, n, 6, 2, 3)
    want = O.dgemm(2.0, A, B, -1.0, C.copy(order="F"))
    dAt = dev(np.asfortranarray(A.T))          # A stored transposed, trans bit set
    dB, dC = dev(B), dev(C)
    oa = F.obj_attach(dAt.ptr.value, k, m, 1, k)
    oa.info |= F.TRANS_BIT
    ob = F.obj_attach(dB.ptr.value, k, n, 1, k)
    oc = F.obj_attach(dC.ptr.value, m, n, 1, m)
    al, s1 = F.obj_scalar(2.0)
    be, s2 = F.obj_scalar(-1.0)
    F.lib().bli_strassen_ab_ex(al, oa, ob, be, oc, fmm(name))
    assert np.array_equal(dC.to_numpy(), want)


@pytest.mark.parametrize("dt", ["f8", "f4"])
def test_dropin_constant_scalars(dt):
    """alpha and beta as BLIS_CONSTANT objects (the form of BLIS_ONE,
    BLIS_MINUS_ONE), whose buffer holds one value per datatype: the double
    value is taken (here 2 and -1; the float fields hold 3 and 5 so a wrong
    offset shows), for double and float operands alike."""
    m, n, k = 130, 100, 90
    A, B, C = O.integers(m, k, 9, 0, 3), O.integers(k, n, 9, 1, 3), O.integers(m, n, 9, 2, 3)
    want = O.dgemm(2.0, A, B, -1.0, C.copy(order="F"))
    code = F.BLIS_DOUBLE if dt == "f8" else F.BLIS_FLOAT
    if dt == "f4":
        A, B, C = (np.asfortranarray(X.astype(np.float32)) for X in (A, B, C))
        want = want.astype(np.float32)
    dA, dB, dC = dev(A), dev(B), dev(C)
    oa = F.obj_attach(dA.ptr.value, m, k, 1, m, dt=code)
    ob = F.obj_attach(dB.ptr.value, k, n, 1, k, dt=code)
    oc = F.obj_attach(dC.ptr.value, m, n, 1, m, dt=code)
    al, s1 = F.obj_constant(2.0, value_s=3.0)
    be, s2 = F.obj_constant(-1.0, value_s=5.0)
    F.lib().bli_strassen_ab_ex(al, oa, ob, be, oc, fmm("222"))
    assert np.array_equal(dC.to_numpy(), want)


def test_dropin_row_stored_host_and_levels():
    m, n, k = 120, 100, 90
    A, B, C = O.integers(m, k, 8, 0, 3), O.integers(k, n, 8, 1, 3), O.integers(m, n, 8, 2, 3)
    want = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    Ar, Br, Cr = np.ascontiguousarray(A), np.ascontiguousarray(B), np.ascontiguousarray(C)
    one, s1 = F.obj_scalar(1.0)
    levels = (F.FmmT * 2)(fmm("222"), fmm("322"))
    F.lib().bli_fmm_ab_ex(one, F.obj_attach(Ar, m, k, k, 1), F.obj_attach(Br, k, n, n, 1), one,
                          F.obj_attach(Cr, m, n, n, 1), levels, 2, F.LEAF_NATIVE)
    assert np.array_equal(Cr, want)


def test_dropin_host_beta_zero_skips_c_and_keeps_gaps():
    """beta = 0 on host objects: C is never read (NaN in C does not reach the
    result), and a C with gaps between its columns (ld > m) keeps the gap
    elements unchanged through the staging round trip."""
    m, n, k = 150, 130, 140
    A, B = O.integers(m, k, 13, 0, 3), O.integers(k, n, 13, 1, 3)
    want = A @ B
    one, s1 = F.obj_scalar(1.0)
    zero, s0 = F.obj_scalar(0.0)
    Cn = np.full((m, n), np.nan, order="F")
    F.lib().bli_strassen_ab_ex(one, F.obj_attach(A, m, k, 1, m), F.obj_attach(B, k, n, 1, k), zero,
                               F.obj_attach(Cn, m, n, 1, m), fmm("222"))
    assert np.array_equal(Cn, want)
    ldc = m + 5
    big = np.full((ldc, n), 7.0, order="F")
    big[:m, :] = np.nan
    F.lib().bli_strassen_ab_ex(one, F.obj_attach(A, m, k, 1, m), F.obj_attach(B, k, n, 1, k), zero,
                               F.obj_attach(big, m, n, 1, ldc), fmm("222"))
    assert np.array_equal(big[:m, :], want)
    assert np.all(big[m:, :] == 7.0)


def test_classical_object_entry_float():
    m, n, k = 64, 80, 96
    A, B, C = O.uniform_f32(m, k, 1, 0), O.uniform_f32(k, n, 1, 1), O.uniform_f32(m, n, 1, 2)
    Cw = C.copy(order="F")
    one, s1 = F.obj_scalar(1.0)
    F.lib().bli_gemm_hip(one, F.obj_attach(A, m, k, 1, m, F.BLIS_FLOAT), F.obj_attach(B, k, n, 1, k, F.BLIS_FLOAT),
                         one, F.obj_attach(Cw, m, n, 1, m, F.BLIS_FLOAT))
    ref = C.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
    err = O.normalized_error(Cw.astype(np.float64), ref, A.astype(np.float64), B.astype(np.float64))
    assert err < 1e-7


@pytest.mark.parametrize("name", ["222", "333"])
def test_dropin_float_objects_fmm(name):
    """bli_strassen_ab_ex on float objects (FP32 FMM through the object API):
    small-integer inputs are exact in FP32, so the result equals the oracle's
    bit for bit; uniform inputs within 1e-6 of the FP64 product of the
    FP32-rounded inputs."""
    m, n, k = 131, 97, 203
    f = fmm(name)
    og = O.parse(alg_path(name))
    A, B, C = O.integers(m, k, 8, 0, 2), O.integers(k, n, 8, 1, 2), O.integers(m, n, 8, 2, 2)
    want = O.fmm_dgemm(og, 1.0, A, B, 1.0, C.copy(order="F")).astype(np.float32)
    As, Bs, Cs = (np.asfortranarray(X.astype(np.float32)) for X in (A, B, C))
    one, s1 = F.obj_scalar(1.0)
    F.lib().bli_strassen_ab_ex(one, F.obj_attach(As, m, k, 1, m, F.BLIS_FLOAT),
                               F.obj_attach(Bs, k, n, 1, k, F.BLIS_FLOAT), one,
                               F.obj_attach(Cs, m, n, 1, m, F.BLIS_FLOAT), f)
    assert np.array_equal(Cs, want)
    U, V, W = O.uniform_f32(m, k, 9, 0), O.uniform_f32(k, n, 9, 1), O.uniform_f32(m, n, 9, 2)
    Wc = W.copy(order="F")
    F.lib().bli_strassen_ab_ex(one, F.obj_attach(U, m, k, 1, m, F.BLIS_FLOAT),
                               F.obj_attach(V, k, n, 1, k, F.BLIS_FLOAT), one,
                               F.obj_attach(Wc, m, n, 1, m, F.BLIS_FLOAT), f)
    ref = W.astype(np.float64) + U.astype(np.float64) @ V.astype(np.float64)
    assert O.normalized_error(Wc.astype(np.float64), ref, U.astype(np.float64), V.astype(np.float64)) < 1e-6


@pytest.mark.parametrize("levels", [1, 2])
def test_fp32_leaf_and_sgemm_fmm(levels):
    m = n = k = 512
    A, B, C = O.uniform(m, k, 0, 0), O.uniform(k, n, 0, 1), O.uniform(m, n, 0, 2)
    ref = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    alg = F.Alg([F.strassen()] * levels)
    got = run_fmm(alg, 1.0, A, B, 1.0, C, leaf=F.LEAF_F32)
    err = O.normalized_error(got, ref, A, B)
    say(f"fp32 leaf L{levels} err {err:.3e}")
    assert err < 1e-7
    As, Bs, Cs = A.astype(np.float32), B.astype(np.float32), C.astype(np.float32)
    dA, dB, dC = dev(np.asfortranarray(As)), dev(np.asfortranarray(Bs)), dev(np.asfortranarray(Cs))
    F.sgemm_fmm(alg, m, n, k, 1.0, dA, dB, 1.0, dC)
    got = dC.to_numpy().astype(np.float64)
    ref32 = Cs.astype(np.float64) + As.astype(np.float64) @ Bs.astype(np.float64)
    assert O.normalized_error(got, ref32, As.astype(np.float64), Bs.astype(np.float64)) < 1e-7


def test_large_integer_fmm_equals_classical_on_device():
    """Size-independent property at 4096^3: with integer inputs both paths are
    exact, so 2-level Strassen must equal the classical kernel bit for bit."""
    n = 4096
    say("large int", n)
    A, B = F.DeviceMatrix(n, n).fill_int(9, 0, 2), F.DeviceMatrix(n, n).fill_int(9, 1, 2)
    C1, C2 = F.DeviceMatrix(n, n).fill_int(9, 2, 2), F.DeviceMatrix(n, n).fill_int(9, 2, 2)
    F.dgemm(n, n, n, 1.0, A, B, 1.0, C1)
    F.dgemm_fmm(F.Alg([F.strassen()] * 2), n, n, n, 1.0, A, B, 1.0, C2)
    F.sync()
    assert F.fro_diff(C1, C2) == 0.0
    assert F.fro(C1) > 0


@pytest.mark.parametrize("m,n,k,levels", [(2048, 2048, 2048, 1), (1001, 999, 1003, 1), (1024, 768, 1536, 2)])
def test_reference_style_c_harness(tmp_path, m, n, k, levels):
    """The reference's harness flow in C (tests/c/strassen_oapi.c) on the
    plain configuration (222, C += A*B, 2048^3) and ragged shapes: host
    objects, 222.txt, alpha = beta = 1; normalised error <= 1e-14 against
    the classical product and 64 sampled entries against host dot products."""
    import os
    import subprocess
    from conftest import ROOT
    exe = tmp_path / "strassen_oapi"
    subprocess.run(["gcc", "-std=c11", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "strassen_oapi.c"), "-o", str(exe),
                    "-L", os.path.dirname(F.LIB_PATH), "-lfmm_blis_amd", "-lm",
                    f"-Wl,-rpath,{os.path.dirname(F.LIB_PATH)}"], check=True)
    r = subprocess.run([str(exe), str(m), str(n), str(k), alg_path("222"), "2", str(levels)],
                       capture_output=True, text=True, timeout=300)
    say("harness:", r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0
    cols = r.stdout.split()
    assert [int(c) for c in cols[:3]] == [m, n, k]
    assert float(cols[6]) < 1e-14


def test_large_host_copies_round_trip():
    """Host <-> device copies of 64 MiB and more go through the pinned,
    multi-threaded staging path: a round trip of an odd-sized buffer (not a
    multiple of the 16 MiB chunk or of the page) returns the same bytes, and
    a device-side product of a large uploaded operand is right."""
    nbytes = (64 << 20) + 3 * (16 << 20) + 4096 * 7 + 24
    rng = np.random.default_rng(11)
    host = rng.integers(0, 256, size=nbytes, dtype=np.uint8)
    buf = F.DeviceBuffer(nbytes)
    F.check(F.lib().fmmhip_memcpy_h2d(buf.ptr, host.ctypes.data_as(ctypes.c_void_p), nbytes), "h2d")
    back = buf.download(np.uint8, nbytes)
    assert np.array_equal(back, host)
    m = 3000                                   # 72 MB operands
    A = O.integers(m, m, 12, 0, 2)
    dA = dev(A)
    assert np.array_equal(dA.to_numpy(), A)
    x = O.integers(m, 1, 12, 1, 2)
    dx, dy = dev(x), dev(np.zeros((m, 1)))
    F.dgemm(m, 1, m, 1.0, dA, dx, 0.0, dy)
    F.sync()
    assert np.array_equal(dy.to_numpy(), A @ x)


def test_concurrent_host_threads_on_one_stream():
    """Four host threads issue pre-summed and product-split FMM calls of
    growing sizes on the library's default stream at once (each call may
    regrow the stream's workspace): every result equals the oracle's FMM
    bit for bit on integer inputs."""
    import threading
    og = O.parse(alg_path("222"))
    og2 = O.compose(og, og)
    cases = []
    for t in range(4):
        for j, (m, n, k) in enumerate([(130, 120, 110), (260, 250, 240), (520, 500, 480)]):
            A, B, C = (O.integers(m, k, 20 + t, 3 * j, 3), O.integers(k, n, 20 + t, 3 * j + 1, 3),
                       O.integers(m, n, 20 + t, 3 * j + 2, 3))
            cases.append((t, A, B, C, dev(A), dev(B), dev(C)))
    errors = []

    def worker(t):
        try:
            a1 = F.Alg([F.strassen()])
            a2 = F.Alg([F.strassen()] * 2)
            a2.set_split(4 if t % 2 else -1)
            for (tt, A, B, C, dA, dB, dC) in cases:
                if tt != t:
                    continue
                m, k = A.shape
                n = B.shape[1]
                F.dgemm_fmm(a2 if m > 200 else a1, m, n, k, 1.0, dA, dB, 1.0, dC)
        except Exception as e:          # reported below, on the main thread
            errors.append(e)

    ths = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    F.sync()
    assert not errors, errors
    for (t, A, B, C, dA, dB, dC) in cases:
        og_used = og2 if A.shape[0] > 200 else og
        want = O.fmm_dgemm(og_used, 1.0, A, B, 1.0, C.copy(order="F"))
        assert np.array_equal(dC.to_numpy(), want), (t, A.shape)


@pytest.mark.parametrize("name", ["333", "444"])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, True), (True, False)])
def test_two_stage_sums_exact_any_layout(name, ta, tb):
    """Two-level 333 and 444 (81 and 256 composed parts per operand) form
    their operand sums in two stages (inner level per outer block, then the
    outer level): exact against the oracle's FMM on integer inputs, with
    row- or column-major operands and uneven parts, in both pre-summed
    modes."""
    og = O.parse(alg_path(name))
    og = O.compose(og, og)
    m, n, k = (163, 151, 145) if name == "333" else (150, 133, 141)
    A, B, C = O.integers(m, k, 9, 0, 3), O.integers(k, n, 9, 1, 3), O.integers(m, n, 9, 2, 3)
    want = O.fmm_dgemm(og, 1.0, A, B, -2.0, C.copy(order="F"))
    rsA, csA = (k, 1) if ta else (1, m)
    rsB, csB = (n, 1) if tb else (1, k)
    for pack in (1, 2):
        alg = F.Alg([fmm(name)] * 2)
        alg.set_pack(pack)
        dA = dev(np.asfortranarray(A.T)) if ta else dev(A)
        dB = dev(np.asfortranarray(B.T)) if tb else dev(B)
        dC = dev(C)
        say("two-stage", name, ta, tb, pack)
        F.check(F.lib().fmmhip_dgemm_fmm(alg.h, 0, m, n, k, 1.0, dA.ptr, rsA, csA, dB.ptr, rsB, csB, -2.0,
                                         dC.ptr, 1, m, None), "dgemm_fmm two-stage")
        F.sync()
        assert np.array_equal(dC.to_numpy(), want), pack
