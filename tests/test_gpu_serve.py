"""GPU tests of the serving form (fmmhip_server_*): continuous batching of
GEMM requests over device-resident operands, admitted by HBM budget onto
several streams.

Bars (stated per test):
  * each result is the same alone and in the mix: a request's C equals, bit
    for bit, the same product run alone through the single-call entry with
    the candidate the server chose (the kernels are deterministic and a
    request's tiles never depend on other requests);
  * against the oracle: normalised error below 5e-15 (FP64) / 1e-6 (FP32);
  * admission: with a budget that holds only a few requests at once, all
    complete, the arena peak stays within the budget and requests overlap;
  * a request larger than the budget fails with FMMHIP_ERR_NOMEM alone;
  * a request is not completed before its arrival time.
"""
import ctypes
import sys

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O

pytestmark = pytest.mark.gpu

NOMEM = 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    rc = F.lib().fmmhip_init()
    if rc != 0:
        pytest.fail(f"no usable gfx950 device: {F.STATUS.get(rc)}")
    yield


def dev(X):
    D = F.DeviceMatrix(X.shape[0], X.shape[1], dtype=X.dtype)
    D.upload(X)
    return D


def fetch(ptr, m, n, ld, dtype):
    out = np.empty(ld * n, dtype=dtype)
    F.check(F.lib().fmmhip_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes), "d2h")
    return out.reshape(n, ld)[:, :m].T.copy()


def alone(cand, dt, m, n, k, alpha, dA, dB, beta, dC):
    """The single-call equivalent of a served request: the server's FMM
    candidates run with the products materialised, which gives the same C
    as pre-summed without product split; the classical one keeps its
    per-call k-split choice."""
    alg = F.Alg(classical=True) if cand == 0 else F.Alg([F.strassen()] * cand)
    if cand > 0:
        alg.set_pack(1)
        alg.set_split(1)
    (F.dgemm_fmm if dt == "f8" else F.sgemm_fmm)(alg, m, n, k, alpha, dA, dB, beta, dC)
    F.sync()
    return dC.to_numpy()


MIX = [("f8", 300, 280, 260, 2), ("f4", 300, 280, 260, 1), ("f8", 1024, 1024, 1024, -1), ("f4", 640, 512, 768, -1),
       ("f8", 129, 65, 1000, 1), ("f4", 33, 1, 17, 0), ("f8", 1, 1, 1, -1), ("f8", 2048, 1536, 1024, 2),
       ("f4", 1536, 2048, 1024, 2), ("f8", 700, 900, 64, 0), ("f8", 96, 96, 1, 2), ("f4", 512, 512, 512, -1)]


def make_inputs(dt, m, n, k, i):
    A, B = O.uniform(m, k, 21, 2 * i), O.uniform(k, n, 21, 2 * i + 1)
    if dt == "f4":
        A, B = np.asfortranarray(A.astype(np.float32)), np.asfortranarray(B.astype(np.float32))
    return A, B


def test_serve_results_same_alone_and_in_mix():
    srv = F.Server(2 << 30, slots=4)
    reqs = []
    for i, (dt, m, n, k, cand) in enumerate(MIX):
        A, B = make_inputs(dt, m, n, k, i)
        dA, dB = dev(A), dev(B)
        rid = srv.submit(m, n, k, 1.0, dA, dB, cand=cand, keep=1)
        reqs.append((rid, dt, m, n, k, A, B, dA, dB))
    srv.drain()
    st = srv.stats()
    say("stats", st)
    assert st["completed"] == len(MIX) and st["failed"] == 0
    for rid, dt, m, n, k, A, B, dA, dB in reqs:
        status, cptr, ldc, lat, cand = srv.result(rid)
        assert status == 0
        got = fetch(cptr, m, n, ldc, np.float64 if dt == "f8" else np.float32)
        ref = O.dgemm(1.0, A.astype(np.float64), B.astype(np.float64), 0.0, np.zeros((m, n), order="F"))
        err = O.normalized_error(got.astype(np.float64), ref, A.astype(np.float64), B.astype(np.float64))
        assert err < (5e-15 if dt == "f8" else 1e-6), (dt, m, n, k, cand, err)
        dC = F.DeviceMatrix(m, n, dtype=dt)
        solo = alone(cand, dt, m, n, k, 1.0, dA, dB, 0.0, dC)
        assert np.array_equal(solo, got), (dt, m, n, k, cand)
        srv.release(rid)
    srv.close()


def test_serve_user_c_with_beta():
    srv = F.Server(256 << 20, slots=2)
    m, n, k = 257, 199, 300
    A, B, C = O.uniform(m, k, 3, 0), O.uniform(k, n, 3, 1), O.uniform(m, n, 3, 2)
    dA, dB, dC = dev(A), dev(B), dev(C)
    rid = srv.submit(m, n, k, 2.0, dA, dB, beta=-0.5, C=dC, cand=1)
    assert srv.wait(rid) == 0
    got = dC.to_numpy()
    dC2 = dev(C)
    assert np.array_equal(alone(1, "f8", m, n, k, 2.0, dA, dB, -0.5, dC2), got)
    srv.close()


def test_serve_budget_admission_and_overlap():
    """12 requests of 1024^3 FP64 with two-level Strassen need about 85 MB
    each (output 8 MB, lists, workspace: the pre-summed leaves 51 MB and the
    49 product slots 26 MB); a 200 MiB budget holds two at a time."""
    m = n = k = 1024
    A, B = make_inputs("f8", m, n, k, 40)
    dA, dB = dev(A), dev(B)
    budget = 200 << 20
    srv = F.Server(budget, slots=4)
    ids = [srv.submit(m, n, k, 1.0, dA, dB, cand=2, keep=0) for _ in range(12)]
    srv.drain()
    st = srv.stats()
    say("stats", st)
    assert st["completed"] == 12 and st["failed"] == 0
    assert st["arena_peak"] <= budget
    assert st["max_inflight"] >= 2
    for rid in ids:
        assert srv.result(rid)[0] == 0
    srv.close()


def test_serve_oversize_request_fails_alone():
    srv = F.Server(8 << 20, slots=2)
    small = make_inputs("f8", 64, 64, 64, 50)
    big = make_inputs("f8", 1024, 1024, 64, 51)
    sA, sB, bA, bB = dev(small[0]), dev(small[1]), dev(big[0]), dev(big[1])
    r1 = srv.submit(64, 64, 64, 1.0, sA, sB, cand=0)
    r2 = srv.submit(1024, 1024, 64, 1.0, bA, bB, cand=0)          # C alone is 8 MiB
    r3 = srv.submit(64, 64, 64, 1.0, sA, sB, cand=1)
    assert srv.wait(r1) == 0
    assert srv.wait(r2) == NOMEM
    assert srv.wait(r3) == 0
    srv.close()


def test_serve_arrival_time_respected():
    srv = F.Server(64 << 20, slots=2)
    A, B = make_inputs("f8", 128, 128, 128, 60)
    dA, dB = dev(A), dev(B)
    t_arr = srv.now() + 0.3
    rid = srv.submit(128, 128, 128, 1.0, dA, dB, cand=0, arrival_s=t_arr)
    assert srv.wait(rid) == 0
    done_at = srv.now()
    status, _, _, lat, _ = srv.result(rid)
    assert done_at >= t_arr and 0.0 <= lat < 1000.0
    srv.close()
