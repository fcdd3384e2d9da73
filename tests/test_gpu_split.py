"""GPU parity of the product split (fmmhip_alg_set_split): G workgroups per
C tile, each taking a contiguous range of the R products (of the k-steps
for the classical algorithm) and scattering into its own zeroed temporary
of C, then C := beta*C + T_0 + ... + T_{G-1}.

Bars (stated per test):
  * integer-valued inputs: every split (1, 2, 4, 8), pack mode, level count
    and C layout equals the oracle's FMM bit for bit, beta applied once;
  * uniform[-1,1] inputs: normalised error against the oracle's classical
    product below the same per-level bounds as the unsplit kernel, and a
    repeated call reproduces the first bit for bit (the reduction order is
    fixed);
  * beta == 0 never reads C (NaN in C does not propagate);
  * FP32 operands: the split result is within 1e-7 of the product of the
    FP32-rounded inputs.
"""
import sys

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O
from conftest import alg_path

pytestmark = pytest.mark.gpu

TOL_FP64 = {1: 1e-15, 2: 5e-15}
SPLITS = [1, 2, 4, 8]


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


def alg(name, levels, pack, split):
    f = F.new_fmm(alg_path(name))
    a = F.Alg([f] * levels)
    a.set_pack(pack)
    a.set_split(split)
    return a


def oracle_alg(name, levels):
    og = O.parse(alg_path(name))
    return og if levels == 1 else O.compose(og, og)


@pytest.mark.parametrize("name,levels", [("222", 1), ("222", 2), ("333", 1), ("232", 2)])
@pytest.mark.parametrize("pack", [0, 1])
def test_split_exact_integers(name, levels, pack):
    og = oracle_alg(name, levels)
    for (m, n, k) in [(97, 101, 103), (300, 280, 36), (17, 1, 19), (258, 260, 262)]:
        A, B, C = O.integers(m, k, 9, 0, 4), O.integers(k, n, 9, 1, 4), O.integers(m, n, 9, 2, 4)
        for alpha, beta in [(1.0, 1.0), (2.0, -3.0), (0.5, 0.0)]:
            want = O.fmm_dgemm(og, alpha, A, B, beta, C.copy(order="F"))
            for g in SPLITS:
                say("split int", name, levels, pack, g, m, n, k, alpha, beta)
                a = alg(name, levels, pack, g)
                dA, dB, dC = dev(A), dev(B), dev(C)
                F.dgemm_fmm(a, m, n, k, alpha, dA, dB, beta, dC)
                F.sync()
                assert np.array_equal(dC.to_numpy(), want), (g, m, n, k, alpha, beta)


@pytest.mark.parametrize("g", [2, 8])
def test_split_row_major_c(g):
    """C stored by rows: the reduction walks C's contiguous dimension."""
    m, n, k = 150, 170, 130
    A, B, C = O.integers(m, k, 4, 0, 3), O.integers(k, n, 4, 1, 3), O.integers(m, n, 4, 2, 3)
    want = O.dgemm(1.0, A, B, -1.0, C.copy(order="F"))
    a = alg("232", 1, 1, g)
    dA, dB, dC = dev(A), dev(B), dev(np.asfortranarray(C.T))
    F.check(F.lib().fmmhip_dgemm_fmm(a.h, 0, m, n, k, 1.0, dA.ptr, 1, m, dB.ptr, 1, k, -1.0, dC.ptr, n, 1, None),
            "dgemm_fmm row-major C")
    F.sync()
    assert np.array_equal(dC.to_numpy().T, want)


@pytest.mark.parametrize("levels", [1, 2])
def test_split_random_error_and_repeatable(levels):
    m = n = k = 432
    A, B, C = O.uniform(m, k, 7, 0), O.uniform(k, n, 7, 1), O.uniform(m, n, 7, 2)
    ref = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    dA, dB = dev(A), dev(B)
    for g in SPLITS:
        a = alg("222", levels, -1, g)
        outs = []
        for _ in range(2):
            dC = dev(C)
            F.dgemm_fmm(a, m, n, k, 1.0, dA, dB, 1.0, dC)
            F.sync()
            outs.append(dC.to_numpy())
        err = O.normalized_error(outs[0], ref, A, B)
        say(f"split {g} L{levels} normalised error {err:.3e}")
        assert err < TOL_FP64[levels]
        assert np.array_equal(outs[0], outs[1]), g


def test_split_beta_zero_ignores_nan():
    m, n, k = 70, 90, 50
    A, B = O.uniform(m, k, 2, 0), O.uniform(k, n, 2, 1)
    C = np.full((m, n), np.nan, order="F")
    for g in (2, 4):
        a = alg("222", 1, 0, g)
        dA, dB, dC = dev(A), dev(B), dev(C)
        F.dgemm_fmm(a, m, n, k, 1.0, dA, dB, 0.0, dC)
        F.sync()
        got = dC.to_numpy()
        assert np.all(np.isfinite(got))
        assert O.normalized_error(got, A @ B, A, B) < 1e-15


def test_split_fp32():
    m = n = k = 384
    A, B, C = (O.uniform(d0, d1, 3, i).astype(np.float32) for i, (d0, d1) in enumerate([(m, k), (k, n), (m, n)]))
    A, B, C = np.asfortranarray(A), np.asfortranarray(B), np.asfortranarray(C)
    ref = C.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
    for g in (1, 4):
        a = alg("222", 2, 1, g)
        dA, dB, dC = dev(A), dev(B), dev(C)
        F.sgemm_fmm(a, m, n, k, 1.0, dA, dB, 1.0, dC)
        F.sync()
        err = O.normalized_error(dC.to_numpy().astype(np.float64), ref, A.astype(np.float64), B.astype(np.float64))
        say(f"fp32 split {g} err {err:.3e}")
        assert err < 1e-7


def test_classical_k_split_exact_and_repeatable():
    """The classical algorithm (R = 1) splits the k-steps instead: G groups
    of at least 8 k-steps each, zeroed temporaries and the ordered reduction.
    Integer inputs are exact for every G; uniform inputs stay within 1e-15
    and a repeated call is bit-identical; beta = 0 does not read C."""
    m, n, k = 300, 260, 1100                   # 69 k-steps: G up to 8
    A, B, C = O.integers(m, k, 31, 0, 4), O.integers(k, n, 31, 1, 4), O.integers(m, n, 31, 2, 4)
    for alpha, beta in [(1.0, 1.0), (2.0, -3.0), (0.5, 0.0)]:
        want = O.dgemm(alpha, A, B, beta, C.copy(order="F"))
        Cn = C.copy(order="F")
        if beta == 0.0:
            Cn[:] = np.nan
        for g in (1, 2, 4, 8):
            a = F.Alg(classical=True)
            a.set_split(g)
            dA, dB, dC = dev(A), dev(B), dev(Cn)
            F.dgemm_fmm(a, m, n, k, alpha, dA, dB, beta, dC)
            F.sync()
            say("classical k-split", g, alpha, beta)
            assert np.array_equal(dC.to_numpy(), want), (g, alpha, beta)
    U, V, W = O.uniform(m, k, 32, 0), O.uniform(k, n, 32, 1), O.uniform(m, n, 32, 2)
    ref = O.dgemm(1.0, U, V, 1.0, W.copy(order="F"))
    a = F.Alg(classical=True)
    a.set_split(4)
    outs = []
    for _ in range(2):
        dU, dV, dW = dev(U), dev(V), dev(W)
        F.dgemm_fmm(a, m, n, k, 1.0, dU, dV, 1.0, dW)
        F.sync()
        outs.append(dW.to_numpy())
    assert O.normalized_error(outs[0], ref, U, V) < 1e-15
    assert np.array_equal(outs[0], outs[1])


def test_classical_k_split_fp32_and_other_layouts():
    """FP32 operands take the k-split too; a row-stored C does not (the
    split is prepared for column-major operands only) and is still exact."""
    m, n, k = 200, 180, 900
    A, B, C = (np.asfortranarray(X.astype(np.float32)) for X in
               (O.integers(m, k, 33, 0, 2), O.integers(k, n, 33, 1, 2), O.integers(m, n, 33, 2, 2)))
    want = (C.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)).astype(np.float32)
    a = F.Alg(classical=True)
    a.set_split(8)
    dA, dB, dC = dev(A), dev(B), dev(C)
    F.sgemm_fmm(a, m, n, k, 1.0, dA, dB, 1.0, dC)
    F.sync()
    assert np.array_equal(dC.to_numpy(), want)
    A64, B64, C64 = (X.astype(np.float64) for X in (A, B, C))
    want64 = O.dgemm(1.0, A64, B64, 1.0, C64.copy(order="F"))
    a.set_split(4)
    dA, dB, dC = dev(np.asfortranarray(A64)), dev(np.asfortranarray(B64)), dev(np.asfortranarray(C64.T))
    F.check(F.lib().fmmhip_dgemm_fmm(a.h, 0, m, n, k, 1.0, dA.ptr, 1, m, dB.ptr, 1, k, 1.0, dC.ptr, n, 1, None),
            "dgemm_fmm row-major C")
    F.sync()
    assert np.array_equal(dC.to_numpy().T, want64)
