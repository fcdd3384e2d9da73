"""GPU parity of the products-materialised pack mode (fmmhip_alg_set_pack 2):
the MFMA kernel stores every product M_r once into a slot array and a W
stage forms each C part as beta*C + sum_r alpha*W[c][r]*M_r, in ascending
product order, the FMA chain of the fused epilogue.

Bars (stated per test):
  * integer-valued inputs: equal to the oracle's FMM bit for bit (every
    algorithm and level, uneven shapes, row-major C, alpha/beta);
  * uniform[-1,1] inputs: equal to the pre-summed mode (pack 1, no split)
    bit for bit, for every group count of the product split, FP64, FP32
    operands and the FP32 leaf -- the two modes run the same sums in the
    same order -- except two-level 333 and 444, whose W stage runs in two
    stages (inner level, then outer): there within the per-level error
    bound, and exact on integers;
  * beta == 0 never reads C (NaN in C does not propagate).
"""
import sys

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O
from conftest import ALG_NAMES, alg_path

pytestmark = pytest.mark.gpu


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


def alg(name, levels, pack, split=1):
    a = F.Alg([F.new_fmm(alg_path(name))] * levels)
    a.set_pack(pack)
    a.set_split(split)
    return a


def run(a, alpha, A, B, beta, C, leaf=F.LEAF_NATIVE):
    m, k = A.shape
    n = B.shape[1]
    dA, dB, dC = dev(A), dev(B), dev(C)
    if A.dtype == np.float32:
        F.sgemm_fmm(a, m, n, k, alpha, dA, dB, beta, dC)
    else:
        F.dgemm_fmm(a, m, n, k, alpha, dA, dB, beta, dC, leaf=leaf)
    F.sync()
    return dC.to_numpy()


def size_for(name, levels):
    return {("444", 2): 150, ("333", 2): 200}.get((name, levels), 333)


TWO_STAGE_W = {("333", 2), ("444", 2)}


@pytest.mark.parametrize("name", ALG_NAMES)
@pytest.mark.parametrize("levels", [1, 2])
def test_wmat_equals_presum_bitwise(name, levels):
    n0 = size_for(name, levels)
    m, n, k = n0, n0 + 7, n0 - 5
    A, B, C = O.uniform(m, k, 41, 0), O.uniform(k, n, 41, 1), O.uniform(m, n, 41, 2)
    want = run(alg(name, levels, 1), 0.7, A, B, -1.3, C)
    got = run(alg(name, levels, 2, -1), 0.7, A, B, -1.3, C)
    say("wmat vs presum", name, levels, m, n, k)
    ref = O.dgemm(0.7, A, B, -1.3, C.copy(order="F"))
    if (name, levels) in TWO_STAGE_W:
        # the W sums re-associated: both within the bound, and close
        assert O.normalized_error(want, ref, A, B) < 5e-15
        assert O.normalized_error(got, want, A, B) < 5e-15
    else:
        assert np.array_equal(got, want)
    assert O.normalized_error(got, ref, A, B) < 5e-15


@pytest.mark.parametrize("groups", [1, 2, 4, 8])
def test_wmat_groups_do_not_change_results(groups):
    m = n = k = 520
    A, B, C = O.uniform(m, k, 42, 0), O.uniform(k, n, 42, 1), O.uniform(m, n, 42, 2)
    want = run(alg("222", 2, 1), 1.0, A, B, 1.0, C)
    got = run(alg("222", 2, 2, groups), 1.0, A, B, 1.0, C)
    assert np.array_equal(got, want), groups


def test_wmat_exact_integers_and_row_major_c():
    og = O.parse(alg_path("322"))
    og2 = O.compose(og, og)
    for (m, n, k) in [(97, 101, 103), (17, 1, 19), (300, 280, 36)]:
        A, B, C = O.integers(m, k, 43, 0, 4), O.integers(k, n, 43, 1, 4), O.integers(m, n, 43, 2, 4)
        for alpha, beta in [(1.0, 1.0), (2.0, -3.0), (0.5, 0.0)]:
            want = O.fmm_dgemm(og2, alpha, A, B, beta, C.copy(order="F"))
            say("wmat int", m, n, k, alpha, beta)
            assert np.array_equal(run(alg("322", 2, 2, -1), alpha, A, B, beta, C), want)
            # the same call with C stored by rows (slots stored by rows too)
            a = alg("322", 2, 2, -1)
            dA, dB, dC = dev(A), dev(B), dev(np.asfortranarray(C.T))
            F.check(F.lib().fmmhip_dgemm_fmm(a.h, 0, m, n, k, alpha, dA.ptr, 1, m, dB.ptr, 1, k, beta, dC.ptr, n, 1,
                                             None), "dgemm_fmm row-major C")
            F.sync()
            assert np.array_equal(dC.to_numpy().T, want)


def test_wmat_beta_zero_ignores_nan():
    m, n, k = 130, 90, 70
    A, B = O.uniform(m, k, 44, 0), O.uniform(k, n, 44, 1)
    C = np.full((m, n), np.nan, order="F")
    got = run(alg("333", 1, 2, -1), 1.0, A, B, 0.0, C)
    assert np.all(np.isfinite(got))
    assert O.normalized_error(got, A @ B, A, B) < 1e-15


def test_wmat_fp32_operands_and_fp32_leaf():
    m = n = k = 300
    A, B, C = (np.asfortranarray(O.uniform(d0, d1, 45, i).astype(np.float32))
               for i, (d0, d1) in enumerate([(m, k), (k, n), (m, n)]))
    want = run(alg("222", 2, 1), 1.0, A, B, 1.0, C)
    got = run(alg("222", 2, 2, -1), 1.0, A, B, 1.0, C)
    assert got.dtype == np.float32
    assert np.array_equal(got, want)
    A64, B64, C64 = (np.asfortranarray(X.astype(np.float64)) for X in (A, B, C))
    want = run(alg("232", 1, 1), 1.0, A64, B64, 1.0, C64, leaf=F.LEAF_F32)
    got = run(alg("232", 1, 2, -1), 1.0, A64, B64, 1.0, C64, leaf=F.LEAF_F32)
    assert np.array_equal(got, want)


def test_pack_for_reports_the_mode():
    """fmmhip_alg_pack_for: the explicit mode as set (2 falls back to 1 for
    the classical algorithm, which has one product), and the automatic one
    within {0, 1, 2}: fused for the classical algorithm, pre-summed or
    materialised for multi-part ones."""
    a = F.Alg([F.strassen()] * 2)
    for p in (0, 1, 2):
        a.set_pack(p)
        assert a.pack_for(1000, 900, 800) == p
    a.set_pack(-1)
    assert a.pack_for(4096, 4096, 4096) in (1, 2)
    c = F.Alg(classical=True)
    assert c.pack_for(4096, 4096, 4096) == 0
    c.set_pack(2)
    assert c.pack_for(4096, 4096, 4096) == 1


@pytest.mark.parametrize("name,levels", [("222", 1), ("222", 2), ("333", 1)])
@pytest.mark.parametrize("leaf", [F.LEAF_BF16X3, F.LEAF_BF16X3_8, F.LEAF_BF16X3_6])
def test_bf16x3_materialised_equals_in_kernel_epilogue(name, levels, leaf):
    """BF16x3 leaf: with the products materialised (pack 2, and -1 by
    default) the kernel only stores each product and the W stage applies W;
    the result equals the in-kernel W epilogue (pack 1, no split) bit for
    bit, whatever the group count."""
    m, n, k = 410, 397, 388
    A, B, C = O.uniform(m, k, 46, 0), O.uniform(k, n, 46, 1), O.uniform(m, n, 46, 2)
    want = run(alg(name, levels, 1), 0.9, A, B, 1.1, C, leaf=leaf)
    for groups in (-1, 1, 4):
        a = alg(name, levels, 2, groups)
        assert a.pack_for(m, n, k, leaf) == 2
        got = run(a, 0.9, A, B, 1.1, C, leaf=leaf)
        say("bf16x3 wmat", name, levels, leaf, groups)
        assert np.array_equal(got, want), groups
    assert alg(name, levels, 1).pack_for(m, n, k, leaf) == 1


@pytest.mark.parametrize("leaf", [F.LEAF_NATIVE, F.LEAF_BF16X3])
def test_wmat_leaves_gaps_of_c_untouched(leaf):
    """C with a leading dimension above m: the W stage writes the parts'
    elements only, so the gap rows keep their contents; the result is the
    oracle's FMM bit for bit on small integers (FP64 and BF16x3 leaves)."""
    m, n, k, ldc = 157, 149, 133, 168
    og = O.parse(alg_path("222"))
    og = O.compose(og, og)
    A, B, C = O.integers(m, k, 47, 0, 3), O.integers(k, n, 47, 1, 3), O.integers(m, n, 47, 2, 3)
    want = O.fmm_dgemm(og, 1.0, A, B, -1.0, C.copy(order="F"))
    full = np.full((n, ldc), 7.25)
    full[:, :m] = C.T
    dC = F.DeviceMatrix(m, n, ld=ldc)
    dC.buf.upload(full)
    F.dgemm_fmm(alg("222", 2, 2, -1), m, n, k, 1.0, dev(A), dev(B), -1.0, dC, leaf=leaf)
    F.sync()
    got = dC.buf.download(np.float64, ldc * n).reshape(n, ldc)
    assert np.all(got[:, m:] == 7.25)
    assert np.array_equal(got[:, :m].T, want)


@pytest.mark.parametrize("name", ["333", "444"])
def test_two_stage_w_exact_integers_row_major_c_and_nan(name):
    """Two-stage W stage (333 and 444 at two levels): exact against the
    oracle's FMM on small integers with C stored by columns and by rows,
    alpha/beta pairs, and beta = 0 with NaN in C."""
    og = O.parse(alg_path(name))
    og = O.compose(og, og)
    m, n, k = (171, 158, 149) if name == "333" else (150, 161, 139)
    A, B, C = O.integers(m, k, 48, 0, 3), O.integers(k, n, 48, 1, 3), O.integers(m, n, 48, 2, 3)
    for alpha, beta in [(1.0, 1.0), (2.0, -3.0), (0.5, 0.0)]:
        want = O.fmm_dgemm(og, alpha, A, B, beta, C.copy(order="F"))
        Cn = C.copy(order="F")
        if beta == 0.0:
            Cn[:] = np.nan
        say("two-stage W", name, alpha, beta)
        assert np.array_equal(run(alg(name, 2, 2, -1), alpha, A, B, beta, Cn), want)
        a = alg(name, 2, 2, -1)
        dA, dB, dC = dev(A), dev(B), dev(np.asfortranarray(Cn.T))
        F.check(F.lib().fmmhip_dgemm_fmm(a.h, 0, m, n, k, alpha, dA.ptr, 1, m, dB.ptr, 1, k, beta, dC.ptr, n, 1,
                                         None), "dgemm_fmm row-major C")
        F.sync()
        assert np.array_equal(dC.to_numpy().T, want)


@pytest.mark.parametrize("pack", [1, 2])
def test_two_stage_paths_fp32_operands_exact(pack):
    """FP32 operands through the two-stage summation (both pre-summed
    modes) and the two-stage W stage (materialised), 333 at two levels:
    small integers keep every sum exact in FP32, so C equals the oracle's
    FMM bit for bit."""
    og = O.parse(alg_path("333"))
    og = O.compose(og, og)
    m, n, k = 170, 163, 155
    A, B, C = O.integers(m, k, 49, 0, 2), O.integers(k, n, 49, 1, 2), O.integers(m, n, 49, 2, 2)
    want = O.fmm_dgemm(og, 1.0, A, B, 1.0, C.copy(order="F")).astype(np.float32)
    A32, B32, C32 = (np.asfortranarray(X.astype(np.float32)) for X in (A, B, C))
    got = run(alg("333", 2, pack, -1), 1.0, A32, B32, 1.0, C32)
    assert got.dtype == np.float32
    assert np.array_equal(got, want)
