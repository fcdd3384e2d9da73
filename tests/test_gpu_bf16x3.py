"""GPU parity of the split BF16 leaves: FP64 operands split into three or
four bf16 planes with per-(row, 256-k chunk) power-of-two scales, plane
products on the bf16 MFMA with FP32 accumulation, FP64 recombination and
FP64 W write-back.  Leaves: FMMHIP_LEAF_BF16X3 (four planes, ten plane
products), FMMHIP_LEAF_BF16X3_8 (three planes, eight) and
FMMHIP_LEAF_BF16X3_6 (three planes, six).

Bars (stated per test; the product split is covered by the split tests):
  * small-integer inputs are split exactly and every chunk sum is an exact
    FP32 integer, so the result equals the oracle's bit for bit, for the
    classical leaf and for 1- and 2-level FMM, on ragged shapes;
  * uniform[-1,1] inputs at 1024^3: normalised error
    ||C - C_ref||_F/(||A||_F ||B||_F) against the oracle's FP64 classical
    product below the bounds in TOL (about 1.5x the error of a float64
    restatement of the split scheme on the same inputs, with the MFMA's
    FP32 rounding after every 8 k-terms of a product as measured on the
    device: 10 terms 1.9e-11 classical, 5.1e-11 one-level and 1.3e-10
    two-level Strassen; 8 terms 2.5e-10, 6.6e-10, 1.7e-9; 6 terms 4.1e-10,
    1.3e-9, 4.3e-9; the error falls as 1/sqrt(k));
  * the same with rows of A and columns of B scaled by 10^U(-4,4) (768^3;
    restatement 10 terms 3.6e-11 classical, 1.4e-10 two-level; 8 terms
    4.6e-10, 2.1e-9; 6 terms 1.0e-9, 5.3e-9), and elementwise for the
    classical leaf relative to the row/column scales (restatement 2.4e-9 /
    2.5e-8 / 5.9e-8);
  * the precision target at full size: the classical algorithm and one-
    and two-level Strassen with the FMMHIP_LEAF_BF16X3 leaf at 8192^3 and
    16384^3, uniform and stress inputs, within 1e-10 of the FP64 classical
    kernel (which the parity tests pin against the oracle);
  * a NaN in A makes exactly its row of C NaN.
"""
import sys

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O
from conftest import alg_path

pytestmark = pytest.mark.gpu

LEAVES = {10: F.LEAF_BF16X3, 8: F.LEAF_BF16X3_8, 6: F.LEAF_BF16X3_6}
TOL = {  # (terms, algorithm) -> bound
    (10, "classical"): 3e-11, (10, "222x1"): 7.5e-11, (10, "222x2"): 2e-10,
    (8, "classical"): 4e-10, (8, "222x1"): 1e-9, (8, "222x2"): 2.6e-9,
    (6, "classical"): 6.5e-10, (6, "222x1"): 2e-9, (6, "222x2"): 6.5e-9,
}
TOL_STRESS = {(10, "classical"): 5.5e-11, (10, "222x2"): 2.2e-10, (8, "classical"): 7e-10, (8, "222x2"): 3.2e-9,
              (6, "classical"): 1.6e-9, (6, "222x2"): 8e-9}
TOL_ELEM = {10: 4e-9, 8: 4e-8, 6: 9e-8}
TARGET = 1e-10     # the precision form's bound against the FP64 product


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


def run(alg, alpha, A, B, beta, C, terms=8):
    """terms: 10, 8 or 6 (the leaf with that many plane products)."""
    m, k = A.shape
    n = B.shape[1]
    dA, dB, dC = dev(A), dev(B), dev(C)
    F.dgemm_fmm(alg, m, n, k, alpha, dA, dB, beta, dC, leaf=LEAVES[terms])
    F.sync()
    return dC.to_numpy()


def algs():
    s = F.strassen()
    return {"classical": (F.Alg(classical=True), None),
            "222x1": (F.Alg([s]), O.parse(alg_path("222"))),
            "222x2": (F.Alg([s, s]), O.compose(O.parse(alg_path("222")), O.parse(alg_path("222")))),
            "333x1": (F.Alg([F.new_fmm(alg_path("333"))]), O.parse(alg_path("333")))}


SHAPES = [(1, 1, 1), (3, 13, 7), (96, 96, 1), (129, 130, 257), (300, 280, 600), (64, 512, 300)]


@pytest.mark.parametrize("terms", [10, 8, 6])
@pytest.mark.parametrize("name", ["classical", "222x1", "222x2", "333x1"])
def test_bf16x3_exact_integers(name, terms):
    alg, og = algs()[name]
    for (m, n, k) in SHAPES:
        say("bf16x3 int", name, terms, m, n, k)
        A, B, C = O.integers(m, k, 5, 0, 4), O.integers(k, n, 5, 1, 4), O.integers(m, n, 5, 2, 4)
        for alpha, beta in [(1.0, 1.0), (-2.0, 0.5), (1.0, 0.0)]:
            want = O.dgemm(alpha, A, B, beta, C.copy(order="F")) if og is None else \
                O.fmm_dgemm(og, alpha, A, B, beta, C.copy(order="F"))
            got = run(alg, alpha, A, B, beta, C, terms)
            assert np.array_equal(got, want), (m, n, k, alpha, beta)


@pytest.mark.parametrize("terms", [10, 8, 6])
@pytest.mark.parametrize("name", ["classical", "222x1", "222x2"])
def test_bf16x3_random_error(name, terms):
    alg, _ = algs()[name]
    m = n = k = 1024
    A, B, C = O.uniform(m, k, 0, 0), O.uniform(k, n, 0, 1), O.uniform(m, n, 0, 2)
    ref = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    got = run(alg, 1.0, A, B, 1.0, C, terms)
    err = O.normalized_error(got, ref, A, B)
    say(f"bf16x3 {name} {terms} terms 1024^3 normalised error {err:.3e}")
    assert err < TOL[(terms, name)]


@pytest.mark.parametrize("terms", [10, 8, 6])
@pytest.mark.parametrize("name", ["classical", "222x2"])
def test_bf16x3_scaled_rows_and_columns(name, terms):
    """Stress variant: rows of A and columns of B scaled by 10^U(-4,4)."""
    alg, _ = algs()[name]
    m = n = k = 768
    rng = np.random.default_rng(7)
    A = np.asfortranarray(O.uniform(m, k, 1, 0) * (10.0 ** rng.uniform(-4, 4, size=(m, 1))))
    B = np.asfortranarray(O.uniform(k, n, 1, 1) * (10.0 ** rng.uniform(-4, 4, size=(1, n))))
    C = np.zeros((m, n), order="F")
    ref = O.dgemm(1.0, A, B, 0.0, C.copy(order="F"))
    got = run(alg, 1.0, A, B, 0.0, C, terms)
    err = O.normalized_error(got, ref, A, B)
    say(f"bf16x3 {name} {terms} terms stress normalised error {err:.3e}")
    assert err < TOL_STRESS[(terms, name)]
    if name == "classical":
        # elementwise, relative to the row/column scales: per-row scaling keeps
        # small rows as accurate as large ones (an FMM mixes rows of different
        # scales in its sums, so this holds for the classical leaf only)
        scale = np.abs(A).max(axis=1)[:, None] * np.abs(B).max(axis=0)[None, :] * np.sqrt(k)
        rel = np.abs(got - ref) / scale
        say(f"  max scaled elementwise error {rel.max():.3e}")
        assert rel.max() < TOL_ELEM[terms]


def stress_scales(n, seed):
    rng = np.random.default_rng(seed)
    return 10.0 ** rng.uniform(-4, 4, size=n)


@pytest.mark.parametrize("n", [8192, 16384])
def test_bf16x3_precision_target_at_full_size(n):
    """The classical algorithm and one- and two-level Strassen with the
    FMMHIP_LEAF_BF16X3 leaf at n^3 (the precision form's sizes),
    uniform[-1,1] inputs and the stress variant (rows of A, columns of B
    scaled by 10^U(-4,4)): normalised error against the FP64 classical
    kernel on the same operands <= 1e-10."""
    A = F.DeviceMatrix(n, n).fill_uniform(0, 0)
    B = F.DeviceMatrix(n, n).fill_uniform(0, 1)
    ref, C = F.DeviceMatrix(n, n), F.DeviceMatrix(n, n)
    s = F.strassen()
    for inputs in ("uniform", "stress"):
        if inputs == "stress":
            rs, cs = F.DeviceBuffer(8 * n), F.DeviceBuffer(8 * n)
            rs.upload(stress_scales(n, 1))
            cs.upload(stress_scales(n, 2))
            F.check(F.lib().fmmhip_scale_rows_cols(A.ptr, n, n, n, rs.ptr, None, None), "scale A rows")
            F.check(F.lib().fmmhip_scale_rows_cols(B.ptr, n, n, n, None, cs.ptr, None), "scale B cols")
        F.dgemm(n, n, n, 1.0, A, B, 0.0, ref)
        F.sync()
        nA, nB = F.fro(A), F.fro(B)
        for levels in (0, 1, 2):
            say(f"bf16x3 precision target: {inputs} 222 L{levels} at {n}^3")
            alg = F.Alg(classical=True) if levels == 0 else F.Alg([s] * levels)
            F.dgemm_fmm(alg, n, n, n, 1.0, A, B, 0.0, C, leaf=F.LEAF_BF16X3)
            F.sync()
            err = F.fro_diff(C, ref) / (nA * nB)
            say(f"  normalised error {err:.3e} (bound {TARGET:g})")
            assert err <= TARGET
    F.check(F.lib().fmmhip_workspace_release(), "workspace_release")


def test_bf16x3_nan_row():
    m, n, k = 200, 150, 300
    A, B, C = O.uniform(m, k, 2, 0), O.uniform(k, n, 2, 1), O.uniform(m, n, 2, 2)
    A[17, 123] = np.nan
    got = run(F.Alg(classical=True), 1.0, A, B, 1.0, C)
    assert np.isnan(got[17, :]).all()
    rest = np.delete(got, 17, axis=0)
    assert np.isfinite(rest).all()


@pytest.mark.parametrize("name", ["222x2", "333x1"])
def test_bf16x3_product_split_exact_integers(name):
    """The product split (G workgroups per tile over contiguous product
    ranges, zeroed temporaries, ordered reduction) keeps integer inputs exact
    for every G, beta applied once and beta = 0 not reading C."""
    alg, og = algs()[name]
    for (m, n, k) in [(129, 130, 257), (300, 280, 600)]:
        A, B, C = O.integers(m, k, 6, 0, 4), O.integers(k, n, 6, 1, 4), O.integers(m, n, 6, 2, 4)
        for alpha, beta in [(1.0, 1.0), (-2.0, 0.5), (1.0, 0.0)]:
            want = O.fmm_dgemm(og, alpha, A, B, beta, C.copy(order="F"))
            Cn = C.copy(order="F")
            if beta == 0.0:
                Cn[:] = np.nan
            for g in (1, 2, 4, 8):
                alg.set_split(g)
                say("bf16x3 split", name, g, m, n, k, alpha, beta)
                got = run(alg, alpha, A, B, beta, Cn)
                assert np.array_equal(got, want), (g, m, n, k, alpha, beta)
    alg.set_split(-1)


def test_bf16x3_product_split_random_repeatable():
    alg, _ = algs()["222x2"]
    m = n = k = 1024
    A, B, C = O.uniform(m, k, 8, 0), O.uniform(k, n, 8, 1), O.uniform(m, n, 8, 2)
    ref = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    for g in (2, 8):
        alg.set_split(g)
        outs = [run(alg, 1.0, A, B, 1.0, C) for _ in range(2)]
        err = O.normalized_error(outs[0], ref, A, B)
        say(f"bf16x3 222x2 split {g} normalised error {err:.3e}")
        assert err < TOL[(8, "222x2")]
        assert np.array_equal(outs[0], outs[1])
