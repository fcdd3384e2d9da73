"""GPU parity of the BF16x3 leaf (FMMHIP_LEAF_BF16X3, eight plane products,
and FMMHIP_LEAF_BF16X3_6, six): FP64 operands split into three bf16 planes
with per-(row, 256-k chunk) power-of-two scales, plane products on the bf16
MFMA with FP32 accumulation, FP64 recombination and FP64 W write-back.

Bars (stated per test; the product split is covered by the last two tests):
  * small-integer inputs are split exactly and every chunk sum is an exact
    FP32 integer, so the result equals the oracle's bit for bit, for the
    classical leaf and for 1- and 2-level FMM, on ragged shapes;
  * uniform[-1,1] inputs at 1024^3: normalised error
    ||C - C_ref||_F/(||A||_F ||B||_F) against the oracle's FP64 classical
    product below the bounds in TOL (about 1.5x the error of a float64
    restatement of the split scheme on the same inputs: 8 terms 2.5e-10
    classical, 6.6e-10 one-level and 1.7e-9 two-level Strassen; 6 terms
    4.1e-10, 1.3e-9, 4.3e-9; the error falls as 1/sqrt(k));
  * the same with rows of A and columns of B scaled by 10^U(-4,4) (768^3;
    restatement 8 terms 4.6e-10 classical, 2.1e-9 two-level; 6 terms
    1.0e-9, 5.3e-9), and elementwise for the classical leaf relative to the
    row/column scales (restatement 2.5e-8 / 5.9e-8);
  * a NaN in A makes exactly its row of C NaN.
"""
import sys

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O
from conftest import alg_path

pytestmark = pytest.mark.gpu

LEAVES = {8: F.LEAF_BF16X3, 6: F.LEAF_BF16X3_6}
TOL = {  # (terms, algorithm) -> bound
    (8, "classical"): 4e-10, (8, "222x1"): 1e-9, (8, "222x2"): 2.6e-9,
    (6, "classical"): 6.5e-10, (6, "222x1"): 2e-9, (6, "222x2"): 6.5e-9,
}
TOL_STRESS = {(8, "classical"): 7e-10, (8, "222x2"): 3.2e-9, (6, "classical"): 1.6e-9, (6, "222x2"): 8e-9}
TOL_ELEM = {8: 4e-8, 6: 9e-8}


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


@pytest.mark.parametrize("terms", [8, 6])
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


@pytest.mark.parametrize("terms", [8, 6])
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


@pytest.mark.parametrize("terms", [8, 6])
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
